// graph.h -- CSR graph of the GpuSolver C++ interface.
// Field names and meaning follow the reference's include/graph.h:24-48 so a
// CSRGraph built by reference code and one built here are interchangeable.
#pragma once
#include "common.h"

struct EdgeList {
    std::string name;
    u32 n = 0;
    std::vector<u32> src, dst, w;
    std::vector<double> x, y;   // coordinates (empty when the graph has none)
};

struct CSRGraph {
    std::string name;
    u32 n = 0;                      // vertices
    u32 m = 0;                      // directed arcs
    std::vector<u32> row_offsets;   // n+1
    std::vector<u32> col_indices;   // m, each row sorted by (w, dst)
    std::vector<u32> weights;       // m, >= 1
    std::vector<double> x, y;       // planar coordinates for A* (may be empty)
    double hScale = 0.0;            // admissible heuristic scale (0: A* has h = 0)
    u32 minW = 0, maxW = 0, maxDeg = 0;
    double avgW = 0.0, avgDeg = 0.0;

    bool hasCoords() const { return !x.empty(); }
    u32 degree(u32 u) const { return row_offsets[u + 1] - row_offsets[u]; }
    size_t bytesCSR() const { return (size_t)(n + 1) * 4 + (size_t)m * 8; }
};

// buildCSR: drop self-loops, w >= 1, rows sorted by (w, dst), statistics, hScale.
CSRGraph buildCSR(EdgeList&& el);
// Benchmark workloads (see include/pf_graph.h for the exact definitions).
EdgeList genGridRoadSpec(u32 rows, u32 cols, u64 seed, int jitter = 20, double wlo = 1.0, double whi = 1.5);
EdgeList genRMATSpec(u32 scale, u32 edgeFactor, double a = 0.57, double b = 0.19, double c = 0.19, u32 maxW = 255,
                     u64 seed = 11);
bool saveBinary(const CSRGraph& g, const std::string& path);
bool loadBinary(const std::string& path, CSRGraph& g);
std::vector<u32> pickSources(const CSRGraph& g, u32 count, u64 seed);
