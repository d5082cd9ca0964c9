// gpu_sssp.h -- C++ host interface of the GPU shortest-path engine.
//
// Same declarations as the reference's include/gpu_sssp.h:11-120 (enums,
// GpuOptions, GpuStats, GpuRunResult, GpuDeviceInfo, TransferStudy,
// GpuSolver) so that the reference's callers compile unchanged against it.
// The implementation (pathfinding_gpu_amd/csrc/gpu_solver.cpp) calls the
// HIP engine through its C-ABI (include/pf_engine.h); no HIP type appears
// here.  GpuOptions carries extra MI355X knobs after the reference's fields;
// PF_GPUSOLVER_EXTENSIONS marks them.
#pragma once
#include "graph.h"

#include <string>
#include <vector>

#define PF_GPUSOLVER_EXTENSIONS 1

enum class GpuAlgo { NaiveRelax, Frontier, DeltaStep };
enum class Mapping { Thread, Warp, Hybrid };   // Warp = one 64-lane wavefront per vertex
enum class AtomicMode { MinRBA, CasRBA, MinNoCheck };
enum class Compaction { WarpAtomic, PrefixScan };
// Who runs the bucket loop: the host (reference behaviour), a replayed hipGraph, or one persistent kernel.
enum class Drive { Host, Graph, Persistent };
// A* heuristic: the reference's scaled Euclidean formula, or floor(Euclid) on integer coordinates.
enum class Heuristic { None, Scaled, FloorEuclid };

const char* toString(Compaction c);
const char* toString(GpuAlgo a);
const char* toString(Mapping m);
const char* toString(AtomicMode a);

struct GpuOptions {
    GpuAlgo algo = GpuAlgo::DeltaStep;
    Mapping mapping = Mapping::Hybrid;
    AtomicMode atomic = AtomicMode::MinRBA;
    u32 delta = 0;            // 0 => Davidson default 32 * avgW / avgDeg
    int blockSize = 256;
    u32 syncEvery = 8;        // launches per host sync (Drive::Host) / steps per replay (Drive::Graph)
    Compaction compaction = Compaction::WarpAtomic;
    // ---- extensions ----
    Drive drive = Drive::Host;
    Heuristic heuristic = Heuristic::Scaled;
    u32 waveCutoff = 0;       // hybrid: degree >= -> one wave per vertex (0 = 64)
    u32 blockCutoff = 0;      // hybrid: degree >= -> whole workgroup (0 = blockSize)
    u32 gridBlocks = 0;       // workgroups per launch (0 = automatic)
    u32 asyncLight = 0;       // Delta light phase: 0 one step per light iteration, 1 one work-queue step per
                              // bucket, 2 the same over 64-vertex tiles (symmetric graphs; else as 1)
};

struct GpuStats {
    u64 edgesScanned = 0;
    u64 atomicsIssued = 0;
    u64 updates = 0;
    u64 pushes = 0;
    u32 iterations = 0;
    u32 hostSyncs = 0;
    u32 launches = 0;
    u32 buckets = 0;
    u32 delta = 0;
};

struct GpuRunResult {
    double ms = 0.0;             // device time of init + loop (hipEvent)
    GpuStats stats;
    std::vector<u32> dist;       // SSSP: all vertices; A*: empty
    std::vector<u32> parent;
    u32 targetDist = INF_DIST;   // A*
    std::vector<u32> path;       // A*: src ... dst
};

struct GpuDeviceInfo {
    bool ok = false;
    std::string name;
    int ccMajor = 0, ccMinor = 0, smCount = 0, clockMHz = 0, memClockMHz = 0, busWidth = 0;
    int driverVersion = 0, runtimeVersion = 0, maxThreadsPerSM = 0, warpSize = 64;
    size_t globalMem = 0, sharedPerBlock = 0, l2Bytes = 0;
    double peakBandwidthGBs = 0.0;
};
GpuDeviceInfo gpuQueryDevice(int dev = 0);
void gpuWarmup();

struct TransferStudy {
    double bytes = 0;
    double pageableMs = 0, pinnedMs = 0, pinnedStreamsMs = 0;
};
TransferStudy gpuTransferStudy(const CSRGraph& g, int reps);

class GpuSolver {
public:
    GpuSolver(const CSRGraph& g, u32 maxBatch = 1);
    ~GpuSolver();
    GpuSolver(const GpuSolver&) = delete;
    GpuSolver& operator=(const GpuSolver&) = delete;

    double uploadMs() const;
    size_t deviceBytes() const;
    u32 autoDelta() const;

    GpuRunResult sssp(u32 src, const GpuOptions& opt, bool copyBack = true);
    GpuRunResult astar(u32 src, u32 dst, const GpuOptions& opt);
    GpuRunResult multiSource(const std::vector<u32>& srcs, const GpuOptions& opt,
                             std::vector<std::vector<u32>>* distOut);

private:
    struct Impl;
    Impl* p_;
};
