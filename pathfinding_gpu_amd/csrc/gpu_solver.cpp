// gpu_solver.cpp -- GpuSolver (the reference's include/gpu_sssp.h:97-120
// interface) over the engine's C-ABI.  Plain C++: builds with g++ against
// either include/gpusolver/gpu_sssp.h or the reference's own header, so the
// reference's harness can link it in place of src/gpu_sssp.cu and
// src/gpu_device.cu.  Errors come back from the C-ABI as status codes and are
// rethrown as std::runtime_error, the reference's convention
// (src/gpu_sssp.cu:108,122-127,463-465).
#include "gpu_sssp.h"
#include "pf_engine.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace {

void check(int rc) {
    if (rc != PF_OK) throw std::runtime_error(std::string("GPU engine: ") + pf_last_error());
}

pf_options toC(const GpuOptions& o) {
    pf_options c;
    pf_default_options(&c);
    c.algo = (int)o.algo;
    c.mapping = (int)o.mapping;
    c.atomic = (int)o.atomic;
    c.compaction = (int)o.compaction;
    c.delta = o.delta;
    c.block_size = o.blockSize;
    c.sync_every = o.syncEvery;
#ifdef PF_GPUSOLVER_EXTENSIONS
    c.drive = (int)o.drive;
    c.heuristic = (int)o.heuristic;
    c.wave_cutoff = o.waveCutoff;
    c.block_cutoff = o.blockCutoff;
    c.grid_blocks = o.gridBlocks;
    c.async_light = o.asyncLight;
#endif
    return c;
}

void fromC(const pf_stats& s, GpuRunResult& r) {
    r.ms = s.ms;
    r.stats.edgesScanned = s.edges_scanned;
    r.stats.atomicsIssued = s.atomics_issued;
    r.stats.updates = s.updates;
    r.stats.pushes = s.pushes;
    r.stats.iterations = s.iterations;
    r.stats.hostSyncs = s.host_syncs;
    r.stats.launches = s.launches;
    r.stats.buckets = s.buckets;
    r.stats.delta = s.delta;
}

pf_graph describe(const CSRGraph& g) {
    pf_graph d{};
    d.n = g.n;
    d.m = g.m;
    d.row_offsets = g.row_offsets.data();
    d.col_indices = g.col_indices.data();
    d.weights = g.weights.data();
    d.x = g.hasCoords() ? g.x.data() : nullptr;
    d.y = g.hasCoords() ? g.y.data() : nullptr;
    d.h_scale = g.hScale;
    d.avg_w = g.avgW;
    d.avg_deg = g.avgDeg;
    d.flags = 0;   // buildCSR already sorted every row by (w, dst)
    return d;
}

}  // namespace

const char* toString(GpuAlgo a) {
    switch (a) {
        case GpuAlgo::NaiveRelax: return "Naive-Relax";
        case GpuAlgo::Frontier: return "Frontier";
        case GpuAlgo::DeltaStep: return "Delta-Step";
    }
    return "?";
}
const char* toString(Compaction c) { return c == Compaction::PrefixScan ? "prefix-sum" : "wave-atomic"; }
const char* toString(Mapping m) {
    switch (m) {
        case Mapping::Thread: return "thread";
        case Mapping::Warp: return "wave64";
        case Mapping::Hybrid: return "hybrid";
    }
    return "?";
}
const char* toString(AtomicMode a) {
    switch (a) {
        case AtomicMode::MinRBA: return "atomicMin+RBA";
        case AtomicMode::CasRBA: return "atomicCAS+RBA";
        case AtomicMode::MinNoCheck: return "atomicMin(no-check)";
    }
    return "?";
}

GpuDeviceInfo gpuQueryDevice(int dev) {
    GpuDeviceInfo info;
    pf_device_info d;
    if (pf_device_query(dev, &d) != PF_OK) return info;
    info.ok = d.ok != 0;
    info.name = std::string(d.name) + " (" + d.arch + ")";
    // gfx950 -> "9.5": the architecture generation stands where CUDA puts its compute capability
    int major = 0, minor = 0;
    if (std::sscanf(d.arch, "gfx%1d%1d", &major, &minor) == 2) {
        info.ccMajor = major;
        info.ccMinor = minor;
    }
    info.smCount = d.cu_count;
    info.clockMHz = d.clock_mhz;
    info.memClockMHz = d.mem_clock_mhz;
    info.busWidth = d.bus_width;
    info.driverVersion = d.driver_version;
    info.runtimeVersion = d.runtime_version;
    info.maxThreadsPerSM = d.max_threads_per_cu;
    info.warpSize = d.wave_size;
    info.globalMem = d.global_mem;
    info.sharedPerBlock = d.lds_per_block;
    info.l2Bytes = d.l2_bytes;
    info.peakBandwidthGBs = d.peak_bandwidth_gbs;
    return info;
}

void gpuWarmup() { check(pf_warmup()); }

TransferStudy gpuTransferStudy(const CSRGraph& g, int reps) {
    TransferStudy t;
    pf_graph d = describe(g);
    double ms[3] = {0, 0, 0};
    check(pf_transfer_study(&d, reps, &t.bytes, ms));
    t.pageableMs = ms[0];
    t.pinnedMs = ms[1];
    t.pinnedStreamsMs = ms[2];
    return t;
}

struct GpuSolver::Impl {
    const CSRGraph& g;   // kept by reference, as the reference does (src/gpu_sssp.cu:64)
    pf_solver* s = nullptr;
    Impl(const CSRGraph& graph, u32 maxBatch) : g(graph) {
        pf_graph d = describe(g);
        check(pf_solver_create(&d, maxBatch, &s));
    }
    ~Impl() { pf_solver_destroy(s); }
};

GpuSolver::GpuSolver(const CSRGraph& g, u32 maxBatch) : p_(new Impl(g, maxBatch)) {}
GpuSolver::~GpuSolver() { delete p_; }
double GpuSolver::uploadMs() const { return pf_solver_upload_ms(p_->s); }
size_t GpuSolver::deviceBytes() const { return (size_t)pf_solver_device_bytes(p_->s); }
u32 GpuSolver::autoDelta() const { return pf_solver_auto_delta(p_->s); }

GpuRunResult GpuSolver::sssp(u32 src, const GpuOptions& opt, bool copyBack) {
    GpuRunResult r;
    pf_options o = toC(opt);
    pf_stats st{};
    if (copyBack) {
        r.dist.resize(p_->g.n);
        r.parent.resize(p_->g.n);
    }
    check(pf_sssp(p_->s, src, &o, copyBack ? r.dist.data() : nullptr, copyBack ? r.parent.data() : nullptr, &st));
    fromC(st, r);
    return r;
}

GpuRunResult GpuSolver::astar(u32 src, u32 dst, const GpuOptions& opt) {
    GpuRunResult r;
    pf_options o = toC(opt);
    pf_stats st{};
    u32 cap = std::min<u32>(p_->g.n, 1u << 20);
    std::vector<u32> path(cap);
    u32 len = 0;
    check(pf_astar(p_->s, src, dst, &o, &r.targetDist, path.data(), cap, &len, &st));
    if (r.targetDist != INF_DIST && len == 0 && cap < p_->g.n) {   // longer than the first buffer
        cap = p_->g.n;
        path.assign(cap, 0);
        check(pf_astar(p_->s, src, dst, &o, &r.targetDist, path.data(), cap, &len, &st));
    }
    path.resize(len);
    r.path = std::move(path);
    fromC(st, r);
    return r;
}

GpuRunResult GpuSolver::multiSource(const std::vector<u32>& srcs, const GpuOptions& opt,
                                    std::vector<std::vector<u32>>* distOut) {
    GpuRunResult r;
    pf_options o = toC(opt);
    pf_stats st{};
    const u32 k = (u32)srcs.size();
    const u32 n = p_->g.n;
    std::vector<u32> flat;
    if (distOut) flat.resize((size_t)k * n);
    if (k == 1) {   // the reference runs k == 1 through its single-source path (src/gpu_sssp.cu:559-560)
        if (distOut) r.parent.resize(n);
        check(pf_sssp(p_->s, srcs[0], &o, distOut ? flat.data() : nullptr, distOut ? r.parent.data() : nullptr, &st));
        if (distOut) r.dist = flat;
    } else {
        check(pf_multi_source(p_->s, srcs.data(), k, &o, distOut ? flat.data() : nullptr, nullptr, &st));
    }
    fromC(st, r);
    if (distOut) {
        distOut->assign(k, std::vector<u32>());
        for (u32 i = 0; i < k; ++i) (*distOut)[i].assign(flat.begin() + (size_t)i * n, flat.begin() + (size_t)(i + 1) * n);
    }
    return r;
}
