"""MI355X (gfx950) SSSP / A* engine behind the reference's GpuSolver interface.

Native parts: libpfgpu.so (HIP kernels + C-ABI, include/pf_engine.h),
libpfgraph.so (host graph layer, include/pf_graph.h), libgpusolver.so (C++
GpuSolver, include/gpusolver/gpu_sssp.h).  This package binds them with ctypes.
"""
from .graph import CSRGraph, build_csr, grid_road, rmat, union, save_binary, load_binary  # noqa: F401
from .solver import (GpuSolver, GpuOptions, GpuAlgo, Mapping, AtomicMode, Compaction, Drive, Heuristic,  # noqa: F401
                     GpuRunResult, GpuStats, INF_DIST, NO_PARENT, device_info, warmup, transfer_study)

__version__ = "0.1.0"
