"""The C-ABI libraries load without a GPU and export every function their headers declare."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pathfinding_gpu_amd")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    names = re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, lib)], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


@pytest.mark.parametrize("header,lib", [("pf_engine.h", "libpfgpu.so"), ("pf_graph.h", "libpfgraph.so")])
def test_every_declared_symbol_is_exported(header, lib):
    names = declared(header)
    assert len(names) > 8
    have = exported(lib)
    missing = [n for n in names if n not in have]
    assert not missing, f"{lib} misses {missing}"


@pytest.mark.parametrize("lib", ["libpfgpu.so", "libpfgraph.so", "libgpusolver.so"])
def test_library_loads_without_gpu(lib):
    C.CDLL(os.path.join(PKG, lib))


def test_python_bindings_cover_the_headers():
    from pathfinding_gpu_amd import _native as N
    assert set(declared("pf_engine.h")) == set(N.ENGINE_SYMBOLS)
    assert set(declared("pf_graph.h")) == set(N.GRAPH_SYMBOLS)
    N.engine()
    N.graphlib()


def test_cpp_gpusolver_exports_reference_interface():
    out = subprocess.run(["nm", "-DC", "--defined-only", os.path.join(PKG, "libgpusolver.so")], capture_output=True,
                         text=True, check=True).stdout
    for sym in ["GpuSolver::GpuSolver(CSRGraph const&, unsigned int)", "GpuSolver::~GpuSolver()",
                "GpuSolver::sssp(unsigned int, GpuOptions const&, bool)",
                "GpuSolver::astar(unsigned int, unsigned int, GpuOptions const&)",
                "GpuSolver::multiSource(std::vector<unsigned int, std::allocator<unsigned int> > const&, GpuOptions const&",
                "GpuSolver::uploadMs() const", "GpuSolver::deviceBytes() const", "GpuSolver::autoDelta() const",
                "gpuQueryDevice(int)", "gpuWarmup()", "gpuTransferStudy(CSRGraph const&, int)", "toString(GpuAlgo)",
                "toString(Mapping)", "toString(AtomicMode)", "toString(Compaction)"]:
        assert sym in out, sym


def test_engine_reports_no_device_as_status():
    """Without a GPU the engine returns a status code and a message; it never exits."""
    from pathfinding_gpu_amd import _native as N
    from pathfinding_gpu_amd.solver import device_info
    lib = N.engine()
    d = N.PfDeviceInfo()
    rc = lib.pf_device_query(0, C.byref(d))
    if rc == 0:
        pytest.skip("a GPU is present")
    assert rc == -6 and b"device" in lib.pf_last_error()
    with pytest.raises(RuntimeError):
        device_info(0)


def test_default_options():
    from pathfinding_gpu_amd import _native as N
    o = N.PfOptions()
    N.engine().pf_default_options(C.byref(o))
    assert (o.algo, o.mapping, o.atomic, o.compaction, o.delta, o.block_size, o.sync_every) == (2, 2, 0, 0, 0, 256, 8)
