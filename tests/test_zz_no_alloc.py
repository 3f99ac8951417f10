"""gala_modularity_noalloc: the modularity of a partition from device arrays
with no allocation (its scratch is the library's static per-GPU array).  Two
calls in a row, the second on a smaller graph, must each equal
gala_phase_modularity (the allocating entry point over the same kernels) and
the oracle bit for bit; sizes past the stated maximum are refused."""
import ctypes

import numpy as np
import pytest

import gala
import oracle

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """Device copies of host arrays, made by the caller through the HIP runtime
    the library itself is linked against (the caller's own allocation)."""

    def __init__(self):
        import ctypes.util
        self.hip = ctypes.CDLL(ctypes.util.find_library("amdhip64") or "libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p.value

    def free(self):
        assert self.hip.hipDeviceSynchronize() == 0
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def call(eng, g, comm):
    d = DeviceArrays()
    try:
        wd = d.put(g.weights) if g.weights is not None else None
        rc, q = eng.modularity_noalloc(g.vertex_num, d.put(g.degrees), d.put(g.neighbors), wd,
                                       g.weight_mode if g.weights is not None else gala.W_NONE, d.put(comm))
    finally:
        d.free()
    return rc, q


def test_noalloc_matches_allocating_entry_point():
    print("noalloc modularity", flush=True)
    rng = np.random.default_rng(7)
    big = gala.gen_rmat(16, 16, seed=9)  # float32 weights, 65,536 vertices
    small = gala.gen_chung_lu(5000, 2.1, 17.35, 800.0, 4)  # unit weights
    eng = gala.Engine()
    for g in (big, small):
        comm = rng.integers(0, max(1, g.vertex_num // 20), g.vertex_num).astype(np.int32)
        rc, q = call(eng, g, comm)
        assert rc == 0, rc
        assert q == eng.phase_modularity(g, comm)
        W = oracle.scale_f32(g.weights)[0] if g.weights is not None else np.ones(len(g.neighbors), np.int64)
        assert q == oracle.modularity(g.vertex_num, g.degrees, g.neighbors, W, comm)
    eng.close()


def test_noalloc_refuses_sizes_past_the_maximum():
    eng = gala.Engine()
    rc, _ = eng.modularity_noalloc(gala.NOALLOC_MAX_N + 1, 1, 1, None, gala.W_NONE, 1)
    assert rc == -6  # GALA_E_INVALID before any device access
    eng.close()
