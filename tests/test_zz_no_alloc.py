"""nb_bh_forces_noalloc (the allocation-free Barnes-Hut force pass) against the
entry point it varies, nb_forces_f64 on a Barnes-Hut context, through the C ABI:
the same kernels on the same FP32 bodies, so the forces must agree bit for bit.
Two calls in a row, the second smaller, then the first size again (the scratch
keeps the previous call's contents), and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS2, DT = 40.0, 0.0035


class Hip:
    """Device buffers for the caller's side of the call (the entry point itself
    allocates nothing)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        self.ptrs.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def download(self, p, a):
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def context_forces(nb, pos, m, leaf):
    with nb.Simulation(len(m), EPS2, DT, solver=nb.SOLVER_BH, leaf_size=leaf) as sim:
        sim.upload(pos, np.zeros_like(pos), m)
        return sim.forces_f64()


def noalloc_forces(nb, hip, pos, m, leaf):
    n = len(m)
    pm = np.zeros((n, 4), np.float32)
    pm[:, :3] = pos
    pm[:, 3] = m
    d_pm = hip.upload(pm)
    d_f = hip.upload(np.full(3 * n, np.nan))
    d_st = hip.upload(np.full(1, -1, np.int32))
    rc = nb.load().nb_bh_forces_noalloc(d_pm, n, EPS2, 1.0, 0.5, leaf, 0, d_f, d_st, None)
    assert rc == 0, nb.load().nb_last_error()
    f = hip.download(d_f, np.empty(3 * n))
    st = hip.download(d_st, np.empty(1, np.int32))
    assert st[0] == 0
    return f.reshape(3, n).T


def test_noalloc_equals_context_forces(nb):
    hip = Hip()
    try:
        cases = [("plummer", 50000, 8), ("ball", 3000, 8), ("plummer", 50000, 8), ("discs", 65536, 16)]
        for ic, n, leaf in cases:
            print(f"no-alloc: {ic} n={n} leaf={leaf}", flush=True)
            if ic == "plummer":
                pos, _, m = nb.ic_plummer(90 + leaf, n, 50.0, 1e4)
            elif ic == "ball":
                pos, _, m = nb.ic_ball(91, n, 100.0, 1.0)
            else:
                pos, _, m = nb.ic_discs(92, n)
            pos = pos.astype(np.float32).astype(np.float64)  # the FP32 view is exact
            want = context_forces(nb, pos, m, leaf)
            got = noalloc_forces(nb, hip, pos, m, leaf)
            np.testing.assert_array_equal(got, want)
    finally:
        hip.free()


def test_noalloc_rejects_sizes_beyond_the_scratch(nb):
    lib = nb.load()
    dummy = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    assert lib.nb_bh_forces_noalloc(dummy, nb.NOALLOC_MAX_BODIES + 1, EPS2, 1.0, 0.5, 8, 0, dummy,
                                    None, None) == nb.NB_E_ARG
    assert b"65536" in lib.nb_last_error()
    assert lib.nb_bh_forces_noalloc(None, 100, EPS2, 1.0, 0.5, 8, 0, dummy, None, None) == nb.NB_E_ARG
