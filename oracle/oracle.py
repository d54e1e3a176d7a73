"""ctypes loader for the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the engine. `fast=True` loads the build
with the reference's -O3 -ffast-math -fopenmp flags (the timed CPU baseline);
the default build is the IEEE checker.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_P = C.c_void_p
_I = C.c_int64
_SIGS = {
    "or_direct_f64": (None, [_I, _P, _P, C.c_double, C.c_double, _I, _P, _P, C.c_int]),
    "or_pp_forces": (None, [_I, _P, _P, C.c_float, C.c_float, _P, _I, _I, C.c_int]),
    "or_kdk_update": (None, [_I, _P, _P, _P, _P, C.c_float, C.c_int]),
    "or_euler_update": (None, [_I, _P, _P, _P, C.c_float]),
    "or_pp_step": (None, [_I, _P, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_int]),
    "or_kdk_update_f64": (None, [_I, _P, _P, _P, _P, C.c_double]),
    "or_euler_update_f64": (None, [_I, _P, _P, _P, _P, C.c_double]),
    "or_energy_f64": (None, [_I, _P, _P, _P, C.c_double, C.c_double, _P, _P, _P, C.c_int]),
    "or_v5_forces": (C.c_int, [_I, _P, _P, C.c_float, C.c_float, C.c_float, _P, C.c_double, _P, C.c_int]),
    "or_v6_forces": (C.c_int, [_I, _P, _P, C.c_float, C.c_float, C.c_float, _P, _I, _I, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "or_bbox": (None, [_I, _P, _P, _P]),
    "or_expand_bits": (C.c_uint32, [C.c_uint32]),
    "or_morton": (None, [_I, _P, _P, _P, C.c_int, _P]),
    "or_sort_codes": (None, [_I, _P, _P, _P]),
    "or_tree_bound": (_I, [_I]),
    "or_build_tree": (_I, [_I, _P, C.c_int, _I, _P, _P, _P, _P, _P]),
    "or_tree_com": (None, [_I, _P, _P, _I, _P, _P]),
    "or_bh_forces": (_I, [_I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, C.c_float, C.c_float,
                          C.c_float, _P, C.c_int]),
    "or_bh_forces_group": (_I, [_I, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, C.c_int,
                                C.c_float, C.c_float, C.c_float, C.c_int, _P, C.c_int]),
    "or_compact_bits": (C.c_uint32, [C.c_uint32]),
    "or_tree_open_r2": (None, [_I, _P, _P, _P, _P, _P, _P, C.c_int, C.c_float, _P]),
    "or_bh_wave_ranges": (_I, [_I, _I, _P, _P, _P, _P, C.c_int, _P, _P]),
    "or_bh_pipeline_group": (_I, [_I, _P, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                  _P, _P, _P, _P, C.c_int]),
    "or_bh_pipeline": (_I, [_I, _P, C.c_float, C.c_float, C.c_float, C.c_int, _P, _P, _P, _P, C.c_int]),
}
_libs: dict[bool, C.CDLL] = {}


def load(fast: bool = False) -> C.CDLL:
    if fast in _libs:
        return _libs[fast]
    p = HERE / ("liboracle_fast.so" if fast else "liboracle.so")
    if not p.exists():
        raise FileNotFoundError(f"{p} missing: run `make -C oracle`")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[fast] = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def direct_f64(pos3, mass, eps2, G=1.0, targets=None, nthreads=None):
    pos3 = np.ascontiguousarray(pos3, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    n = pos3.shape[0]
    t = None if targets is None else np.ascontiguousarray(targets, np.int64)
    nt = n if t is None else t.shape[0]
    out = np.empty((nt, 3), np.float64)
    load().or_direct_f64(n, _p(pos3), _p(mass), eps2, G, nt, _p(t), _p(out), nthreads or threads())
    return out


def pp_forces(pos4, mass4, eps2, G=1.0, t0=0, t1=None, nthreads=1, fast=False):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    mass4 = np.ascontiguousarray(mass4, np.float32)
    n = pos4.shape[0]
    t1 = n if t1 is None else t1
    out = np.zeros((n, 4), np.float32)
    load(fast).or_pp_forces(n, _p(pos4), _p(mass4), eps2, G, _p(out), t0, t1, nthreads)
    return out


def pp_step(pos4, vel4, acc4, mass4, eps2, dt, G=1.0, nthreads=1, fast=False):
    """One v1/v2 step in place on float32 (n,4) arrays; returns the forces."""
    n = pos4.shape[0]
    force = np.zeros((n, 4), np.float32)
    load(fast).or_pp_step(n, _p(pos4), _p(vel4), _p(acc4), _p(mass4), _p(force), eps2, G, dt, nthreads)
    return force


def kdk_update(pos4, vel4, acc4, force4, dt):
    load().or_kdk_update(pos4.shape[0], _p(pos4), _p(vel4), _p(acc4), _p(force4), dt, 1)


def euler_update(pos4, vel4, force4, dt):
    load().or_euler_update(pos4.shape[0], _p(pos4), _p(vel4), _p(force4), dt)


def kdk_update_f64(pos3, vel3, acc3, force3, dt):
    load().or_kdk_update_f64(pos3.shape[0], _p(pos3), _p(vel3), _p(acc3), _p(force3), dt)


def euler_update_f64(pos3, vel3, acc3, force3, dt):
    load().or_euler_update_f64(pos3.shape[0], _p(pos3), _p(vel3), _p(acc3), _p(force3), dt)


def energy_f64(pos3, vel3, mass, eps2, G=1.0, nthreads=None):
    pos3 = np.ascontiguousarray(pos3, np.float64)
    vel3 = np.ascontiguousarray(vel3, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    ek = C.c_double()
    ep = C.c_double()
    p = np.empty(3, np.float64)
    load().or_energy_f64(pos3.shape[0], _p(pos3), _p(vel3), _p(mass), eps2, G, C.byref(ek),
                         C.byref(ep), _p(p), nthreads or threads())
    return ek.value, ep.value, p


def v5_forces(pos4, mass4, eps2, theta=0.5, G=1.0, pool_factor=0.0, nthreads=None, fast=False):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    mass4 = np.ascontiguousarray(mass4, np.float32)
    n = pos4.shape[0]
    out = np.zeros((n, 4), np.float32)
    cnt = C.c_int64()
    rc = load(fast).or_v5_forces(n, _p(pos4), _p(mass4), eps2, G, theta, _p(out), pool_factor,
                                 C.byref(cnt), nthreads or threads())
    return rc, out, cnt.value


def v6_forces(pos4, mass4, eps2, theta=0.5, G=1.0, t0=0, t1=None, nthreads=None, fast=False):
    """v6 restatement: returns (rc, force4, build seconds, force seconds)."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    mass4 = np.ascontiguousarray(mass4, np.float32)
    n = pos4.shape[0]
    t1 = n if t1 is None else t1
    out = np.zeros((n, 4), np.float32)
    b = C.c_double()
    f = C.c_double()
    rc = load(fast).or_v6_forces(n, _p(pos4), _p(mass4), eps2, G, theta, _p(out), t0, t1,
                                 nthreads or threads(), C.byref(b), C.byref(f))
    return rc, out, b.value, f.value


def bbox(pos4):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    mn = np.empty(3, np.float32)
    mx = np.empty(3, np.float32)
    load().or_bbox(pos4.shape[0], _p(pos4), _p(mn), _p(mx))
    return mn, mx


def expand_bits(v: int) -> int:
    return int(load().or_expand_bits(v))


def morton(pos4, bmin, bmax, cubic=False):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    codes = np.empty(pos4.shape[0], np.uint32)
    load().or_morton(pos4.shape[0], _p(pos4), _p(np.asarray(bmin, np.float32)),
                     _p(np.asarray(bmax, np.float32)), 1 if cubic else 0, _p(codes))
    return codes


def sort_codes(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    perm = np.empty_like(codes)
    srt = np.empty_like(codes)
    load().or_sort_codes(codes.shape[0], _p(codes), _p(perm), _p(srt))
    return perm, srt


def build_tree(codes_sorted, leaf_size):
    codes_sorted = np.ascontiguousarray(codes_sorted, np.uint32)
    n = codes_sorted.shape[0]
    bound = int(load().or_tree_bound(n))
    arrs = [np.empty(bound, np.int32) for _ in range(5)]
    cnt = load().or_build_tree(n, _p(codes_sorted), leaf_size, bound, *[_p(a) for a in arrs])
    fc, nx, st, nb, lv = (a[:cnt].copy() for a in arrs)
    return {"first_child": fc, "next": nx, "start": st, "count": nb, "level": lv}


def tree_com(tree, pos4s):
    pos4s = np.ascontiguousarray(pos4s, np.float32)
    k = tree["start"].shape[0]
    com = np.empty((k, 4), np.float32)
    load().or_tree_com(k, _p(tree["start"]), _p(tree["count"]), pos4s.shape[0], _p(pos4s), _p(com))
    return com


def bh_forces(pos4s, perm, bmin, bmax, tree, com4, eps2, theta=0.5, G=1.0, nthreads=None):
    n = pos4s.shape[0]
    out = np.zeros((n, 4), np.float32)
    inter = load().or_bh_forces(
        n, _p(np.ascontiguousarray(pos4s, np.float32)), _p(np.ascontiguousarray(perm, np.uint32)),
        _p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)), tree["start"].shape[0],
        _p(tree["first_child"]), _p(tree["next"]), _p(tree["start"]), _p(tree["count"]),
        _p(tree["level"]), _p(np.ascontiguousarray(com4, np.float32)), eps2, G, theta, _p(out),
        nthreads or threads())
    return out, int(inter)


def bh_pipeline(pos4, eps2, theta=0.5, leaf_size=8, G=1.0, nthreads=None, group=0, cubic=False):
    """The engine's Barnes-Hut force pass on the CPU. The GPU's defaults are
    group=NB_BH_WAVE_BODIES (32, the wave traversal) and cubic=True; group=0 is
    the per-body traversal."""
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = pos4.shape[0]
    out = np.zeros((n, 4), np.float32)
    codes = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    cnt = C.c_int64()
    inter = load().or_bh_pipeline_group(n, _p(pos4), eps2, G, theta, leaf_size, group,
                                        1 if cubic else 0, _p(out), _p(codes), _p(perm), C.byref(cnt),
                                        nthreads or threads())
    return out, codes, perm, cnt.value, int(inter)


def bh_wave_ranges(tree, n, cap=64):
    k = tree["start"].shape[0]
    rs = np.empty(max(n, 1), np.int32)
    rc = np.empty(max(n, 1), np.int32)
    r = load().or_bh_wave_ranges(n, k, _p(tree["first_child"]), _p(tree["start"]), _p(tree["count"]),
                                 _p(tree["level"]), cap, _p(rs), _p(rc))
    return rs[:r].copy(), rc[:r].copy()


def compact_bits(v: int) -> int:
    return int(load().or_compact_bits(v))


def tree_open_r2(tree, com4, codes_sorted, bmin, bmax, theta=0.5, cubic=True):
    """Squared opening radius (s_l / theta + |COM - cell centre|)^2 per node."""
    k = tree["start"].shape[0]
    r2 = np.empty(k, np.float32)
    load().or_tree_open_r2(k, _p(tree["start"]), _p(tree["level"]),
                           _p(np.ascontiguousarray(com4, np.float32)),
                           _p(np.ascontiguousarray(codes_sorted, np.uint32)),
                           _p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)),
                           1 if cubic else 0, theta, _p(r2))
    return r2


def bh_forces_group(pos4s, perm, bmin, bmax, tree, com4, eps2, theta=0.5, G=1.0, group=64,
                    nthreads=None, *, codes_sorted, cubic=True):
    """The engine's default traversal on the CPU (groups of `group` sorted
    bodies, the modified opening test of or_tree_open_r2)."""
    n = pos4s.shape[0]
    out = np.zeros((n, 4), np.float32)
    inter = load().or_bh_forces_group(
        n, _p(np.ascontiguousarray(pos4s, np.float32)), _p(np.ascontiguousarray(perm, np.uint32)),
        _p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)), tree["start"].shape[0],
        _p(tree["first_child"]), _p(tree["next"]), _p(tree["start"]), _p(tree["count"]),
        _p(tree["level"]), _p(np.ascontiguousarray(com4, np.float32)),
        _p(np.ascontiguousarray(codes_sorted, np.uint32)), 1 if cubic else 0, eps2, G, theta, group,
        _p(out), nthreads or threads())
    return out, int(inter)
