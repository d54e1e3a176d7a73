"""The C-ABI library loads on the CPU and exports every symbol its headers
declare; host-only entry points (initial conditions, argument checks) work
without a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_exports_every_declared_symbol(nb):
    lib = ctypes.CDLL(str(nb.LIB_PATH))
    syms = declared_symbols()
    assert "nb_create" in syms and "nb_ic_plummer" in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_abi(nb):
    assert set(declared_symbols()) <= set(nb._SIGS)


def _gpu_visible():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU error path")
def test_create_without_gpu_fails_cleanly(nb):
    with pytest.raises(nb.NBodyError) as e:
        nb.Simulation(16, 1.0, 0.01)
    assert e.value.code in (nb.NB_E_NOGPU, nb.NB_E_HIP)


def test_argument_errors(nb):
    lib = nb.load()
    ctx = ctypes.c_void_p()
    bad = [dict(n=0), dict(n=1 << 30), dict(n=8, eps2=-1.0), dict(n=8, leaf_size=65),
           dict(n=8, bodies_per_lane=3), dict(n=8, j_split=32), dict(n=8, j_split=3),
           dict(n=8, bodies_per_lane=4, j_split=8), dict(n=8, bodies_per_lane=2, j_split=16),
           dict(n=8, solver=7),
           dict(n=1 << 27, solver=1),
           # direct sum: the kernels index the [3][n] state in int32, so 3 n must fit
           dict(n=715827883, solver=0)]
    for kw in bad:
        cfg = nb.Config(**{"eps2": 1.0, "dt": 0.01, **kw})
        assert lib.nb_create(ctypes.byref(ctx), ctypes.byref(cfg)) == nb.NB_E_ARG, kw
        assert lib.nb_last_error()
    assert lib.nb_step(None, 1) == nb.NB_E_ARG
    cfg = nb.Config(n=715827883, eps2=1.0, dt=0.01)
    lib.nb_create(ctypes.byref(ctx), ctypes.byref(cfg))
    assert b"715827882" in lib.nb_last_error()
    # the allocation-free pass checks its arguments before any device call
    dummy = ctypes.c_void_p(16)
    for args in ((dummy, nb.NOALLOC_MAX_BODIES + 1, 1.0, 1.0, 0.5, 8, 0, dummy, None, None),
                 (dummy, 0, 1.0, 1.0, 0.5, 8, 0, dummy, None, None),
                 (None, 100, 1.0, 1.0, 0.5, 8, 0, dummy, None, None),
                 (dummy, 100, -1.0, 1.0, 0.5, 8, 0, dummy, None, None),
                 (dummy, 100, 1.0, 1.0, 0.5, 65, 0, dummy, None, None),
                 (dummy, 100, 1.0, 1.0, 0.5, 8, nb.F_NO_GRAPH, dummy, None, None)):
        assert lib.nb_bh_forces_noalloc(*args) == nb.NB_E_ARG, args
        assert lib.nb_last_error()
    # entry points added for diagnostics, tuning, quotas and serving statistics
    # reject null handles before any device call
    assert lib.nb_count_interactions_per_body(None, None) == nb.NB_E_ARG
    assert lib.nb_direct_f64(None, None, 1, None) == nb.NB_E_ARG
    assert lib.nb_tune_leaf(None, 1, None) == nb.NB_E_ARG
    assert lib.nb_batch_set_quotas(None, None, None, 1) == nb.NB_E_ARG
    assert lib.nb_server_get_stats(None, None) == nb.NB_E_ARG
    assert lib.nb_debug_set_walk_stack(None, 0) == nb.NB_E_ARG


def test_initial_conditions_are_seeded(nb):
    a = nb.ic_plummer(7, 2000)
    b = nb.ic_plummer(7, 2000)
    c = nb.ic_plummer(8, 2000)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    pos, vel, m = a
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 10 * 50 * 1.01  # truncated at 10 a (plus the centring shift)
    assert abs(m.astype(np.float64).sum() - 1e4) < 1e-2
    np.testing.assert_allclose(pos.mean(0), 0, atol=1e-9)
    np.testing.assert_allclose(vel.mean(0), 0, atol=1e-9)
    # Plummer: half-mass radius ~ 1.305 a (truncation at 10 a moves it slightly)
    assert 1.2 * 50 < np.median(r) < 1.4 * 50


def test_ball_and_discs(nb):
    pos, vel, m = nb.ic_ball(3, 5000, 100.0, 1.0)
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 100.0 and np.all(vel == 0) and np.all(m == 1)
    assert abs(np.mean(r ** 3) / 100 ** 3 - 0.5) < 0.03  # r^3 is uniform on [0, R^3]
    pos, vel, m = nb.ic_discs(5, 10000)
    assert m[0] == pytest.approx(1e4 * m[1], rel=1e-6) and m[5000] == m[0]
    left, right = pos[:5000], pos[5000:]
    assert abs(np.median(left[:, 0]) + 150) < 5 and abs(np.median(right[:, 0]) - 150) < 5
    # closing at 0.5 units per step of 0.0035
    assert vel[0, 0] - vel[5000, 0] == pytest.approx(0.5 / 0.0035, rel=1e-9)
    R = np.linalg.norm(left[1:, :2] - np.array([-150, 0]), axis=1)
    assert 50 < np.mean(R) < 70  # Gamma(2, 30) has mean 60


def test_lagrange_matches_reference(nb):
    pos, vel, m = nb.ic_lagrange(5.0)
    assert np.all(m == np.float32(10.5))
    c = pos.mean(0)
    d = np.linalg.norm(pos - c, axis=1)
    np.testing.assert_allclose(d, d[0], rtol=1e-6)
    side = np.linalg.norm(pos[0] - pos[1])
    assert side == pytest.approx(np.sqrt(75) / 10, rel=1e-6)
    # velocities are tangential: (p - c) . v = 0
    np.testing.assert_allclose(((pos - c) * vel).sum(1), 0, atol=1e-6)


def test_sizes_and_uniform(nb):
    s = nb.log_uniform_sizes(1, 512, 256, 32768)
    assert s.min() >= 256 and s.max() <= 32768
    assert 4000 < np.mean(s) < 9000  # E = (32768 - 256) / ln(128) ~ 6701
    u = nb.uniform(2, 1000, 1.0, 40.0)
    assert u.min() >= 1.0 and u.max() < 40.0


def test_default_shape_fills_small_systems(nb):
    # K = 16 waves per i-block up to 16384 bodies, 8 above
    for n, k in ((1024, 16), (4096, 16), (16384, 16), (16385, 8), (65536, 8), (1 << 18, 8)):
        t = nb.get_tuning(n)
        assert (t["bodies_per_lane"], t["j_split"]) == (1, k), (n, t)
