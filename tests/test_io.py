"""Particle-file reader/writer (reference format, ParticleSystemFile.cpp:10-90;
writer ParticleSystem.cpp:205-215) — host-only, runs on the CPU."""
import ctypes
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def read(nb, path):
    lib = nb.load()
    lib.nb_ic_read_file.restype = ctypes.c_int64
    lib.nb_ic_read_file.argtypes = [ctypes.c_char_p, ctypes.c_int64] + [ctypes.c_void_p] * 4
    n = lib.nb_ic_read_file(str(path).encode(), 0, None, None, None, None)
    if n < 0:
        return n, None
    p, v, a = (np.zeros((n, 4), np.float32) for _ in range(3))
    m = np.zeros(n, np.float32)
    lib.nb_ic_read_file(str(path).encode(), n, *(x.ctypes.data_as(ctypes.c_void_p) for x in (p, v, a, m)))
    return n, (p, v, a, m)


def write(nb, path, p, v, a, m):
    lib = nb.load()
    lib.nb_ic_write_file.restype = ctypes.c_int
    lib.nb_ic_write_file.argtypes = [ctypes.c_char_p, ctypes.c_int64] + [ctypes.c_void_p] * 4
    return lib.nb_ic_write_file(str(path).encode(), len(m),
                                *(np.ascontiguousarray(x).ctypes.data_as(ctypes.c_void_p) for x in (p, v, a, m)))


def test_reference_readme_example(nb):
    # README.md:150-174 declares 3 particles but lists 4; the reader stops at 3.
    n, (p, v, a, m) = read(nb, GOLDEN / "readme_particles.txt")
    assert n == 3
    np.testing.assert_array_equal(p[:, :3], np.float32([[5.8337, 4.6008, 3.35599],
                                                        [2.13544, 2.37607, 3.12727],
                                                        [2.74389, 2.46503, 3.42229]]))
    np.testing.assert_array_equal(v[1, :3], np.float32([1.01254, -1.94254, 0]))
    np.testing.assert_array_equal(a[:, :3], np.float32([[0, 0, 0], [0.4, 3.0, 0], [0, 0, -8.0]]))
    np.testing.assert_array_equal(m, np.float32([0.625, 0.875, 0.875]))
    assert np.all(p[:, 3] == 0)


def test_round_trip_is_exact(nb, tmp_path):
    rng = np.random.default_rng(1)
    p, v, a = (rng.normal(size=(257, 4)).astype(np.float32) * 1e3 for _ in range(3))
    for x in (p, v, a):
        x[:, 3] = 0
    m = rng.uniform(1e-6, 1e6, 257).astype(np.float32)
    f = tmp_path / "s.txt"
    assert write(nb, f, p, v, a, m) == 0
    n, (p2, v2, a2, m2) = read(nb, f)
    assert n == 257
    for x, y in ((p, p2), (v, v2), (a, a2)):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(m, m2)


def test_missing_and_malformed(nb, tmp_path):
    assert read(nb, tmp_path / "nope.txt")[0] == nb.NB_E_ARG
    bad = tmp_path / "bad.txt"
    bad.write_text("hello\n")
    assert read(nb, bad)[0] == nb.NB_E_ARG
    short = tmp_path / "short.txt"
    short.write_text("Particle System with 5 particles:\nPosition: (1, 2, 3)\nMass: 2\n")
    n, (p, v, a, m) = read(nb, short)
    assert n == 1 and m[0] == 2 and np.all(p[0, :3] == [1, 2, 3])
