"""nbody_amd — Python binding (ctypes) of libnbody, the MI355X N-body engine.

This module is a thin host-side mirror of the C ABI in include/nbody.h and
include/nbody_ic.h; every call goes straight to lib/libnbody.so. There is no
fallback: if the library is missing, `load()` raises.

The Simulation class follows the reference's ParticleSolver contract
(include/ParticleSolver.h:8-15): `step()` is updateParticlePositions,
`squared_softening` is getSquaredSoftening(), `uses_gpu` is usesGPU().
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libnbody.so"

NB_OK = 0
NB_E_ARG = -1
NB_E_HIP = -2
NB_E_OOM = -3
NB_E_BUDGET = -4
NB_E_TREE_OVERFLOW = -5
NB_E_NONFINITE = -6
NB_E_NOGPU = -7

SOLVER_DIRECT = 0
SOLVER_BH = 1
INTEG_KDK_LAGGED = 0
INTEG_EULER = 1
F_NO_GRAPH = 0x1
F_BH_PER_BODY = 0x4
F_BH_AXIS_BOX = 0x8
F_BH_SINGLE_RANGES = 0x10
F_BH_PAIRS = 0x20
BH_PAIR_MIN_BODIES = 1 << 20  # NB_BH_PAIR_MIN_BODIES
BH_WAVE_BODIES = 32  # NB_BH_WAVE_BODIES
NOALLOC_MAX_BODIES = 65536  # NB_NOALLOC_MAX_BODIES
BATCH_UNLIMITED = 0x7FFFFFFF  # NB_BATCH_UNLIMITED
SERVE_ARRIVAL, SERVE_SHORTEST = 0, 1  # NB_SERVE_*


class NBodyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("eps2", C.c_float),
        ("dt", C.c_float),
        ("G", C.c_float),
        ("theta", C.c_float),
        ("solver", C.c_int32),
        ("integrator", C.c_int32),
        ("leaf_size", C.c_int32),
        ("bodies_per_lane", C.c_int32),
        ("j_split", C.c_int32),
        ("flags", C.c_int32),
    ]


class NodeView(C.Structure):
    _fields_ = [
        ("com", C.c_float * 3),
        ("mass", C.c_float),
        ("first_child", C.c_int32),
        ("next", C.c_int32),
        ("body_start", C.c_int32),
        ("body_count", C.c_int32),
        ("level", C.c_int32),
        ("open_r2", C.c_float),
    ]


NODE_DTYPE = np.dtype(
    [
        ("com", np.float32, 3),
        ("mass", np.float32),
        ("first_child", np.int32),
        ("next", np.int32),
        ("body_start", np.int32),
        ("body_count", np.int32),
        ("level", np.int32),
        ("open_r2", np.float32),
    ]
)
assert NODE_DTYPE.itemsize == C.sizeof(NodeView)


class Timing(C.Structure):
    _fields_ = [
        ("force_ms", C.c_double),
        ("force_calls", C.c_int64),
        ("step_ms", C.c_double),
        ("step_calls", C.c_int64),
        ("build_ms", C.c_double),
        ("build_calls", C.c_int64),
        ("interactions", C.c_int64),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("max_bodies", C.c_int64),
        ("max_clusters", C.c_int32),
        ("integrator", C.c_int32),
        ("bodies_per_lane", C.c_int32),
        ("j_split", C.c_int32),
        ("flags", C.c_int32),
        ("pad_", C.c_int32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_int64),
        ("direct_pool", C.c_int64),
        ("max_jobs", C.c_int32),
        ("bh_threshold", C.c_int32),
        ("theta", C.c_float),
        ("leaf_size", C.c_int32),
        ("ic_threads", C.c_int32),
        ("lookahead", C.c_int32),
        ("tick_budget_ms", C.c_float),
        ("policy", C.c_int32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("ticks", C.c_int64),
        ("steps", C.c_int64),
        ("pooled_contexts", C.c_int64),
        ("select_ms", C.c_double),
        ("idle_wait_ms", C.c_double),
        ("start_ms", C.c_double),
        ("enqueue_ms", C.c_double),
        ("device_wait_ms", C.c_double),
        ("retire_ms", C.c_double),
    ]


class Job(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("steps", C.c_int32),
        ("ic", C.c_int32),
        ("seed", C.c_uint64),
        ("eps2", C.c_float),
        ("dt", C.c_float),
        ("G", C.c_float),
        ("pad_", C.c_int32),
    ]


class JobStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("solver", C.c_int32),
        ("steps_done", C.c_int32),
        ("steps", C.c_int32),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]


IC_PLUMMER = 0
IC_BALL = 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, -1

class Tuning(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("bodies_per_lane", C.c_int32),
        ("j_split", C.c_int32),
        ("leaf_size", C.c_int32),
        ("bh_single_ranges", C.c_int32),
        ("direct_ms", C.c_double),
        ("bh_ms", C.c_double),
    ]


TUNE_DIRECT = 0x1
TUNE_BH = 0x2

_lib = None

_P = C.c_void_p
_SIGS = {
    "nb_create": (C.c_int, [C.POINTER(_P), C.POINTER(Config)]),
    "nb_destroy": (None, [_P]),
    "nb_last_error": (C.c_char_p, []),
    "nb_upload": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64]),
    "nb_upload_f64": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64]),
    "nb_step": (C.c_int, [_P, C.c_int32]),
    "nb_sync": (C.c_int, [_P]),
    "nb_forces": (C.c_int, [_P, _P]),
    "nb_forces_f64": (C.c_int, [_P, _P]),
    "nb_download": (C.c_int, [_P, _P, _P, _P]),
    "nb_download_f64": (C.c_int, [_P, _P, _P, _P]),
    "nb_energy": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), _P]),
    "nb_save_snapshot": (C.c_int, [_P, C.c_char_p]),
    "nb_load_snapshot": (C.c_int, [_P, C.c_char_p]),
    "nb_debug_morton": (C.c_int, [_P, _P, _P]),
    "nb_debug_tree": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64), _P]),
    "nb_debug_set_walk_stack": (C.c_int, [_P, C.c_int32]),
    "nb_bh_forces_noalloc": (C.c_int, [_P, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_int32,
                                       C.c_int32, _P, _P, _P]),
    "nb_reorder": (C.c_int, [_P]),
    "nb_render_view": (C.c_int, [_P, _P]),
    "nb_get_order": (C.c_int, [_P, _P]),
    "nb_set_timing": (C.c_int, [_P, C.c_int32]),
    "nb_get_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "nb_count_interactions": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "nb_tune_leaf": (C.c_int, [_P, C.c_int32, C.POINTER(Tuning)]),
    "nb_count_interactions_per_body": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "nb_direct_f64": (C.c_int, [_P, _P, C.c_int64, _P]),
    "nb_device_info": (C.c_int, [C.c_char_p, C.c_int32]),
    "nb_batch_create": (C.c_int, [C.POINTER(_P), C.POINTER(BatchConfig)]),
    "nb_batch_destroy": (None, [_P]),
    "nb_batch_add": (C.c_int, [_P, _P, _P, _P, _P, C.c_int64, C.c_float, C.c_float, C.c_float,
                               C.POINTER(C.c_int32)]),
    "nb_batch_remove": (C.c_int, [_P, C.c_int32]),
    "nb_batch_step": (C.c_int, [_P, C.c_int32]),
    "nb_batch_set_quotas": (C.c_int, [_P, _P, _P, C.c_int32]),
    "nb_batch_sync": (C.c_int, [_P]),
    "nb_batch_forces": (C.c_int, [_P, C.c_int32, _P]),
    "nb_batch_download": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "nb_batch_energy": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "nb_batch_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "nb_batch_tune": (C.c_int, [_P, C.c_int32, C.POINTER(Tuning)]),
    "nb_batch_set_timing": (C.c_int, [_P, C.c_int32]),
    "nb_batch_get_timing": (C.c_int, [_P, C.POINTER(Timing)]),
    "nb_server_create": (C.c_int, [C.POINTER(_P), C.POINTER(ServerConfig)]),
    "nb_server_destroy": (None, [_P]),
    "nb_server_submit": (C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_int64)]),
    "nb_server_tick": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "nb_server_poll": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "nb_server_status": (C.c_int, [_P, C.c_int64, C.POINTER(JobStatus)]),
    "nb_server_result": (C.c_int, [_P, C.c_int64, _P, _P]),
    "nb_server_forget": (C.c_int, [_P, C.c_int64]),
    "nb_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "nb_server_now": (C.c_double, [_P]),
    "nb_autotune": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(Tuning)]),
    "nb_get_tuning": (C.c_int, [C.c_int64, C.POINTER(Tuning)]),
    "nb_ic_ball": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_double, _P, _P, _P]),
    "nb_ic_plummer": (
        C.c_int,
        [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double, _P, _P, _P],
    ),
    "nb_ic_discs": (
        C.c_int,
        [C.c_uint64, C.c_int64] + [C.c_double] * 8 + [_P, _P, _P],
    ),
    "nb_ic_lagrange": (C.c_int, [C.c_double, _P, _P, _P]),
    "nb_ic_cube": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, _P, _P, _P]),
    "nb_ic_cube_surface": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, _P, _P, _P]),
    "nb_ic_ball_world": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, _P, _P, _P]),
    "nb_ic_sphere_surface": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_int, _P, _P, _P]),
    "nb_ic_galaxy": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_int, _P, _P, _P]),
    "nb_ic_read_file": (C.c_int64, [C.c_char_p, C.c_int64, _P, _P, _P, _P]),
    "nb_ic_write_file": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, _P, _P]),
    "nb_ic_log_uniform_sizes": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, _P]),
    "nb_ic_uniform": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, _P]),
}


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libnbody.so (raises FileNotFoundError if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is missing: build it with `make -C n-body_amd`")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int):
    if rc != NB_OK:
        raise NBodyError(rc, load().nb_last_error().decode(errors="replace"))


def device_info() -> str:
    buf = C.create_string_buffer(256)
    rc = load().nb_device_info(buf, 256)
    _check(rc)
    return buf.value.decode()


def autotune(n: int, what: int = TUNE_DIRECT | TUNE_BH, reps: int = 3) -> dict:
    """Time candidate launch shapes on the device and install the fastest as the
    defaults for systems of n's size class (nb_autotune)."""
    t = Tuning()
    _check(load().nb_autotune(n, what, reps, C.byref(t)))
    return {k: getattr(t, k) for k, _ in Tuning._fields_ if k != "pad_"}


def get_tuning(n: int) -> dict:
    t = Tuning()
    _check(load().nb_get_tuning(n, C.byref(t)))
    return {k: getattr(t, k) for k, _ in Tuning._fields_ if k != "pad_"}


# ---- seeded initial conditions -------------------------------------------------
def ic_ball(seed: int, n: int, radius: float = 100.0, mass_each: float = 1.0):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    m = np.empty(n, np.float32)
    _check(load().nb_ic_ball(seed, n, radius, mass_each, _ptr(pos), _ptr(vel), _ptr(m)))
    return pos, vel, m


def ic_plummer(seed: int, n: int, a: float = 50.0, total_mass: float = 1e4,
               rmax_over_a: float = 10.0, G: float = 1.0):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    m = np.empty(n, np.float32)
    _check(load().nb_ic_plummer(seed, n, a, total_mass, rmax_over_a, G, _ptr(pos), _ptr(vel), _ptr(m)))
    return pos, vel, m


def ic_discs(seed: int, n: int, scale_length: float = 30.0, scale_height: float = 2.0,
             central_factor: float = 1e4, total_mass: float = 1e4, separation: float = 300.0,
             closing_speed: float = 0.5 / 0.0035, eps2: float = 40.0, G: float = 1.0):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    m = np.empty(n, np.float32)
    _check(load().nb_ic_discs(seed, n, scale_length, scale_height, central_factor, total_mass,
                              separation, closing_speed, eps2, G, _ptr(pos), _ptr(vel), _ptr(m)))
    return pos, vel, m


def ic_lagrange(world_size: float = 5.0):
    pos = np.empty((3, 3), np.float64)
    vel = np.empty((3, 3), np.float64)
    m = np.empty(3, np.float32)
    _check(load().nb_ic_lagrange(world_size, _ptr(pos), _ptr(vel), _ptr(m)))
    return pos, vel, m


def _ic_world(fn, seed, n, world_size, *mode):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    m = np.empty(n, np.float32)
    _check(fn(seed, n, world_size, *mode, _ptr(pos), _ptr(vel), _ptr(m)))
    return pos, vel, m


def ic_cube(seed: int, n: int, world_size: float = 5.0):
    """Reference -i 1 (ParticleSystemCubeInitializer.cpp:8-29), seeded."""
    return _ic_world(load().nb_ic_cube, seed, n, world_size)


def ic_cube_surface(seed: int, n: int, world_size: float = 5.0):
    """Reference -i 6 (ParticleSystemCubeSurface.cpp:8-49), seeded."""
    return _ic_world(load().nb_ic_cube_surface, seed, n, world_size)


def ic_ball_world(seed: int, n: int, world_size: float = 5.0):
    """Reference -i 5 (ParticleSystemBall.cpp:10-92), seeded."""
    return _ic_world(load().nb_ic_ball_world, seed, n, world_size)


def ic_sphere_surface(seed: int, n: int, world_size: float = 5.0, pattern: int = -1):
    """Reference -i 4 (ParticleSystemSphere.cpp:20-133), seeded; pattern 0 outward,
    1 inward, 2 swirl, -1 drawn from the seed."""
    return _ic_world(load().nb_ic_sphere_surface, seed, n, world_size, pattern)


def ic_galaxy(seed: int, n: int, world_size: float = 5.0, style: int = -1):
    """Reference -i 2 (ParticleSystemGalaxyInitializer.cpp:12-210), seeded; style 0 or
    1 selects the reference's two branches, -1 draws it from the seed."""
    return _ic_world(load().nb_ic_galaxy, seed, n, world_size, style)


def read_particle_file(path):
    """Reference text format (ParticleSystemFile.cpp:10-90) -> pos4, vel4, acc4, mass."""
    lib = load()
    n = lib.nb_ic_read_file(str(path).encode(), 0, None, None, None, None)
    if n < 0:
        raise NBodyError(int(n), f"cannot read particle file {path}")
    p, v, a = (np.zeros((n, 4), np.float32) for _ in range(3))
    m = np.zeros(n, np.float32)
    lib.nb_ic_read_file(str(path).encode(), n, _ptr(p), _ptr(v), _ptr(a), _ptr(m))
    return p, v, a, m


def write_particle_file(path, pos4, vel4, acc4, mass):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (pos4, vel4, acc4, mass)]
    _check(load().nb_ic_write_file(str(path).encode(), arrs[3].shape[0], *(_ptr(x) for x in arrs)))


def log_uniform_sizes(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(count, np.int32)
    _check(load().nb_ic_log_uniform_sizes(seed, count, lo, hi, _ptr(out)))
    return out


def uniform(seed: int, count: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty(count, np.float64)
    _check(load().nb_ic_uniform(seed, count, lo, hi, _ptr(out)))
    return out


def to_vec4(a3: np.ndarray) -> np.ndarray:
    """(n,3) -> (n,4) float32 with w = 0 (the reference's vec4 layout)."""
    out = np.zeros((a3.shape[0], 4), np.float32)
    out[:, :3] = a3
    return out


# ---- the solver --------------------------------------------------------------------
class Simulation:
    """One device-resident particle system and its solver (an nb_ctx).

    solver: SOLVER_DIRECT (drop-in for the reference's v4) or SOLVER_BH (v7).
    """

    def __init__(self, n: int, eps2: float, dt: float, solver: int = SOLVER_DIRECT,
                 integrator: int = INTEG_KDK_LAGGED, G: float = 1.0, theta: float = 0.5,
                 leaf_size: int = 0, bodies_per_lane: int = 0, j_split: int = 0, flags: int = 0):
        self._lib = load()
        self.n = int(n)
        cfg = Config(n=self.n, eps2=eps2, dt=dt, G=G, theta=theta, solver=solver,
                     integrator=integrator, leaf_size=leaf_size, bodies_per_lane=bodies_per_lane,
                     j_split=j_split, flags=flags)
        self.cfg = cfg
        self._ctx = C.c_void_p()
        _check(self._lib.nb_create(C.byref(self._ctx), C.byref(cfg)))

    # ParticleSolver contract
    @property
    def squared_softening(self) -> float:
        return float(self.cfg.eps2)

    uses_gpu = True

    def close(self):
        if self._ctx:
            self._lib.nb_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, pos, vel, mass, acc=None):
        """pos/vel/acc: (n,3) float64 or (n,4) float32 (reference vec4 layout); mass: (n,)."""
        mass = np.ascontiguousarray(mass, np.float32)
        if pos.shape[1] == 4:
            pos = np.ascontiguousarray(pos, np.float32)
            vel = np.ascontiguousarray(vel, np.float32)
            acc = None if acc is None else np.ascontiguousarray(acc, np.float32)
            _check(self._lib.nb_upload(self._ctx, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(mass), self.n))
        else:
            pos = np.ascontiguousarray(pos, np.float64)
            vel = np.ascontiguousarray(vel, np.float64)
            acc = None if acc is None else np.ascontiguousarray(acc, np.float64)
            _check(self._lib.nb_upload_f64(self._ctx, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(mass), self.n))

    def step(self, nsteps: int = 1):
        _check(self._lib.nb_step(self._ctx, nsteps))

    def reorder(self):
        """Permute the device state into Morton order (nb_reorder); the caller's
        body order is kept at every copy."""
        _check(self._lib.nb_reorder(self._ctx))

    def render_view(self) -> np.ndarray:
        """(x, y, z, |v|) per body, float32 (nb_render_view)."""
        out = np.empty((self.n, 4), np.float32)
        _check(self._lib.nb_render_view(self._ctx, _ptr(out)))
        return out

    def order(self) -> np.ndarray:
        o = np.empty(self.n, np.int32)
        _check(self._lib.nb_get_order(self._ctx, _ptr(o)))
        return o

    def sync(self):
        _check(self._lib.nb_sync(self._ctx))

    def forces(self) -> np.ndarray:
        out = np.empty((self.n, 4), np.float32)
        _check(self._lib.nb_forces(self._ctx, _ptr(out)))
        return out

    def forces_f64(self) -> np.ndarray:
        out = np.empty((self.n, 3), np.float64)
        _check(self._lib.nb_forces_f64(self._ctx, _ptr(out)))
        return out

    def download(self):
        p = np.empty((self.n, 4), np.float32)
        v = np.empty((self.n, 4), np.float32)
        a = np.empty((self.n, 4), np.float32)
        _check(self._lib.nb_download(self._ctx, _ptr(p), _ptr(v), _ptr(a)))
        return p, v, a

    def download_f64(self):
        p = np.empty((self.n, 3), np.float64)
        v = np.empty((self.n, 3), np.float64)
        a = np.empty((self.n, 3), np.float64)
        _check(self._lib.nb_download_f64(self._ctx, _ptr(p), _ptr(v), _ptr(a)))
        return p, v, a

    def energy(self):
        ek = C.c_double()
        ep = C.c_double()
        p = np.empty(3, np.float64)
        _check(self._lib.nb_energy(self._ctx, C.byref(ek), C.byref(ep), _ptr(p)))
        return ek.value, ep.value, p

    def save_snapshot(self, path):
        _check(self._lib.nb_save_snapshot(self._ctx, str(path).encode()))

    def load_snapshot(self, path):
        _check(self._lib.nb_load_snapshot(self._ctx, str(path).encode()))

    def debug_set_walk_stack(self, entries: int):
        """Cap the traversal's LDS stack region (0 restores the full 512)."""
        _check(self._lib.nb_debug_set_walk_stack(self._ctx, entries))

    def debug_morton(self):
        codes = np.empty(self.n, np.uint32)
        perm = np.empty(self.n, np.uint32)
        _check(self._lib.nb_debug_morton(self._ctx, _ptr(codes), _ptr(perm)))
        return codes, perm

    def debug_tree(self):
        cnt = C.c_int64()
        bbox = np.empty(6, np.float32)
        _check(self._lib.nb_debug_tree(self._ctx, None, 0, C.byref(cnt), _ptr(bbox)))
        nodes = np.empty(cnt.value, NODE_DTYPE)
        _check(self._lib.nb_debug_tree(self._ctx, _ptr(nodes), cnt.value, C.byref(cnt), _ptr(bbox)))
        return nodes, bbox

    def set_timing(self, enable: bool):
        _check(self._lib.nb_set_timing(self._ctx, 1 if enable else 0))

    def timing(self) -> dict:
        t = Timing()
        _check(self._lib.nb_get_timing(self._ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def count_interactions(self) -> int:
        v = C.c_int64()
        _check(self._lib.nb_count_interactions(self._ctx, C.byref(v)))
        return v.value

    def count_interactions_per_body(self) -> int:
        v = C.c_int64()
        _check(self._lib.nb_count_interactions_per_body(self._ctx, C.byref(v)))
        return v.value

    def direct_f64(self, targets) -> np.ndarray:
        """FP64 direct-sum accelerations of the given bodies, on the device."""
        t = np.ascontiguousarray(targets, np.int64)
        out = np.empty((len(t), 3), np.float64)
        _check(self._lib.nb_direct_f64(self._ctx, _ptr(t), len(t), _ptr(out)))
        return out

    def tune_leaf(self, reps: int = 3) -> dict:
        """Barnes-Hut leaf size timed on this system's own bodies (nb_tune_leaf)."""
        t = Tuning()
        _check(self._lib.nb_tune_leaf(self._ctx, reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_ if k != "pad_"}


class Batch:
    """Many independent clusters in one persistent launch (nb_batch_*)."""

    def __init__(self, max_bodies: int, max_clusters: int, integrator: int = INTEG_KDK_LAGGED,
                 bodies_per_lane: int = 0, j_split: int = 0, flags: int = 0):
        self._lib = load()
        cfg = BatchConfig(max_bodies=max_bodies, max_clusters=max_clusters, integrator=integrator,
                          bodies_per_lane=bodies_per_lane, j_split=j_split, flags=flags)
        self._b = C.c_void_p()
        _check(self._lib.nb_batch_create(C.byref(self._b), C.byref(cfg)))
        self.sizes: dict[int, int] = {}

    def close(self):
        if self._b:
            self._lib.nb_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, pos, vel, mass, eps2: float, dt: float, G: float = 1.0, acc=None) -> int:
        pos = np.ascontiguousarray(pos, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        mass = np.ascontiguousarray(mass, np.float32)
        acc = None if acc is None else np.ascontiguousarray(acc, np.float64)
        cid = C.c_int32()
        _check(self._lib.nb_batch_add(self._b, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(mass),
                                      mass.shape[0], eps2, dt, G, C.byref(cid)))
        self.sizes[cid.value] = mass.shape[0]
        return cid.value

    def try_add(self, *a, **k):
        """add() that returns None instead of raising when the pool is full."""
        try:
            return self.add(*a, **k)
        except NBodyError as e:
            if e.code == NB_E_BUDGET:
                return None
            raise

    def remove(self, cid: int):
        _check(self._lib.nb_batch_remove(self._b, cid))
        self.sizes.pop(cid, None)

    def step(self, nsteps: int = 1):
        _check(self._lib.nb_batch_step(self._b, nsteps))

    def set_quotas(self, quotas: dict):
        """Steps per step() call for the given clusters ({id: steps};
        BATCH_UNLIMITED lifts the limit)."""
        ids = np.ascontiguousarray(list(quotas.keys()), np.int32)
        st = np.ascontiguousarray(list(quotas.values()), np.int32)
        _check(self._lib.nb_batch_set_quotas(self._b, _ptr(ids), _ptr(st), len(ids)))

    def tune(self, reps: int = 3) -> dict:
        """Time every launch shape on the current clusters and keep the fastest."""
        t = Tuning()
        _check(self._lib.nb_batch_tune(self._b, reps, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_ if k != "pad_"}

    def sync(self):
        _check(self._lib.nb_batch_sync(self._b))

    def forces(self, cid: int) -> np.ndarray:
        out = np.empty((self.sizes[cid], 3), np.float64)
        _check(self._lib.nb_batch_forces(self._b, cid, _ptr(out)))
        return out

    def download(self, cid: int):
        n = self.sizes[cid]
        p, v, a = (np.empty((n, 3), np.float64) for _ in range(3))
        _check(self._lib.nb_batch_download(self._b, cid, _ptr(p), _ptr(v), _ptr(a)))
        return p, v, a

    def energy(self, cid: int):
        ek = C.c_double()
        ep = C.c_double()
        _check(self._lib.nb_batch_energy(self._b, cid, C.byref(ek), C.byref(ep)))
        return ek.value, ep.value

    def info(self):
        c = C.c_int32()
        b = C.c_int64()
        i = C.c_int64()
        _check(self._lib.nb_batch_info(self._b, C.byref(c), C.byref(b), C.byref(i)))
        return c.value, b.value, i.value

    def set_timing(self, enable: bool):
        _check(self._lib.nb_batch_set_timing(self._b, 1 if enable else 0))

    def timing(self) -> dict:
        t = Timing()
        _check(self._lib.nb_batch_get_timing(self._b, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}


class Server:
    """Continuous-batching request loop (nb_server_*)."""

    def __init__(self, hbm_budget: int = 0, direct_pool: int = 0, max_jobs: int = 0,
                 bh_threshold: int = 0, theta: float = 0.0, leaf_size: int = 0, ic_threads: int = 0,
                 lookahead: int = 0, tick_budget_ms: float = 0.0, policy: int = 0):
        self._lib = load()
        cfg = ServerConfig(hbm_budget=hbm_budget, direct_pool=direct_pool, max_jobs=max_jobs, policy=policy,
                           bh_threshold=bh_threshold, theta=theta, leaf_size=leaf_size,
                           ic_threads=ic_threads, lookahead=lookahead,
                           tick_budget_ms=tick_budget_ms)
        self._s = C.c_void_p()
        _check(self._lib.nb_server_create(C.byref(self._s), C.byref(cfg)))
        self.sizes: dict[int, int] = {}

    def close(self):
        if self._s:
            self._lib.nb_server_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, n: int, steps: int, ic: int = IC_PLUMMER, seed: int = 0, eps2: float = 40.0,
               dt: float = 0.0035, G: float = 1.0) -> int:
        job = Job(n=n, steps=steps, ic=ic, seed=seed, eps2=eps2, dt=dt, G=G)
        jid = C.c_int64()
        _check(self._lib.nb_server_submit(self._s, C.byref(job), C.byref(jid)))
        self.sizes[jid.value] = n
        return jid.value

    def tick(self):
        r = C.c_int32()
        f = C.c_int32()
        _check(self._lib.nb_server_tick(self._s, C.byref(r), C.byref(f)))
        return r.value, f.value

    def poll(self, cap: int = 4096) -> list:
        """Ids of the jobs retired since the last poll (nb_server_poll)."""
        out, buf, cnt = [], np.empty(cap, np.int64), C.c_int32()
        while True:
            _check(self._lib.nb_server_poll(self._s, _ptr(buf), cap, C.byref(cnt)))
            out.extend(int(x) for x in buf[:cnt.value])
            if cnt.value < cap:
                return out

    def status(self, jid: int) -> dict:
        st = JobStatus()
        _check(self._lib.nb_server_status(self._s, jid, C.byref(st)))
        return {k: getattr(st, k) for k, _ in JobStatus._fields_}

    def result(self, jid: int):
        n = self.sizes[jid]
        p = np.empty((n, 3), np.float64)
        v = np.empty((n, 3), np.float64)
        _check(self._lib.nb_server_result(self._s, jid, _ptr(p), _ptr(v)))
        return p, v

    def forget(self, jid: int):
        _check(self._lib.nb_server_forget(self._s, jid))

    def now(self) -> float:
        return self._lib.nb_server_now(self._s)

    def stats(self) -> dict:
        """Per-phase host time of the serving loop (nb_server_get_stats)."""
        st = ServerStats()
        _check(self._lib.nb_server_get_stats(self._s, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
