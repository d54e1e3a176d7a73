"""Shared test helpers (not collected)."""
import numpy as np


def vec4(a3):
    out = np.zeros((a3.shape[0], 4), np.float32)
    out[:, :3] = a3
    return out


def mass4(m):
    out = np.zeros((m.shape[0], 4), np.float32)
    out[:, 0] = m
    return out


def rel_err(a, b):
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    den = np.linalg.norm(b, axis=1)
    den = np.where(den > 0, den, 1.0)
    return np.linalg.norm(a - b, axis=1) / den


def seeded_ball(n, seed=1, radius=100.0):
    """Uniform ball from numpy (the oracle tests do not depend on libnbody)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    r = radius * np.cbrt(u)
    z = 2 * rng.random(n) - 1
    phi = 2 * np.pi * rng.random(n)
    s = np.sqrt(1 - z * z)
    pos = np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * z], 1)
    return pos
