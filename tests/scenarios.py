"""Synthetic inputs shared by the parity tests and the benchmark.

Each builder returns a plain dict (n, box, periodic, columns) so the same
state can be loaded into the device system and into the CPU oracle.
"""
import math

import numpy as np


def plane_disks(n, phi=0.8, radius=1.0, seed=1, polydisperse=None, lz=10.0):
    """Disks uniform in a periodic square box at area fraction phi (directors uniform)."""
    rng = np.random.default_rng(seed)
    if polydisperse:
        r = rng.uniform(polydisperse[0], polydisperse[1], n)
    else:
        r = np.full(n, radius)
    L = math.sqrt(float(np.sum(math.pi * r * r)) / phi)
    th = rng.uniform(0.0, 2.0 * math.pi, n)
    cols = {
        "x": rng.uniform(-L / 2, L / 2, n), "y": rng.uniform(-L / 2, L / 2, n), "z": np.zeros(n),
        "nx": np.cos(th), "ny": np.sin(th), "nz": np.zeros(n), "radius": r,
    }
    return {"n": n, "box": (L, L, lz), "periodic": True, "cols": cols}


def sphere_disks(n, R=None, phi=1.0, seed=2, box=None):
    """Points uniform on a sphere with random tangent directors."""
    rng = np.random.default_rng(seed)
    if R is None:
        R = math.sqrt(n / (4.0 * phi))
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    p = R * v
    t = rng.normal(size=(n, 3))
    t -= np.sum(t * v, axis=1)[:, None] * v
    t /= np.linalg.norm(t, axis=1)[:, None]
    L = box if box else 2.0 * R + 8.0
    cols = {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2], "nx": t[:, 0], "ny": t[:, 1], "nz": t[:, 2]}
    return {"n": n, "box": (L, L, L), "periodic": False, "cols": cols, "R": R}


def lj_lattice(nside, rho=0.7, jitter=0.1, T=1.0, seed=3, lz=10.0):
    """Jittered square lattice at number density rho with Maxwell velocities at T."""
    rng = np.random.default_rng(seed)
    n = nside * nside
    a = 1.0 / math.sqrt(rho)
    L = nside * a
    i, j = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    x = -L / 2 + (i.ravel() + 0.5) * a + rng.uniform(-jitter, jitter, n)
    y = -L / 2 + (j.ravel() + 0.5) * a + rng.uniform(-jitter, jitter, n)
    v = rng.normal(0.0, math.sqrt(T), size=(n, 2))
    v -= v.mean(axis=0)
    cols = {"x": x, "y": y, "z": np.zeros(n), "vx": v[:, 0], "vy": v[:, 1], "vz": np.zeros(n),
            "nx": np.ones(n), "ny": np.zeros(n), "nz": np.zeros(n)}
    return {"n": n, "box": (L, L, lz), "periodic": True, "cols": cols}


def load(cls, st, **kw):
    s = cls(st["n"], st["box"], st["periodic"], **kw)
    for k, v in st["cols"].items():
        s.set(k, v)
    return s


def min_image(d, L):
    return d - L * np.round(d / L)


def rel_force_err(a, b):
    """max_i |a_i - b_i| / max_i |b_i| over stacked (n, 3) arrays."""
    scale = max(float(np.max(np.linalg.norm(b, axis=1))), 1e-300)
    return float(np.max(np.linalg.norm(a - b, axis=1))) / scale
