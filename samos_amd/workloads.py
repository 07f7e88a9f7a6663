"""Synthetic workloads: particle configurations and replica parameter sets.

Each builder returns a plain dict (n, box, periodic, columns) so the same state
can be loaded into a System, an Ensemble or any object with the same set()
interface.  Used by the benchmark, the warm-start training script and the tests.
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


def surface_disks(kind, n, params, seed=5, jitter=0.05, lz=None):
    """Unit disks near a cylinder {r}, torus {c, a} or ellipsoid {a, b, c},
    displaced off the surface by up to `jitter` along x, y, z, with random
    directors and small random velocities (neither tangent: the first enforce
    projects them).  The cylinder box is periodic in z (length lz, default
    n / (2 r) so the surface density is about 1/pi); the others are open boxes."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 2.0 * math.pi, n)
    if kind == "cylinder":
        r = params[0]
        Lz = lz if lz else max(4.0, n / (2.0 * r))
        p = np.stack([r * np.cos(u), r * np.sin(u), rng.uniform(-Lz / 2, Lz / 2, n)], 1)
        box, periodic = (2.0 * r + 8.0, 2.0 * r + 8.0, Lz), True
    elif kind == "torus":
        c, a = params
        w = rng.uniform(0.0, 2.0 * math.pi, n)
        rho = c + a * np.cos(w)
        p = np.stack([rho * np.cos(u), rho * np.sin(u), a * np.sin(w)], 1)
        L = 2.0 * (c + a) + 8.0
        box, periodic = (L, L, 2.0 * a + 8.0), False
    else:
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        p = v * np.asarray(params, dtype=float)[None, :]
        L = 2.0 * max(params) + 8.0
        box, periodic = (L, L, L), False
    p = p + rng.uniform(-jitter, jitter, size=p.shape)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    vel = rng.normal(0.0, 0.1, size=(n, 3))
    cols = {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2], "nx": d[:, 0], "ny": d[:, 1], "nz": d[:, 2],
            "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]}
    return {"n": n, "box": box, "periodic": periodic, "cols": cols}


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


def sphere_replicas(counts, seed=40, phi_range=(0.5, 1.2)):
    """One sphere_disks state per replica, packing fraction uniform in phi_range."""
    rng = np.random.default_rng(seed)
    phis = rng.uniform(phi_range[0], phi_range[1], len(counts))
    return [sphere_disks(int(n), phi=float(p), seed=seed * 1000 + i) for i, (n, p) in enumerate(zip(counts, phis))]


def plane_replicas(counts, seed=50, phi=0.88, polydisperse=(0.8, 1.2)):
    """Periodic polydisperse plane_disks states, one per replica."""
    return [plane_disks(int(n), phi=phi, seed=seed * 1000 + i, polydisperse=polydisperse)
            for i, n in enumerate(counts)]


def active_replica_params(nrep, seed=60):
    """Per-replica active-matter parameters: v0 log-uniform in [0.01, 1], J
    uniform in [0, 1], rotational diffusion nu log-uniform in [0.01, 1]."""
    rng = np.random.default_rng(seed)
    return {
        "v0": np.exp(rng.uniform(math.log(0.01), math.log(1.0), nrep)),
        "J": rng.uniform(0.0, 1.0, nrep),
        "nu": np.exp(rng.uniform(math.log(0.01), math.log(1.0), nrep)),
        "seed": rng.integers(1, 2**31 - 1, nrep).astype(np.float64),
    }


ENSEMBLE_COLUMNS = ["x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "radius"]


def load_ensemble(cls, states, constraint):
    """Concatenate per-replica states into an Ensemble (replica-major columns)."""
    counts = [st["n"] for st in states]
    boxes = [st["box"] for st in states]
    periodic = [st["periodic"] for st in states]
    con = [st["R"] for st in states] if constraint == "sphere" else None
    e = cls(counts, boxes, periodic, constraint, con)
    for k in ENSEMBLE_COLUMNS:
        parts = [st["cols"][k] if k in st["cols"] else (np.ones(st["n"]) if k == "radius" else np.zeros(st["n"]))
                 for st in states]
        e.set(k, np.concatenate(parts))
    return e


def c4_ensemble_workload(nrep=256, seed=2025, nmin=1024, nmax=65536):
    """Batched active-sphere workload: per-replica N log-uniform in [nmin, nmax],
    packing fraction uniform in [0.5, 1.2], R = sqrt(N / (4 phi)); v0, J and nu
    as active_replica_params; k = 10, a = 2, polar rcut 2.4, dt = 0.005."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nrep))).astype(int)
    states = sphere_replicas(counts, seed=seed)
    p = active_replica_params(nrep, seed=seed + 1)
    params = dict(k=10.0, a=2.0, rcut_align=2.4, mu=1.0, mur=1.0, dt=0.005, **p)
    return states, params
