"""Cylinder, torus and ellipsoid constraints.

Oracle pins (CPU): the torus and ellipsoid use the generic SHAKE projection
(src/constraints/constraint.cpp:40-87), which moves a point along the gradient
taken where it started.  With tol = 0 the iteration runs to machine precision,
so the result must equal closed forms:
  * ellipsoid a = b = c = R: the sphere constraint's radial scaling, with the
    same velocity and director projections;
  * torus: the gradient at p points from the tube's centre circle to p, so the
    projection lands at centre + a (p - centre) / |p - centre|;
  * ellipsoid a != b != c: the root of the quadratic g(p - l grad g(p)) = 0
    nearest l = 0 (a mistake that re-took the gradient each step lands elsewhere);
  * cylinder: radial scaling in xy, z unchanged and wrapped without an image flag.
Plus invariants after the default tolerance (|g| <= 1e-6, v.N = n.N = 0,
|n| = 1), max_iter = 0 (positions untouched, projections still done), rotations
by phi then -phi, and projected torques against the closed-form normals.

GPU parity: one enforce (positions, velocities, directors <= 1e-12) and
noise-off Brownian, noise-on Langevin (same Philox stream) and NVE
trajectories against the oracle (<= 1e-9).
"""
import math

import numpy as np
import pytest

import oracle
from tests import scenarios as sc

SURF = {"cylinder": [3.0], "torus": [6.0, 2.0], "ellipsoid": [7.0, 5.0, 4.0]}


def cols(s, keys):
    return np.stack([s.get(k) for k in keys], 1)


def P(s):
    return cols(s, ("x", "y", "z"))


def V(s):
    return cols(s, ("vx", "vy", "vz"))


def Nd(s):
    return cols(s, ("nx", "ny", "nz"))


def loaded(kind, n=200, seed=5, jitter=0.3, **kw):
    st = sc.surface_disks(kind, n, SURF[kind], seed=seed, jitter=jitter, **kw)
    return st, sc.load(oracle.OracleSystem, st)


def normals(kind, p):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    if kind == "cylinder":
        r = SURF[kind][0]
        return np.stack([x / r, y / r, 0 * z], 1)
    if kind == "torus":
        c, _ = SURF[kind]
        rho = np.hypot(x, y)
        d = np.stack([x - c * x / rho, y - c * y / rho, z], 1)
    else:
        a, b, c = SURF[kind]
        d = np.stack([x / a ** 2, y / b ** 2, z / c ** 2], 1)
    return d / np.linalg.norm(d, axis=1)[:, None]


def g_value(kind, p):
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    if kind == "torus":
        c, a = SURF[kind]
        return (c - np.hypot(x, y)) ** 2 + z ** 2 - a * a
    a, b, c = SURF[kind]
    return x * x / a / a + y * y / b / b + z * z / c / c - 1.0


def test_ellipsoid_with_equal_axes_is_the_sphere():
    st = sc.sphere_disks(300, R=5.0, seed=6)
    rng = np.random.default_rng(0)
    for k in ("x", "y", "z"):
        st["cols"][k] = st["cols"][k] * (1.0 + rng.uniform(-0.05, 0.05, 300))
    st["cols"]["vx"] = rng.normal(size=300)
    st["cols"]["vz"] = rng.normal(size=300)
    a = sc.load(oracle.OracleSystem, st)
    b = sc.load(oracle.OracleSystem, st)
    a.constraint("sphere", 5.0)
    b.constraint_surface("ellipsoid", [5.0, 5.0, 5.0], max_iter=100, tol=0.0)
    assert np.abs(P(a) - P(b)).max() < 1e-12
    assert np.abs(V(a) - V(b)).max() < 1e-12
    assert np.abs(Nd(a) - Nd(b)).max() < 1e-12


def test_torus_projection_closed_form():
    st, s = loaded("torus")
    p0 = np.stack([st["cols"][k] for k in ("x", "y", "z")], 1)
    c, a = SURF["torus"]
    rho = np.hypot(p0[:, 0], p0[:, 1])
    centre = np.stack([c * p0[:, 0] / rho, c * p0[:, 1] / rho, 0 * rho], 1)
    d = p0 - centre
    expect = centre + a * d / np.linalg.norm(d, axis=1)[:, None]
    s.constraint_surface("torus", SURF["torus"], max_iter=100, tol=0.0)
    assert np.abs(P(s) - expect).max() < 1e-12


def test_ellipsoid_projection_is_along_the_entry_gradient():
    st, s = loaded("ellipsoid")
    p0 = np.stack([st["cols"][k] for k in ("x", "y", "z")], 1)
    ax = np.asarray(SURF["ellipsoid"])
    g0 = 2.0 * p0 / ax ** 2
    # sum ((p - l g0) / ax)^2 = 1  ->  A l^2 - 2 B l + C = 0
    A = np.sum((g0 / ax) ** 2, 1)
    B = np.sum(p0 * g0 / ax ** 2, 1)
    C = np.sum((p0 / ax) ** 2, 1) - 1.0
    lam = (B - np.sqrt(B * B - A * C)) / A
    expect = p0 - lam[:, None] * g0
    s.constraint_surface("ellipsoid", SURF["ellipsoid"], max_iter=100, tol=0.0)
    assert np.abs(P(s) - expect).max() < 1e-12
    # and not the nearest point on the surface (that would need a varying direction)
    assert np.abs(np.cross(P(s) - p0, normals("ellipsoid", P(s)))).max() > 1e-4


def test_cylinder_closed_form_and_z_wrap():
    st = sc.surface_disks("cylinder", 200, SURF["cylinder"], seed=7, jitter=0.3, lz=20.0)
    st["cols"]["z"][:5] = [10.2, -10.3, 9.9, 10.0, -10.0]
    s = sc.load(oracle.OracleSystem, st)
    p0 = np.stack([st["cols"][k] for k in ("x", "y", "z")], 1)
    s.constraint_surface("cylinder", SURF["cylinder"])
    r = SURF["cylinder"][0]
    rho = np.hypot(p0[:, 0], p0[:, 1])
    assert np.abs(P(s)[:, :2] - r * p0[:, :2] / rho[:, None]).max() < 1e-13
    z = p0[:, 2].copy()
    z[z > 10.0] -= 20.0
    z[z < -10.0] += 20.0
    assert np.array_equal(P(s)[:, 2], z)
    assert np.all(s.get("iz") == 0)


@pytest.mark.parametrize("kind", ["cylinder", "torus", "ellipsoid"])
def test_invariants_after_enforce(kind):
    st, s = loaded(kind, jitter=0.1)
    s.constraint_surface(kind, SURF[kind])
    p = P(s)
    N = normals(kind, p)
    if kind == "cylinder":
        assert np.abs(np.hypot(p[:, 0], p[:, 1]) - SURF[kind][0]).max() < 1e-12
    else:
        assert np.abs(g_value(kind, p)).max() <= 1e-6
    assert np.abs(np.sum(V(s) * N, 1)).max() < 1e-12
    assert np.abs(np.sum(Nd(s) * N, 1)).max() < 1e-12
    assert np.abs(np.linalg.norm(Nd(s), axis=1) - 1.0).max() < 1e-14


def test_zero_iterations_keep_positions():
    st, s = loaded("torus")
    s.constraint_surface("torus", SURF["torus"], max_iter=0)
    p0 = np.stack([st["cols"][k] for k in ("x", "y", "z")], 1)
    assert np.array_equal(P(s), p0)
    assert np.abs(np.sum(Nd(s) * normals("torus", p0), 1)).max() < 1e-12


@pytest.mark.parametrize("kind", ["cylinder", "torus", "ellipsoid"])
def test_rotation_inverse_and_projected_torque(kind):
    st, s = loaded(kind, jitter=0.1)
    s.constraint_surface(kind, SURF[kind], tol=0.0)
    n0 = Nd(s)
    rng = np.random.default_rng(2)
    phi = rng.uniform(-1.0, 1.0, st["n"])
    s.rotate_directors(phi)
    n1 = Nd(s)
    N = normals(kind, P(s))
    # rotation about N by phi: the angle between n0 and n1 is |phi|
    cosang = np.sum(n0 * n1, 1)
    assert np.abs(cosang - np.cos(phi)).max() < 1e-12
    assert np.abs(np.sum(np.cross(n0, n1) * N, 1) - np.sin(phi)).max() < 1e-12
    s.rotate_directors(-phi)
    assert np.abs(Nd(s) - n0).max() < 1e-12
    tau = rng.normal(size=(st["n"], 3))
    for k, key in enumerate(("tau_x", "tau_y", "tau_z")):
        s.set(key, tau[:, k])
    assert np.abs(s.project_torques() - np.sum(tau * N, 1)).max() < 1e-12


def test_torus_rejects_thick_tube():
    st, s = loaded("torus")
    with pytest.raises(RuntimeError, match="Tube radius has to be smaller"):
        s.constraint_surface("torus", [2.0, 2.0])


# ------------------------------------------------------------------ GPU ---
def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


def pos_err(a, b, box, periodic):
    d = a - b
    if periodic:
        for k in range(3):
            d[:, k] = sc.min_image(d[:, k], box[k])
    return float(np.abs(d).max())


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["cylinder", "torus", "ellipsoid"])
def test_gpu_enforce(gpu, kind):
    st = sc.surface_disks(kind, 2000, SURF[kind], seed=8, jitter=0.3)
    d, o = both(gpu, st)
    for s in (d, o):
        s.constraint_surface(kind, SURF[kind])
    assert np.abs(P(d) - P(o)).max() <= 1e-12
    assert np.abs(V(d) - V(o)).max() <= 1e-12
    assert np.abs(Nd(d) - Nd(o)).max() <= 1e-12
    assert np.array_equal(d.get("iz"), o.get("iz"))


def setup_active(s, kind):
    s.nlist(2.4, 0.3)
    s.pair_soft(1.0, 2.0)
    s.align_polar(0.5, 2.4)
    s.constraint_surface(kind, SURF[kind])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["cylinder", "torus", "ellipsoid"])
def test_gpu_brownian_noise_off(gpu, kind):
    n = {"cylinder": 1200, "torus": 1500, "ellipsoid": 1500}[kind]
    st = sc.surface_disks(kind, n, [2.0 * q for q in SURF[kind]], seed=9, jitter=0.05)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.5, 2.4)
        s.constraint_surface(kind, [2.0 * q for q in SURF[kind]])
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.3, nu=0.0, mu=1.0, seed=1)
        s.run(299)
    assert pos_err(P(d), P(o), st["box"], st["periodic"]) <= 1e-9
    assert np.abs(Nd(d) - Nd(o)).max() <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["torus", "ellipsoid"])
def test_gpu_langevin_noise_on_same_stream(gpu, kind):
    st = sc.surface_disks(kind, 1500, [2.0 * q for q in SURF[kind]], seed=10, jitter=0.05)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint_surface(kind, [2.0 * q for q in SURF[kind]])
        s.set_timestep(0.005)
        s.integrator_langevin(gamma=1.0, seed=4, temp=(0, 0.5, 0.5, 1))
        s.run(99)
    assert pos_err(P(d), P(o), st["box"], False) <= 1e-9
    assert np.abs(V(d) - V(o)).max() <= 1e-9


@pytest.mark.gpu
def test_gpu_nve_cylinder(gpu):
    st = sc.surface_disks("cylinder", 1200, [6.0], seed=11, jitter=0.05)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(2.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.constraint_surface("cylinder", [6.0])
        s.set_timestep(0.005)
        s.integrator_nve(limit=0.01, angle_limit=0.05)
        s.run(199)
    assert pos_err(P(d), P(o), st["box"], True) <= 1e-9
    assert np.abs(V(d) - V(o)).max() <= 1e-9
    assert np.abs(Nd(d) - Nd(o)).max() <= 1e-9
