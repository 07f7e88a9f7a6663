"""Further pair potentials: morse, gaussian, yukawa, soft_attractive.

Oracle pins (CPU): two-particle closed forms of energy and force written from
the textbook potentials (Morse D((1 - e^{-a(r-re)})^2 - 1), double Gaussian,
screened Coulomb g e^{-kappa r}/r, piecewise soft attraction), the energy
shift at rcut, Newton's third law and the finite-difference gradient of the
energy on random disks; for soft_attractive the three distance branches and
the coordination recount against a brute-force count.

GPU parity: forces, the energy and the coordination equal the oracle's on the
same inputs (relative force error <= 1e-12, energy rel 1e-12, coordination
exact), and a noise-off Brownian trajectory with two of them stays within 1e-9.
"""
import math

import numpy as np
import pytest

import oracle
from tests import scenarios as sc

PARAMS = {
    "morse": [2.0, 1.5, 1.2, 3.0],                       # D, a, re, rcut
    "gaussian": [2.0, -1.0, 1.0, 3.0, 0.0, 1.5, 3.0],    # A, B, alpha, beta, rA, rB, rcut
    "yukawa": [1.5, 0.8, 3.0],                           # g, kappa, rcut
    "soft_attractive": [2.0, 2.0, 1.25],                 # k, a, re_fact
}


def two(dx):
    s = oracle.OracleSystem(2, (50.0, 50.0, 10.0), True)
    s.set("x", [0.0, dx])
    s.set("y", [0.0, 0.0])
    s.set("z", [0.0, 0.0])
    return s


def closed_form(kind, r, shifted=False):
    """(energy, x-force on the particle at +r) from the potential's formula."""
    q = PARAMS[kind]
    if kind == "morse":
        D, a, re, rc = q
        V = lambda x: D * ((1.0 - math.exp(-a * (x - re))) ** 2 - 1.0)  # noqa: E731
        dV = 2.0 * D * a * math.exp(-a * (r - re)) * (1.0 - math.exp(-a * (r - re)))
        return V(r) - (V(rc) if shifted else 0.0), -dV
    if kind == "gaussian":
        A, B, al, be, rA, rB, rc = q
        V = lambda x: A * math.exp(-al * (x - rA) ** 2) + B * math.exp(-be * (x - rB) ** 2)  # noqa: E731
        dV = (-2.0 * al * (r - rA) * A * math.exp(-al * (r - rA) ** 2)
              - 2.0 * be * (r - rB) * B * math.exp(-be * (r - rB) ** 2))
        return V(r) - (V(rc) if shifted else 0.0), -dV
    if kind == "yukawa":
        g, ka, rc = q
        V = g * math.exp(-ka * r) / r
        return V, V * (ka + 1.0 / r)
    k, a, fact = q
    r_turn = fact * a
    r_f = 2.0 * r_turn - a
    if r <= r_turn:
        return 0.5 * k * (a - r) ** 2, k * (a - r)
    if r < r_f:
        return -0.5 * k * (r_f - r) ** 2, -k * (r_f - r)
    return 0.0, 0.0


CASES = [("morse", r) for r in (0.9, 1.2, 2.2)] + [("gaussian", r) for r in (0.4, 1.5, 2.6)] + \
        [("yukawa", r) for r in (0.7, 1.9, 2.9)] + [("soft_attractive", r) for r in (1.5, 2.7, 3.2)]


@pytest.mark.parametrize("kind,r", CASES)
@pytest.mark.parametrize("shifted", [False, True])
def test_two_particle_closed_form(kind, r, shifted):
    s = two(r)
    s.nlist(3.5, 0.3)
    s.pair_extra(kind, PARAMS[kind], shifted=shifted)
    e = s.compute_forces()
    # yukawa and soft_attractive have no energy shift (their headers take no "shifted")
    E, F = closed_form(kind, r, shifted and kind in ("morse", "gaussian"))
    fx = s.get("fx")
    assert fx[1] == pytest.approx(F, rel=1e-13, abs=1e-15)
    assert fx[0] == -fx[1]
    assert e["pair"] == pytest.approx(E, rel=1e-13, abs=1e-15)
    assert s.extra_energy() == e["pair"]


@pytest.mark.parametrize("kind", ["morse", "gaussian", "yukawa"])
def test_beyond_cutoff_is_zero(kind):
    s = two(PARAMS[kind][-1] + 0.05)
    s.nlist(3.5, 0.3)
    s.pair_extra(kind, PARAMS[kind])
    e = s.compute_forces()
    assert np.all(s.get("fx") == 0.0) and e["pair"] == 0.0


def test_shifted_morse_energy_vanishes_at_cutoff():
    s = two(3.0 - 1e-9)
    s.nlist(3.5, 0.3)
    s.pair_extra("morse", PARAMS["morse"], shifted=True)
    assert abs(s.compute_forces()["pair"]) < 1e-8


def test_use_radii_sets_the_distance_scale():
    # morse re -> a_i + a_j, gaussian rB -> a_i + a_j, soft_attractive a -> a_i + a_j
    for kind, idx in [("morse", 2), ("gaussian", 5), ("soft_attractive", 1)]:
        q = list(PARAMS[kind])
        s = two(1.3)
        s.set("radius", [0.6, 0.9])
        s.nlist(3.5, 0.3)
        s.pair_extra(kind, q, use_radii=True)
        s.compute_forces()
        q2 = list(q)
        q2[idx] = 1.5
        s2 = two(1.3)
        s2.nlist(3.5, 0.3)
        s2.pair_extra(kind, q2)
        s2.compute_forces()
        assert s.get("fx")[1] == s2.get("fx")[1], kind


def test_soft_attractive_branches_and_coordination():
    st = sc.plane_disks(400, phi=0.7, seed=31)
    s = sc.load(oracle.OracleSystem, st)
    s.nlist(3.2, 0.3)
    s.pair_extra("soft_attractive", PARAMS["soft_attractive"])
    s.compute_forces()
    x, y = s.get("x"), s.get("y")
    L = st["box"][0]
    dx = sc.min_image(x[:, None] - x[None, :], L)
    dy = sc.min_image(y[:, None] - y[None, :], L)
    r = np.hypot(dx, dy)
    np.fill_diagonal(r, np.inf)
    assert np.array_equal(s.get("coordination"), np.sum(r < 3.0, axis=1))
    # both branches occur in this configuration
    assert np.any(r < 2.5) and np.any((r > 2.5) & (r < 3.0))


def energy_and_forces(st, setup):
    s = sc.load(oracle.OracleSystem, st)
    setup(s)
    e = s.compute_forces()
    return e["pair"], np.stack([s.get("fx"), s.get("fy"), s.get("fz")], 1)


@pytest.mark.parametrize("kind", ["morse", "gaussian", "yukawa", "soft_attractive"])
def test_newton_third_law_and_gradient(kind):
    st = sc.plane_disks(300, phi=0.6, seed=32, polydisperse=(0.8, 1.2))

    def setup(s):
        s.nlist(3.2, 0.3)
        s.pair_extra(kind, PARAMS[kind], use_radii=kind != "yukawa", shifted=kind in ("morse", "gaussian"))

    E0, F = energy_and_forces(st, setup)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()
    rng = np.random.default_rng(1)
    h = 1e-6
    for i in rng.choice(st["n"], 6, replace=False):
        for ax, key in [(0, "x"), (1, "y")]:
            g = []
            for sgn in (+1, -1):
                st2 = {**st, "cols": {k: v.copy() for k, v in st["cols"].items()}}
                st2["cols"][key][i] += sgn * h
                g.append(energy_and_forces(st2, setup)[0])
            grad = (g[0] - g[1]) / (2 * h)
            assert -grad == pytest.approx(F[i, ax], abs=1e-5 * max(1.0, abs(F[i, ax])))


# ------------------------------------------------------------------ GPU ---
def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


def forces(s):
    return np.stack([s.get("fx"), s.get("fy"), s.get("fz")], 1)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["morse", "gaussian", "yukawa", "soft_attractive"])
@pytest.mark.parametrize("radii", [False, True])
def test_gpu_forces_energy_coordination(gpu, kind, radii):
    st = sc.plane_disks(3000, phi=0.75, seed=33, polydisperse=(0.8, 1.2))
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(3.2, 0.3)
        s.pair_extra(kind, PARAMS[kind], use_radii=radii and kind != "yukawa", shifted=True)
        s.constraint("plane")
    d.compute_forces()
    o.compute_forces()
    assert sc.rel_force_err(forces(d), forces(o)) <= 1e-12
    assert d.extra_energy() == pytest.approx(o.extra_energy(), rel=1e-12)
    if kind == "soft_attractive":
        assert np.array_equal(d.get("coordination"), o.get("coordination"))


@pytest.mark.gpu
def test_gpu_extras_with_soft_and_polar(gpu):
    # extras add to the soft / aligner pass; two extras at once
    st = sc.plane_disks(2500, phi=0.8, seed=34)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(3.2, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.3, 2.4)
        s.pair_extra("morse", PARAMS["morse"], shifted=True)
        s.pair_extra("yukawa", PARAMS["yukawa"])
        s.constraint("plane")
    d.compute_forces()
    o.compute_forces()
    assert sc.rel_force_err(forces(d), forces(o)) <= 1e-12
    assert d.extra_energy() == pytest.approx(o.extra_energy(), rel=1e-12)


@pytest.mark.gpu
def test_gpu_brownian_trajectory_with_extras(gpu):
    st = sc.plane_disks(3000, phi=0.7, seed=35)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(3.2, 0.3)
        s.pair_extra("soft_attractive", PARAMS["soft_attractive"])
        s.pair_extra("gaussian", PARAMS["gaussian"], shifted=True)
        s.align_nematic(0.2, 2.4)
        s.constraint("plane")
        s.set_timestep(0.005)
        s.integrator_brownian(v0=0.2, nu=0.0, mu=1.0, seed=3)
        s.run(299)
    dd = np.stack([d.get("x") - o.get("x"), d.get("y") - o.get("y")], 1)
    dd = sc.min_image(dd, st["box"][0])
    assert np.abs(dd).max() <= 1e-9
    assert np.array_equal(d.get("coordination"), o.get("coordination"))
