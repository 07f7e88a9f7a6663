"""Pins for the CPU oracle (oracle/samos_oracle.cpp).

The reference ships no tests or golden vectors for this path, so each oracle
function is pinned by checks chosen so that a plausible mistake fails one:

  cell list / nlist build   brute-force O(N^2) pair set on random, boundary and
                            example inputs; N^2 fallback; duplicate-cell quirk
  soft, lj                  two-particle closed forms, energy shift at rc,
                            Newton's third law, finite-difference gradient
  polar, nematic            2-D closed forms J sin(dtheta), J' sin(2 dtheta),
                            pairwise torque cancellation, energies
  constraints               |x| = R, n.N = 0, |n| = 1, v.N = 0, plane z/vz/nz,
                            rotate by phi then -phi
  integrators               free ABP drift, BAOAB O-step decay at T = 0,
                            SPV/BBK free flight, FIRE two-particle relaxation,
                            Langevin equipartition with noise on, NVE free
                            flight / limits / LJ-pair energy conservation /
                            angle limit, the linear temperature ramp
  (further pair potentials: tests/test_pair_extra.py; cylinder, torus and
   ellipsoid constraints: tests/test_surfaces.py)
  RNG                       Random123 philox4x32-10 known-answer vectors,
                            Box-Muller moments
  golden files              regression pins of the example-input outputs
"""
import math
import os

import numpy as np
import pytest

import oracle
from tests import scenarios as sc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def two(dx, dy=0.0, box=(50.0, 50.0, 10.0), periodic=True):
    s = oracle.OracleSystem(2, box, periodic)
    s.set("x", [0.0, dx])
    s.set("y", [0.0, dy])
    s.set("z", [0.0, 0.0])
    return s


# ---------------------------------------------------------------- RNG ----
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_gaussian_moments():
    g = np.array([oracle.gauss3(7, 2, pid, call) for pid in range(3000) for call in range(4)]).ravel()
    assert abs(g.mean()) < 4.0 / math.sqrt(g.size)
    assert abs(g.var() - 1.0) < 6.0 * math.sqrt(2.0 / g.size)
    assert abs(np.mean(g ** 4) - 3.0) < 0.15


def test_box_muller_transcendentals_pinned_to_libm():
    """The noise transform's own log and sin/cos of 2 pi u (the device
    evaluates the same expression) against the x87 extended-precision libm:
    log within 2 ulp, sin and cos within 2 ulp where |value| > 1/64 and within
    2^-53 absolute everywhere (the reduction is exact, so values near the
    zeros keep their relative accuracy too)."""
    rng = np.random.default_rng(11)
    words = np.concatenate([rng.integers(0, 2 ** 32, 400000), [0, 1, 2 ** 31, 2 ** 32 - 1],
                            np.arange(2 ** 29 - 40, 2 ** 29 + 40), np.arange(3 * 2 ** 29 - 40, 3 * 2 ** 29 + 40),
                            np.arange(2 ** 31 - 40, 2 ** 31 + 40), np.arange(int(0.70710678 * 2 ** 32) - 400,
                                                                             int(0.70710678 * 2 ** 32) + 400)])
    u = (words.astype(np.float64) + 0.5) * 2.0 ** -32
    lg, sn, cs = oracle.bm_eval(u)
    ul = u.astype(np.longdouble)
    two_pi = np.longdouble("6.283185307179586476925286766559005768394")
    ref_log = np.log(ul)
    ref_sin, ref_cos = np.sin(two_pi * ul), np.cos(two_pi * ul)

    def ulps(got, ref):
        return np.abs(got.astype(np.longdouble) - ref) / np.spacing(np.abs(ref.astype(np.float64)))

    assert ulps(lg, ref_log).max() <= 2.0
    for got, ref in ((sn, ref_sin), (cs, ref_cos)):
        big = np.abs(ref) > 1.0 / 64
        assert ulps(got[big], ref[big]).max() <= 2.0
        assert float(np.abs(got.astype(np.longdouble) - ref).max()) <= 2.0 ** -52
    # quadrant signs and the exact reduction: u and 1/2 + u, 1/4 + u
    assert np.array_equal(oracle.bm_eval([0.5 + 2.0 ** -33])[1], -oracle.bm_eval([2.0 ** -33])[1])


# ---------------------------------------------------------- pair sets ----
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("periodic", [True, False])
def test_cell_pairs_equal_brute_force(seed, periodic):
    rng = np.random.default_rng(seed)
    n = 700
    box = (31.0 + seed, 27.5, 12.0)
    s = oracle.OracleSystem(n, box, periodic)
    s.set("x", rng.uniform(-box[0] / 2, box[0] / 2, n))
    s.set("y", rng.uniform(-box[1] / 2, box[1] / 2, n))
    s.set("z", rng.uniform(-box[2] / 2, box[2] / 2, n))
    s.nlist(2.4, 0.3)
    assert s.uses_cells()
    assert np.array_equal(s.pairs(), s.brute_pairs(2.7))


def test_points_on_cell_and_box_boundaries():
    # lattice exactly on cell faces and on the low box face
    L = 27.0
    g = np.arange(-L / 2, L / 2, 1.35)
    x, y = np.meshgrid(g, g)
    n = x.size
    s = oracle.OracleSystem(n, (L, L, 10.0), True)
    s.set("x", x.ravel())
    s.set("y", y.ravel())
    s.nlist(2.4, 0.3)
    assert s.cell_dims() == (10, 10, 3)
    assert np.array_equal(s.pairs(), s.brute_pairs(2.7))


def test_thin_box_uses_n_squared():
    rng = np.random.default_rng(5)
    n = 300
    s = oracle.OracleSystem(n, (30.0, 30.0, 4.0), True)  # lz <= 2 (rcut + pad)
    s.set("x", rng.uniform(-15, 15, n))
    s.set("y", rng.uniform(-15, 15, n))
    s.nlist(2.4, 0.3)
    assert not s.uses_cells()
    assert np.array_equal(s.pairs(), s.brute_pairs(2.7))


def test_two_cell_side_duplicates_pairs():
    # 2 r_l < L < 3 r_l gives two cells per side: every stencil holds the
    # neighbour cell twice, so the reference lists cross-cell pairs twice
    # (src/system/cell_list.cpp:52-63).
    rng = np.random.default_rng(9)
    n = 60
    L = 7.0
    s = oracle.OracleSystem(n, (L, L, L), True)
    for k in "xyz":
        s.set(k, rng.uniform(-L / 2, L / 2, n))
    s.nlist(2.4, 0.6)
    assert s.cell_dims() == (2, 2, 2)
    p = s.pairs()
    b = s.brute_pairs(3.0)
    uniq = np.unique(p, axis=0)
    assert np.array_equal(uniq, b)
    assert len(p) > len(b)


def test_need_update_threshold():
    s = two(1.0)
    s.nlist(2.4, 0.5)
    assert not s.need_update()
    s.set("x", [0.0, 1.0 + 0.2499999])
    assert not s.need_update()
    s.set("x", [0.0, 1.25])  # |dr|^2 = 0.0625 = pad^2 / 4 exactly -> rebuild (strict <)
    assert s.need_update()


# ---------------------------------------------------------- potentials ----
@pytest.mark.parametrize("r", [0.5, 1.3, 1.999])
def test_soft_two_particle_closed_form(r):
    k, a = 3.0, 2.0
    s = two(r)
    s.nlist(2.4, 0.3)
    s.pair_soft(k, a)
    e = s.compute_forces()
    fx = s.get("fx")
    assert fx[1] == pytest.approx(k * (a - r), rel=1e-14)
    assert fx[0] == pytest.approx(-k * (a - r), rel=1e-14)
    assert e["pair"] == pytest.approx(0.5 * k * (a - r) ** 2, rel=1e-14)


def test_soft_outside_range_is_zero():
    s = two(2.0)
    s.nlist(2.4, 0.3)
    s.pair_soft(1.0, 2.0)
    s.compute_forces()
    assert np.all(s.get("fx") == 0.0)


@pytest.mark.parametrize("r", [0.95, 1.12, 1.7, 2.5])
def test_lj_two_particle_closed_form(r):
    s = two(r)
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    e = s.compute_forces()
    f = 48.0 * (r ** -13 - 0.5 * r ** -7)  # sigma = eps = 1
    shift = 4.0 * (2.5 ** -12 - 2.5 ** -6)
    assert s.get("fx")[1] == pytest.approx(f, rel=1e-13)
    assert e["pair"] == pytest.approx(4.0 * (r ** -12 - r ** -6) - shift, rel=1e-12, abs=1e-15)


def test_lj_sigma_quirk():
    # force factor uses sigma^2 / r^2 in place of 1 / r^2 (pair_lj_potential.cpp:96,108)
    r, sig = 2.0, 1.5
    s = two(r)
    s.nlist(4.0, 0.3)
    s.pair_lj(1.0, sig, 4.0)
    s.compute_forces()
    s2 = sig * sig / (r * r)
    s6 = s2 ** 3
    assert s.get("fx")[1] == pytest.approx(48.0 * s6 * (s6 - 0.5) * s2 * r, rel=1e-13)


def energy_and_forces(st, setup):
    s = sc.load(oracle.OracleSystem, st)
    setup(s)
    e = s.compute_forces()
    return e["pair"], np.stack([s.get("fx"), s.get("fy"), s.get("fz")], 1)


@pytest.mark.parametrize("pot", ["soft", "lj"])
def test_newton_third_law_and_gradient(pot):
    st = sc.plane_disks(300, phi=0.6, seed=4) if pot == "soft" else sc.lj_lattice(16, seed=4)

    def setup(s):
        s.nlist(2.5 if pot == "lj" else 2.4, 0.3)
        if pot == "soft":
            s.pair_soft(1.0, 2.0)
        else:
            s.pair_lj(1.0, 1.0, 2.5, shifted=True)

    E0, F = energy_and_forces(st, setup)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).sum()
    rng = np.random.default_rng(0)
    h = 1e-6
    for i in rng.choice(st["n"], 5, replace=False):
        for ax, key in [(0, "x"), (1, "y")]:
            g = []
            for sgn in (+1, -1):
                st2 = {**st, "cols": {k: v.copy() for k, v in st["cols"].items()}}
                st2["cols"][key][i] += sgn * h
                g.append(energy_and_forces(st2, setup)[0])
            grad = (g[0] - g[1]) / (2 * h)
            assert -grad == pytest.approx(F[i, ax], abs=1e-5 * max(1.0, abs(F[i, ax])))


@pytest.mark.parametrize("kind", ["polar", "nematic"])
def test_aligner_two_particle_closed_form(kind):
    J = 0.7
    for dth in [0.3, 1.2, 2.5, -0.8]:
        s = two(1.5)
        s.set("nx", [1.0, math.cos(dth)])
        s.set("ny", [0.0, math.sin(dth)])
        s.nlist(2.4, 0.3)
        getattr(s, "align_" + kind)(J, 2.4)
        e = s.compute_forces()
        tz = s.get("tau_z")
        if kind == "polar":
            assert tz[0] == pytest.approx(J * math.sin(dth), rel=1e-13)
            assert e["align"] == pytest.approx(-2 * J * math.cos(dth), rel=1e-13)
        else:  # J' = 2J without per-type parameters
            assert tz[0] == pytest.approx(2 * J * math.cos(dth) * math.sin(dth), rel=1e-13)
            assert e["align"] == pytest.approx(-4 * J * math.cos(2 * dth), rel=1e-12, abs=1e-14)
        assert tz[0] + tz[1] == pytest.approx(0.0, abs=1e-15)


def test_nematic_per_type_coupling_not_doubled():
    s = two(1.5)
    dth = 0.4
    s.set("nx", [1.0, math.cos(dth)])
    s.set("ny", [0.0, math.sin(dth)])
    s.nlist(2.4, 0.3)
    s.align_nematic(0.7, 2.4)
    s.align_type_param("nematic", 1, 1, 0.7, 2.4)
    s.compute_forces()
    assert s.get("tau_z")[0] == pytest.approx(0.7 * math.cos(dth) * math.sin(dth), rel=1e-13)


# ---------------------------------------------------------- constraints ---
def test_sphere_invariants_and_rotation_inverse():
    st = sc.sphere_disks(500, R=9.0, seed=3)
    rng = np.random.default_rng(1)
    for k in ["x", "y", "z"]:
        st["cols"][k] = st["cols"][k] * rng.uniform(0.9, 1.1, 500)
    s = sc.load(oracle.OracleSystem, st)
    s.set("vx", rng.normal(size=500))
    s.set("vy", rng.normal(size=500))
    s.set("vz", rng.normal(size=500))
    s.constraint("sphere", 9.0)
    x = np.stack([s.get("x"), s.get("y"), s.get("z")], 1)
    n = np.stack([s.get("nx"), s.get("ny"), s.get("nz")], 1)
    v = np.stack([s.get("vx"), s.get("vy"), s.get("vz")], 1)
    N = x / 9.0
    assert np.abs(np.linalg.norm(x, axis=1) - 9.0).max() < 1e-13
    assert np.abs(np.sum(n * N, 1)).max() < 1e-14
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-15
    assert np.abs(np.sum(v * N, 1)).max() < 1e-13
    phi = rng.uniform(-3, 3, 500)
    s.rotate_directors(phi)
    s.rotate_directors(-phi)
    n2 = np.stack([s.get("nx"), s.get("ny"), s.get("nz")], 1)
    assert np.abs(n2 - n).max() < 1e-14


def test_plane_invariants():
    st = sc.plane_disks(200, seed=8)
    st["cols"]["z"] = np.linspace(-1, 1, 200)
    st["cols"]["nz"] = np.full(200, 0.5)
    s = sc.load(oracle.OracleSystem, st)
    s.set("vz", np.ones(200))
    s.constraint("plane", 0.0)
    assert np.all(s.get("z") == 0.0) and np.all(s.get("vz") == 0.0) and np.all(s.get("nz") == 0.0)
    nrm = np.hypot(s.get("nx"), s.get("ny"))
    assert np.abs(nrm - 1).max() < 1e-15


# ---------------------------------------------------------- integrators ---
def test_free_abp_drift():
    s = oracle.OracleSystem(1, (100.0, 100.0, 10.0), True)
    th = 0.7
    s.set("nx", [math.cos(th)])
    s.set("ny", [math.sin(th)])
    s.set_timestep(0.01)
    s.constraint("plane")
    s.integrator_brownian(v0=0.3, nu=0.0, mu=1.0)
    s.run(99)  # 100 integrate calls
    assert s.get("x")[0] == pytest.approx(0.3 * 1.0 * math.cos(th), rel=1e-12)
    assert s.get("y")[0] == pytest.approx(0.3 * 1.0 * math.sin(th), rel=1e-12)


def test_baoab_free_decay_at_zero_temperature():
    s = oracle.OracleSystem(1, (1000.0, 1000.0, 10.0), True)
    s.set("vx", [1.0])
    s.set_timestep(0.002)
    s.constraint("plane")
    s.integrator_langevin(gamma=1.5, temp=(0, 0.0, 1.0, 1))
    s.run(499)
    assert s.get("vx")[0] == pytest.approx(math.exp(-1.5 * 0.002 * 500), rel=1e-12)
    # positions: sum of half-step drifts
    v = np.exp(-1.5 * 0.002 * np.arange(501))
    x = 0.001 * (v[:-1] * math.exp(0) + v[1:]).sum()
    assert s.get("x")[0] == pytest.approx(x, rel=1e-12)


@pytest.mark.parametrize("method", ["spv", "bbk"])
def test_langevin_methods_frictionless_free_flight(method):
    s = oracle.OracleSystem(1, (1000.0, 1000.0, 10.0), True)
    s.set("vx", [0.5])
    s.set_timestep(0.01)
    s.constraint("plane")
    s.integrator_langevin(gamma=0.0, method=method, temp=(0, 0.0, 1.0, 1))
    s.run(99)
    assert s.get("vx")[0] == pytest.approx(0.5, rel=1e-14)
    assert s.get("x")[0] == pytest.approx(0.5 * 1.0, rel=1e-12)


def test_fire_finds_lj_pair_minimum():
    # unique minimum of the LJ pair at r = 2^(1/6) sigma
    s = two(1.3, box=(40.0, 40.0, 10.0))
    s.set_timestep(0.01)
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.integrator_fire(dt_max=0.1, f_tol=1e-13, e_tol=1e-20)
    s.run(2000)
    r = s.get("x")[1] - s.get("x")[0]
    assert r == pytest.approx(2.0 ** (1.0 / 6.0), abs=1e-9)
    assert s.fire_state()["Fnorm"] < 1e-9


def test_langevin_equipartition():
    n = 4000
    s = oracle.OracleSystem(n, (400.0, 400.0, 10.0), True)
    s.set_timestep(0.01)
    s.constraint("plane")
    s.integrator_langevin(gamma=2.0, seed=3, temp=(0, 1.7, 1.0, 1))
    s.run(400)
    v2 = np.concatenate([s.get("vx"), s.get("vy")]) ** 2
    assert abs(v2.mean() - 1.7) < 4 * 1.7 * math.sqrt(2.0 / v2.size) + 0.02


def test_brownian_rotational_variance():
    # dtheta per step has variance nu * dt (integrator_brownian.hpp:159)
    n, nu, dt, steps = 3000, 0.2, 0.01, 49
    s = oracle.OracleSystem(n, (1e4, 1e4, 10.0), True)
    s.set("nx", np.ones(n))
    s.set_timestep(dt)
    s.constraint("plane")
    s.integrator_brownian(v0=0.0, nu=nu, seed=11)
    s.run(steps)
    th = np.arctan2(s.get("ny"), s.get("nx"))
    var = th.var()
    expect = nu * dt * (steps + 1)
    assert abs(var - expect) < 5 * expect * math.sqrt(2.0 / n)


def test_linear_temperature_ramp_variance():
    # ValueLinear (src/system/value.hpp:140-172): T(step) = min + step (max - min) / steps,
    # max beyond; Brownian noise sqrt(2 mu T dt) per step, so after run steps
    # 0..99 of a 0 -> 1 ramp over 100 steps the variance per axis is
    # 2 mu dt sum_s s/100 = 0.99 (an off-by-one step gives 1.01, 6 sigma away)
    n = 100000
    s = oracle.OracleSystem(n, (1e6, 1e6, 10.0), True)
    s.set("nx", np.ones(n))
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_brownian(v0=0.0, nu=0.0, mu=1.0, seed=9, temp=(1, 0.0, 1.0, 100))
    s.run(99)
    var = 0.5 * float(np.mean(s.get("x") ** 2 + s.get("y") ** 2))
    assert var == pytest.approx(0.99, abs=3.5 * 0.99 * math.sqrt(1.0 / n))


def test_nve_free_flight():
    # no forces: x(t) = x0 + t v exactly step by step (integrator_nve.cpp:47-83)
    n = 5
    s = oracle.OracleSystem(n, (1000.0, 1000.0, 10.0), True)
    v = np.linspace(-1.0, 1.0, n)
    s.set("vx", v)
    s.set("vy", 0.5 * v)
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_nve()
    x = np.zeros(n)
    for _ in range(20):
        x = x + 0.01 * v
    s.run(19)
    assert np.array_equal(s.get("x"), x)
    assert np.array_equal(s.get("vx"), v)


def test_nve_limit_caps_displacement_and_speed():
    # |dt v| > limit: each step moves exactly `limit` along v, and the speed is
    # then cut to limit / dt (integrator_nve.cpp:66-76, 115-124)
    s = oracle.OracleSystem(1, (1000.0, 1000.0, 10.0), True)
    s.set("vx", [3.0])
    s.set("vy", [4.0])
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_nve(limit=0.002)
    s.integrate()
    assert s.get("x")[0] == pytest.approx(0.002 * 0.6, rel=1e-14)
    assert s.get("y")[0] == pytest.approx(0.002 * 0.8, rel=1e-14)
    assert math.hypot(s.get("vx")[0], s.get("vy")[0]) == pytest.approx(0.2, rel=1e-14)
    s.integrate()
    assert s.get("x")[0] == pytest.approx(2 * 0.002 * 0.6, rel=1e-13)


def test_nve_lj_pair_conserves_energy():
    # velocity Verlet: total energy of an LJ pair oscillating about r = 2^(1/6)
    # drifts only by O(dt^2); a dropped or doubled half kick breaks this at O(dt)
    s = two(1.2)
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=False)
    s.constraint("plane")
    s.set_timestep(0.001)
    s.integrator_nve()
    e0 = s.compute_forces()["pair"]
    s.run(1999)
    e = s.compute_forces()["pair"]
    ke = 0.5 * float(np.sum(s.get("vx") ** 2 + s.get("vy") ** 2))
    assert abs(e + ke - e0) < 1e-5 * abs(e0)
    assert abs(s.get("x")[1] - s.get("x")[0] - 1.2) > 1e-3  # it moved


def test_nve_angle_limit_caps_rotation():
    # two aligned-against-each-other polar disks: J n_i x n_j torques; with
    # angle_limit each step's director rotation is at most the limit
    s = two(1.5)
    s.set("nx", [1.0, 0.0])
    s.set("ny", [0.0, 1.0])
    s.nlist(2.4, 0.3)
    s.align_polar(5.0, 2.4)
    s.constraint("plane")
    s.set_timestep(0.1)
    s.integrator_nve(angle_limit=0.01)
    th0 = np.arctan2(s.get("ny"), s.get("nx"))
    s.compute_forces()
    s.integrate()
    th1 = np.arctan2(s.get("ny"), s.get("nx"))
    assert np.all(np.abs(th1 - th0) <= 0.01 + 1e-15)
    assert np.all(np.abs(th1 - th0) > 0.0)


# ---------------------------------------------------------- golden pins ---
@pytest.mark.parametrize("name", ["plane", "sphere", "lj", "relax"])
def test_golden_example_inputs(name):
    import json
    import tests.golden.make_golden as mg
    man = json.load(open(os.path.join(GOLD, "MANIFEST.json")))[name]
    inp = dict(np.load(os.path.join(GOLD, f"inputs_{name}.npz")))
    gold = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    s = mg.build_oracle(inp, mg.CASES[name])
    # pair set equals the brute-force set on the example input
    cut = mg.CASES[name]["nlist"][0] + mg.CASES[name]["nlist"][1]
    assert np.array_equal(s.pairs(), s.brute_pairs(cut))
    e = s.compute_forces()
    assert np.array_equal(s.pairs(), gold["pairs"])
    assert len(gold["pairs"]) == man["npairs"]
    for k in ["fx", "fy", "fz", "tau_x", "tau_y", "tau_z"]:
        assert np.array_equal(s.get(k), gold[k]), k
    assert e["pair"] == gold["energy_pair"] and e["align"] == gold["energy_align"]
