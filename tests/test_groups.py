"""Particle groups and per-group integrators (the `group` command,
src/samos.cpp:1108-1140, System::make_group src/system/system.cpp:462-570, and
integrators keyed type_group, src/samos.cpp:654-690).

Oracle pins (CPU): an integrator moves only its group's particles (the others
keep their state up to the loop-head enforce); two Brownian integrators on
disjoint groups drift each particle with its own group's v0 (closed form);
FIRE's power and norms run over its group only; integrator keys collide per
(type, group) as the reference's do.

Device against the oracle (GPU): the physics of the reference's
configurations/two_component_plane (passive type 1 at T = 0.3, active type 2
with v0 and rotational noise, polar alignment among active particles only)
through the fused multi-group Brownian step and the faithful step sequence;
Langevin, SPV, NVE and FIRE on one group.  Bars as in test_gpu_parity.py: noise
on with the same Philox stream, |dx|, |dn|, |dv| <= 1e-9; ages bit-exact.
"""
import math

import numpy as np
import pytest

import oracle
from tests import scenarios as sc


def two_types(st, frac=0.5, seed=3):
    rng = np.random.default_rng(seed)
    st["cols"]["type"] = np.where(rng.uniform(size=st["n"]) < frac, 2, 1).astype(np.int32)
    return st


# ------------------------------------------------------------ oracle pins ---
def test_integrator_moves_only_its_group():
    n = 6
    s = oracle.OracleSystem(n, (100.0, 100.0, 10.0), True)
    s.set("x", np.arange(n) * 5.0 - 12.0)
    s.set("nx", np.ones(n))
    member = np.array([1, 0, 1, 0, 0, 1], bool)
    s.group("active", member)
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_brownian(v0=0.4, nu=0.0, mu=1.0, group="active")
    x0 = s.get("x").copy()
    s.run(99)
    x = s.get("x")
    assert np.all(x[~member] == x0[~member])
    np.testing.assert_allclose(x[member] - x0[member], 0.4 * 100 * 0.01, rtol=1e-12)
    assert np.all(s.get("age")[~member] == 0.0)
    assert np.all(s.get("age")[member] > 0.0)


def test_disjoint_groups_drift_with_their_own_v0():
    n = 8
    s = oracle.OracleSystem(n, (100.0, 100.0, 10.0), True)
    s.set("y", np.arange(n) * 4.0 - 14.0)
    th = np.linspace(0.0, 2.0, n)
    s.set("nx", np.cos(th))
    s.set("ny", np.sin(th))
    a = np.arange(n) % 2 == 0
    s.group("a", a)
    s.group("b", ~a)
    s.constraint("plane")
    s.set_timestep(0.02)
    s.integrator_brownian(v0=0.3, nu=0.0, group="a")
    s.integrator_brownian(v0=0.7, nu=0.0, group="b")
    y0 = s.get("y").copy()
    s.run(49)
    v0 = np.where(a, 0.3, 0.7)
    np.testing.assert_allclose(s.get("x"), v0 * 50 * 0.02 * np.cos(th), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s.get("y") - y0, v0 * 50 * 0.02 * np.sin(th), rtol=1e-12, atol=1e-15)


def test_fire_norms_run_over_its_group():
    # an LJ pair inside the group and a third particle outside it
    s = oracle.OracleSystem(3, (40.0, 40.0, 10.0), True)
    s.set("x", [-0.65, 0.65, 1.6])
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(0.005)
    s.group("pair", [True, True, False])
    s.integrator_fire(f_tol=0.0, e_tol=0.0, group="pair")
    s.run(0)
    st = s.fire_state()
    fx = s.get("fx")
    assert st["Fnorm"] == pytest.approx(math.hypot(fx[0], fx[1]), rel=1e-14)
    assert s.get("x")[2] == 1.6 and s.get("vx")[2] == 0.0


def test_integrator_keys_per_group():
    s = oracle.OracleSystem(4, (50.0, 50.0, 10.0), True)
    s.set_timestep(0.01)
    s.constraint("plane")
    s.group("g", [True, False, True, False])
    s.integrator_brownian(v0=0.1, group="g")
    s.integrator_brownian(v0=0.2)  # type "brownian" for group all: a different key
    with pytest.raises(oracle.OracleError, match="already exists"):
        s.integrator_brownian(v0=0.3, group="g")
    s.disable("brownian", "g")
    s.integrator_brownian(v0=0.3, group="g")


# ------------------------------------------------------------ device ---------
def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


def stack(s, names):
    return np.stack([s.get(k) for k in names], 1)


def wrapped(a, b, box):
    d = a - b
    for k in range(d.shape[1]):
        d[:, k] = sc.min_image(d[:, k], box[k])
    return float(np.abs(d).max())


def two_component(s, st):
    """configurations/two_component_plane/two_component_plane.conf:55-120 physics."""
    types = st["cols"]["type"]
    s.group("passive", types == 1)
    s.group("active", types == 2)
    s.nlist(2.5, 0.5)
    s.pair_soft(1.0, 2.0)
    s.align_polar(1.0, 2.5)
    s.align_type_param("polar", 1, 1, 0.0, 2.5)
    s.align_type_param("polar", 1, 2, 0.0, 2.5)
    s.align_type_param("polar", 2, 2, 1.0, 2.5)
    s.constraint("plane")
    s.set_timestep(0.005)
    s.integrator_brownian(v0=0.5, nu=0.05, mu=1.0, seed=1, group="active")
    s.integrator_brownian(v0=0.0, nu=0.0, mu=1.0, seed=4, temp=(0, 0.3, 0.3, 1000), group="passive")


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_two_component_plane_against_oracle(gpu, flags):
    st = two_types(sc.plane_disks(3000, phi=0.7, seed=71))
    d, o = both(gpu, st)
    for s in (d, o):
        two_component(s, st)
    assert d.group_size("active") == int(np.sum(st["cols"]["type"] == 2))
    d.run_range(0, 40, 150, flags)
    d.run_range(40, 151, 150, flags)
    o.run(150)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["nx", "ny"]) - stack(o, ["nx", "ny"])).max() <= 1e-9
    assert np.abs(stack(d, ["vx", "vy"]) - stack(o, ["vx", "vy"])).max() <= 1e-9
    assert np.array_equal(d.get("age"), o.get("age"))
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["baoab", "spv"])
@pytest.mark.parametrize("flags", [0, 1])
def test_langevin_on_one_group(gpu, method, flags):
    st = two_types(sc.lj_lattice(36, rho=0.7, seed=72), frac=0.6)
    d, o = both(gpu, st)
    for s in (d, o):
        s.group("hot", st["cols"]["type"] == 2)
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=6, method=method, temp=(0, 1.2, 1.0, 1), group="hot")
    d.run(80, flags)
    o.run(80)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["vx", "vy"]) - stack(o, ["vx", "vy"])).max() <= 1e-9
    assert np.array_equal(d.get("age"), o.get("age"))
    d.close()


@pytest.mark.gpu
def test_nve_then_groups_after_disable(gpu):
    # the example conf's schedule: NVE with a limit on everyone, disable, then
    # the two Brownian groups
    st = two_types(sc.plane_disks(1500, phi=0.85, seed=73))
    d, o = both(gpu, st)
    for s in (d, o):
        s.group("passive", st["cols"]["type"] == 1)
        s.group("active", st["cols"]["type"] == 2)
        s.nlist(2.5, 0.5)
        s.pair_soft(1.0, 2.0)
        s.align_polar(1.0, 2.5)
        s.constraint("plane")
        s.set_timestep(0.001)
        s.integrator_nve(limit=0.0001)
        s.run(30)
        s.disable("nve", "all")
        s.set_timestep(0.005)
        s.integrator_brownian(v0=0.5, nu=0.01, seed=1, group="active")
        s.integrator_brownian(v0=0.0, nu=0.0, seed=4, temp=(0, 0.3, 0.3, 1000), group="passive")
        s.run(60)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["nx", "ny"]) - stack(o, ["nx", "ny"])).max() <= 1e-9
    d.close()


@pytest.mark.gpu
def test_fire_on_one_group(gpu):
    st = sc.plane_disks(800, phi=0.9, seed=74, polydisperse=(0.8, 1.2))
    rng = np.random.default_rng(2)
    member = rng.uniform(size=st["n"]) < 0.7
    d, o = both(gpu, st)
    for s in (d, o):
        s.group("relax", member)
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0, use_radii=True)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_fire(dt_max=0.1, f_tol=0.0, e_tol=0.0, group="relax")
        s.run(199)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    fd, fo = d.fire_state(), o.fire_state()
    assert fd["dt"] == pytest.approx(fo["dt"], rel=1e-12)
    assert fd["Fnorm"] == pytest.approx(fo["Fnorm"], rel=1e-9)
    # particles outside the group never move
    assert np.array_equal(d.get("x")[~member], o.get("x")[~member])
    assert np.array_equal(o.get("x")[~member], st["cols"]["x"][~member])
    d.close()
