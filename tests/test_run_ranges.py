"""The drop-in run loop cut into ranges (a conf with frequent log/dump stops):
short ranges replay cached step-chunk graphs, ages stay on the device while
they are all equal, and none of it may change a bit of the trajectory.

Bars: bit equality against the same run launched step by step
(SAMOS_RUN_NO_GRAPH, one range); ages bit-exact against the oracle.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "vx", "vy", "nx", "ny", "fx", "fy", "age", "ix", "iy")


def setup(s, kind):
    s.constraint("plane")
    if kind == "langevin":
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.set_timestep(0.003)
        s.integrator_langevin(gamma=1.0, seed=9, temp=(0, 1.0, 1.0, 1))
    elif kind == "brownian":
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.1, 2.4)
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.3, nu=0.1, seed=9, temp=(0, 0.02, 0.02, 1))
    elif kind == "fire":  # captured FIRE iterations with the scalar state on the device
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0, use_radii=True)
        s.set_timestep(0.01)
        s.integrator_fire(dt_max=0.1, f_tol=0.0, e_tol=0.0)
    else:  # generic captured loop (Langevin SPV)
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.set_timestep(0.01)
        s.integrator_langevin(gamma=1.0, seed=4, method="spv", temp=(0, 0.3, 0.3, 1))


def state(kind):
    if kind == "fire":
        return sc.plane_disks(1500, phi=0.88, seed=41, polydisperse=(0.8, 1.2))
    return sc.lj_lattice(36, rho=0.7, seed=41) if kind == "langevin" else sc.plane_disks(1800, phi=0.7, seed=41)


@pytest.mark.parametrize("kind", ["langevin", "brownian", "spv", "fire"])
def test_short_ranges_equal_step_by_step(gpu, kind):
    st = state(kind)
    steps = 160
    # ragged ranges: single steps, two-step ranges, odd lengths, one longer
    # than a chunk, and repeats that reuse the instantiated graphs
    cuts = [0, 1, 3, 4, 9, 11, 14, 31, 33, 36, 52, 53, 55, 90, 92, 95, 130, 131, 161]
    ref = sc.load(gpu.System, st)
    setup(ref, kind)
    ref.run(steps, gpu.RUN_NO_GRAPH)
    got = sc.load(gpu.System, st)
    setup(got, kind)
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        got.run_range(lo, hi, steps)
    # a second run on the same context continues from the first
    ref.run(40, gpu.RUN_NO_GRAPH)
    for lo, hi in [(0, 2), (2, 5), (5, 41)]:
        got.run_range(lo, hi, 40)
    for k in KEYS:
        assert np.array_equal(ref.get(k), got.get(k)), k
    assert np.array_equal(ref.pairs(), got.pairs())
    if kind == "fire":
        assert ref.fire_state() == got.fire_state()
    ref.close()
    got.close()


@pytest.mark.parametrize("kind", ["langevin", "brownian", "spv"])
def test_unequal_ages_advance_per_particle(gpu, kind):
    st = state(kind)
    rng = np.random.default_rng(5)
    st["cols"]["age"] = rng.uniform(0.0, 50.0, st["n"])
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        setup(s, kind)
    d.run_range(0, 7, 60)
    d.run_range(7, 61, 60)
    o.run(60)
    assert np.array_equal(d.get("age"), o.get("age"))
    # equal ages again: the common value path takes over, still exact
    same = np.full(st["n"], 2.5)
    d.set("age", same)
    o.set("age", same)
    d.run(33)
    o.run(33)
    assert np.array_equal(d.get("age"), o.get("age"))
    d.close()
