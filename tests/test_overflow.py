"""Neighbour-list overflow inside a run.  The reference's lists are vectors
and cannot overflow (src/system/neighbour_list.cpp:216-273); the device list
has a fixed capacity per particle.  A device rebuild that finds more
neighbours than the capacity must not let any step read the truncated list:
captured chunks stop at it, the capacity grows, the list is rebuilt from the
same positions, and the run continues.

Bar: a run started with no spare capacity (so that rebuilds overflow) is bit
for bit the run with room to spare: positions, velocities, directors, pair
sets and build counts; at least one overflow was recovered.
"""
import numpy as np
import pytest

from tests import scenarios as sc

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "vx", "vy", "nx", "ny", "fx", "fy", "ix", "iy", "age")


def run_pair(gpu, st, setup, ranges, flags=0):
    out = []
    for tight in (False, True):
        d = sc.load(gpu.System, st)
        setup(d)
        if tight:
            d.set_option("neighbour_capacity", 0)
        b0 = d.nlist_info()["builds"]
        steps = ranges[-1][1] - 1
        for lo, hi in ranges:
            d.run_range(lo, hi, steps, flags)
        out.append(({k: d.get(k) for k in KEYS}, d.pairs(), d.nlist_info()["builds"] - b0, d.run_stats()))
        d.close()
    (a, pa, ba, sa), (c, pc, bc, scc) = out
    for k in KEYS:
        assert np.array_equal(a[k], c[k]), k
    assert np.array_equal(pa, pc)
    assert ba == bc and ba > 3
    return sa, scc


def lj(d):
    d.nlist(2.5, 0.3)
    d.pair_lj(1.0, 1.0, 2.5, shifted=True)
    d.constraint("plane")
    d.set_timestep(0.004)
    d.integrator_langevin(gamma=1.0, seed=8, temp=(0, 1.5, 1.5, 1))


def active(d):
    d.nlist(2.4, 0.3)
    d.pair_soft(2.0, 2.0)
    d.align_polar(1.0, 2.4)
    d.constraint("plane")
    d.set_timestep(0.01)
    d.integrator_brownian(v0=1.0, nu=0.02, seed=5)


@pytest.mark.parametrize("flags", [0, 2])
def test_fused_langevin_recovers(gpu, flags):
    st = sc.lj_lattice(48, rho=0.75, seed=61)
    normal, tight = run_pair(gpu, st, lj, [(0, 1), (1, 150), (150, 401)], flags)
    assert tight["overflow_recoveries"] > 0 and normal["overflow_recoveries"] == 0
    assert tight["capacity"] > 0


@pytest.mark.parametrize("flags", [0, 2])
def test_fused_brownian_recovers(gpu, flags):
    # a lattice start: every particle has the same count, which motion raises
    st = sc.lj_lattice(48, rho=0.6, seed=62)
    th = np.random.default_rng(62).uniform(0.0, 2.0 * np.pi, st["n"])
    st["cols"]["nx"], st["cols"]["ny"] = np.cos(th), np.sin(th)
    normal, tight = run_pair(gpu, st, active, [(0, 100), (100, 301)], flags)
    assert tight["overflow_recoveries"] > 0 and normal["overflow_recoveries"] == 0


def test_step_by_step_generic_sequence_recovers(gpu):
    def spv(d):
        d.nlist(2.4, 0.3)
        d.pair_soft(1.0, 2.0)
        d.constraint("plane")
        d.set_timestep(0.01)
        d.integrator_langevin(gamma=1.0, seed=4, method="spv", temp=(0, 1.0, 1.0, 1))

    st = sc.plane_disks(2000, phi=0.8, seed=63)
    normal, tight = run_pair(gpu, st, spv, [(0, 201)], gpu.RUN_NO_GRAPH)
    assert tight["overflow_recoveries"] > 0


def test_captured_generic_sequence_keeps_headroom(gpu):
    def spv(d):
        d.nlist(2.4, 0.3)
        d.pair_soft(1.0, 2.0)
        d.constraint("plane")
        d.set_timestep(0.01)
        d.integrator_langevin(gamma=1.0, seed=4, method="spv", temp=(0, 1.0, 1.0, 1))

    st = sc.plane_disks(2000, phi=0.8, seed=64)
    # no spare capacity at the start: it grows before the captured iterations,
    # which then never overflow
    normal, tight = run_pair(gpu, st, spv, [(0, 201)])
    assert tight["overflow_recoveries"] == 0


def test_graphs_are_reused_across_ranges(gpu):
    st = sc.lj_lattice(32, rho=0.7, seed=65)
    d = sc.load(gpu.System, st)
    lj(d)
    d.run_range(0, 20, 400)
    first = d.run_stats()["graphs_instantiated"]
    for lo in range(20, 400, 20):
        d.run_range(lo, lo + 20, 400)
    st2 = d.run_stats()
    assert st2["graphs_instantiated"] == first and st2["graphs_reused"] >= 19
    d.close()
