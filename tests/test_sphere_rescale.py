"""Sphere radius schedule (ConstraintSphere::rescale,
src/constraints/constraint_sphere.cpp:87-104, parameters
src/constraints/constraint.hpp:85-108, called at every loop head before the
enforce, src/samos.cpp:882-886, followed by a neighbour-list rebuild).

Oracle pins (CPU): the radius after a run is R0 * rescale^(k freq / steps)
with k the loop heads whose time step is a multiple of freq below steps
(counted over runs, since the time step carries on), and every particle sits
on the current sphere.  Device against the oracle (GPU): active Brownian and
Langevin runs across the rescale window, trajectories to 1e-9, pair sets
after the run exact, and the radius equal.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc


def expected_radius(R0, rescale, steps, freq, time_steps):
    k = sum(1 for t in range(time_steps) if t % freq == 0 and t < steps)
    return R0 * (rescale ** (freq / steps)) ** k


def test_radius_schedule_and_projection():
    st = sc.sphere_disks(400, R=10.0, seed=91)
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.4, 0.3)
    o.pair_soft(1.0, 2.0)
    o.constraint("sphere", 10.0)
    o.constraint_rescale(1.5, steps=100, freq=10)
    o.set_timestep(0.01)
    o.integrator_brownian(v0=0.2, nu=0.1, seed=3)
    o.run(44)  # loop heads at time steps 0..44: rescales at 0, 10, 20, 30, 40
    R = o.constraint_radius()
    assert R == pytest.approx(expected_radius(10.0, 1.5, 100, 10, 45), rel=1e-14)
    x = np.stack([o.get("x"), o.get("y"), o.get("z")], 1)
    assert np.abs(np.linalg.norm(x, axis=1) - R).max() < 1e-12
    o.run(100)  # time steps 45..145: rescales at 50..90, none from 100 on
    assert o.constraint_radius() == pytest.approx(10.0 * 1.5, rel=1e-13)


def test_no_schedule_keeps_the_radius():
    st = sc.sphere_disks(200, R=8.0, seed=92)
    o = sc.load(oracle.OracleSystem, st)
    o.constraint("sphere", 8.0)
    o.constraint_rescale(1.0, steps=100, freq=10)
    o.set_timestep(0.01)
    o.integrator_brownian(v0=0.2, nu=0.0)
    o.run(30)
    assert o.constraint_radius() == 8.0


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 1, 2])
def test_rescaling_brownian_sphere_against_oracle(gpu, flags):
    st = sc.sphere_disks(1500, phi=0.9, seed=93)
    R0 = st["R"]
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(10.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.constraint("sphere", R0)
        s.constraint_rescale(1.3, steps=90, freq=10)
        s.set_timestep(0.005)
        s.integrator_brownian(v0=0.5, nu=0.05, seed=5, temp=(0, 0.01, 0.01, 1))
    d.run_range(0, 25, 120, flags)
    d.run_range(25, 121, 120, flags)
    o.run(120)
    assert d.constraint_radius() == pytest.approx(o.constraint_radius(), rel=1e-15)
    X = np.stack([d.get(k) for k in "xyz"], 1)
    Y = np.stack([o.get(k) for k in "xyz"], 1)
    assert np.abs(X - Y).max() <= 1e-9
    N = np.stack([d.get(k) for k in ("nx", "ny", "nz")], 1)
    M = np.stack([o.get(k) for k in ("nx", "ny", "nz")], 1)
    assert np.abs(N - M).max() <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())
    d.close()


@pytest.mark.gpu
def test_rescaling_langevin_sphere_against_oracle(gpu):
    st = sc.sphere_disks(1200, phi=0.8, seed=94)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(10.0, 2.0)  # random points on the sphere overlap: LJ would blow up
        s.constraint("sphere", st["R"])
        s.constraint_rescale(0.9, steps=60, freq=5)
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=7, temp=(0, 0.5, 0.5, 1))
    d.run(80)
    o.run(80)
    assert d.constraint_radius() == pytest.approx(o.constraint_radius(), rel=1e-15)
    X = np.stack([d.get(k) for k in "xyz"], 1)
    Y = np.stack([o.get(k) for k in "xyz"], 1)
    assert np.abs(X - Y).max() <= 1e-9
    V = np.stack([d.get(k) for k in ("vx", "vy", "vz")], 1)
    W = np.stack([o.get(k) for k in ("vx", "vy", "vz")], 1)
    assert np.abs(V - W).max() <= 1e-9
    d.close()
