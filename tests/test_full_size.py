"""Parity at the sizes the benchmark runs (list rows of 2^20 slots, gathers
with 32-bit byte offsets into 16 MiB columns, pairs across periodic faces
that lie far apart in the Morton order):

  * the LJ Langevin benchmark system itself: N = 1,048,576 disks at density
    0.7 on a plane, rcut 2.5, skin 0.5, dt 0.002, T = 1 (captured step
    chunks with device-side rebuilds), against the oracle after 12 steps:
    positions and velocities within 1e-9, the pair set of the last build
    equal;
  * an active nematic Brownian system on a sphere with 40,000 particles (the
    3-D gather layout) against the oracle after 30 steps: positions and
    directors within 1e-9, pair sets equal.

The oracle takes a few seconds per run at these sizes.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def wrapped_max(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


def test_lj_langevin_benchmark_system_1m(gpu):
    st = sc.lj_lattice(1024, rho=0.7, jitter=0.1, T=1.0, seed=2024)
    assert st["n"] == 1 << 20
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        s.nlist(2.5, 0.5)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=1, temp=(0, 1.0, 1.0, 1))
    d.run(12)
    o.run(12)
    L = st["box"][0]
    assert wrapped_max(d.get("x"), o.get("x"), L) <= 1e-9
    assert wrapped_max(d.get("y"), o.get("y"), L) <= 1e-9
    assert np.abs(d.get("vx") - o.get("vx")).max() <= 1e-9
    assert np.abs(d.get("vy") - o.get("vy")).max() <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())
    d.close()
    o.close()


def test_nematic_brownian_sphere_40k(gpu):
    st = sc.sphere_disks(40000, phi=0.9, seed=95)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        s.nlist(2.4, 0.4)
        s.pair_soft(10.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.constraint("sphere", st["R"])
        s.set_timestep(0.005)
        s.integrator_brownian(v0=0.5, nu=0.05, seed=9, temp=(0, 0.01, 0.01, 1))
    d.run(30)
    o.run(30)
    for k in ("x", "y", "z", "nx", "ny", "nz"):
        assert np.abs(d.get(k) - o.get(k)).max() <= 1e-9, k
    assert np.array_equal(d.pairs(), o.pairs())
    d.close()
    o.close()
