"""The device's noise stream against the oracle's, bit for bit.

The draws are Philox4x32-10 words (integer arithmetic) turned into Gaussians
by the Box-Muller transform, whose log and sin/cos of 2 pi u are evaluated by
the same IEEE/fma expression on both sides (pinned against libm in
test_oracle.py).  Free overdamped diffusion exposes the draws directly: with
no pair forces, no self-propulsion and no alignment each step adds
sqrt(2 T mu dt) g to each coordinate and a rotational kick to the director.

Bar: positions, velocities and image flags identical to the oracle's after
every range (no tolerance), for the Brownian step (fused and step by step)
and the Langevin step.  Directors are turned by the drawn angle through cos
and sin of that angle (the device's and the C library's, which may differ in
the last bit), so they are held to 1e-13.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


@pytest.mark.parametrize("flags", [0, 2])
def test_free_brownian_diffusion_bit_exact(gpu, flags):
    st = sc.plane_disks(3000, phi=0.3, seed=71)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.0, nu=0.3, mu=1.0, seed=12, temp=(0, 0.7, 0.7, 1))
    for steps in (1, 17, 60):
        d.run(steps, flags)
        o.run(steps)
        for k in ("x", "y", "ix", "iy"):
            a, b = d.get(k), o.get(k)
            assert np.array_equal(a, b), (k, steps, float(np.abs(a - b).max()))
        for k in ("nx", "ny"):
            assert np.abs(d.get(k) - o.get(k)).max() <= 1e-13, k
    d.close()


def test_langevin_noise_bit_exact_without_forces(gpu):
    st = sc.plane_disks(2000, phi=0.3, seed=72)
    d, o = both(gpu, st)
    for s in (d, o):
        s.constraint("plane")
        s.set_timestep(0.005)
        s.integrator_langevin(gamma=0.8, seed=5, temp=(0, 1.3, 1.3, 1))
    d.run(40)
    o.run(40)
    for k in ("x", "y", "vx", "vy"):
        a, b = d.get(k), o.get(k)
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))
    d.close()
