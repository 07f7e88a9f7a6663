"""Morton re-sort inside long runs (option "resort_every": every k rebuilds
of the captured steps the particles' storage is re-sorted along the Morton
order of their cells, and the list is rebuilt from the permuted snapshot).

Only storage order changes: against the oracle (which never re-sorts) the
trajectory stays within 1e-9 and the pair set of the last build is equal,
for the fused Langevin and the fused Brownian schedules; the number of
counted builds is the oracle's; the re-sorts did happen.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def langevin(s):
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(0.004)
    s.integrator_langevin(gamma=1.0, seed=3, temp=(0, 1.2, 1.2, 1))


def brownian(s):
    s.nlist(2.4, 0.3)
    s.pair_soft(2.0, 2.0)
    s.align_polar(1.0, 2.4)
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_brownian(v0=1.0, nu=0.05, seed=6, temp=(0, 0.05, 0.05, 1))


def wrapped_max(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


@pytest.mark.parametrize("kind", ["langevin", "brownian"])
def test_resorted_run_matches_the_oracle(gpu, kind):
    st = sc.lj_lattice(40, rho=0.7, seed=111) if kind == "langevin" else sc.plane_disks(1600, phi=0.8, seed=112)
    if kind == "brownian":
        th = np.random.default_rng(113).uniform(0.0, 2.0 * np.pi, st["n"])
        st["cols"]["nx"], st["cols"]["ny"] = np.cos(th), np.sin(th)
    setup = langevin if kind == "langevin" else brownian
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    setup(d)
    setup(o)
    d.set_option("resort_every", 2)
    b0, ob0 = d.nlist_info()["builds"], o.builds()
    d.run(300)
    o.run(300)
    assert d.run_stats()["resorts"] >= 2
    L = st["box"][0]
    keys = ("x", "y", "vx", "vy") if kind == "langevin" else ("x", "y", "nx", "ny")
    for k in keys:
        err = wrapped_max(d.get(k), o.get(k), L) if k in ("x", "y") else float(np.abs(d.get(k) - o.get(k)).max())
        assert err <= 1e-9, (k, err)
    assert np.array_equal(d.get("ix"), o.get("ix")) and np.array_equal(d.get("iy"), o.get("iy"))
    assert np.array_equal(d.pairs(), o.pairs())
    # the re-sorts' list rebuilds are not counted: the builds are the run's own
    assert d.nlist_info()["builds"] - b0 == o.builds() - ob0
    d.close()


def test_resort_off_keeps_storage_order(gpu):
    st = sc.lj_lattice(32, rho=0.7, seed=114)
    d = sc.load(gpu.System, st)
    langevin(d)
    d.set_option("resort_every", 0)
    d.run(300)
    assert d.run_stats()["resorts"] == 0
    d.close()
