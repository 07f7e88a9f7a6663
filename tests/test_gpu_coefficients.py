"""Device against the oracle where mass, mu and mur differ from 1: the
coefficients dt/(2m), B/sqrt(m), mu F, sqrt(2 mu T) and mur tau of the fused
Langevin and Brownian steps, the generic (faithful) step sequence and the
ensemble's per-replica parameters.

Bars: noise on, the same Philox stream on both sides, |dx|, |dv|, |dn| <= 1e-9
after 50 steps (device sin/cos/log and the host libm differ in the last bit).
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


def stack(s, names):
    return np.stack([s.get(k) for k in names], 1)


def wrapped(a, b, box):
    d = a - b
    for k in range(d.shape[1]):
        d[:, k] = sc.min_image(d[:, k], box[k])
    return float(np.abs(d).max())


@pytest.mark.parametrize("masses", ["uniform", "per_particle"])
@pytest.mark.parametrize("flags", [0, 1])
def test_langevin_with_mass(gpu, masses, flags):
    st = sc.lj_lattice(40, rho=0.7, seed=61)
    n = st["n"]
    rng = np.random.default_rng(7)
    st["cols"]["mass"] = np.full(n, 2.5) if masses == "uniform" else rng.uniform(0.5, 3.0, n)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.003)
        s.integrator_langevin(gamma=0.7, seed=3, temp=(0, 1.3, 1.0, 1))
    d.run(49, flags)
    o.run(49)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["vx", "vy"]) - stack(o, ["vx", "vy"])).max() <= 1e-9
    assert sc.rel_force_err(stack(d, ["fx", "fy", "fz"]), stack(o, ["fx", "fy", "fz"])) <= 1e-9
    d.close()


@pytest.mark.parametrize("method", ["spv", "bbk"])
def test_langevin_methods_with_mass(gpu, method):
    st = sc.lj_lattice(30, rho=0.6, seed=62)
    st["cols"]["mass"] = np.random.default_rng(8).uniform(0.7, 2.2, st["n"])
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.4, seed=5, method=method, temp=(0, 0.6, 1.0, 1))
    d.run(49)
    o.run(49)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["vx", "vy"]) - stack(o, ["vx", "vy"])).max() <= 1e-9
    d.close()


@pytest.mark.parametrize("case", ["plane_polar", "sphere_nematic"])
@pytest.mark.parametrize("flags", [0, 1])
def test_brownian_mu_mur(gpu, case, flags):
    if case == "plane_polar":
        st = sc.plane_disks(2000, phi=0.75, seed=63)
    else:
        st = sc.sphere_disks(1500, phi=0.8, seed=63)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(3.0, 2.0)
        if case == "plane_polar":
            s.align_polar(0.4, 2.4)
            s.constraint("plane")
            s.self_propulsion(0.3)
        else:
            s.align_nematic(0.5, 2.4)
            s.constraint("sphere", st["R"])
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.2, nu=0.05, mu=0.6 if case == "plane_polar" else 1.7,
                              mur=0.4 if case == "plane_polar" else 2.2, seed=5, temp=(0, 0.05, 0.05, 1))
    d.run(49, flags)
    o.run(49)
    box, per = st["box"], st["periodic"]
    X, Y = stack(d, ["x", "y", "z"]), stack(o, ["x", "y", "z"])
    assert (wrapped(X, Y, box) if per else float(np.abs(X - Y).max())) <= 1e-9
    assert np.abs(stack(d, ["nx", "ny", "nz"]) - stack(o, ["nx", "ny", "nz"])).max() <= 1e-9
    assert np.abs(stack(d, ["vx", "vy", "vz"]) - stack(o, ["vx", "vy", "vz"])).max() <= 1e-9
    assert np.abs(d.get("omega") - o.get("omega")).max() <= 1e-9
    d.close()


def test_brownian_align_pos_pair_mu_mur(gpu):
    st = sc.plane_disks(1800, phi=0.7, seed=64)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.6, 2.4)
        s.constraint("plane")
        s.self_propulsion(0.5)
        s.set_timestep(0.01)
        s.integrator_brownian(which="brownian_align", nu=0.1, mur=0.35, seed=2)
        s.integrator_brownian(which="brownian_pos", mu=0.45, seed=3, temp=(0, 0.02, 0.02, 1))
    d.run(49)
    o.run(49)
    assert wrapped(stack(d, ["x", "y"]), stack(o, ["x", "y"]), st["box"]) <= 1e-9
    assert np.abs(stack(d, ["nx", "ny"]) - stack(o, ["nx", "ny"])).max() <= 1e-9
    d.close()


def test_ensemble_per_replica_mu_mur(gpu):
    counts = [500, 1300, 260, 900]
    states = sc.sphere_replicas(counts, seed=65)
    rng = np.random.default_rng(66)
    p = dict(k=10.0, a=2.0, rcut_align=2.4, dt=0.005, **sc.active_replica_params(len(counts), seed=67))
    p["mu"] = rng.uniform(0.3, 1.8, len(counts))
    p["mur"] = rng.uniform(0.2, 2.5, len(counts))
    p["J"] = np.maximum(p["J"], 0.2)
    e = sc.load_ensemble(gpu.Ensemble, states, "sphere")
    e.set_params(**p)
    e.nlist(2.4, 0.3)
    e.run_brownian(49)
    X, N = stack(e, ["x", "y", "z"]), stack(e, ["nx", "ny", "nz"])
    for r, st in enumerate(states):
        o = sc.load(oracle.OracleSystem, st)
        o.nlist(2.4, 0.3)
        o.pair_soft(p["k"], p["a"])
        o.align_polar(float(p["J"][r]), p["rcut_align"])
        o.constraint("sphere", st["R"])
        o.set_timestep(p["dt"])
        o.integrator_brownian(v0=float(p["v0"][r]), nu=float(p["nu"][r]), mu=float(p["mu"][r]),
                              mur=float(p["mur"][r]), seed=int(p["seed"][r]))
        o.run(49)
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert np.abs(X[sl] - stack(o, ["x", "y", "z"])).max() <= 1e-9, f"replica {r}"
        assert np.abs(N[sl] - stack(o, ["nx", "ny", "nz"])).max() <= 1e-9, f"replica {r}"
    e.close()
