"""Replica ensembles (libsamos_hip.so, samos_ens_*) against each replica run
alone in the CPU oracle on the same inputs.

Bars (as for single systems):
  pair sets per replica          bit-exact (sorted half-pair rows, replica input order)
  forces, torques                max_i |dF_i| / max_i |F_i| <= 1e-12
  soft energy per replica        relative 1e-12
  noise-off trajectories         |dx|, |dn| <= 1e-9 after 1000 steps
  noise-on, same Philox stream   |dx|, |dn| <= 1e-9 after 50 steps
  FIRE, 200 iterations           |dx| <= 1e-9; dt, alpha relative 1e-12
  FIRE to max|F| < 1e-10         every replica stops below the threshold
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

RCUT, PAD = 2.4, 0.3


def stack(s, names):
    return np.stack([s.get(k) for k in names], 1)


def rep_stack(e, r, cols):
    return np.stack([e.replica(r, c) for c in cols], 1)


def at(p, key, r):
    """Replica r's value of a per-replica or broadcast parameter."""
    a = np.asarray(p[key], dtype=np.float64)
    return float(a if a.ndim == 0 else a[r])


def oracle_active(st, p, r, constraint="sphere"):
    """Replica r as its own system in the oracle: soft + polar, Brownian at T = 0."""
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(RCUT, PAD)
    o.pair_soft(p["k"], p["a"])
    if at(p, "J", r) != 0.0:
        o.align_polar(at(p, "J", r), p["rcut_align"])
    if constraint == "sphere":
        o.constraint("sphere", st["R"])
    else:
        o.constraint(constraint)
    o.set_timestep(p["dt"])
    o.integrator_brownian(v0=at(p, "v0", r), nu=at(p, "nu", r), mu=p["mu"], mur=p["mur"],
                          seed=int(at(p, "seed", r)))
    return o


def active_ensemble(gpu, counts, seed, **over):
    states = sc.sphere_replicas(counts, seed=seed)
    p = dict(k=10.0, a=2.0, rcut_align=2.4, mu=1.0, mur=1.0, dt=0.005,
             **sc.active_replica_params(len(counts), seed=seed + 1))
    p.update(over)
    e = sc.load_ensemble(gpu.Ensemble, states, "sphere")
    e.set_params(**p)
    e.nlist(RCUT, PAD)
    return e, states, p


COUNTS = [300, 1500, 517, 2048, 96, 1200]


def test_pairs_per_replica_sphere(gpu):
    e, states, p = active_ensemble(gpu, COUNTS, seed=41)
    for r, st in enumerate(states):
        o = sc.load(oracle.OracleSystem, st)
        o.nlist(RCUT, PAD)
        assert o.uses_cells()
        assert np.array_equal(e.pairs(r), o.pairs()), f"replica {r}"
    info = e.nlist_info()
    assert info["particles"] == sum(COUNTS) and info["replicas"] == len(COUNTS)


def test_pairs_per_replica_plane_periodic(gpu):
    states = sc.plane_replicas([400, 1000, 257, 2000], seed=51)
    e = sc.load_ensemble(gpu.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0)
    e.nlist(RCUT, PAD)
    for r, st in enumerate(states):
        o = sc.load(oracle.OracleSystem, st)
        o.nlist(RCUT, PAD)
        assert np.array_equal(e.pairs(r), o.pairs()), f"replica {r}"


def test_forces_torques_energy_per_replica(gpu):
    e, states, p = active_ensemble(gpu, COUNTS, seed=42)
    en = e.forces()
    F = stack(e, ["fx", "fy", "fz"])
    T = stack(e, ["tau_x", "tau_y", "tau_z"])
    for r, st in enumerate(states):
        o = oracle_active(st, p, r)
        eo = o.compute_forces()
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert sc.rel_force_err(F[sl], stack(o, ["fx", "fy", "fz"])) <= 1e-12
        if at(p, "J", r) != 0.0:
            assert sc.rel_force_err(T[sl], stack(o, ["tau_x", "tau_y", "tau_z"])) <= 1e-12
        assert en[r] == pytest.approx(eo["pair"], rel=1e-12)


def test_brownian_noise_off_1000_steps(gpu):
    e, states, p = active_ensemble(gpu, [400, 900, 250, 1300], seed=43, nu=0.0)
    e.run_brownian(999)
    X = stack(e, ["x", "y", "z"])
    N = stack(e, ["nx", "ny", "nz"])
    age = e.get("age")
    for r, st in enumerate(states):
        o = oracle_active(st, p, r)
        o.run(999)
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert np.abs(X[sl] - stack(o, ["x", "y", "z"])).max() <= 1e-9, f"replica {r}"
        assert np.abs(N[sl] - stack(o, ["nx", "ny", "nz"])).max() <= 1e-9, f"replica {r}"
        assert np.array_equal(age[sl], o.get("age"))


def test_brownian_noise_on_same_stream(gpu):
    e, states, p = active_ensemble(gpu, [600, 333, 1000], seed=44)
    e.run_brownian(49)
    X = stack(e, ["x", "y", "z"])
    N = stack(e, ["nx", "ny", "nz"])
    V = stack(e, ["vx", "vy", "vz"])
    for r, st in enumerate(states):
        o = oracle_active(st, p, r)
        o.run(49)
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert np.abs(X[sl] - stack(o, ["x", "y", "z"])).max() <= 1e-9
        assert np.abs(N[sl] - stack(o, ["nx", "ny", "nz"])).max() <= 1e-9
        assert np.abs(V[sl] - stack(o, ["vx", "vy", "vz"])).max() <= 1e-9


def test_brownian_graph_and_direct_launches_agree(gpu):
    a, _, _ = active_ensemble(gpu, [700, 1400, 300], seed=45)
    b, _, _ = active_ensemble(gpu, [700, 1400, 300], seed=45)
    a.run_brownian(99)
    b.run_brownian(99, gpu.RUN_NO_GRAPH)
    for k in ["x", "y", "z", "nx", "ny", "nz", "fx", "tau_z", "omega"]:
        assert np.array_equal(a.get(k), b.get(k)), k
    # a second run continues the same noise stream
    a.run_brownian(40)
    b.run_brownian(40, gpu.RUN_NO_GRAPH)
    assert np.array_equal(a.get("nx"), b.get("nx"))


def test_brownian_split_runs_match_oracle(gpu):
    e, states, p = active_ensemble(gpu, [500, 800], seed=46)
    e.run_brownian(70)
    e.run_brownian(29)
    for r, st in enumerate(states):
        o = oracle_active(st, p, r)
        o.run(70)
        o.run(29)
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert np.abs(stack(e, ["x", "y", "z"])[sl] - stack(o, ["x", "y", "z"])).max() <= 1e-9


# f_tolerance = 0 switches off the reference's RMS-force / energy-change test,
# which would otherwise stop the velocity mixing (IntegratorFIRE::integrate
# returns before it) long before max|F| reaches 1e-10.
def test_autotune_leaves_state_and_physics(gpu):
    e, states, p = active_ensemble(gpu, [900, 400, 1600], seed=47)
    before = {k: e.get(k) for k in ["x", "y", "z", "nx", "ny", "nz", "age"]}
    best = e.autotune(trial_steps=40, pads=(0.2, 0.45))
    assert best["pad"] in (0.2, 0.45) and best["gather_batch"] in (4, 8) and best["ms_per_step"] > 0
    for k, v in before.items():
        assert np.array_equal(e.get(k), v), k
    e.run_brownian(99)
    for r, st in enumerate(states):
        o = oracle_active(st, p, r)
        o.run(99)
        sl = slice(e.offsets[r], e.offsets[r + 1])
        assert np.abs(stack(e, ["x", "y", "z"])[sl] - stack(o, ["x", "y", "z"])).max() <= 1e-9
        assert np.abs(stack(e, ["nx", "ny", "nz"])[sl] - stack(o, ["nx", "ny", "nz"])).max() <= 1e-9


def fire_ensemble(gpu, counts, seed):
    states = sc.plane_replicas(counts, seed=seed)
    e = sc.load_ensemble(gpu.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    e.nlist(RCUT, PAD)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    return e, states


def oracle_fire(st):
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(RCUT, PAD)
    o.pair_soft(1.0, 2.0, use_radii=True)
    o.constraint("plane")
    o.set_timestep(0.01)
    o.integrator_fire(dt_max=0.1, alpha=0.1, f_tol=0.0)
    return o


def test_fire_200_iterations_per_replica(gpu):
    e, states = fire_ensemble(gpu, [512, 1024, 300], seed=52)
    rem, it = e.fire_run(200, check=50)
    assert it == 200 and rem == len(states)
    fs = e.fire_state()
    for r, st in enumerate(states):
        o = oracle_fire(st)
        o.run(199)
        fo = o.fire_state()
        sl = slice(e.offsets[r], e.offsets[r + 1])
        d = stack(e, ["x", "y", "z"])[sl] - stack(o, ["x", "y", "z"])
        L = st["box"]
        for k in range(2):
            d[:, k] = sc.min_image(d[:, k], L[k])
        assert np.abs(d).max() <= 1e-9, f"replica {r}"
        assert fs["dt"][r] == pytest.approx(fo["dt"], rel=1e-12)
        assert fs["alpha"][r] == pytest.approx(fo["alpha"], rel=1e-12)
        assert fs["energy"][r] == pytest.approx(fo["E"], rel=1e-9)


def test_fire_converges_every_replica(gpu):
    e, states = fire_ensemble(gpu, [256, 400, 300, 512], seed=53)
    rem, it = e.fire_run(200000, check=100)
    assert rem == 0
    fs = e.fire_state()
    assert np.all(fs["stopped"] == 1)
    assert np.all(fs["max_force"] < 1e-10)
    en = e.forces()
    F = stack(e, ["fx", "fy", "fz"])
    assert np.linalg.norm(F, axis=1).max() < 1e-10
    np.testing.assert_allclose(en, fs["energy"], rtol=1e-12)


def test_ensemble_rejects_small_boxes(gpu):
    st = sc.plane_disks(50, phi=0.5, seed=3, lz=6.0)
    e = sc.load_ensemble(gpu.Ensemble, [st], "plane")
    with pytest.raises(gpu.SamosError, match="three cells"):
        e.nlist(RCUT, PAD)


def test_ensemble_rejects_bad_params(gpu):
    st = sc.plane_disks(100, phi=0.5, seed=3)
    e = sc.load_ensemble(gpu.Ensemble, [st], "plane")
    with pytest.raises(gpu.SamosError, match="timestep"):
        e.set_params(dt=0.0)
