"""The trained warm-start network (samos_amd/data/warm_start_mpnn_n4096_phi088.npz)
against the method without it, on configurations it was not trained on
(2-D polydisperse soft disks, N = 4096, packing 0.88).

What it gains: its predicted displacement explains part of the relaxation
(loss below that of predicting no displacement).  What every answer must
satisfy: warm-started FIRE stops at max|F| < 1e-10 on every configuration.
The same-minimum rate (energy equal to plain FIRE's to 1e-12) is asserted
at its stated expected value: the shipped network does not keep plain FIRE's
minimum, and bench.py charges every such configuration the plain-FIRE fallback.

No restart keeps it either: FIRE restarted (v = 0, fresh dt and alpha) from
the exact positions plain FIRE itself reached after one iteration ends in
another minimum of these jammed packings (0 of 128 configurations kept the
minimum on the bench workload; restarts after 5 to 2000 iterations, with or
without the velocities, gave the same).  The minimum reached depends on the
whole FIRE history, so a warm start that only supplies positions cannot keep
it, and its fallback-charged speedup stays below 1.
"""
import json
import os

import numpy as np
import pytest

from tests import scenarios as sc

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "samos_amd", "data",
                       "warm_start_mpnn_n4096_phi088.npz")


def fire_ensemble(gpu, states):
    e = sc.load_ensemble(gpu.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    e.nlist(2.4, 0.3)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    return e


@pytest.fixture(scope="module")
def trained(gpu):
    if not os.path.exists(WEIGHTS):
        pytest.skip("trained weights not present")
    z = np.load(WEIGHTS, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    mp = gpu.Mpnn(seed=0)
    mp.set_params(z["params"])
    return mp, z["params"], meta


def test_weights_metadata(trained):
    _, params, meta = trained
    assert params.dtype == np.float32 and np.all(np.isfinite(params))
    assert meta["train_configs"] >= 20000
    assert meta["history"][-1]["val_loss"] < meta["val_loss_zero_prediction"]


def test_trained_network_beats_no_prediction_on_new_configurations(gpu, trained):
    mp, params, _ = trained
    states = sc.plane_replicas([4096] * 8, seed=777)
    e = fire_ensemble(gpu, states)
    mp.dataset_reserve(8, 4096)
    mp.dataset_begin(e)
    rem, _ = e.fire_run(400000, check=200)
    assert rem == 0
    mp.dataset_end(e)
    work = sc.load_ensemble(gpu.Ensemble, sc.plane_replicas([4096] * 8, seed=778), "plane")
    work.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    work.nlist(2.4, 0.3)
    base = mp.dataset_size() - 8
    loss = mp.evaluate(work, base, 8)
    mp.set_params(np.zeros_like(params))
    loss_zero = mp.evaluate(work, base, 8)
    mp.set_params(params)
    assert loss < 0.9 * loss_zero


def test_warm_started_fire_meets_the_residual_bar(gpu, trained):
    mp, _, _ = trained
    states = sc.plane_replicas([4096] * 8, seed=779)
    plain = fire_ensemble(gpu, states)
    rem, _ = plain.fire_run(400000, check=200)
    assert rem == 0
    fp = plain.fire_state()
    warm = sc.load_ensemble(gpu.Ensemble, states, "plane")
    warm.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    warm.nlist(2.4, 0.3)
    mp.warm_start(warm, 1.0)
    warm.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    rem, _ = warm.fire_run(400000, check=200)
    assert rem == 0
    fw = warm.fire_state()
    assert np.all(fw["max_force"] < 1e-10)
    en = warm.forces()
    np.testing.assert_allclose(en, fw["energy"], rtol=1e-12)
    same = np.abs(fw["energy"] - fp["energy"]) <= 1e-12 * np.abs(fp["energy"])
    # Stated expectation for the shipped network (0 of 512 on the bench
    # workload): its starting points lead FIRE into other minima of this
    # jammed packing, so the bench charges those configurations the plain-FIRE
    # fallback.  A network that kept the basin would fail this line and the
    # accounting text would have to change with it.
    assert same.mean() <= 0.25
    assert np.all(fw["iters"] > 0)


def test_restart_from_fire_own_path_changes_the_minimum(gpu):
    states = sc.plane_replicas([4096] * 16, seed=780)
    plain = fire_ensemble(gpu, states)
    rem, _ = plain.fire_run(400000, check=200)
    assert rem == 0
    fp = plain.fire_state()
    plain.close()
    out = {}
    for k in (0, 1):
        e = fire_ensemble(gpu, states)
        if k:
            e.fire_run(k, check=1)
            zeros = np.zeros(16 * 4096)
            e.set("vx", zeros)
            e.set("vy", zeros)
            e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
        rem, _ = e.fire_run(400000, check=200)
        assert rem == 0
        fr = e.fire_state()
        e.close()
        out[k] = np.abs(fr["energy"] - fp["energy"]) <= 1e-12 * np.abs(fp["energy"])
    assert out[0].all()          # the same start and state: the same minimum
    assert out[1].sum() <= 2     # one iteration, then a restart: another minimum
