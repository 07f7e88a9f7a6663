"""Warm-start network on the GPU (libsamos_hip.so, samos_mpnn_*) against the
CPU oracle (oracle/mpnn.py, bf16 rounding emulated at the same points).

Bars:
  forward displacements          |d_gpu - d_cpu| <= 2e-3 |d_cpu| (norms over all particles)
  loss                           relative 1e-4
  gradient, every block          |g_gpu - g_cpu| <= 2e-2 |g_cpu|; whole vector 5e-3
  training                       loss falls on (start, relaxed) pairs from batched FIRE
  warm start                     x moves by scale * delta, then FIRE reaches max|F| < 1e-10
Differences come from FP32 summation order and the rare bf16 rounding tie that
an FP32 difference moves across.
"""
import numpy as np
import pytest

from oracle import mpnn as om
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

RCUT, PAD = 2.4, 0.3
CUT = RCUT + PAD


def ensemble_of(gpu, states):
    e = sc.load_ensemble(gpu.Ensemble, states, "plane")
    e.set_params(k=1.0, a=2.0, use_radii=1.0, dt=0.01)
    e.nlist(RCUT, PAD)
    return e


def params_f32(seed):
    p = om.unflatten(om.init_params(seed))
    rng = np.random.default_rng(seed + 100)
    for name, r, cc in om.LAYOUT:
        if r == 1:
            p[name] = p[name] + rng.normal(0, 0.1, size=p[name].shape)
    return om.flatten(p).astype(np.float32)


def test_forward_matches_oracle(gpu):
    st = sc.plane_disks(90, phi=0.88, seed=31, polydisperse=(0.8, 1.2))
    e = ensemble_of(gpu, [st])
    mp = gpu.Mpnn(seed=1)
    flat = params_f32(1)
    mp.set_params(flat)
    d = mp.predict(e)
    c = st["cols"]
    ref = om.Net(flat.astype(np.float64)).forward(c["x"], c["y"], c["radius"], st["box"][0], CUT)
    assert np.linalg.norm(d - ref) <= 2e-3 * np.linalg.norm(ref)


def test_gradients_match_oracle(gpu):
    st = sc.plane_disks(90, phi=0.88, seed=32, polydisperse=(0.8, 1.2))
    e = ensemble_of(gpu, [st])
    mp = gpu.Mpnn(seed=2)
    flat = params_f32(2)
    mp.set_params(flat)
    target = np.random.default_rng(9).normal(0, 0.7, size=(st["n"], 2))
    loss, g = mp.gradients(e, target)
    c = st["cols"]
    net = om.Net(flat.astype(np.float64))
    d = net.forward(c["x"], c["y"], c["radius"], st["box"][0], CUT)
    loss_ref, g_ref = net.backward(d, target)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    assert np.linalg.norm(g - g_ref) <= 5e-3 * np.linalg.norm(g_ref)
    o = 0
    for name, rows, cols in om.LAYOUT:
        sl = slice(o, o + rows * cols)
        o += rows * cols
        ref = g_ref[sl]
        if np.linalg.norm(ref) == 0.0:
            assert np.all(g[sl] == 0.0), name
            continue
        assert np.linalg.norm(g[sl] - ref) <= 2e-2 * np.linalg.norm(ref), name


def test_gradients_several_replicas_sum_like_one_batch(gpu):
    sts = [sc.plane_disks(70, phi=0.88, seed=40 + i, polydisperse=(0.8, 1.2)) for i in range(3)]
    e = ensemble_of(gpu, sts)
    mp = gpu.Mpnn(seed=3)
    flat = params_f32(3)
    mp.set_params(flat)
    target = np.random.default_rng(11).normal(0, 0.5, size=(e.n, 2))
    loss, g = mp.gradients(e, target)
    # the loss is the mean over every particle of the batch
    n_tot = e.n
    acc_loss, acc_g = 0.0, np.zeros_like(g, dtype=np.float64)
    for i, st in enumerate(sts):
        c = st["cols"]
        net = om.Net(flat.astype(np.float64))
        d = net.forward(c["x"], c["y"], c["radius"], st["box"][0], CUT)
        t = target[e.offsets[i]:e.offsets[i + 1]]
        l_i, g_i = net.backward(d, t)
        w = st["n"] / n_tot
        acc_loss += w * l_i
        acc_g += w * g_i
    assert loss == pytest.approx(acc_loss, rel=1e-4)
    assert np.linalg.norm(g - acc_g) <= 5e-3 * np.linalg.norm(acc_g)


def relaxed_dataset(gpu, mp, nconf, n, seed):
    states = sc.plane_replicas([n] * nconf, seed=seed)
    e = ensemble_of(gpu, states)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    mp.dataset_reserve(nconf, n)
    mp.dataset_begin(e)
    rem, _ = e.fire_run(300000, check=200)
    assert rem == 0
    mp.dataset_end(e)
    return states


def test_training_reduces_loss(gpu):
    mp = gpu.Mpnn(seed=5)
    relaxed_dataset(gpu, mp, 24, 256, seed=61)
    assert mp.dataset_size() == 24
    work = ensemble_of(gpu, sc.plane_replicas([256] * 4, seed=62))
    before = mp.evaluate(work, 20, 4)
    losses = mp.train(work, 150, lr=2e-3, seed=1, first=0, count=20)
    after_train = mp.evaluate(work, 0, 20)
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-20:]) < 0.8 * np.mean(losses[:10])
    assert after_train < before
    # parameters stayed finite and padding rows stayed zero
    p = om.unflatten(mp.get_params().astype(np.float64))
    for name, used in om.USED_ROWS.items():
        assert np.all(p[name][used:] == 0.0), name


def test_warm_start_moves_by_prediction_then_fire_converges(gpu):
    states = sc.plane_replicas([256, 300], seed=71)
    e = ensemble_of(gpu, states)
    mp = gpu.Mpnn(seed=6)
    mp.set_params(params_f32(6))
    x0, y0 = e.get("x"), e.get("y")
    d = mp.predict(e)
    mp.warm_start(e, 0.5)
    x1, y1 = e.get("x"), e.get("y")
    for r, st in enumerate(states):
        sl = slice(e.offsets[r], e.offsets[r + 1])
        L = st["box"][0]
        np.testing.assert_allclose(sc.min_image(x1[sl] - x0[sl] - 0.5 * d[sl, 0], L), 0.0, atol=1e-12)
        np.testing.assert_allclose(sc.min_image(y1[sl] - y0[sl] - 0.5 * d[sl, 1], L), 0.0, atol=1e-12)
    e.fire_setup(dt=0.01, alpha=0.1, dt_max=0.1, f_tol=0.0, stop_fmax=1e-10)
    rem, _ = e.fire_run(300000, check=200)
    assert rem == 0
    assert np.all(e.fire_state()["max_force"] < 1e-10)
