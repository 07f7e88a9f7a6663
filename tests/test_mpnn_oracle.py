"""Pins of the CPU oracle of the warm-start network (oracle/mpnn.py).

  backward pass    analytic gradient == central finite differences (float64 mode)
  equivariance     rotating the configuration by 90 degrees rotates the prediction;
                   translating it leaves the prediction unchanged
  bf16 rounding    round-to-nearest-even on known values, ties included
  padding          padded weight rows receive exactly zero gradient
"""
import numpy as np
import pytest

from oracle import mpnn as om
from tests import scenarios as sc

CUT = 2.7


def small_config(n=40, seed=4):
    st = sc.plane_disks(n, phi=0.88, seed=seed, polydisperse=(0.8, 1.2))
    c = st["cols"]
    return c["x"].copy(), c["y"].copy(), c["radius"].copy(), st["box"][0]


def perturbed_params(seed=1):
    p = om.unflatten(om.init_params(seed))
    rng = np.random.default_rng(seed + 100)
    for name, r, cc in om.LAYOUT:
        if r == 1:
            p[name] = p[name] + rng.normal(0, 0.1, size=p[name].shape)
    return om.flatten(p)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7 + 2 ** -9, 3.0e-3], dtype=np.float32)
    want = np.array([1.0, 1.0, 1.015625, -2.5, 1.0078125, 0.0029983520507812], dtype=np.float32)
    got = om.bf16(x)
    assert np.array_equal(got[:5], want[:5])
    # 3e-3 = 1.536 * 2^-9: mantissa rounded to 7 bits
    assert got[5] == np.float32(np.round(1.536 * 128) / 128 * 2 ** -9)


def test_backward_matches_finite_differences():
    x, y, r, L = small_config()
    flat = perturbed_params()
    net = om.Net(flat, exact=True)
    target = np.random.default_rng(7).normal(0, 0.5, size=(len(x), 2))
    delta = net.forward(x, y, r, L, CUT)
    loss, grad = net.backward(delta, target)
    rng = np.random.default_rng(3)
    # a sample of parameters from every block, used rows only
    idx, o = [], 0
    for name, rows, cols in om.LAYOUT:
        used = om.USED_ROWS.get(name, rows) if rows > 1 else 1
        for _ in range(4):
            rr = rng.integers(0, used)
            idx.append(o + rr * cols + rng.integers(0, cols))
        o += rows * cols
    eps = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp = om.Net(fp, exact=True)
        lm = om.Net(fm, exact=True)
        Lp = lp.backward(lp.forward(x, y, r, L, CUT), target)[0]
        Lm = lm.backward(lm.forward(x, y, r, L, CUT), target)[0]
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_padding_rows_get_no_gradient():
    x, y, r, L = small_config(seed=5)
    flat = perturbed_params(2)
    net = om.Net(flat, exact=True)
    d = net.forward(x, y, r, L, CUT)
    _, grad = net.backward(d, np.zeros_like(d))
    g = om.unflatten(grad)
    for name, used in om.USED_ROWS.items():
        assert np.all(g[name][used:] == 0.0), name


def test_prediction_rotates_and_translates_with_the_configuration():
    x, y, r, L = small_config(seed=6)
    flat = perturbed_params(3)
    d0 = om.Net(flat, exact=True).forward(x, y, r, L, CUT)
    # rotation by 90 degrees maps the square periodic box onto itself
    d1 = om.Net(flat, exact=True).forward(-y, x, r, L, CUT)
    np.testing.assert_allclose(d1[:, 0], -d0[:, 1], atol=1e-12)
    np.testing.assert_allclose(d1[:, 1], d0[:, 0], atol=1e-12)
    xs = sc.min_image(x + 3.3, L)
    ys = sc.min_image(y - 1.7, L)
    d2 = om.Net(flat, exact=True).forward(xs, ys, r, L, CUT)
    np.testing.assert_allclose(d2, d0, atol=1e-12)


def test_bf16_mode_close_to_exact():
    x, y, r, L = small_config(seed=8)
    flat = perturbed_params(4)
    de = om.Net(flat, exact=True).forward(x, y, r, L, CUT)
    db = om.Net(flat, exact=False).forward(x, y, r, L, CUT)
    assert np.linalg.norm(db - de) <= 0.05 * np.linalg.norm(de)
