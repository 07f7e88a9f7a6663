"""Noise-on parity in distribution.

The reference draws its noise from one sequential GSL MT19937 stream per
integrator (src/utils/rng.cpp:34-61) in a per-particle order
(src/integrators/integrator_brownian.cpp:57-118,
src/integrators/integrator_langevin.cpp:65-94); the device draws a
counter-based Philox stream keyed by particle id and call.  The oracle can run
either stream, so the checks here are:

  * the reference-order stream itself: MT19937 known answers (the 1000th word
    of GSL's default seed 4357, 1186927261, GSL's own generator test; the
    10000th word of seed 5489, 4123659995, the C++ standard's) and Gaussian
    moments of the polar method;
  * oracle(Philox) against oracle(MT19937), on the CPU, and device (Philox)
    against oracle(MT19937), on the GPU: velocity autocorrelation at three
    lags, the polar or nematic order and <|v|^2>, averaged per seed over M
    seeds from one start configuration, must agree within 2 sigma as a
    vector (Hotelling's T^2 turned into a normal deviate, tests/noise_stats.py);
  * the comparison has power: a run with the rotational noise variance nu or
    the friction gamma doubled fails it by a wide margin.
"""
import math

import numpy as np
import pytest

import oracle
from tests import noise_stats as ns
from tests import scenarios as sc

LAGS = [1, 3, 6]


def _stream(s, stream):
    if stream is not None:
        s.rng_stream(stream)


def active_plane(cls, seed, n, kind="polar", nu=None, stream=None, chunks=10, per=100, **kw):
    """C1 physics (polar) or C2 physics on a plane (nematic), T = 0: the only
    noise is rotational, variance nu*dt per step."""
    st = sc.plane_disks(n, phi=0.8, seed=31)
    s = sc.load(cls, st, **kw)
    _stream(s, stream)
    s.nlist(2.4, 0.3)
    if kind == "polar":
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.1, 2.4)
        s.set_timestep(0.01)
        s.constraint("plane")
        s.integrator_brownian(v0=0.1, nu=0.1 if nu is None else nu, seed=seed)
    else:
        s.pair_soft(10.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.set_timestep(0.005)
        s.constraint("plane")
        s.integrator_brownian(v0=0.5, nu=0.05 if nu is None else nu, seed=seed)
    V, D = ns.snapshots(s, chunks, per)
    s.close()
    return ns.statistics(V, D, LAGS, nematic=(kind != "polar"))


def active_sphere(cls, seed, n, stream=None, chunks=10, per=100, **kw):
    """C2 physics: nematic soft disks on a sphere, rotational noise only; the
    order column is the polar order (the sphere has no global nematic axis)."""
    st = sc.sphere_disks(n, phi=1.0, seed=33)
    s = sc.load(cls, st, **kw)
    _stream(s, stream)
    s.nlist(2.4, 0.3)
    s.pair_soft(10.0, 2.0)
    s.align_nematic(0.5, 2.4)
    s.constraint("sphere", st["R"])
    s.set_timestep(0.005)
    s.integrator_brownian(v0=0.5, nu=0.05, seed=seed)
    V, D = ns.snapshots(s, chunks, per)
    s.close()
    return ns.statistics(V, D, LAGS)


def lj_langevin(cls, seed, nside, gamma=1.0, stream=None, chunks=10, per=50, **kw):
    """C3 physics: LJ Langevin BAOAB at T = 1 from a jittered lattice."""
    st = sc.lj_lattice(nside, rho=0.7, seed=32)
    s = sc.load(cls, st, **kw)
    _stream(s, stream)
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(0.002)
    s.integrator_langevin(gamma=gamma, seed=seed, temp=(0, 1.0, 1.0, 1))
    V, D = ns.snapshots(s, chunks, per)
    s.close()
    return ns.statistics(V, D, [1, 2, 4])[[0, 1, 2, 4]]  # directors do not move


def oracle_runs(fn, seeds, stream, **kw):
    return np.array([fn(oracle.OracleSystem, sd, stream=stream, fast=True, **kw) for sd in seeds])


# ------------------------------------------------------------------ CPU ---
def test_reference_stream_known_answers():
    u, _ = oracle.gsl_stream_sample(0, 1000)          # seed 0 means GSL's default 4357
    assert int(round(u[-1] * 2.0 ** 32)) == 1186927261
    u, _ = oracle.gsl_stream_sample(5489, 10000)
    assert int(round(u[-1] * 2.0 ** 32)) == 4123659995
    u, _ = oracle.gsl_stream_sample(7, 200000)
    _, g = oracle.gsl_stream_sample(8, 0, 200000)
    assert np.all((u >= 0.0) & (u < 1.0))
    se = 1.0 / math.sqrt(g.size)
    assert abs(g.mean()) < 4 * se
    assert abs(g.var() - 1.0) < 4 * math.sqrt(2.0) * se
    assert abs((g ** 4).mean() - 3.0) < 4 * math.sqrt(96.0) * se


def test_reference_stream_draw_order():
    # Langevin BAOAB draws three Gaussians per particle, in particle order
    # (integrator_langevin.cpp:79-83): with gamma*dt huge the O step leaves
    # v = sqrt(T) * (g_3i, g_3i+1) for particle i (v_z is dropped by the plane).
    n = 5
    s = oracle.OracleSystem(n, (100.0, 100.0, 10.0), True)
    s.set("x", np.arange(n) * 10.0)
    s.rng_stream("mt19937")
    s.set_timestep(1.0)
    s.constraint("plane")
    s.integrator_langevin(gamma=1000.0, seed=11, temp=(0, 4.0, 1.0, 1))
    s.integrate()
    _, g = oracle.gsl_stream_sample(11, 0, 3 * n)
    np.testing.assert_allclose(s.get("vx"), 2.0 * g[0::3], rtol=1e-12)
    np.testing.assert_allclose(s.get("vy"), 2.0 * g[1::3], rtol=1e-12)


@pytest.mark.parametrize("scenario", ["polar_plane", "lj_langevin"])
def test_philox_stream_statistics_match_reference_stream(scenario):
    m = 48
    if scenario == "polar_plane":
        a = oracle_runs(active_plane, range(100, 100 + m), "philox", n=256)
        b = oracle_runs(active_plane, range(200, 200 + m), "mt19937", n=256)
    else:
        a = oracle_runs(lj_langevin, range(100, 100 + m), "philox", nside=16)
        b = oracle_runs(lj_langevin, range(200, 200 + m), "mt19937", nside=16)
    zj = ns.joint_z(a, b)
    assert ns.same_distribution(a, b), (zj, ns.zscores(a, b))


def test_comparison_detects_wrong_noise():
    m = 48
    a = oracle_runs(active_plane, range(100, 100 + m), "philox", n=256)
    b = oracle_runs(active_plane, range(300, 300 + m), "mt19937", n=256, nu=0.2)
    assert ns.zscores(a, b).max() > 8.0 and ns.joint_z(a, b) > 8.0
    a = oracle_runs(lj_langevin, range(100, 100 + m), "philox", nside=16)
    b = oracle_runs(lj_langevin, range(300, 300 + m), "mt19937", nside=16, gamma=2.0)
    assert ns.zscores(a, b).max() > 4.0 and ns.joint_z(a, b) > 4.0


# ------------------------------------------------------------------ GPU ---
GPU_CASES = {
    "polar_plane": (active_plane, {"n": 1024}),
    "nematic_plane": (active_plane, {"n": 1024, "kind": "nematic"}),
    "nematic_sphere": (active_sphere, {"n": 1000}),
    "lj_langevin": (lj_langevin, {"nside": 32}),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(GPU_CASES))
def test_device_noise_statistics_match_reference_stream(gpu, case):
    fn, kw = GPU_CASES[case]
    m = 64
    print("noise statistics: %s, %d seeds per side" % (case, m), flush=True)
    dev = np.array([fn(gpu.System, sd, **kw) for sd in range(500, 500 + m)])
    ref = oracle_runs(fn, range(600, 600 + m), "mt19937", **kw)
    z, zj = ns.zscores(dev, ref), ns.joint_z(dev, ref)
    print("  device means %s\n  oracle means %s\n  z %s, joint %.3f" % (dev.mean(0), ref.mean(0), z, zj), flush=True)
    assert ns.same_distribution(dev, ref), (zj, z)
