"""Statistics of noisy runs, used to compare two noise streams.

The device draws its noise from a counter-based Philox stream; the reference
draws from one sequential GSL MT19937 stream per integrator
(src/utils/rng.cpp:34-61).  Trajectories therefore cannot agree draw by draw,
only in distribution.  Each scenario is run from one fixed start configuration
with M independent seeds on each side, and every statistic is averaged per
seed.  The two sides agree when their vectors of K statistics agree within
2 sigma: Hotelling's two-sample T^2 of the per-seed rows (pooled covariance,
F distribution with K and Ma + Mb - K - 1 degrees of freedom) gives a p-value,
and its two-sided normal deviate must be below 2 (p > 0.0455).  The
statistics of one case are strongly correlated (autocorrelations at
neighbouring lags); a separate 2-sigma test of each would reject a correct
stream in 1 - 0.9545^K of cases (21% for K = 5).  The per-statistic z-scores
are reported with it.

Observables:
  velocity autocorrelation  utils/Glassy.py:274-311 (getVelAuto): each
                            snapshot's velocities are normalised by that
                            snapshot's sqrt(<v^2>), then C(u) averages
                            v(t).v(t+u) over particles and time origins t;
  polar order               |sum_i n_i| / N (utils/Configuration.py:173-175);
  nematic order (2-D)       largest eigenvalue of Q = <2 n n - I> over the
                            plane's two components, sqrt(<cos 2t>^2 + <sin 2t>^2)
                            (the reference defines no nematic order parameter;
                            this is the textbook one);
  mean square speed         <|v|^2> over particles and snapshots.
"""
import math

import numpy as np


def vacf(V):
    """C(u) for u = 0 .. T-1 from snapshots V[T, N, 3] (utils/Glassy.py:274-311)."""
    V = np.asarray(V, dtype=np.float64)
    T, N = V.shape[0], V.shape[1]
    v2av = np.sum(V * V, axis=(1, 2)) / N
    Vn = V / np.sqrt(v2av)[:, None, None]
    out = np.empty(T)
    for u in range(T):
        smax = T - u
        out[u] = np.sum(Vn[:smax] * Vn[u:]) / (N * smax)
    return out


def polar_order(n):
    n = np.asarray(n, dtype=np.float64)
    return float(np.linalg.norm(n.sum(axis=0)) / n.shape[0])


def nematic_order_2d(n):
    th = np.arctan2(n[:, 1], n[:, 0])
    return float(math.hypot(np.cos(2 * th).mean(), np.sin(2 * th).mean()))


def snapshots(s, chunks, steps_per_chunk):
    """Run `chunks` chunks of `steps_per_chunk` steps (the run command executes
    N + 1 steps, so each chunk asks for steps_per_chunk - 1) and return the
    velocity and director snapshots taken after each chunk."""
    V, D = [], []
    for _ in range(chunks):
        s.run(steps_per_chunk - 1)
        V.append(np.stack([s.get("vx"), s.get("vy"), s.get("vz")], 1))
        D.append(np.stack([s.get("nx"), s.get("ny"), s.get("nz")], 1))
    return np.array(V), np.array(D)


def statistics(V, D, lags, nematic=False):
    """Per-run statistic vector: C(lag) for each lag, the order parameter
    averaged over the second half of the snapshots, and <|v|^2>."""
    c = vacf(V)
    order = nematic_order_2d if nematic else polar_order
    half = D.shape[0] // 2
    op = float(np.mean([order(d) for d in D[half:]]))
    v2 = float(np.mean(np.sum(V * V, axis=2)))
    return np.array([c[l] for l in lags] + [op, v2])


def zscores(a, b):
    """|mean(a) - mean(b)| / sqrt(var(a)/Ma + var(b)/Mb), per statistic, for
    per-seed statistic rows a[Ma, K], b[Mb, K]."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    se = np.sqrt(a.var(axis=0, ddof=1) / a.shape[0] + b.var(axis=0, ddof=1) / b.shape[0])
    diff = np.abs(a.mean(axis=0) - b.mean(axis=0))
    return np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(diff > 0, np.inf, 0.0))


def joint_z(a, b):
    """Two-sided normal deviate of Hotelling's two-sample T^2 test of equal
    means of the per-seed statistic rows a[Ma, K] and b[Mb, K] (pooled
    covariance).  Statistics with zero variance on both sides are left out
    when their means agree, and make the deviate infinite when they differ."""
    from scipy import stats

    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = a.shape[0], b.shape[0]
    d = a.mean(axis=0) - b.mean(axis=0)
    pooled = ((na - 1) * np.cov(a, rowvar=False, ddof=1) + (nb - 1) * np.cov(b, rowvar=False, ddof=1)) / (na + nb - 2)
    pooled = np.atleast_2d(pooled)
    keep = np.diag(pooled) > 0
    if np.any(~keep & (d != 0)):
        return math.inf
    d, pooled = d[keep], pooled[np.ix_(keep, keep)]
    k = d.size
    if k == 0:
        return 0.0
    t2 = na * nb / (na + nb) * float(d @ np.linalg.solve(pooled, d))
    f = (na + nb - k - 1) / (k * (na + nb - 2)) * t2
    p = float(stats.f.sf(f, k, na + nb - k - 1))
    return float(stats.norm.isf(p / 2.0)) if p > 0.0 else math.inf


def same_distribution(a, b):
    """Acceptance rule of a case: the statistic vectors agree within 2 sigma."""
    return joint_z(a, b) < 2.0
