"""Statistics of noisy runs, used to compare two noise streams.

The device draws its noise from a counter-based Philox stream; the reference
draws from one sequential GSL MT19937 stream per integrator
(src/utils/rng.cpp:34-61).  Trajectories therefore cannot agree draw by draw,
only in distribution.  Each scenario is run from one fixed start configuration
with M independent seeds on each side; every statistic is averaged per seed,
and the two seed means must agree within two standard errors of their
difference (z < 2).  Under the null hypothesis about 4.6% of such statistics
land outside 2 sigma by sampling noise alone, so a case with K statistics
passes when every z < 3 and at most one z lies in [2, 3) (false-alarm rate
about 2% per case of K = 4 or 5 statistics; see `same_distribution`).

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


def same_distribution(z):
    """Acceptance rule of a case: all z < 3 and at most one z >= 2."""
    z = np.asarray(z)
    return bool(np.all(z < 3.0) and np.count_nonzero(z >= 2.0) <= 1)
