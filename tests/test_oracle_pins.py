"""Closed-form pins for the oracle's integrator coefficients.

Each test fixes a coefficient that the reference's formula carries and that a
plausible slip (a dropped 1/m, dt for dt/2, a wrong sign, mu or mur left out)
would change, using a result fixed by the mathematics rather than by the
oracle's own expressions:

  BAOAB, gamma = 0, constant force    velocity Verlet under constant force is
  (integrator_langevin.cpp:54-118)    exact: v = a t, x = a t^2 / 2, a = F / m
  BAOAB, gamma > 0, constant force    the geometric series of the B-A-O-A-B map
  BAOAB, gamma = 0, soft pair         energy conservation: exit speed
                                      (a - r0) sqrt(k / (2 m)) within O(dt^2)
  BAOAB noise with mass               equipartition <m v^2> = T per degree of freedom
  SPV, constant force                 v_k = (a / gamma)(1 - exp(-gamma k dt)) exactly;
  (integrator_langevin.cpp:126-182)   the reference's eta = 0 at gamma = 0 ignores forces
  BBK, constant force                 v' = r v + dt a / (1 + g), r = (1 - g) / (1 + g),
  (integrator_langevin.cpp:192-258)   g = gamma dt / 2; at gamma = 0 velocity Verlet
  plane rotate_director               (n0 x n1) . z = sin(phi), n0 . n1 = cos(phi)
  (constraint_plane.cpp:96-113)
  Brownian mu (integrator_brownian    drift (v0 + mu alpha) n t under self propulsion;
  .cpp:42-124)                        positional variance 2 mu T dt per step
  Brownian mur                        one step of two aligned disks turns each
                                      director by dt mur J sin(theta_j - theta_i)
  FIRE (integrator_fire.cpp:34-148)   the first step of an LJ pair from rest with
                                      the textbook LJ force; dt grows by f_inc only
                                      after N_min positive-power steps, up to
                                      dt_max, alpha by alpha_scale; P <= 0 halves
                                      dt, resets alpha and zeroes every velocity
"""
import math

import numpy as np
import pytest

import oracle


def one_particle(box=1e4):
    return oracle.OracleSystem(1, (box, box, 10.0), True)


def director(s, th):
    s.set("nx", [math.cos(th)])
    s.set("ny", [math.sin(th)])


def lj_force(r, eps=1.0):
    """Textbook LJ pair force magnitude at sigma = 1 (positive = repulsive)."""
    return 24.0 * eps * (2.0 / r ** 13 - 1.0 / r ** 7)


# ------------------------------------------------------------ BAOAB --------
@pytest.mark.parametrize("m", [2.0, 0.5])
def test_baoab_constant_force_without_friction(m):
    alpha, th, dt, K = 0.6, 0.9, 0.01, 200
    s = one_particle()
    director(s, th)
    s.set("mass", [m])
    s.self_propulsion(alpha)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_langevin(gamma=0.0, temp=(0, 0.0, 1.0, 1))
    s.run(K - 1)  # K integrator calls
    a, t = alpha / m, K * dt
    assert s.get("vx")[0] == pytest.approx(a * t * math.cos(th), rel=1e-12)
    assert s.get("vy")[0] == pytest.approx(a * t * math.sin(th), rel=1e-12)
    assert s.get("x")[0] == pytest.approx(0.5 * a * t * t * math.cos(th), rel=1e-11)
    assert s.get("y")[0] == pytest.approx(0.5 * a * t * t * math.sin(th), rel=1e-11)


def test_baoab_constant_force_with_friction():
    # per step: v' = e (v + h a) + h a, x' = x + h (1 + e)(v + h a), with
    # h = dt/2, e = exp(-gamma dt): v_K = v* (1 - e^K), v* = (1 + e) h a / (1 - e)
    alpha, m, gamma, dt, K = 0.8, 2.0, 1.3, 0.02, 150
    s = one_particle()
    director(s, 0.0)
    s.set("mass", [m])
    s.self_propulsion(alpha)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_langevin(gamma=gamma, temp=(0, 0.0, 1.0, 1))
    s.run(K - 1)
    h, e, a = 0.5 * dt, math.exp(-gamma * dt), alpha / m
    vs = (1.0 + e) * h * a / (1.0 - e)
    vK = vs * (1.0 - e ** K)
    sum_v = vs * (K - (1.0 - e ** K) / (1.0 - e))  # sum of v_k, k < K
    xK = h * (1.0 + e) * (sum_v + K * h * a)
    assert s.get("vx")[0] == pytest.approx(vK, rel=1e-11)
    assert s.get("x")[0] == pytest.approx(xK, rel=1e-11)
    assert s.get("vy")[0] == 0.0


# ------------------------------------------------------------ SPV, BBK -----
def langevin_constant_force(method, gamma, m=2.0, alpha=0.8, dt=0.02, K=150):
    s = one_particle()
    director(s, 0.0)
    s.set("mass", [m])
    s.self_propulsion(alpha)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_langevin(gamma=gamma, method=method, temp=(0, 0.0, 1.0, 1))
    s.run(K - 1)
    assert s.get("vy")[0] == 0.0
    return s.get("vx")[0], s.get("x")[0]


def test_spv_constant_force_velocity_is_the_exact_exponential():
    # integrator_langevin.cpp:126-182: v' = c v + eta a with c = exp(-gamma dt),
    # eta = (1 - c) / gamma, so v_k = (a / gamma)(1 - c^k) is the continuous
    # solution at t = k dt; x' = x + h (v + v') with h = dt / 2
    gamma, m, alpha, dt, K = 0.7, 2.0, 0.8, 0.02, 150
    v, x = langevin_constant_force("spv", gamma, m, alpha, dt, K)
    a, c, h = alpha / m, math.exp(-gamma * dt), 0.5 * dt
    vK = a / gamma * (1.0 - c ** K)
    S = a / gamma * ((K + 1) - (1.0 - c ** (K + 1)) / (1.0 - c))  # sum of v_k, k = 0..K
    assert v == pytest.approx(a / gamma * (1.0 - math.exp(-gamma * K * dt)), rel=1e-12)
    assert v == pytest.approx(vK, rel=1e-12)
    assert x == pytest.approx(h * (2.0 * S - vK), rel=1e-11)


def test_spv_without_friction_ignores_the_force():
    # the reference sets eta = 0 at gamma = 0 (integrator_langevin.cpp:132-133),
    # so a constant force never reaches the velocity
    v, x = langevin_constant_force("spv", 0.0)
    assert v == 0.0 and x == 0.0


@pytest.mark.parametrize("gamma", [0.7, 0.0])
def test_bbk_constant_force_geometric_velocity(gamma):
    # integrator_langevin.cpp:192-258: v_half = (1 - g) v + h a, x' = x + dt v_half,
    # v' = (v_half + h a) / (1 + g) with g = gamma dt / 2, h = dt / 2:
    # v' = r v + dt a / (1 + g), r = (1 - g) / (1 + g)
    m, alpha, dt, K = 2.0, 0.8, 0.02, 150
    v, x = langevin_constant_force("bbk", gamma, m, alpha, dt, K)
    a, h, g = alpha / m, 0.5 * dt, 0.5 * gamma * dt
    if gamma == 0.0:
        t = K * dt
        assert v == pytest.approx(a * t, rel=1e-12)
        assert x == pytest.approx(0.5 * a * t * t, rel=1e-11)
        return
    r = (1.0 - g) / (1.0 + g)
    vs = a / gamma
    vK = vs * (1.0 - r ** K)
    sum_v = vs * (K - (1.0 - r ** K) / (1.0 - r))  # sum of v_k, k < K
    assert v == pytest.approx(vK, rel=1e-12)
    assert x == pytest.approx(dt * ((1.0 - g) * sum_v + K * h * a), rel=1e-11)


def soft_pair_exit_speed(dt, m):
    s = oracle.OracleSystem(2, (100.0, 100.0, 10.0), True)
    s.set("x", [-0.7, 0.7])
    s.set("mass", [m, m])
    s.nlist(2.4, 0.3)
    s.pair_soft(1.0, 2.0)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_langevin(gamma=0.0, temp=(0, 0.0, 1.0, 1))
    s.run(int(round(4.0 / dt)))  # well past separation at r = a
    assert s.get("x")[1] - s.get("x")[0] > 2.0
    return s.get("vx")[1], -s.get("vx")[0]


@pytest.mark.parametrize("m", [2.0, 1.0])
def test_baoab_soft_pair_conserves_energy(m):
    # U0 = k (a - r0)^2 / 2 becomes the kinetic energy 2 (m v^2 / 2) of the pair
    exact = 0.6 * math.sqrt(1.0 / (2.0 * m))
    for dt in (0.004, 0.002):
        v1, v0 = soft_pair_exit_speed(dt, m)
        assert v1 == v0  # Newton's third law, symmetric start
        # velocity Verlet's energy error is O(dt^2) (the force's kink at r = a
        # makes its constant depend on where the exit falls inside a step)
        assert abs(v1 - exact) < 0.05 * dt * dt


def test_baoab_equipartition_with_mass():
    n, m, T = 6000, 3.0, 1.4
    s = oracle.OracleSystem(n, (600.0, 600.0, 10.0), True)
    s.set("mass", np.full(n, m))
    s.constraint("plane")
    s.set_timestep(0.01)
    s.integrator_langevin(gamma=2.0, seed=5, temp=(0, T, 1.0, 1))
    s.run(300)
    mv2 = m * np.concatenate([s.get("vx"), s.get("vy")]) ** 2
    assert abs(mv2.mean() - T) < 5 * T * math.sqrt(2.0 / mv2.size) + 0.02


# ------------------------------------------------------------ plane ---------
def test_plane_rotation_sign():
    rng = np.random.default_rng(4)
    n = 400
    s = oracle.OracleSystem(n, (100.0, 100.0, 10.0), True)
    th = rng.uniform(0, 2 * math.pi, n)
    s.set("nx", np.cos(th))
    s.set("ny", np.sin(th))
    s.constraint("plane")
    phi = rng.uniform(-3.0, 3.0, n)
    s.rotate_directors(phi)
    nx, ny = s.get("nx"), s.get("ny")
    cross = np.cos(th) * ny - np.sin(th) * nx
    dot = np.cos(th) * nx + np.sin(th) * ny
    assert np.abs(cross - np.sin(phi)).max() < 1e-14
    assert np.abs(dot - np.cos(phi)).max() < 1e-14


# ------------------------------------------------------------ Brownian ------
def test_brownian_drift_with_mu_and_self_propulsion():
    v0, mu, alpha, th, dt, K = 0.2, 0.5, 0.6, 2.1, 0.01, 100
    s = one_particle()
    director(s, th)
    s.self_propulsion(alpha)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_brownian(v0=v0, nu=0.0, mu=mu, mur=0.7)
    s.run(K - 1)
    t = K * dt
    assert s.get("x")[0] == pytest.approx((v0 + mu * alpha) * t * math.cos(th), rel=1e-12)
    assert s.get("y")[0] == pytest.approx((v0 + mu * alpha) * t * math.sin(th), rel=1e-12)
    assert s.get("vx")[0] == pytest.approx((v0 + mu * alpha) * math.cos(th), rel=1e-14)


def test_brownian_pos_drift_is_mu_times_force():
    mu, alpha, th, dt, K = 0.35, 1.2, -0.4, 0.02, 50
    s = one_particle()
    director(s, th)
    s.self_propulsion(alpha)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_brownian(which="brownian_pos", mu=mu)
    s.run(K - 1)
    assert s.get("x")[0] == pytest.approx(mu * alpha * K * dt * math.cos(th), rel=1e-12)


@pytest.mark.parametrize("mur", [0.5, 1.7])
def test_brownian_mur_turns_aligned_pair(mur):
    J, dt, th0, th1 = 0.8, 0.05, 0.3, 1.4
    s = oracle.OracleSystem(2, (50.0, 50.0, 10.0), True)
    s.set("x", [0.0, 1.5])
    s.set("nx", [math.cos(th0), math.cos(th1)])
    s.set("ny", [math.sin(th0), math.sin(th1)])
    s.nlist(2.4, 0.3)
    s.align_polar(J, 2.4)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_brownian(v0=0.0, nu=0.0, mu=1.0, mur=mur)
    s.run(0)  # one loop head (torques) and one integrator call
    got = np.arctan2(s.get("ny"), s.get("nx"))
    d = dt * mur * J * math.sin(th1 - th0)
    assert got[0] == pytest.approx(th0 + d, abs=1e-14)
    assert got[1] == pytest.approx(th1 - d, abs=1e-14)
    assert s.get("x")[1] == 1.5  # no force, no drift


def test_brownian_positional_variance_with_mu():
    n, mu, T, dt, K = 40000, 0.5, 0.8, 0.01, 20
    s = oracle.OracleSystem(n, (1e6, 1e6, 10.0), True)
    s.set("nx", np.ones(n))
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_brownian(v0=0.0, nu=0.0, mu=mu, seed=3, temp=(0, T, T, 1))
    s.run(K - 1)
    var = np.concatenate([s.get("x"), s.get("y")]).var()
    expect = 2.0 * mu * T * dt * K
    assert abs(var - expect) < 5 * expect * math.sqrt(2.0 / (2 * n))


# ------------------------------------------------------------ FIRE ----------
def lj_pair(r0, dt, **fire):
    s = oracle.OracleSystem(2, (40.0, 40.0, 10.0), True)
    s.set("x", [-0.5 * r0, 0.5 * r0])
    s.nlist(2.5, 0.3)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(dt)
    s.integrator_fire(**fire)
    return s


def test_fire_first_step_of_lj_pair():
    r0, dt = 1.3, 0.01
    s = lj_pair(r0, dt, f_tol=0.0, e_tol=0.0)
    s.run(0)  # loop-head forces F(r0), then one FIRE step
    f0 = lj_force(r0)  # on the right particle along +x
    v_half = 0.5 * dt * f0
    r1 = r0 + 2.0 * dt * v_half
    f1 = lj_force(r1)
    v = v_half + 0.5 * dt * f1  # same direction as f1: the mixing keeps it
    assert s.get("x")[1] == pytest.approx(0.5 * r1, rel=1e-13)
    assert s.get("vx")[1] == pytest.approx(v, rel=1e-11)
    assert s.get("vx")[0] == pytest.approx(-v, rel=1e-11)
    st = s.fire_state()
    assert st["P"] == pytest.approx(2.0 * v * f1, rel=1e-11)
    assert st["Fnorm"] == pytest.approx(math.sqrt(2.0) * abs(f1), rel=1e-11)
    assert st["dt"] == dt and st["last_neg"] == 1 and st["alpha"] == 0.1


def test_fire_dt_and_alpha_schedule():
    dt0, dt_max, a0 = 0.01, 0.05, 0.1
    s = lj_pair(1.3, dt0, alpha=a0, alpha_scale=0.99, f_tol=0.0, e_tol=0.0, dt_inc=1.1, dt_dec=0.5,
                dt_max=dt_max, min_neg_steps=5)
    s.compute_forces()
    dt, alpha, last = dt0, a0, 0
    seen = {"grow": 0, "cap": 0, "reset": 0, "wait": 0}
    for it in range(400):
        s.integrate()
        st = s.fire_state()
        if st["P"] > 0.0:
            last += 1
            if last > 5:
                new = dt * 1.1
                seen["cap" if new > dt_max else "grow"] += 1
                dt, alpha = min(new, dt_max), alpha * 0.99
            else:
                seen["wait"] += 1
        else:
            dt, alpha, last = dt * 0.5, a0, 0
            seen["reset"] += 1
            assert np.all(s.get("vx") == 0.0) and np.all(s.get("vy") == 0.0)
        assert st["dt"] == dt, it
        assert st["alpha"] == alpha, it
        assert st["last_neg"] == last, it
    assert all(v > 0 for v in seen.values()), seen
