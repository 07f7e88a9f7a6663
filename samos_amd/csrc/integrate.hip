// Integrator steps with the surface constraints fused in, FP64 throughout.
//
// Restates, per particle:
//   Brownian         src/integrators/integrator_brownian.cpp:42-124
//   Brownian pos     src/integrators/integrator_brownian_pos.cpp:37-84
//   Brownian align   src/integrators/integrator_brownian_align.cpp:41-76
//   Langevin BAOAB   src/integrators/integrator_langevin.cpp:54-118
//   Langevin SPV     src/integrators/integrator_langevin.cpp:128-183
//   Langevin BBK     src/integrators/integrator_langevin.cpp:193-253
//   FIRE             src/integrators/integrator_fire.cpp:34-148 (per-particle parts)
//   sphere           src/constraints/constraint_sphere.cpp:38-81
//   plane            src/constraints/constraint_plane.cpp:36-160
//   none             src/constraints/constraint_none.hpp:58-80
//   rotate/project   src/constraints/constraint.cpp:96-173
//   periodic wrap    src/system/system.cpp:1119-1155
// The reference's per-particle virtual calls (enforce, project_torque,
// rotate_director) become inline device functions.  The noise stream is the
// counter-based Philox stream of samos_device.hpp, one independent draw block
// per (particle id, integrator call).
#include <algorithm>

#include "kernels.hpp"
#include "pair_sum.hpp"
#include "constraint.hpp"

namespace samos {

namespace {

__device__ __forceinline__ void load_ps(const Fields& f, int s, PS& p) {
  p.x = f.x[s]; p.y = f.y[s]; p.z = f.z[s];
  p.vx = f.vx[s]; p.vy = f.vy[s]; p.vz = f.vz[s];
  p.nx = f.nx[s]; p.ny = f.ny[s]; p.nz = f.nz[s];
  p.dix = p.diy = p.diz = 0;
}

__device__ __forceinline__ void store_ps(const Fields& f, int s, const PS& p) {
  f.x[s] = p.x; f.y[s] = p.y; f.z[s] = p.z;
  f.vx[s] = p.vx; f.vy[s] = p.vy; f.vz[s] = p.vz;
  f.nx[s] = p.nx; f.ny[s] = p.ny; f.nz[s] = p.nz;
  if (p.dix) f.ix[s] += p.dix;
  if (p.diy) f.iy[s] += p.diy;
  if (p.diz) f.iz[s] += p.diz;
}


__device__ __forceinline__ bool differs(double a, double b) { return __double_as_longlong(a) != __double_as_longlong(b); }

// Constraint::enforce over every particle.  Only the values the enforce
// changed are stored (bit comparison), and with f.xy set the interleaved
// (x, y) column is written as well (the planar step's layout).
__global__ void k_enforce(Fields f, BoxD b, ConstraintD c, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  PS p; load_ps(f, s, p);
  const PS q = p;
  if (enforce(b, c, p)) f.fz[s] = 0.0;
  if (differs(p.x, q.x)) f.x[s] = p.x;
  if (differs(p.y, q.y)) f.y[s] = p.y;
  if (differs(p.z, q.z)) f.z[s] = p.z;
  if (differs(p.vx, q.vx)) f.vx[s] = p.vx;
  if (differs(p.vy, q.vy)) f.vy[s] = p.vy;
  if (differs(p.vz, q.vz)) f.vz[s] = p.vz;
  if (differs(p.nx, q.nx)) f.nx[s] = p.nx;
  if (differs(p.ny, q.ny)) f.ny[s] = p.ny;
  if (differs(p.nz, q.nz)) f.nz[s] = p.nz;
  if (p.dix) f.ix[s] += p.dix;
  if (p.diy) f.iy[s] += p.diy;
  if (p.diz) f.iz[s] += p.diz;
  if (f.xy) f.xy[s] = make_double2(p.x, p.y);
}

// which: 0 brownian, 1 brownian_pos, 2 brownian_align
template <int WHICH>
__global__ void k_brownian(Fields f, BoxD b, ConstraintD c, int n, double dt, double v0, double mu, double mur,
                           double stoch, double flip_p, int nematic, int velocity, TempD temp, uint32_t seed,
                           unsigned long long call0, const ClockD* clk, int slot, int t_off, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  const unsigned long long call = eff_call(clk, slot, call0);
  const uint32_t pid = (uint32_t)f.id[s];
  const uint32_t tag = WHICH == 0 ? TAG_BROWNIAN : (WHICH == 1 ? TAG_BROWNIAN_POS : TAG_BROWNIAN_ALIGN);
  PS p; load_ps(f, s, p);
  bool fz0 = false;
  double g_rot = 0.0, u_flip = 1.0;
  if (WHICH != 1) gauss1_u1(seed, tag, pid, call, g_rot, u_flip);
  if (WHICH != 1 && nematic && u_flip < flip_p) {
    p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz;
    if (WHICH == 0 && velocity) { p.vx = -p.vx; p.vy = -p.vy; p.vz = -p.vz; }
  }
  if (WHICH != 2) {
    const double T = temp.at(eff_step(clk, t_off));
    const double B = sqrt(2.0 * mu * T);
    const double sqrt_dt = sqrt(dt);
    const double fx = f.fx[s], fy = f.fy[s], fz = f.fz[s];
    double fdx, fdy, fdz;
    if (WHICH == 0) { fdx = v0 * p.nx + mu * fx; fdy = v0 * p.ny + mu * fy; fdz = v0 * p.nz + mu * fz; }
    else { fdx = mu * fx; fdy = mu * fy; fdz = mu * fz; }
    p.vx = fdx; p.vy = fdy; p.vz = fdz;
    p.x += dt * fdx; p.y += dt * fdy; p.z += dt * fdz;
    if (T > 0.0) {
      double g[3]; gauss3(seed, tag, pid, call, g);
      const double frx = B * g[0], fry = B * g[1], frz = B * g[2];
      p.vx += frx; p.vy += fry; p.vz += frz;
      p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
    }
    fz0 = enforce(b, c, p);
  }
  if (WHICH != 1) {
    const double om = mur * project_torque(c, p, f.tx[s], f.ty[s], f.tz[s]);
    f.omega[s] = om;
    const double dtheta = dt * om + stoch * g_rot;
    rotate_director(c, p, dtheta);
    if (WHICH == 0 && velocity) rotate_velocity(c, p, dtheta);
  }
  if (WHICH != 2) f.age[s] += dt;
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// First per-particle pass of a Langevin step (before the internal force call).
template <int METHOD>
__global__ void k_langevin_a(Fields f, BoxD b, ConstraintD c, int n, double dt, double gamma, TempD temp,
                             uint32_t seed, unsigned long long call0, const ClockD* clk, int slot, int t_off,
                             const double* R, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  PS p; load_ps(f, s, p);
  const double dt2 = 0.5 * dt;
  const double m = f.mass[s];
  bool fz0 = false;
  if (METHOD == 0) {
    const double T = temp.at(eff_step(clk, t_off));
    const double B = sqrt(T * (1.0 - exp(-2.0 * gamma * dt)));
    const double exp_dt = exp(-gamma * dt);
    const double fact = dt2 / m;
    p.vx += fact * f.fx[s]; p.vy += fact * f.fy[s]; p.vz += fact * f.fz[s];
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    fz0 |= enforce(b, c, p);
    p.vx *= exp_dt; p.vy *= exp_dt; p.vz *= exp_dt;
    if (B != 0.0) {
      const double sf = B / sqrt(m);
      double g[3]; gauss3(seed, TAG_LANGEVIN, (uint32_t)f.id[s], eff_call(clk, slot, call0), g);
      p.vx += sf * g[0]; p.vy += sf * g[1]; p.vz += sf * g[2];
    }
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    fz0 |= enforce(b, c, p);
  } else if (METHOD == 1) {
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    fz0 |= enforce(b, c, p);
  } else {
    const double T = temp.at(eff_step(clk, t_off));
    const double B = sqrt(2.0 * T * dt * gamma);
    const double one_m_dt2 = 1.0 - gamma * dt2;
    const double dm = dt2 / m;
    p.vx = one_m_dt2 * p.vx + dm * f.fx[s];
    p.vy = one_m_dt2 * p.vy + dm * f.fy[s];
    p.vz = one_m_dt2 * p.vz + dm * f.fz[s];
    if (B != 0.0) {
      const double sf = 0.5 * B / sqrt(m);
      const int id = f.id[s];
      p.vx += sf * R[3 * id]; p.vy += sf * R[3 * id + 1]; p.vz += sf * R[3 * id + 2];
    }
    p.x += dt * p.vx; p.y += dt * p.vy; p.z += dt * p.vz;
    fz0 |= enforce(b, c, p);
  }
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// Second per-particle pass (after the internal force call).
template <int METHOD>
__global__ void k_langevin_b(Fields f, int n, double dt, double gamma, TempD temp, uint32_t seed,
                             unsigned long long call0, const ClockD* clk, int slot, int t_off, double* R, BoxD b,
                             ConstraintD c, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  const double dt2 = 0.5 * dt;
  const double m = f.mass[s];
  if (METHOD == 0) {
    const double fact = dt2 / m;
    f.vx[s] += fact * f.fx[s]; f.vy[s] += fact * f.fy[s]; f.vz[s] += fact * f.fz[s];
    f.age[s] += dt;
  } else if (METHOD == 1) {
    const double T = temp.at(eff_step(clk, t_off));
    const double zeta = sqrt(T * (1.0 - exp(-2.0 * gamma * dt)));
    const double eta = (gamma == 0.0) ? 0.0 : (1.0 - exp(-dt * gamma)) / gamma;
    const double exp_dt = exp(-dt * gamma);
    PS p; load_ps(f, s, p);
    const double fact = eta / m;
    p.vx = exp_dt * p.vx + fact * f.fx[s];
    p.vy = exp_dt * p.vy + fact * f.fy[s];
    p.vz = exp_dt * p.vz + fact * f.fz[s];
    if (zeta != 0.0) {
      const double sf = zeta / sqrt(m);
      double g[3]; gauss3(seed, TAG_LANGEVIN, (uint32_t)f.id[s], eff_call(clk, slot, call0), g);
      p.vx += sf * g[0]; p.vy += sf * g[1]; p.vz += sf * g[2];
    }
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    if (enforce(b, c, p)) f.fz[s] = 0.0;
    store_ps(f, s, p);
    f.age[s] += dt;
  } else {
    const double T = temp.at(eff_step(clk, t_off));
    const double B = sqrt(2.0 * T * dt * gamma);
    const double one_div = 1.0 / (1.0 + gamma * dt2);
    const double dm = dt2 / m;
    double vx = one_div * (f.vx[s] + dm * f.fx[s]);
    double vy = one_div * (f.vy[s] + dm * f.fy[s]);
    double vz = one_div * (f.vz[s] + dm * f.fz[s]);
    if (B != 0.0) {
      const double sf = 0.5 * B * one_div / sqrt(m);
      const int id = f.id[s];
      double g[3]; gauss3(seed, TAG_LANGEVIN, (uint32_t)id, eff_call(clk, slot, call0), g);
      R[3 * id] = g[0]; R[3 * id + 1] = g[1]; R[3 * id + 2] = g[2];
      vx += sf * g[0]; vy += sf * g[1]; vz += sf * g[2];
      f.age[s] += dt;
    }
    f.vx[s] = vx; f.vy[s] = vy; f.vz[s] = vz;
  }
}

__global__ void k_bbk_init(Fields f, int n, uint32_t seed, double* R) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int id = f.id[s];
  double g[3]; gauss3(seed, TAG_BBK_INIT, (uint32_t)id, 0ULL, g);
  R[3 * id] = g[0]; R[3 * id + 1] = g[1]; R[3 * id + 2] = g[2];
}

// FIRE first pass: half kick, drift, enforce, director half kick and rotation.
__global__ void k_fire_a(Fields f, BoxD b, ConstraintD c, int n, const FireD* fs, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  const double dt = fs->dt;
  PS p; load_ps(f, s, p);
  const double dt_2 = 0.5 * dt;
  p.vx += dt_2 * f.fx[s]; p.vy += dt_2 * f.fy[s]; p.vz += dt_2 * f.fz[s];
  p.x += dt * p.vx; p.y += dt * p.vy; p.z += dt * p.vz;
  const bool fz0 = enforce(b, c, p);
  double om = f.omega[s] + dt_2 * project_torque(c, p, f.tx[s], f.ty[s], f.tz[s]);
  f.omega[s] = om;
  rotate_director(c, p, dt * om);
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// IntegratorNVE::integrate (src/integrators/integrator_nve.cpp:37-141), first
// part: the reference's first three loops touch one particle each, so they run
// fused per particle: half kick, enforce, omega half kick; the displacement
// (limited to `limit` when lim_on); enforce; the director rotation by dt omega
// (limited to `theta` when theta_on).  limit < 0 / theta < 0: not given.
__device__ __forceinline__ double sign_of(double x) { return x >= 0.0 ? 1.0 : -1.0; }

__global__ void k_nve_a(Fields f, BoxD b, ConstraintD c, int n, double dt, double limit, double theta, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  PS p; load_ps(f, s, p);
  const double dt_2 = 0.5 * dt;
  p.vx += dt_2 * f.fx[s]; p.vy += dt_2 * f.fy[s]; p.vz += dt_2 * f.fz[s];
  bool fz0 = enforce(b, c, p);
  double om = f.omega[s] + dt_2 * project_torque(c, p, f.tx[s], f.ty[s], f.tz[s]);
  double dx = dt * p.vx, dy = dt * p.vy, dz = dt * p.vz;
  if (limit >= 0.0) {
    const double dr = sqrt(dx * dx + dy * dy + dz * dz);
    if (dr > limit) {
      dx = limit * dx / dr;
      dy = limit * dy / dr;
      dz = limit * dz / dr;
    }
  }
  p.x += dx; p.y += dy; p.z += dz;
  fz0 |= enforce(b, c, p);
  double dtheta = dt * om;
  if (theta >= 0.0 && fabs(dtheta) > theta) dtheta = sign_of(dtheta) * theta;
  rotate_director(c, p, dtheta);
  f.omega[s] = om;
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// IntegratorNVE::integrate, last loop (after the force and torque evaluation):
// half kick, speed limit, enforce, omega half kick and its limit (applied when
// both limits are given; the reference reads an unset member otherwise), age.
__global__ void k_nve_b(Fields f, BoxD b, ConstraintD c, int n, double dt, double limit, double theta, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit)) return;
  PS p; load_ps(f, s, p);
  const double dt_2 = 0.5 * dt;
  p.vx += dt_2 * f.fx[s]; p.vy += dt_2 * f.fy[s]; p.vz += dt_2 * f.fz[s];
  if (limit >= 0.0) {
    const double v = sqrt(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
    if (v * dt > limit) {
      p.vx = p.vx / v * limit / dt;
      p.vy = p.vy / v * limit / dt;
      p.vz = p.vz / v * limit / dt;
    }
  }
  const bool fz0 = enforce(b, c, p);
  double om = f.omega[s] + dt_2 * project_torque(c, p, f.tx[s], f.ty[s], f.tz[s]);
  if (limit >= 0.0 && theta >= 0.0 && fabs(om) * dt > theta) om = sign_of(om) * theta / dt;
  f.omega[s] = om;
  f.age[s] += dt;
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// FIRE second pass: half kick, per-slot P = v.F, |F|^2, |v|^2 (before the
// enforce, as in the reference), enforce, director half kick.
__global__ void k_fire_b(Fields f, BoxD b, ConstraintD c, int n, const FireD* fs, double* cols, long ld,
                         unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double dt = fs->dt;
  if (!in_group(f, s, gbit)) {
    cols[s] = 0.0; cols[ld + s] = 0.0; cols[2 * ld + s] = 0.0;
    return;
  }
  PS p; load_ps(f, s, p);
  const double dt_2 = 0.5 * dt;
  const double fx = f.fx[s], fy = f.fy[s], fz = f.fz[s];
  p.vx += dt_2 * fx; p.vy += dt_2 * fy; p.vz += dt_2 * fz;
  cols[s] = p.vx * fx + p.vy * fy + p.vz * fz;
  cols[ld + s] = fx * fx + fy * fy + fz * fz;
  cols[2 * ld + s] = p.vx * p.vx + p.vy * p.vy + p.vz * p.vz;
  const bool fz0 = enforce(b, c, p);
  f.omega[s] += dt_2 * project_torque(c, p, f.tx[s], f.ty[s], f.tz[s]);
  if (fz0) f.fz[s] = 0.0;
  store_ps(f, s, p);
}

// IntegratorFIRE::integrate, scalar part (src/integrators/integrator_fire.cpp:97-147),
// in one thread with the reference's operation order.
__global__ void k_fire_update(FireD* fs, const double* red, int has_extra) {
  if (threadIdx.x != 0) return;
  FireD q = *fs;
  const double P = red[0];
  const double Fn = sqrt(red[1]);
  const double Vn = sqrt(red[2]);
  double E = 0.5 * red[3] + 0.5 * red[4];
  if (has_extra) E += 0.5 * red[5];
  const double sqrt_ndof = sqrt((double)q.ndof3);
  q.last[0] = P; q.last[1] = Fn; q.last[2] = Vn; q.last[3] = E;
  if (Fn / sqrt_ndof < q.F_tol && fabs(E - q.old_energy) < q.E_tol) {
    q.converged = 1;
    q.mix = 0;
    *fs = q;
    return;
  }
  const double inv_F = 1.0 / Fn;
  q.f1 = 1.0 - q.alpha;
  q.f2 = q.alpha * inv_F * Vn;
  if (P > 0.0) {
    q.last_neg++;
    if (q.last_neg > q.N_min) { q.dt = fmin(q.dt * q.f_inc, q.dt_max); q.alpha *= q.f_alpha; }
    q.zero_v = 0;
  } else {
    q.dt *= q.f_dec;
    q.alpha = q.alpha_init;
    q.last_neg = 0;
    q.zero_v = 1;
  }
  q.mix = 1;
  q.converged = 0;
  q.old_energy = E;
  *fs = q;
}

__global__ void k_fire_mix(Fields f, int n, const FireD* fs, unsigned gbit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !in_group(f, s, gbit) || !fs->mix) return;
  const double f1 = fs->f1, f2 = fs->f2;
  const int zero_v = fs->zero_v;
  double vx = f1 * f.vx[s] + f2 * f.fx[s];
  double vy = f1 * f.vy[s] + f2 * f.fy[s];
  double vz = f1 * f.vz[s] + f2 * f.fz[s];
  if (zero_v) { vx = 0.0; vy = 0.0; vz = 0.0; }
  f.vx[s] = vx; f.vy[s] = vy; f.vz[s] = vz;
}

__global__ void k_clock_advance(ClockD* clk, int nsteps) {
  if (threadIdx.x < 8) clk->calls[threadIdx.x] += (unsigned long long)nsteps;
  if (threadIdx.x == 0) clk->run_step += nsteps;
}

// ---------------------------------------------------------------------------
// Fused Langevin BAOAB step (pair forces + the whole per-particle update).
// ---------------------------------------------------------------------------
// exp_dt = exp(-gamma dt) and, for a constant temperature, Bc = sqrt(T (1 -
// exp(-2 gamma dt))) come from the host (the same libm the reference uses);
// Bc < 0 selects the per-step schedule value instead.
template <int POT, int PHASE, bool F32, bool PLANAR, bool MT, int BATCH>
__global__ void __launch_bounds__(kBlock, (PLANAR && !MT) ? 7 : 1) k_baoab_fused(Fields cur, Fields nxt, NListD nl, BoxD b, ConstraintD c,
                                                        InteractD ip, int n, FusedArgs fa) {
  const double dt = fa.dt, gamma = fa.gamma, mass_u = fa.mass_u, exp_dt = fa.exp_dt, Bc = fa.Bc;
  const TempD temp = fa.temp;
  const uint32_t seed = fa.seed;
  const ClockD* clk = fa.clock;
  const int slot = fa.clock_slot, t_off = fa.t_off, parity = fa.parity;
  int head_enforce = fa.head_enforce;
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  ChunkCtl* const ck = fa.chunk;
  if (ck) {
    // Every wavefront reads the same words, written by earlier kernels only:
    // the whole grid either returns here or runs the step.  A list that
    // overflowed at a rebuild stops every later slot before any step reads it.
    const int j = fa.slot;
    const SlotGate g = slot_gate(ck, nl.flags, j);
    if (g.stop) {
      // the first test that holds decides, in this order
      if (g.stop & kGateOverflow) return;  // the list overflowed at a rebuild: no step may read it
      if (g.stop & kGateResort) {  // a storage re-sort is due before this slot
        if (blockIdx.x == 0 && threadIdx.x == 0) ck->stop |= 2;
        return;
      }
      if (g.stop & kGateHalt) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ck->halt = 1;
        return;
      }
      return;  // past the range
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ck->last = j + 1;
    head_enforce = g.head;
  } else if (blockIdx.x == 0 && threadIdx.x == 0 && PHASE != 2) {
    nl.flags[1 - parity] = 0;
  }
  bool need = false;
  if (s < n) {
    double xi, yi, zi;
    if (PLANAR) {
      const double2 p = cur.xy[s];
      xi = p.x; yi = p.y; zi = c.zpos;
    } else {
      xi = cur.x[s]; yi = cur.y[s]; zi = cur.z[s];
    }
    const int ti = MT ? cur.type[s] - 1 : 0;
    double ri = 0.0;
    if (((POT & POT_SOFT) && ip.soft_radii) || ((POT & POT_LJ) && ip.lj_radii)) ri = cur.radius[s];
    // outside the integrator's group a slot is not integrated; it only takes
    // the next loop head's enforce
    const bool member = in_group(cur, s, fa.gbit);
    // The O step's Gaussian draw does not depend on the forces: it is made
    // before the pair sum, so its arithmetic overlaps the neighbour gathers.
    double B = 0.0, gn0 = 0.0, gn1 = 0.0, gn2 = 0.0;
    if (PHASE <= 1) {
      B = Bc;
      if (B < 0.0) {
        const double T = temp.at(eff_step(clk, t_off));
        B = sqrt(T * (1.0 - exp(-2.0 * gamma * dt)));
      }
      if (member && B != 0.0) {
        const unsigned long long call = clk ? clk->calls[slot] + (unsigned long long)t_off : fa.call0;
        double g[3];
        if (PLANAR) {  // the z draw only reaches v_z, which the next enforce resets
          gauss2(seed, TAG_LANGEVIN, (uint32_t)cur.id[s], call, g);
          g[2] = 0.0;
        } else {
          gauss3(seed, TAG_LANGEVIN, (uint32_t)cur.id[s], call, g);
        }
        gn0 = g[0]; gn1 = g[1]; gn2 = g[2];
      }
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    pair_force_sum<POT, F32, PLANAR, MT, BATCH>(cur, nl, b, ip, s, xi, yi, zi, ti, ri, fx, fy, fz);
    if (ip.has_sp) { fx += ip.sp_alpha * cur.nx[s]; fy += ip.sp_alpha * cur.ny[s]; fz += ip.sp_alpha * cur.nz[s]; }

    // PLANAR keeps z = zpos, v_z = n_z = 0 in storage (every enforce resets
    // them and planar pair forces have f_z = 0), so those columns are not
    // streamed.
    const double dt2 = 0.5 * dt;
    const double m = mass_u > 0.0 ? mass_u : cur.mass[s];
    const double fact = mass_u > 0.0 ? fa.fact_u : dt2 / m;
    double vx = cur.vx[s], vy = cur.vy[s], vz = PLANAR ? 0.0 : cur.vz[s];
    if (PHASE >= 1 && member) {  // closing B of step t-1
      vx += fact * fx; vy += fact * fy; vz += fact * fz;
      if (!(fa.age && fa.age->uniform)) cur.age[s] += dt;
    }
    if (PHASE <= 1) {
      PS p;
      p.x = xi; p.y = yi; p.z = zi; p.vx = vx; p.vy = vy; p.vz = vz;
      // On a plane only the enforce's normalisation changes a director, and
      // the first step of every range (phase 0) has normalised them all; the
      // later steps of the range neither read nor write them.
      constexpr bool kDirs = !(PLANAR && PHASE == 1);
      const double nx0 = kDirs ? cur.nx[s] : 0.0, ny0 = kDirs ? cur.ny[s] : 0.0;
      const double nz0 = PLANAR ? 0.0 : cur.nz[s];
      p.nx = nx0; p.ny = ny0; p.nz = nz0;
      p.dix = p.diy = p.diz = 0;
      constexpr int kSet = PLANAR ? kSurfPlane : kSurfAll;
      bool fz0 = false;
      if (member) {
        if (!PLANAR && PHASE == 1) {
          // The loop-head enforce of step t acts on the velocity after step t-1's
          // closing kick.  The positions already carry it (applied when written);
          // the velocity projection is repeated here on the kicked velocity.  On a
          // plane the kick has no normal component, so this is skipped there.
          PS q = p;
          enforce<kSet>(b, c, q);
          p.vx = q.vx; p.vy = q.vy; p.vz = q.vz;
        }
        p.vx += fact * fx; p.vy += fact * fy; p.vz += fact * fz;
        p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
        fz0 = enforce<kSet>(b, c, p, kDirs);
        p.vx *= exp_dt; p.vy *= exp_dt; p.vz *= exp_dt;
        if (B != 0.0) {
          const double sf = fa.sf_u >= 0.0 ? fa.sf_u : B / sqrt(m);
          p.vx += sf * gn0; p.vy += sf * gn1; p.vz += sf * gn2;
        }
        p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
        // on a plane the director was normalised by the first enforce
        fz0 |= enforce<kSet>(b, c, p, !PLANAR);
      }
      if (head_enforce) fz0 |= enforce<kSet>(b, c, p, !PLANAR || (kDirs && !member));
      if (PLANAR) {
        nxt.xy[s] = make_double2(p.x, p.y);
      } else {
        nxt.x[s] = p.x; nxt.y[s] = p.y; nxt.z[s] = p.z;
      }
      cur.vx[s] = p.vx; cur.vy[s] = p.vy;
      if (!PLANAR) cur.vz[s] = p.vz;
      // the enforce renormalises the director; store only what changed
      if (kDirs && (p.nx != nx0 || p.ny != ny0)) { cur.nx[s] = p.nx; cur.ny[s] = p.ny; }
      if (!PLANAR && p.nz != nz0) cur.nz[s] = p.nz;
      if (p.dix) cur.ix[s] += p.dix;
      if (p.diy) cur.iy[s] += p.diy;
      if (p.diz) cur.iz[s] += p.diz;
      double ddx = nl.ox[s] - p.x, ddy = nl.oy[s] - p.y, ddz = PLANAR ? 0.0 : nl.oz[s] - p.z;
      min_image(b, ddx, ddy, ddz);
      need = !(ddx * ddx + ddy * ddy + ddz * ddz < nl.pad2q);
    } else {
      // closing phase: forces F(x_t) are the state the reference leaves behind;
      // a planar run's positions go back to the x, y columns here
      if (PLANAR && fa.out_x) { fa.out_x[s] = xi; fa.out_y[s] = yi; }
      cur.vx[s] = vx; cur.vy[s] = vy;
      if (!PLANAR) cur.vz[s] = vz;
      cur.fx[s] = fx; cur.fy[s] = fy; cur.fz[s] = fz;
    }
  }
  if (PHASE != 2) {
    if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomic_set_once(ck ? &ck->req[fa.slot & 1] : &nl.flags[parity]);
  }
}

__global__ void __launch_bounds__(kBlock) k_chunk_settle(ChunkCtl* ck, Fields f0, Fields f1, int n, int planar,
                                                         int directors) {
  const int last = ck->last;
  if (blockIdx.x == 0 && threadIdx.x == 0) ck->rebuild = ck->halt | ck->req[0] | ck->req[1];
  if (!(last & 1)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (planar) {
      f0.xy[i] = f1.xy[i];
    } else {
      f0.x[i] = f1.x[i]; f0.y[i] = f1.y[i]; f0.z[i] = f1.z[i];
    }
    if (directors) {
      f0.nx[i] = f1.nx[i]; f0.ny[i] = f1.ny[i];
      if (!planar) f0.nz[i] = f1.nz[i];
    }
  }
}

__global__ void k_chunk_finish(ChunkCtl* ck, ClockD* clk, const int* nl_flags) {
  const int last = ck->last;
  __syncthreads();
  if (threadIdx.x < 8) clk->calls[threadIdx.x] += (unsigned long long)last;
  if (threadIdx.x == 0) {
    clk->run_step += last;
    ck->base += last;
    if (nl_flags[2]) ck->stop |= 1;
    ck->halt = 0; ck->req[0] = 0; ck->req[1] = 0; ck->last = 0; ck->rebuild = 0;
  }
}

// pending: a rebuild request left by the step before the chunks (slot 0 then
// stops at once and the chunk's rebuild serves it), or null.
__global__ void k_chunk_init(ChunkCtl* ck, int base, int limit, int steps, const int* pending, ClockD* clk, int slot,
                             unsigned long long calls, int resort_at) {
  if (threadIdx.x == 0) {
    if (clk) { clk->calls[slot] = calls; clk->run_step = base; }
    ck->halt = 0; ck->req[0] = 0; ck->req[1] = pending ? *pending : 0; ck->last = 0; ck->rebuild = 0;
    ck->stop = 0; ck->resort_at = resort_at;
    ck->base = base; ck->limit = limit; ck->steps = steps;
  }
}

template <int POT, int PHASE, bool PLANAR, bool MT, int BATCH>
void fused_dispatch(hipStream_t st, int fp32, Fields cur, Fields nxt, NListD nl, BoxD b, ConstraintD c, InteractD ip,
                    int n, const FusedArgs& fa) {
  dim3 g(grid1(n)), blk(kBlock);
  if (fp32)
    hipLaunchKernelGGL((k_baoab_fused<POT, PHASE, true, PLANAR, MT, BATCH>), g, blk, 0, st, cur, nxt, nl, b, c, ip, n, fa);
  else
    hipLaunchKernelGGL((k_baoab_fused<POT, PHASE, false, PLANAR, MT, BATCH>), g, blk, 0, st, cur, nxt, nl, b, c, ip, n, fa);
}

// The gather batch is a tuning choice for the planar single-type layout (the
// benchmarked one); the other layouts use batches of 4.
template <int POT, int PHASE>
void fused_layout(hipStream_t st, int fp32, int planar, int mt, Fields cur, Fields nxt, NListD nl, BoxD b,
                  ConstraintD c, InteractD ip, int n, const FusedArgs& fa) {
  if (planar) {
    if (mt) fused_dispatch<POT, PHASE, true, true, 4>(st, fp32, cur, nxt, nl, b, c, ip, n, fa);
    else if (fa.batch == 8) fused_dispatch<POT, PHASE, true, false, 8>(st, fp32, cur, nxt, nl, b, c, ip, n, fa);
    else fused_dispatch<POT, PHASE, true, false, 4>(st, fp32, cur, nxt, nl, b, c, ip, n, fa);
  } else {
    if (mt) fused_dispatch<POT, PHASE, false, true, 4>(st, fp32, cur, nxt, nl, b, c, ip, n, fa);
    else fused_dispatch<POT, PHASE, false, false, 4>(st, fp32, cur, nxt, nl, b, c, ip, n, fa);
  }
}

template <int POT>
void fused_phase(hipStream_t st, int phase, int fp32, int planar, int mt, Fields cur, Fields nxt, NListD nl, BoxD b,
                 ConstraintD c, InteractD ip, int n, const FusedArgs& fa) {
  if (phase == 0) fused_layout<POT, 0>(st, fp32, planar, mt, cur, nxt, nl, b, c, ip, n, fa);
  else if (phase == 1) fused_layout<POT, 1>(st, fp32, planar, mt, cur, nxt, nl, b, c, ip, n, fa);
  else fused_layout<POT, 2>(st, fp32, planar, mt, cur, nxt, nl, b, c, ip, n, fa);
}

__global__ void k_age_init(AgeCtl* ctl, const double* age, int n) {
  if (threadIdx.x == 0) {
    ctl->uniform = 1;
    ctl->age = n > 0 ? age[0] : 0.0;
  }
}

// Any age that differs from age[0] (a NaN included) clears the flag; every
// writer stores the same 0.
__global__ void k_age_check(AgeCtl* ctl, const double* age, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const bool differs = s < n && !(age[s] == ctl->age);
  if (__ballot(differs) != 0 && (threadIdx.x & 63) == 0) ctl->uniform = 0;
}

// The reference adds dt to each age once per step: the same additions, in
// order, on the common value.
__global__ void k_age_advance(AgeCtl* ctl, double dt, int count) {
  if (threadIdx.x == 0 && ctl->uniform) {
    double a = ctl->age;
    for (int k = 0; k < count; ++k) a += dt;
    ctl->age = a;
  }
}

__global__ void k_age_fill(const AgeCtl* ctl, double* age, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n && ctl->uniform) age[s] = ctl->age;
}

__global__ void k_pack_xy(Fields f, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) f.xy[s] = make_double2(f.x[s], f.y[s]);
}

__global__ void k_unpack_xy(Fields f, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) {
    const double2 p = f.xy[s];
    f.x[s] = p.x;
    f.y[s] = p.y;
  }
}

}  // namespace

void launch_enforce(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n) {
  if (n <= 0 || con.kind < 0) return;
  hipLaunchKernelGGL(k_enforce, dim3(grid1(n)), dim3(kBlock), 0, st, f, box, con, n);
}

void launch_brownian(hipStream_t st, int which, Fields f, BoxD box, ConstraintD con, int n, double dt, double v0,
                     double mu, double mur, double stoch, double flip_p, int nematic, int velocity, TempD temp,
                     uint32_t seed, unsigned long long call, const ClockD* clock, int clock_slot, int t_off,
                     unsigned gbit) {
  dim3 g(grid1(n)), blk(kBlock);
  if (which == 0)
    hipLaunchKernelGGL(k_brownian<0>, g, blk, 0, st, f, box, con, n, dt, v0, mu, mur, stoch, flip_p, nematic, velocity,
                       temp, seed, call, clock, clock_slot, t_off, gbit);
  else if (which == 1)
    hipLaunchKernelGGL(k_brownian<1>, g, blk, 0, st, f, box, con, n, dt, v0, mu, mur, stoch, flip_p, nematic, velocity,
                       temp, seed, call, clock, clock_slot, t_off, gbit);
  else
    hipLaunchKernelGGL(k_brownian<2>, g, blk, 0, st, f, box, con, n, dt, v0, mu, mur, stoch, flip_p, nematic, velocity,
                       temp, seed, call, clock, clock_slot, t_off, gbit);
}

void launch_langevin_a(hipStream_t st, int method, Fields f, BoxD box, ConstraintD con, int n, double dt, double gamma,
                       TempD temp, uint32_t seed, unsigned long long call, const ClockD* clock, int clock_slot,
                       int t_off, double* R, unsigned gbit) {
  dim3 g(grid1(n)), blk(kBlock);
  if (method == 0)
    hipLaunchKernelGGL(k_langevin_a<0>, g, blk, 0, st, f, box, con, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, gbit);
  else if (method == 1)
    hipLaunchKernelGGL(k_langevin_a<1>, g, blk, 0, st, f, box, con, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, gbit);
  else
    hipLaunchKernelGGL(k_langevin_a<2>, g, blk, 0, st, f, box, con, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, gbit);
}

void launch_langevin_b(hipStream_t st, int method, Fields f, int n, double dt, double gamma, TempD temp, uint32_t seed,
                       unsigned long long call, const ClockD* clock, int clock_slot, int t_off, double* R, BoxD box,
                       ConstraintD con, unsigned gbit) {
  dim3 g(grid1(n)), blk(kBlock);
  if (method == 0)
    hipLaunchKernelGGL(k_langevin_b<0>, g, blk, 0, st, f, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, box, con, gbit);
  else if (method == 1)
    hipLaunchKernelGGL(k_langevin_b<1>, g, blk, 0, st, f, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, box, con, gbit);
  else
    hipLaunchKernelGGL(k_langevin_b<2>, g, blk, 0, st, f, n, dt, gamma, temp, seed, call, clock, clock_slot, t_off, R, box, con, gbit);
}

void launch_bbk_init(hipStream_t st, Fields f, int n, uint32_t seed, double* R) {
  hipLaunchKernelGGL(k_bbk_init, dim3(grid1(n)), dim3(kBlock), 0, st, f, n, seed, R);
}

void launch_fire_a(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, const FireD* fs, unsigned gbit) {
  hipLaunchKernelGGL(k_fire_a, dim3(grid1(n)), dim3(kBlock), 0, st, f, box, con, n, fs, gbit);
}

void launch_fire_b(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, const FireD* fs, double* cols, long ld,
                   unsigned gbit) {
  hipLaunchKernelGGL(k_fire_b, dim3(grid1(n)), dim3(kBlock), 0, st, f, box, con, n, fs, cols, ld, gbit);
}

void launch_fire_update(hipStream_t st, FireD* fs, const double* red, int has_extra) {
  hipLaunchKernelGGL(k_fire_update, dim3(1), dim3(64), 0, st, fs, red, has_extra);
}

void launch_nve_a(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, double dt, double limit, double theta,
                  unsigned gbit) {
  hipLaunchKernelGGL(k_nve_a, dim3(grid1(n)), dim3(kBlock), 0, st, f, box, con, n, dt, limit, theta, gbit);
}

void launch_nve_b(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, double dt, double limit, double theta,
                  unsigned gbit) {
  hipLaunchKernelGGL(k_nve_b, dim3(grid1(n)), dim3(kBlock), 0, st, f, box, con, n, dt, limit, theta, gbit);
}

void launch_fire_mix(hipStream_t st, Fields f, int n, const FireD* fs, unsigned gbit) {
  hipLaunchKernelGGL(k_fire_mix, dim3(grid1(n)), dim3(kBlock), 0, st, f, n, fs, gbit);
}

void launch_chunk_settle(hipStream_t st, ChunkCtl* ck, Fields f0, Fields f1, int n, int planar, int directors) {
  const int g = std::max(1, std::min(grid1(n), 1024));
  hipLaunchKernelGGL(k_chunk_settle, dim3(g), dim3(kBlock), 0, st, ck, f0, f1, n, planar, directors);
}

void launch_chunk_finish(hipStream_t st, ChunkCtl* ck, ClockD* clock, const int* nl_flags) {
  hipLaunchKernelGGL(k_chunk_finish, dim3(1), dim3(64), 0, st, ck, clock, nl_flags);
}

void launch_chunk_init(hipStream_t st, ChunkCtl* ck, int base, int limit, int steps, int resort_at,
                       const int* pending, ClockD* clk, int slot, unsigned long long calls) {
  hipLaunchKernelGGL(k_chunk_init, dim3(1), dim3(64), 0, st, ck, base, limit, steps, pending, clk, slot, calls,
                     resort_at);
}

void launch_clock_advance(hipStream_t st, ClockD* clock, int nsteps) {
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, st, clock, nsteps);
}

void launch_baoab_fused(hipStream_t st, int pot_mask, int phase, Fields cur, Fields nxt, NListD nl, BoxD box,
                        ConstraintD con, InteractD ip, int n, const FusedArgs& fa, int fp32, int planar,
                        int multitype) {
  switch (pot_mask) {
    case 1: fused_phase<1>(st, phase, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, fa); break;
    case 2: fused_phase<2>(st, phase, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, fa); break;
    default: fused_phase<3>(st, phase, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, fa); break;
  }
}

void launch_age_check(hipStream_t st, AgeCtl* ctl, const double* age, int n) {
  hipLaunchKernelGGL(k_age_init, dim3(1), dim3(64), 0, st, ctl, age, n);
  if (n > 0) hipLaunchKernelGGL(k_age_check, dim3(grid1(n)), dim3(kBlock), 0, st, ctl, age, n);
}

void launch_age_finish(hipStream_t st, AgeCtl* ctl, double* age, int n, double dt, int count) {
  hipLaunchKernelGGL(k_age_advance, dim3(1), dim3(64), 0, st, ctl, dt, count);
  if (n > 0) hipLaunchKernelGGL(k_age_fill, dim3(grid1(n)), dim3(kBlock), 0, st, ctl, age, n);
}

void launch_pack_xy(hipStream_t st, Fields f, int n) {
  hipLaunchKernelGGL(k_pack_xy, dim3(grid1(n)), dim3(kBlock), 0, st, f, n);
}

void launch_unpack_xy(hipStream_t st, Fields f, int n) {
  hipLaunchKernelGGL(k_unpack_xy, dim3(grid1(n)), dim3(kBlock), 0, st, f, n);
}

}  // namespace samos
