// Fused Brownian step: one launch computes the loop-head forces and torques
// of step t from the neighbour list and applies IntegratorBrownian::integrate
// to each particle.
//
//   loop head          src/samos.cpp:887-894 (enforce, reset, Potential and
//                      Aligner compute)
//   pair terms         pair_sum.hpp (soft, lj, polar, nematic)
//   self propulsion    src/potentials/external/external_self_propulsion.cpp:33-50
//   integrator         src/integrators/integrator_brownian.cpp:42-117
//
// Positions and directors are read from `cur` and written to `nxt`, so
// neighbour gathers never see a half-updated step; velocities, forces,
// torques, omega, age and image flags are per slot and updated in place.
// The forces and torques written are those of step t's loop head (the state
// the reference holds while it integrates step t).  With head_enforce the
// next loop head's enforce is applied to the positions written.  The
// rebuild request follows the fused Langevin step: flags[parity] in direct
// launches, the chunk's request words in captured chunks.
#include <algorithm>

#include "kernels.hpp"
#include "pair_sum.hpp"
#include "constraint.hpp"

namespace samos {

namespace {

template <int POT, int AL, bool F32, bool PLANAR, bool MT>
__global__ void __launch_bounds__(kBlock) k_brownian_fused(Fields cur, Fields nxt, NListD nl, BoxD b, ConstraintD c,
                                                           InteractD ip, int n, BrownArgs ba) {
  ChunkCtl* const ck = ba.chunk;
  int head_enforce = ba.head_enforce;
  if (ck) {
    const int j = ba.slot;
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
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    nl.flags[1 - ba.parity] = 0;
  }
  // ba.lanes (a power of two dividing 64) consecutive threads share one slot:
  // each sums every lanes-th batch of the list, the partial sums are added
  // in a fixed butterfly order, and lane 0 integrates and writes.
  const int gt = xcd_block() * blockDim.x + threadIdx.x;
  const int s = gt >> ba.lane_shift, lane = gt & (ba.lanes - 1);
  bool need = false;
  if (s < n) {
    double xi, yi, zi;
    if (PLANAR) {
      const double2 q = cur.xy[s];
      xi = q.x; yi = q.y; zi = c.zpos;
    } else {
      xi = cur.x[s]; yi = cur.y[s]; zi = cur.z[s];
    }
    const double nxi = cur.nx[s], nyi = cur.ny[s], nzi = PLANAR ? 0.0 : cur.nz[s];
    const int ti = MT ? cur.type[s] - 1 : 0;
    double ri = 0.0;
    if (((POT & POT_SOFT) && ip.soft_radii) || ((POT & POT_LJ) && ip.lj_radii)) ri = cur.radius[s];
    bool wrap = b.periodic != 0;
    if (wrap) {
      const double R = 2.0 * sqrt(nl.cut2);
      bool inner = xi - b.xlo > R && b.xhi - xi > R && yi - b.ylo > R && b.yhi - yi > R;
      if (!PLANAR) inner = inner && zi - b.zlo > R && b.zhi - zi > R;
      wrap = !inner;
    }
    double fx = 0.0, fy = 0.0, fz = 0.0, tq[3] = {0.0, 0.0, 0.0};
    pair_force_loop<POT, F32, PLANAR, MT, 4, AL>(cur, nl, b, ip, s, xi, yi, zi, ti, ri, fx, fy, fz, wrap, nxi, nyi,
                                                 nzi, tq, lane, ba.lanes);
    for (int off = 1; off < ba.lanes; off <<= 1) {
      fx += __shfl_xor(fx, off); fy += __shfl_xor(fy, off);
      if (!PLANAR) fz += __shfl_xor(fz, off);
      if (AL) {
        if (!PLANAR) { tq[0] += __shfl_xor(tq[0], off); tq[1] += __shfl_xor(tq[1], off); }
        tq[2] += __shfl_xor(tq[2], off);
      }
    }
    if (ip.has_sp) { fx += ip.sp_alpha * nxi; fy += ip.sp_alpha * nyi; fz += ip.sp_alpha * nzi; }

    const uint32_t pid = (uint32_t)cur.id[s];
    PS p;
    p.x = xi; p.y = yi; p.z = zi;
    p.vx = cur.vx[s]; p.vy = cur.vy[s]; p.vz = PLANAR ? 0.0 : cur.vz[s];
    p.nx = nxi; p.ny = nyi; p.nz = nzi;
    p.dix = p.diy = p.diz = 0;
    constexpr int kSet = PLANAR ? kSurfPlane : kSurfAll;
    bool fz0 = false;
    double om, age_dt;
    // The integrator of this slot: with groups, the one whose group holds it
    // (the groups are disjoint); a slot in none of them is not integrated.
    bool member = true;
    double q_dt = ba.dt, q_v0 = ba.v0, q_mu = ba.mu, q_mur = ba.mur, q_stoch = ba.stoch, q_flip = ba.flip_p;
    int q_nem = ba.nematic, q_vel = ba.velocity, q_cs = ba.clock_slot;
    TempD q_temp = ba.temp;
    uint32_t q_seed = ba.seed;
    if (MT && ba.ngroups > 0) {
      const unsigned gm = (unsigned)cur.groups[s];
      int k = 0;
      while (k < ba.ngroups && !(gm & ba.grp[k].gbit)) ++k;
      member = k < ba.ngroups;
      if (member) {
        const BrownArgs::Group& g = ba.grp[k];
        q_dt = g.dt; q_v0 = g.v0; q_mu = g.mu; q_mur = g.mur; q_stoch = g.stoch; q_flip = g.flip_p;
        q_nem = g.nematic; q_vel = g.velocity; q_cs = g.clock_slot; q_temp = g.temp; q_seed = g.seed;
      }
    }
    if (!member) {
      om = cur.omega[s];
      age_dt = 0.0;
    } else if (ba.mode == 0) {
      // IntegratorBrownian::integrate for this particle, with the forces and
      // torques just computed
      const unsigned long long call = eff_call(ba.clock, q_cs, 0ULL) + (unsigned long long)ba.t_off;
      double g_rot = 0.0, u_flip = 1.0;
      gauss1_u1(q_seed, TAG_BROWNIAN, pid, call, g_rot, u_flip);
      if (q_nem && u_flip < q_flip) {
        p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz;
        if (q_vel) { p.vx = -p.vx; p.vy = -p.vy; p.vz = -p.vz; }
      }
      const double T = q_temp.at(eff_step(ba.clock, ba.t_off));
      const double B = sqrt(2.0 * q_mu * T);
      const double sqrt_dt = sqrt(q_dt);
      const double fdx = q_v0 * p.nx + q_mu * fx, fdy = q_v0 * p.ny + q_mu * fy, fdz = q_v0 * p.nz + q_mu * fz;
      p.vx = fdx; p.vy = fdy; p.vz = fdz;
      p.x += q_dt * fdx; p.y += q_dt * fdy; p.z += q_dt * fdz;
      if (T > 0.0) {
        double g[3];
        gauss3(q_seed, TAG_BROWNIAN, pid, call, g);
        const double frx = B * g[0], fry = B * g[1], frz = B * g[2];
        p.vx += frx; p.vy += fry; p.vz += frz;
        p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
      }
      fz0 = enforce<kSet>(b, c, p);
      om = q_mur * project_torque<kSet>(c, p, tq[0], tq[1], tq[2]);
      const double dtheta = q_dt * om + q_stoch * g_rot;
      rotate_director<kSet>(c, p, dtheta);
      if (q_vel) rotate_velocity<kSet>(c, p, dtheta);
      age_dt = q_dt;
    } else {
      // IntegratorBrownianAlign (directors, at the loop-head position), then
      // IntegratorBrownianPos (positions), each with its own noise stream
      const unsigned long long call_a = eff_call(ba.clock, ba.clock_slot, 0ULL) + (unsigned long long)ba.t_off;
      double g_rot = 0.0, u_flip = 1.0;
      gauss1_u1(ba.seed, TAG_BROWNIAN_ALIGN, pid, call_a, g_rot, u_flip);
      if (ba.nematic && u_flip < ba.flip_p) { p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz; }
      om = ba.mur * project_torque<kSet>(c, p, tq[0], tq[1], tq[2]);
      rotate_director<kSet>(c, p, ba.dt * om + ba.stoch * g_rot);
      const unsigned long long call_p = eff_call(ba.clock, ba.pos_slot, 0ULL) + (unsigned long long)ba.t_off;
      const double T = ba.pos_temp.at(eff_step(ba.clock, ba.t_off));
      const double B = sqrt(2.0 * ba.pos_mu * T);
      const double sqrt_dt = sqrt(ba.pos_dt);
      const double fdx = ba.pos_mu * fx, fdy = ba.pos_mu * fy, fdz = ba.pos_mu * fz;
      p.vx = fdx; p.vy = fdy; p.vz = fdz;
      p.x += ba.pos_dt * fdx; p.y += ba.pos_dt * fdy; p.z += ba.pos_dt * fdz;
      if (T > 0.0) {
        double g[3];
        gauss3(ba.pos_seed, TAG_BROWNIAN_POS, pid, call_p, g);
        const double frx = B * g[0], fry = B * g[1], frz = B * g[2];
        p.vx += frx; p.vy += fry; p.vz += frz;
        p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
      }
      fz0 = enforce<kSet>(b, c, p);
      age_dt = ba.pos_dt;
    }
    const bool writer = lane == 0;
    if (head_enforce) enforce<kSet>(b, c, p);
    if (writer) {
      cur.fx[s] = fx; cur.fy[s] = fy; cur.fz[s] = fz0 ? 0.0 : fz;
      if (AL) { cur.tx[s] = tq[0]; cur.ty[s] = tq[1]; cur.tz[s] = tq[2]; }
      cur.omega[s] = om;
      if (member && !(ba.age && ba.age->uniform)) cur.age[s] += age_dt;
      cur.vx[s] = p.vx; cur.vy[s] = p.vy;
      if (!PLANAR) cur.vz[s] = p.vz;
      if (PLANAR) {
        nxt.xy[s] = make_double2(p.x, p.y);
      } else {
        nxt.x[s] = p.x; nxt.y[s] = p.y; nxt.z[s] = p.z;
      }
      nxt.nx[s] = p.nx; nxt.ny[s] = p.ny;
      if (!PLANAR) nxt.nz[s] = p.nz;
      if (p.dix) cur.ix[s] += p.dix;
      if (p.diy) cur.iy[s] += p.diy;
      if (p.diz) cur.iz[s] += p.diz;
      double ddx = nl.ox[s] - p.x, ddy = nl.oy[s] - p.y, ddz = PLANAR ? 0.0 : nl.oz[s] - p.z;
      min_image(b, ddx, ddy, ddz);
      need = !(ddx * ddx + ddy * ddy + ddz * ddz < nl.pad2q);
    }
  }
  if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomic_set_once(ck ? &ck->req[ba.slot & 1] : &nl.flags[ba.parity]);
}

template <int POT, int AL, bool F32, bool PLANAR>
void brownian_mt(hipStream_t st, int multitype, Fields cur, Fields nxt, NListD nl, BoxD b, ConstraintD c, InteractD ip,
                 int n, const BrownArgs& ba) {
  dim3 g(grid1((long)n * ba.lanes)), blk(kBlock);
  if (multitype)
    hipLaunchKernelGGL((k_brownian_fused<POT, AL, F32, PLANAR, true>), g, blk, 0, st, cur, nxt, nl, b, c, ip, n, ba);
  else
    hipLaunchKernelGGL((k_brownian_fused<POT, AL, F32, PLANAR, false>), g, blk, 0, st, cur, nxt, nl, b, c, ip, n, ba);
}

template <int POT, int AL>
void brownian_pa(hipStream_t st, int fp32, int planar, int multitype, Fields cur, Fields nxt, NListD nl, BoxD b,
                 ConstraintD c, InteractD ip, int n, const BrownArgs& ba) {
  if (fp32) {
    if (planar) brownian_mt<POT, AL, true, true>(st, multitype, cur, nxt, nl, b, c, ip, n, ba);
    else brownian_mt<POT, AL, true, false>(st, multitype, cur, nxt, nl, b, c, ip, n, ba);
  } else {
    if (planar) brownian_mt<POT, AL, false, true>(st, multitype, cur, nxt, nl, b, c, ip, n, ba);
    else brownian_mt<POT, AL, false, false>(st, multitype, cur, nxt, nl, b, c, ip, n, ba);
  }
}

template <int POT>
void brownian_pot(hipStream_t st, int al_mask, int fp32, int planar, int multitype, Fields cur, Fields nxt, NListD nl,
                  BoxD b, ConstraintD c, InteractD ip, int n, const BrownArgs& ba) {
  switch (al_mask) {
    case 0: brownian_pa<POT, 0>(st, fp32, planar, multitype, cur, nxt, nl, b, c, ip, n, ba); break;
    case AL_POLAR: brownian_pa<POT, AL_POLAR>(st, fp32, planar, multitype, cur, nxt, nl, b, c, ip, n, ba); break;
    default: brownian_pa<POT, AL_NEMATIC>(st, fp32, planar, multitype, cur, nxt, nl, b, c, ip, n, ba); break;
  }
}

}  // namespace

bool brownian_fused_supported(int pot_mask, int al_mask) {
  return pot_mask >= 1 && pot_mask <= 3 && (al_mask == 0 || al_mask == AL_POLAR || al_mask == AL_NEMATIC);
}

void launch_brownian_fused(hipStream_t st, int pot_mask, int al_mask, Fields cur, Fields nxt, NListD nl, BoxD box,
                           ConstraintD con, InteractD ip, int n, const BrownArgs& ba, int fp32, int planar,
                           int multitype) {
  switch (pot_mask) {
    case 1: brownian_pot<1>(st, al_mask, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, ba); break;
    case 2: brownian_pot<2>(st, al_mask, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, ba); break;
    default: brownian_pot<3>(st, al_mask, fp32, planar, multitype, cur, nxt, nl, box, con, ip, n, ba); break;
  }
}

}  // namespace samos
