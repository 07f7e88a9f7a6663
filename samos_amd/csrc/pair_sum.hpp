// Pair-force sum of one slot over its neighbour list, shared by the force and
// fused step kernels.  Neighbours are consumed in batches of kBatch: all
// indices of a batch are loaded first, then all neighbour positions, then the
// arithmetic, so each wavefront keeps kBatch independent gathers in flight
// instead of one dependent chain per neighbour.
//
// Per-pair arithmetic restates src/potentials/pair/pair_soft_potential.cpp:66-109
// and src/potentials/pair/pair_lj_potential.cpp:48-121 (LJ with sigma^2/r^2 in
// the force factor, inclusive cutoff; soft with strict a_ij).
//
// PLANAR: every particle lies on the same plane z = const (plane constraint),
// so neighbour positions come from the interleaved (x, y) column and dz = 0
// exactly, which leaves r^2 bit-identical to the 3-D expression.
// MT: the general form: per-type table lookups and per-particle radii
// (use_particle_radii).  Without MT the tables' (1, 1) entries hold for every
// pair and no radius is read.
#pragma once
#include "samos_device.hpp"

namespace samos {

// base[byte_off / sizeof(T)] with a 32-bit byte offset.
template <typename T>
__device__ __forceinline__ T at32(const T* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// wrap_i: apply the minimum image to each pair.  A slot farther than
// 2 (rcut + pad) from every periodic face has all its listed neighbours
// (within rcut + pad at the build, each moved by less than pad / 2 since)
// inside the box at their unwrapped offsets, so for it the image step is an
// exact no-op and is skipped (a per-lane test inside one loop: whole
// wavefronts away from the faces branch over it).  PLANAR pairs have dz = 0
// and no z image.
//
// AL (AL_POLAR, AL_NEMATIC or 0) adds the aligner torques on the slot's
// director n_i (src/aligner/pair/pair_polar_aligner.cpp:58-103,
// src/aligner/pair/pair_nematic_aligner.cpp:53-92): neighbour directors are
// gathered with the positions.
//
// lane, lanes: this thread takes batches lane, lane + lanes, ... of the
// slot's list (several threads per slot for small systems; the caller adds
// the partial sums).
template <int POT, bool F32, bool PLANAR, bool MT, int kBatch, int AL = 0>
__device__ __forceinline__ void pair_force_loop(const Fields& f, const NListD& nl, const BoxD& b, const InteractD& ip,
                                                int s, double xi, double yi, double zi, int ti, double ri,
                                                double& fx, double& fy, double& fz, bool wrap_i,
                                                double nxi = 0.0, double nyi = 0.0, double nzi = 0.0,
                                                double* tq = nullptr, int lane = 0, int lanes = 1) {
  const bool radii = MT && (((POT & POT_SOFT) && ip.soft_radii) || ((POT & POT_LJ) && ip.lj_radii));
  // One particle type: the table entries of the pair (1, 1) are loop constants,
  // and so are the first products of each expression (sigma^2, rcut^2,
  // 48 epsilon), which round exactly as when formed per pair.
  const int nt2s = ip.soft_nt * ip.soft_nt, nt2l = ip.lj_nt * ip.lj_nt;
  double c_kk = 0.0, c_a = 0.0, c_rc2 = 0.0, c_eps = 0.0, c_e48 = 0.0, c_sig2 = 0.0;
  if (!MT) {
    if (POT & POT_SOFT) { c_kk = ip.soft_tab[0]; c_a = ip.soft_tab[nt2s]; }
    if (POT & POT_LJ) {
      const double rc = ip.lj_tab[2 * nt2l], sig = ip.lj_tab[nt2l];
      c_eps = ip.lj_tab[0]; c_rc2 = rc * rc; c_e48 = 48.0 * c_eps; c_sig2 = sig * sig;
    }
  }
  const int nt2p = ip.pol_nt * ip.pol_nt, nt2n = ip.nem_nt * ip.nem_nt;
  double c_pj = 0.0, c_prc = 0.0, c_nj = 0.0, c_nrc = 0.0;
  if (!MT) {
    if (AL & AL_POLAR) { c_pj = ip.pol_tab[0]; c_prc = ip.pol_tab[nt2p]; }
    if (AL & AL_NEMATIC) { c_nj = ip.nem_tab[0]; c_nrc = ip.nem_tab[nt2n]; }
  }
  double tx = 0.0, ty = 0.0, tz = 0.0;
  const int kn = nl.nnb[s];
  // A list is padded with the slot's own index to a multiple of 8 entries
  // (the builds write the padding), so whole batches load without bounds
  // tests; the arithmetic below skips the padding.  The next batch's indices
  // are loaded while the current batch is gathered and summed.  Gather offsets are 32-bit byte offsets from the column bases (np < 2^28,
  // which the host checks before choosing these kernels), so each gather
  // takes its address from a scalar base and one vector register.
  static_assert(kBatch % 4 == 0, "batches of whole four-entry groups");
  const int step = lanes * kBatch;
  const unsigned s16 = (unsigned)s << 4;
  const uint4* G = reinterpret_cast<const uint4*>(nl.nbr);
  unsigned t[kBatch];  // neighbour slots
  if (lane * kBatch < kn) {
#pragma unroll
    for (int g = 0; g < kBatch / 4; ++g) {
      const uint4 q = at32(G + (long)((lane * kBatch >> 2) + g) * nl.np, s16);
      t[4 * g] = q.x; t[4 * g + 1] = q.y; t[4 * g + 2] = q.z; t[4 * g + 3] = q.w;
    }
  }
  for (int k0 = lane * kBatch; k0 < kn; k0 += step) {
    double xj[kBatch], yj[kBatch], zj[kBatch], rj[kBatch];
    double nxj[kBatch], nyj[kBatch], nzj[kBatch];
    int tj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (PLANAR) {
        const double2 p = at32(f.xy, t[u] << 4);
        xj[u] = p.x; yj[u] = p.y; zj[u] = zi;
      } else {
        xj[u] = at32(f.x, t[u] << 3); yj[u] = at32(f.y, t[u] << 3); zj[u] = at32(f.z, t[u] << 3);
      }
      tj[u] = MT ? at32(f.type, t[u] << 2) - 1 : 0;
      if (AL) {
        nxj[u] = at32(f.nx, t[u] << 3); nyj[u] = at32(f.ny, t[u] << 3);
        nzj[u] = PLANAR ? 0.0 : at32(f.nz, t[u] << 3);
      }
      rj[u] = radii ? at32(f.radius, t[u] << 3) : 0.0;
    }
    if (k0 + step < kn) {
#pragma unroll
      for (int g = 0; g < kBatch / 4; ++g) {
        const uint4 q = at32(G + (long)(((k0 + step) >> 2) + g) * nl.np, s16);
        t[4 * g] = q.x; t[4 * g + 1] = q.y; t[4 * g + 2] = q.z; t[4 * g + 3] = q.w;
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (k0 + u >= kn) break;
      // d = x_j - x_i; LJ's reference convention x_i - x_j is -d
      double dx = xj[u] - xi, dy = yj[u] - yi, dz = PLANAR ? 0.0 : zj[u] - zi;
      if (wrap_i) {
        if (PLANAR) min_image_xy(b, dx, dy);
        else min_image(b, dx, dy, dz);
      }
      // planar: dz = 0, and adding +0 to a sum of squares changes no bit
      const double r_sq = PLANAR ? dx * dx + dy * dy : dx * dx + dy * dy + dz * dz;
      if (POT & POT_SOFT) {
        const int tt = MT ? ti * ip.soft_nt + tj[u] : 0;
        const double kk = MT ? ip.soft_tab[tt] : c_kk;
        const double a = (MT && ip.soft_radii) ? ri + rj[u] : (MT ? ip.soft_tab[nt2s + tt] : c_a);
        if (F32) {
          // hardware square root and reciprocal (each within one float ulp)
          const float r = __builtin_amdgcn_sqrtf((float)r_sq);
          if (r < (float)a) {
            const float diff = (float)a - r;
            const float ff = (r > 0.0f) ? (float)kk * diff * __builtin_amdgcn_rcpf(r) : (float)kk * diff;
            fx = fma(-(double)ff, dx, fx); fy = fma(-(double)ff, dy, fy);
            if (!PLANAR) fz = fma(-(double)ff, dz, fz);
          }
        } else {
          const double r = sqrt(r_sq);
          if (r < a) {
            const double diff = a - r;
            const double ff = (r > 0.0) ? kk * diff / r : kk * diff;
            fx = fma(-ff, dx, fx); fy = fma(-ff, dy, fy);
            if (!PLANAR) fz = fma(-ff, dz, fz);
          }
        }
      }
      if (POT & POT_LJ) {
        const int tt = MT ? ti * ip.lj_nt + tj[u] : 0;
        double rc2 = c_rc2;
        if (MT) { const double rc = ip.lj_tab[2 * nt2l + tt]; rc2 = rc * rc; }
        if (r_sq <= rc2) {
          const double eps = MT ? ip.lj_tab[tt] : c_eps;
          const double e48 = MT ? 48.0 * eps : c_e48;
          double sig2 = c_sig2;
          if (MT && ip.lj_radii) { const double sg = ri + rj[u]; sig2 = sg * sg; }
          else if (MT) { const double sg = ip.lj_tab[nt2l + tt]; sig2 = sg * sg; }
          if (F32) {
            // sigma^2 / r^2 from the hardware reciprocal (within one float ulp);
            // the factor in FP32, each product accumulated in FP64
            const float s2 = (float)sig2 * __builtin_amdgcn_rcpf((float)r_sq);
            const float s6 = s2 * s2 * s2;
            const float ff = 48.0f * (float)eps * s6 * (s6 - 0.5f) * s2;
            fx = fma(-(double)ff, dx, fx); fy = fma(-(double)ff, dy, fy);
            if (!PLANAR) fz = fma(-(double)ff, dz, fz);
          } else {
            const double inv_r_sq = lj_inv_r_sq(sig2, r_sq);
            const double inv_r_6 = inv_r_sq * inv_r_sq * inv_r_sq;
            const double ff = e48 * inv_r_6 * (inv_r_6 - 0.5) * inv_r_sq;
            // one rounding per accumulation (fused multiply-add)
            fx = fma(-ff, dx, fx); fy = fma(-ff, dy, fy);
            if (!PLANAR) fz = fma(-ff, dz, fz);
          }
        }
      }
      if (AL) {
        // n_i x n_j; on a plane both directors have n_z = 0 and only the z
        // component is non-zero (the x and y terms would add exact zeros)
        const double cz = nxi * nyj[u] - nyi * nxj[u];
        const double cx = PLANAR ? 0.0 : nyi * nzj[u] - nzi * nyj[u];
        const double cy = PLANAR ? 0.0 : nzi * nxj[u] - nxi * nzj[u];
        if (AL & AL_POLAR) {
          const int tt = MT ? ti * ip.pol_nt + tj[u] : 0;
          const double J = MT ? ip.pol_tab[tt] : c_pj;
          const double rc = MT ? ip.pol_tab[nt2p + tt] : c_prc;
          if (r_sq <= rc * rc) {
            if (!PLANAR) { tx += J * cx; ty += J * cy; }
            tz += J * cz;
          }
        }
        if (AL & AL_NEMATIC) {
          const int tt = MT ? ti * ip.nem_nt + tj[u] : 0;
          const double J = MT ? ip.nem_tab[tt] : c_nj;
          const double rc = MT ? ip.nem_tab[nt2n + tt] : c_nrc;
          if (r_sq <= rc * rc) {
            const double nn = PLANAR ? nxi * nxj[u] + nyi * nyj[u] : nxi * nxj[u] + nyi * nyj[u] + nzi * nzj[u];
            if (!PLANAR) { tx += J * nn * cx; ty += J * nn * cy; }
            tz += J * nn * cz;
          }
        }
      }
    }
  }
  if (AL) { tq[0] = tx; tq[1] = ty; tq[2] = tz; }
}

template <int POT, bool F32, bool PLANAR, bool MT, int kBatch = 4>
__device__ __forceinline__ void pair_force_sum(const Fields& f, const NListD& nl, const BoxD& b, const InteractD& ip,
                                               int s, double xi, double yi, double zi, int ti, double ri,
                                               double& fx, double& fy, double& fz) {
  bool wrap = b.periodic != 0;
  if (wrap) {
    const double R = 2.0 * sqrt(nl.cut2);
    bool inner = xi - b.xlo > R && b.xhi - xi > R && yi - b.ylo > R && b.yhi - yi > R;
    if (!PLANAR) inner = inner && zi - b.zlo > R && b.zhi - zi > R;
    wrap = !inner;
  }
  pair_force_loop<POT, F32, PLANAR, MT, kBatch>(f, nl, b, ip, s, xi, yi, zi, ti, ri, fx, fy, fz, wrap);
}

}  // namespace samos
