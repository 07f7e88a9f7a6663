// Pair forces and alignment torques over the full neighbour list.
//
// One thread owns one slot and sums every neighbour's contribution in
// registers, so no atomics are needed and the result is deterministic.  The
// per-pair formulas restate:
//   soft    src/potentials/pair/pair_soft_potential.cpp:66-109
//   lj      src/potentials/pair/pair_lj_potential.cpp:48-121
//   polar   src/aligner/pair/pair_polar_aligner.cpp:58-103
//   nematic src/aligner/pair/pair_nematic_aligner.cpp:53-92
// The reference visits each pair once (half list) and scatters to both ends;
// here each pair is visited from both ends and each end keeps its own share,
// which is the same sum in a different order.  Per-slot energies accumulate
// the reference's per-pair energy; the host halves the total.
#include "kernels.hpp"

namespace samos {

namespace {

template <int POT, int AL, bool EN, bool F32>
__global__ void __launch_bounds__(kBlock) k_interact(int mode, Fields f, NListD nl, BoxD b, InteractD ip, int n,
                                                     double* e_out) {
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double xi = f.x[s], yi = f.y[s], zi = f.z[s];
  const int ti = f.type[s] - 1;
  double ri = 0.0;
  if (((POT & POT_SOFT) && ip.soft_radii) || ((POT & POT_LJ) && ip.lj_radii)) ri = f.radius[s];
  double nxi = 0.0, nyi = 0.0, nzi = 0.0;
  if (AL) { nxi = f.nx[s]; nyi = f.ny[s]; nzi = f.nz[s]; }
  double fx = 0.0, fy = 0.0, fz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
  double e_soft = 0.0, e_lj = 0.0, e_pol = 0.0, e_nem = 0.0;
  const int kn = nl.nnb[s];
  for (int k = 0; k < kn; ++k) {
    const int t = nl.nbr[nbr_at(nl, k, s)];
    // d = x_j - x_i (the soft and aligner convention); LJ uses -d.
    double dx = f.x[t] - xi, dy = f.y[t] - yi, dz = f.z[t] - zi;
    min_image(b, dx, dy, dz);
    const double r_sq = dx * dx + dy * dy + dz * dz;
    const int tj = f.type[t] - 1;
    if (POT & POT_SOFT) {
      const int tt = ti * ip.soft_nt + tj;
      const int nt2 = ip.soft_nt * ip.soft_nt;
      const double kk = ip.soft_tab[tt];
      const double a = ip.soft_radii ? ri + f.radius[t] : ip.soft_tab[nt2 + tt];
      if (F32) {
        const float r = sqrtf((float)r_sq);
        if (r < (float)a) {
          const float diff = (float)a - r;
          if (EN) e_soft += (double)(0.5f * (float)kk * diff * diff);
          const float ff = (r > 0.0f) ? (float)kk * diff / r : (float)kk * diff;
          fx -= (double)(ff * (float)dx); fy -= (double)(ff * (float)dy); fz -= (double)(ff * (float)dz);
        }
      } else {
        const double r = sqrt(r_sq);
        if (r < a) {
          const double diff = a - r;
          if (EN) e_soft += 0.5 * kk * diff * diff;
          const double ff = (r > 0.0) ? kk * diff / r : kk * diff;
          fx -= ff * dx; fy -= ff * dy; fz -= ff * dz;
        }
      }
    }
    if (POT & POT_LJ) {
      const int tt = ti * ip.lj_nt + tj;
      const int nt2 = ip.lj_nt * ip.lj_nt;
      const double rc = ip.lj_tab[2 * nt2 + tt];
      const double rc_sq = rc * rc;
      if (r_sq <= rc_sq) {
        const double eps = ip.lj_tab[tt];
        const double sig = ip.lj_radii ? ri + f.radius[t] : ip.lj_tab[nt2 + tt];
        const double sig_sq = sig * sig;
        if (F32) {
          const float s2 = (float)sig_sq / (float)r_sq;
          const float s6 = s2 * s2 * s2;
          if (EN) {
            float e = 4.0f * (float)eps * s6 * (s6 - 1.0f);
            if (ip.lj_shifted) {
              const float c2 = (float)(sig_sq / rc_sq);
              const float c6 = c2 * c2 * c2;
              e -= 4.0f * (float)eps * c6 * (c6 - 1.0f);
            }
            e_lj += (double)e;
          }
          const float ff = 48.0f * (float)eps * s6 * (s6 - 0.5f) * s2;
          fx -= (double)(ff * (float)dx); fy -= (double)(ff * (float)dy); fz -= (double)(ff * (float)dz);
        } else {
          const double inv_r_sq = lj_inv_r_sq(sig_sq, r_sq);
          const double inv_r_6 = inv_r_sq * inv_r_sq * inv_r_sq;
          if (EN) {
            double e = 4.0 * eps * inv_r_6 * (inv_r_6 - 1.0);
            if (ip.lj_shifted) {
              const double c2 = sig_sq / rc_sq;
              const double c6 = c2 * c2 * c2;
              e -= 4.0 * eps * c6 * (c6 - 1.0);
            }
            e_lj += e;
          }
          const double ff = 48.0 * eps * inv_r_6 * (inv_r_6 - 0.5) * inv_r_sq;
          fx -= ff * dx; fy -= ff * dy; fz -= ff * dz;
        }
      }
    }
    if (AL) {
      const double nxj = f.nx[t], nyj = f.ny[t], nzj = f.nz[t];
      const double cx = nyi * nzj - nzi * nyj;
      const double cy = nzi * nxj - nxi * nzj;
      const double cz = nxi * nyj - nyi * nxj;
      if (AL & AL_POLAR) {
        const int tt = ti * ip.pol_nt + tj;
        const double J = ip.pol_tab[tt];
        const double rc = ip.pol_tab[ip.pol_nt * ip.pol_nt + tt];
        if (r_sq <= rc * rc) {
          tx += J * cx; ty += J * cy; tz += J * cz;
          if (EN) e_pol += -2.0 * J * (nxi * nxj + nyi * nyj + nzi * nzj);
        }
      }
      if (AL & AL_NEMATIC) {
        const int tt = ti * ip.nem_nt + tj;
        const double J = ip.nem_tab[tt];
        const double rc = ip.nem_tab[ip.nem_nt * ip.nem_nt + tt];
        if (r_sq <= rc * rc) {
          const double nn = nxi * nxj + nyi * nyj + nzi * nzj;
          tx += J * nn * cx; ty += J * nn * cy; tz += J * nn * cz;
          if (EN) e_nem += -2.0 * J * (2.0 * nn * nn - 1.0);
        }
      }
    }
  }
  if (ip.has_sp) {
    fx += ip.sp_alpha * f.nx[s]; fy += ip.sp_alpha * f.ny[s]; fz += ip.sp_alpha * f.nz[s];
  }
  if ((POT || ip.has_sp) && mode != 2) {
    if (mode == 0) { f.fx[s] = fx; f.fy[s] = fy; f.fz[s] = fz; }
    else { f.fx[s] += fx; f.fy[s] += fy; f.fz[s] += fz; }
  }
  if (AL && mode != 2) {
    if (mode == 0) { f.tx[s] = tx; f.ty[s] = ty; f.tz[s] = tz; }
    else { f.tx[s] += tx; f.ty[s] += ty; f.tz[s] += tz; }
  }
  if (EN) {
    e_out[s] = e_soft;
    e_out[(long)nl.np + s] = e_lj;
    e_out[2L * nl.np + s] = e_pol;
    e_out[3L * nl.np + s] = e_nem;
  }
}

// Further pair potentials over the full list, added to the forces.  Per-pair
// terms restate
//   morse            src/potentials/pair/pair_morse_potential.cpp:71-97
//   gaussian         src/potentials/pair/pair_gaussian_potential.cpp:71-101
//   yukawa           src/potentials/pair/pair_yukawa_potential.cpp:56-72
//   soft_attractive  src/potentials/pair/pair_soft_attractive_potential.cpp:72-110
// (phase_in not supported).  morse, gaussian and yukawa use d = x_i - x_j, the
// single-wrap image of which is exactly minus that of x_j - x_i.
// soft_attractive recounts the coordination (neighbours closer than r_f).
__global__ void __launch_bounds__(kBlock) k_interact_extra(Fields f, NListD nl, BoxD b, ExtraD ex, int n,
                                                           double* e_col, int recount, int add_forces) {
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double xi = f.x[s], yi = f.y[s], zi = f.z[s], ai = f.radius[s];
  double fx = 0.0, fy = 0.0, fz = 0.0, en = 0.0;
  int coord = 0;
  const int kn = nl.nnb[s];
  for (int k = 0; k < kn; ++k) {
    const int t = nl.nbr[nbr_at(nl, k, s)];
    double dx = f.x[t] - xi, dy = f.y[t] - yi, dz = f.z[t] - zi;   // x_j - x_i
    min_image(b, dx, dy, dz);
    const double r_sq = dx * dx + dy * dy + dz * dz;
    const double aj = f.radius[t];
    for (int e = 0; e < ex.count; ++e) {
      const double* q = ex.q[e];
      const int kind = ex.kind[e];
      if (kind == EXTRA_SOFT_ATTRACTIVE) {
        const double r = sqrt(r_sq);
        const double kk = q[0], alpha = 1.0;
        const double ai_p_aj = ex.radii[e] ? ai + aj : q[1];
        const double r_turn = q[2] * ai_p_aj;
        const double r_f = 2.0 * r_turn - ai_p_aj;
        if (r < r_f) {
          ++coord;
          double ff = 0.0, pe = 0.0;
          if (r <= r_turn) {
            const double diff = ai_p_aj - r;
            pe = 0.5 * kk * alpha * diff * diff;
            ff = (r > 0.0) ? kk * alpha * diff / r : kk * diff;
          } else if (r > r_turn && r < r_f) {
            const double diff = r_f - r;
            pe = -0.5 * kk * alpha * diff * diff;
            ff = -kk * alpha * diff / r;
          }
          en += pe;
          fx -= ff * dx; fy -= ff * dy; fz -= ff * dz;
        }
        continue;
      }
      double ff = 0.0, pe = 0.0;
      if (kind == EXTRA_MORSE) {
        const double rcut = q[3];
        if (!(r_sq <= rcut * rcut)) continue;
        const double D = q[0], a = q[1], alpha = 1.0;
        const double re = ex.radii[e] ? ai + aj : q[2];
        const double r = sqrt(r_sq);
        const double exp_fact = exp(-a * (r - re));
        const double pot_fact = exp_fact - 1.0;
        pe = D * alpha * (pot_fact * pot_fact - 1.0);
        if (ex.shifted[e]) {
          const double shift_fact = exp(-a * (rcut - re)) - 1.0;
          pe -= D * alpha * (shift_fact * shift_fact - 1.0);
        }
        ff = 2.0 * D * a * alpha * exp_fact * pot_fact / r;
      } else if (kind == EXTRA_GAUSSIAN) {
        const double rcut = q[6];
        if (!(r_sq <= rcut * rcut)) continue;
        const double A = q[0], B = q[1], al = q[2], be = q[3], rA = q[4], pf = 1.0;
        const double rB = ex.radii[e] ? ai + aj : q[5];
        const double r = sqrt(r_sq);
        const double r_m_rA = r - rA, r_m_rB = r - rB;
        pe = A * pf * exp(-al * r_m_rA * r_m_rA) + B * exp(-be * r_m_rB * r_m_rB);
        if (ex.shifted[e]) {
          const double rcut_m_rA = rcut - rA, rcut_m_rB = rcut - rB;
          pe -= A * pf * exp(-al * rcut_m_rA * rcut_m_rA) + B * exp(-be * rcut_m_rB * rcut_m_rB);
        }
        ff = (2.0 / r) * pf * (A * al * r_m_rA * exp(-al * r_m_rA * r_m_rA) + B * be * r_m_rB * exp(-be * r_m_rB * r_m_rB));
      } else {  // yukawa
        const double rcut = q[2];
        if (!(r_sq <= rcut * rcut)) continue;
        const double g = q[0], kappa = q[1];
        const double r = sqrt(r_sq);
        const double exp_fact = g * exp(-kappa * r);
        pe = exp_fact / r;
        ff = exp_fact * (1 + kappa * r) / (r_sq * r);
      }
      en += pe;
      // force on i along x_i - x_j = -(x_j - x_i)
      fx += ff * -dx; fy += ff * -dy; fz += ff * -dz;
    }
  }
  if (add_forces) { f.fx[s] += fx; f.fy[s] += fy; f.fz[s] += fz; }
  if (recount) f.coord[s] = coord;
  if (e_col) e_col[s] = en;
}

template <int POT, int AL>
void dispatch_en(hipStream_t st, int mode, Fields f, NListD nl, BoxD b, InteractD ip, int n, double* e_out, bool f32) {
  dim3 g(grid1(n)), blk(kBlock);
  if (e_out) {
    if (f32) hipLaunchKernelGGL((k_interact<POT, AL, true, true>), g, blk, 0, st, mode, f, nl, b, ip, n, e_out);
    else hipLaunchKernelGGL((k_interact<POT, AL, true, false>), g, blk, 0, st, mode, f, nl, b, ip, n, e_out);
  } else {
    if (f32) hipLaunchKernelGGL((k_interact<POT, AL, false, true>), g, blk, 0, st, mode, f, nl, b, ip, n, e_out);
    else hipLaunchKernelGGL((k_interact<POT, AL, false, false>), g, blk, 0, st, mode, f, nl, b, ip, n, e_out);
  }
}

template <int POT>
void dispatch_al(hipStream_t st, int al, int mode, Fields f, NListD nl, BoxD b, InteractD ip, int n, double* e_out, bool f32) {
  switch (al) {
    case 0: dispatch_en<POT, 0>(st, mode, f, nl, b, ip, n, e_out, f32); break;
    case 1: dispatch_en<POT, 1>(st, mode, f, nl, b, ip, n, e_out, f32); break;
    case 2: dispatch_en<POT, 2>(st, mode, f, nl, b, ip, n, e_out, f32); break;
    default: dispatch_en<POT, 3>(st, mode, f, nl, b, ip, n, e_out, f32); break;
  }
}

constexpr int kRedBlocksMax = 512;

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Pass 1: one partial per (column, block) with a fixed grid, so the summation
// tree depends only on n.
__global__ void __launch_bounds__(kBlock) k_reduce_partial(const double* cols, long ld, int n, double* partial) {
  __shared__ double lds[kBlock / 64];
  const double* c = cols + (long)blockIdx.y * ld;
  double acc = 0.0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) acc += c[s];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_reduce_final(const double* partial, int nb, double* out) {
  __shared__ double lds[kBlock / 64];
  const double* p = partial + (long)blockIdx.x * nb;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += p[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    out[blockIdx.x] = t;
  }
}

}  // namespace

void launch_interact(hipStream_t st, int pot_mask, int al_mask, int mode, Fields f, NListD nl, BoxD box,
                     InteractD ip, int n, double* e_out) {
  const bool f32 = false;
  switch (pot_mask) {
    case 0: dispatch_al<0>(st, al_mask, mode, f, nl, box, ip, n, e_out, f32); break;
    case 1: dispatch_al<1>(st, al_mask, mode, f, nl, box, ip, n, e_out, f32); break;
    case 2: dispatch_al<2>(st, al_mask, mode, f, nl, box, ip, n, e_out, f32); break;
    default: dispatch_al<3>(st, al_mask, mode, f, nl, box, ip, n, e_out, f32); break;
  }
}

void launch_interact_f32(hipStream_t st, int pot_mask, int al_mask, int mode, Fields f, NListD nl, BoxD box,
                         InteractD ip, int n, double* e_out) {
  switch (pot_mask) {
    case 0: dispatch_al<0>(st, al_mask, mode, f, nl, box, ip, n, e_out, true); break;
    case 1: dispatch_al<1>(st, al_mask, mode, f, nl, box, ip, n, e_out, true); break;
    case 2: dispatch_al<2>(st, al_mask, mode, f, nl, box, ip, n, e_out, true); break;
    default: dispatch_al<3>(st, al_mask, mode, f, nl, box, ip, n, e_out, true); break;
  }
}

void launch_interact_extra(hipStream_t st, Fields f, NListD nl, BoxD box, const ExtraD& ex, int n, double* e_col,
                           int add_forces) {
  if (ex.count == 0) return;
  int recount = 0;
  for (int e = 0; e < ex.count; ++e) recount |= ex.kind[e] == EXTRA_SOFT_ATTRACTIVE;
  hipLaunchKernelGGL(k_interact_extra, dim3(grid1(n)), dim3(kBlock), 0, st, f, nl, box, ex, n, e_col, recount,
                     add_forces);
}

int reduce_blocks(int n) {
  int g = div_up(n, kBlock);
  return g < 1 ? 1 : (g > kRedBlocksMax ? kRedBlocksMax : g);
}

void launch_reduce_sum(hipStream_t st, const double* cols, int ncols, long ld, int n, double* partial, double* out) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_reduce_partial, dim3(nb, ncols), dim3(kBlock), 0, st, cols, ld, n, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(ncols), dim3(kBlock), 0, st, partial, nb, out);
}

}  // namespace samos
