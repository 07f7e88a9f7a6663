// Replica ensembles: hundreds of independent particle systems of different
// sizes and parameters advanced together, one fused step launch for all of
// them per step.
//
// Layout.  The replicas' particles are concatenated in HBM: replica r owns the
// slots [n0_r, n0_r + n_r), Morton-ordered on its own cell grid.  Positions
// are 32-byte records (x, y, z, radius) and directors (nx, ny, nz, 0), so a
// neighbour gather touches one 32-byte sector instead of one cache line per
// coordinate; both are double-buffered for the fused step.  Each replica's
// cell grid is hashed into its own bucket table of 2^h >= n_r buckets (plus
// one end entry), so a sphere of radius R costs O(n) buckets instead of
// (2R / r_l)^3 grid cells.  The build visits the distinct buckets of the 27
// stencil cells and keeps only candidates whose own cell lies in the stencil,
// which gives exactly the pair set of the reference's cell list when every
// side has three or more cells (src/system/neighbour_list.cpp:216-273,
// src/system/cell_list.cpp:37-104).
//
// Each replica decides its own rebuild from its own particles' displacements
// (src/system/neighbour_list.cpp:278-302), so a replica advances exactly as it
// would alone.  Rebuild passes and FIRE passes run over work lists (selected
// replicas with prefix sums of their particle and bucket counts, built on the
// device), so their cost follows the replicas that need them, not the
// ensemble.  Noise is the single-system Philox stream keyed by the replica's
// seed and the particle's index inside the replica.
//
// Physics per replica:
//   Brownian step (T = 0, rotational noise)  src/integrators/integrator_brownian.cpp:42-124
//   soft repulsion                            src/potentials/pair/pair_soft_potential.cpp:33-119
//   polar alignment                           src/aligner/pair/pair_polar_aligner.cpp:32-107
//   FIRE                                      src/integrators/integrator_fire.cpp:34-148
//   run loop order                            src/samos.cpp:878-970
#include "constraint.hpp"
#include "ensemble_impl.hpp"
#include "kernels.hpp"
#include "samos_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace samos {

namespace {

constexpr int kSelThreads = 1024;
constexpr int kTileThreads = 256;
constexpr int kTileItems = 16;
constexpr int kTile = kTileThreads * kTileItems;

// Parameters of replica r, read through a wave-uniform index when the whole
// wavefront belongs to one replica (the common case: replicas are contiguous
// slot ranges), which lets the compiler keep them in scalar registers.
template <class F>
__device__ __forceinline__ void with_rep(const RepD* __restrict__ reps, int r, F&& f) {
  const int r0 = __builtin_amdgcn_readfirstlane(r);
  if (__ballot(r != r0) == 0ull) f(reps[r0]);
  else f(reps[r]);
}

__device__ __forceinline__ int bucket_of(const RepD& R, int c) {
  const unsigned long long h = (unsigned long long)(unsigned)c * 0x9E3779B97F4A7C15ull;
  return R.b0 + (int)(h >> (64 - R.hbits));
}

__device__ __forceinline__ void cell_ijk(const RepD& R, double x, double y, double z, int& i, int& j, int& k) {
  i = (int)((x - R.box.xlo) / R.wx);
  j = (int)((y - R.box.ylo) / R.wy);
  k = (int)((z - R.box.zlo) / R.wz);
  i = min(max(i, 0), R.ncx - 1);
  j = min(max(j, 0), R.ncy - 1);
  k = min(max(k, 0), R.ncz - 1);
}

// Largest i < nsel with off[i] <= v (off strictly increasing, off[0] = 0).
__device__ __forceinline__ int seg_of(const int* off, int nsel, int v) {
  int lo = 0, hi = nsel - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Exclusive scan across the workgroup (blockDim a multiple of 64, <= 1024).
__device__ __forceinline__ int block_scan(int v, int* lds, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  int off = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    const int x = lds[w];
    if (w < wave) off += x;
    tot += x;
  }
  __syncthreads();
  total = tot;
  return off + incl - v;
}

// Work list of the replicas with rep_need set (MODE 0) or not yet stopped
// (MODE 1).  One workgroup.
// clear_rep_need / clear_need / zero_me (may be null): also clear the
// per-replica rebuild requests, the global request and a counter, which the
// FIRE iteration needs at this point (it saves separate launches).
template <int MODE>
__global__ void __launch_bounds__(kSelThreads) k_ens_select(const RepD* __restrict__ reps, int nrep,
                                                            const int* rep_need, Work w, const int* need,
                                                            int* clear_rep_need, int* clear_need, int* zero_me) {
  if (need && *need == 0) return;
  __shared__ int lds[kSelThreads / 64];
  if (clear_rep_need)
    for (int r = threadIdx.x; r < nrep; r += blockDim.x) clear_rep_need[r] = 0;
  if (threadIdx.x == 0) {
    if (clear_need) *clear_need = 0;
    if (zero_me) *zero_me = 0;
  }
  int cs = 0, cp = 0, cb = 0;
  for (int base = 0; base < nrep; base += blockDim.x) {
    const int r = base + threadIdx.x;
    int sel = 0, np_ = 0, nb_ = 0;
    if (r < nrep) {
      sel = MODE == 0 ? (rep_need[r] != 0) : (reps[r].frozen == 0);
      if (sel) { np_ = reps[r].n; nb_ = reps[r].nb + 1; }
    }
    int ts, tp, tb;
    const int os = block_scan(sel, lds, ts);
    const int op = block_scan(np_, lds, tp);
    const int ob = block_scan(nb_, lds, tb);
    if (sel) {
      w.rep[cs + os] = r;
      w.poff[cs + os] = cp + op;
      w.boff[cs + os] = cb + ob;
    }
    cs += ts; cp += tp; cb += tb;
  }
  if (threadIdx.x == 0) {
    w.poff[cs] = cp;
    w.boff[cs] = cb;
    w.cnt[0] = cs; w.cnt[1] = cp; w.cnt[2] = cb;
  }
}

// ---- neighbour-list rebuild of the replicas in the work list ---------------
// CellList::populate (src/system/cell_list.cpp:86-104): periodic wrap, then the
// clamped cell index; the bucket count is raised.
__global__ void k_ens_bin(EState e, int b, const RepD* __restrict__ reps, Work w, NListD nl, const int* need) {
  if (*need == 0) return;
  const int total = w.cnt[1], nsel = w.cnt[0];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    const int i = seg_of(w.poff, nsel, v);
    const RepD& R = reps[w.rep[i]];
    const int s = R.n0 + v - w.poff[i];
    D4 q = e.p[b][s];
    if (R.box.periodic) {
      const BoxD& B = R.box;
      bool wr = false;
      if (q.x < B.xlo) { q.x += B.Lx; wr = true; } else if (q.x > B.xhi) { q.x -= B.Lx; wr = true; }
      if (q.y < B.ylo) { q.y += B.Ly; wr = true; } else if (q.y > B.yhi) { q.y -= B.Ly; wr = true; }
      if (q.z < B.zlo) { q.z += B.Lz; wr = true; } else if (q.z > B.zhi) { q.z -= B.Lz; wr = true; }
      if (wr) e.p[b][s] = q;
    }
    int ci, cj, ck;
    cell_ijk(R, q.x, q.y, q.z, ci, cj, ck);
    e.cellc[s] = (ci << 20) | (cj << 10) | ck;
    const int bk = bucket_of(R, R.ncy * R.ncz * ci + R.ncz * cj + ck);
    nl.cell_of[s] = bk;
    atomicAdd(&nl.count[bk], 1);
  }
}

// Bucket entry v of the work list -> (list index, bucket index).
__device__ __forceinline__ int bucket_at(const RepD* __restrict__ reps, const Work& w, int i, int v) {
  return reps[w.rep[i]].b0 + (v - w.boff[i]);
}

// Tile-local exclusive scan of the listed replicas' bucket counts (bucket
// entries concatenated in list order); tile totals into bsum.
__global__ void __launch_bounds__(kTileThreads) k_ens_scan_tiles(const RepD* __restrict__ reps, Work w, NListD nl,
                                                                 const int* need) {
  if (*need == 0) return;
  __shared__ int lds[kTileThreads / 64];
  const int total = w.cnt[2], nsel = w.cnt[0];
  const int ntile = (total + kTile - 1) / kTile;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int v0 = tile * kTile + threadIdx.x * kTileItems;
    int cnt[kTileItems];
    int sum = 0;
    int i = v0 < total ? seg_of(w.boff, nsel, v0) : 0;
#pragma unroll
    for (int q = 0; q < kTileItems; ++q) {
      const int v = v0 + q;
      cnt[q] = 0;
      if (v < total) {
        while (i + 1 < nsel && w.boff[i + 1] <= v) ++i;
        cnt[q] = nl.count[bucket_at(reps, w, i, v)];
      }
      sum += cnt[q];
    }
    int tsum;
    int off = block_scan(sum, lds, tsum);
    i = v0 < total ? seg_of(w.boff, nsel, v0) : 0;
#pragma unroll
    for (int q = 0; q < kTileItems; ++q) {
      const int v = v0 + q;
      if (v < total) {
        while (i + 1 < nsel && w.boff[i + 1] <= v) ++i;
        nl.start[bucket_at(reps, w, i, v)] = off;
      }
      off += cnt[q];
    }
    if (threadIdx.x == 0) nl.bsum[tile] = tsum;
  }
}

// Adds the preceding tiles' totals and turns the list-wide prefix x into the
// replica's member position n0 + x - poff (each replica's entries sum to n,
// its end entry holds n0 + n).
__global__ void __launch_bounds__(kTileThreads) k_ens_scan_add(const RepD* __restrict__ reps, Work w, NListD nl,
                                                               const int* need) {
  if (*need == 0) return;
  __shared__ int lds[kTileThreads / 64];
  const int total = w.cnt[2], nsel = w.cnt[0];
  const int ntile = (total + kTile - 1) / kTile;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    int part = 0;
    for (int t = threadIdx.x; t < tile; t += blockDim.x) part += nl.bsum[t];
    int add;
    block_scan(part, lds, add);
    const int v0 = tile * kTile + threadIdx.x * kTileItems;
    int i = v0 < total ? seg_of(w.boff, nsel, v0) : 0;
    for (int q = 0; q < kTileItems; ++q) {
      const int v = v0 + q;
      if (v >= total) break;
      while (i + 1 < nsel && w.boff[i + 1] <= v) ++i;
      const int bk = bucket_at(reps, w, i, v);
      nl.start[bk] = reps[w.rep[i]].n0 + nl.start[bk] + add - w.poff[i];
    }
  }
}

__global__ void k_ens_scatter(const RepD* __restrict__ reps, Work w, NListD nl, const int* need) {
  if (*need == 0) return;
  const int total = w.cnt[1], nsel = w.cnt[0];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    const int i = seg_of(w.poff, nsel, v);
    const int s = reps[w.rep[i]].n0 + v - w.poff[i];
    const int c = nl.cell_of[s];
    nl.members[nl.start[c] + atomicSub(&nl.count[c], 1) - 1] = s;
  }
}

// Ascending slot order inside each bucket (deterministic lists).
__global__ void k_ens_sort(const RepD* __restrict__ reps, Work w, NListD nl, const int* need) {
  if (*need == 0) return;
  const int total = w.cnt[2], nsel = w.cnt[0];
  int local_max = 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    const int i = seg_of(w.boff, nsel, v);
    const RepD& R = reps[w.rep[i]];
    const int loc = v - w.boff[i];
    if (loc == R.nb) continue;  // end entry
    const int c = R.b0 + loc;
    const int a = nl.start[c], z = nl.start[c + 1];
    local_max = max(local_max, z - a);
    for (int m = a + 1; m < z; ++m) {
      const int key = nl.members[m];
      int j = m - 1;
      while (j >= a && nl.members[j] > key) { nl.members[j + 1] = nl.members[j]; --j; }
      nl.members[j + 1] = key;
    }
  }
  for (int o = 32; o > 0; o >>= 1) local_max = max(local_max, __shfl_xor(local_max, o, 64));
  if ((threadIdx.x & 63) == 0 && local_max > 0) atomic_max_if_above(&nl.flags[3], local_max);
}

// Cell a is in the periodic stencil of cell b along one axis of nc >= 3 cells.
__device__ __forceinline__ bool adjacent(int a, int b, int nc) {
  const int d = a - b;
  return (d >= -1 && d <= 1) || (a == 0 && b == nc - 1) || (a == nc - 1 && b == 0);
}

// Full list of every slot of a listed replica: the distinct buckets of its 27
// stencil cells, candidates restricted to stencil cells, d^2 < (rcut + pad)^2.
__global__ void __launch_bounds__(kBlock) k_ens_build(EState e, int b, const RepD* __restrict__ reps, Work w,
                                                      NListD nl, const int* need) {
  if (*need == 0) return;
  const int total = w.cnt[1], nsel = w.cnt[0];
  bool overflow = false;
  int kmax = 0;
  const D4* P = e.p[b];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    const int i = seg_of(w.poff, nsel, v);
    const RepD& R = reps[w.rep[i]];
    const int s = R.n0 + v - w.poff[i];
    const D4 pi = P[s];
    const int cc = e.cellc[s];
    const int ci = cc >> 20, cj = (cc >> 10) & 1023, ck = cc & 1023;
    int bk[27];
    int nbk = 0;
    for (int ix = -1; ix <= 1; ix++) {
      int iix = ci + ix;
      if (iix < 0) iix = R.ncx - 1; else if (iix == R.ncx) iix = 0;
      for (int iy = -1; iy <= 1; iy++) {
        int iiy = cj + iy;
        if (iiy < 0) iiy = R.ncy - 1; else if (iiy == R.ncy) iiy = 0;
        for (int iz = -1; iz <= 1; iz++) {
          int iiz = ck + iz;
          if (iiz < 0) iiz = R.ncz - 1; else if (iiz == R.ncz) iiz = 0;
          const int q = bucket_of(R, R.ncy * R.ncz * iix + R.ncz * iiy + iiz);
          bool dup = false;
          for (int u = 0; u < nbk; ++u) dup |= bk[u] == q;
          if (!dup) bk[nbk++] = q;
        }
      }
    }
    int k = 0;
    uint4 qb{0u, 0u, 0u, 0u};
    for (int u = 0; u < nbk; ++u) {
      const int a0 = nl.start[bk[u]], a1 = nl.start[bk[u] + 1];
      for (int m = a0; m < a1; ++m) {
        const int t = nl.members[m];
        if (t == s) continue;
        const int tc = e.cellc[t];
        if (!adjacent(tc >> 20, ci, R.ncx) || !adjacent((tc >> 10) & 1023, cj, R.ncy) ||
            !adjacent(tc & 1023, ck, R.ncz))
          continue;
        const D4 pt = P[t];
        double dx = pi.x - pt.x, dy = pi.y - pt.y, dz = pi.z - pt.z;
        min_image(R.box, dx, dy, dz);
        if (dx * dx + dy * dy + dz * dz < R.cut2) {
          list_push(nl, s, k, (unsigned)t, qb);
          ++k;
        }
      }
    }
    list_close(nl, s, k, qb);
    overflow |= k > nl.stride;
    kmax = max(kmax, k);
    nl.nnb[s] = min(k, nl.stride);
    nl.ox[s] = pi.x; nl.oy[s] = pi.y; nl.oz[s] = pi.z;
  }
  for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
  const bool any_over = __ballot(overflow) != 0;
  if ((threadIdx.x & 63) == 0) {
    if (any_over) atomicOr(&nl.flags[2], 1);
    if (kmax > 0) atomic_max_if_above(&nl.flags[5], kmax);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&nl.flags[4], 1);
}

// Start of a step: no replica needs a rebuild yet.
__global__ void k_ens_clear_need(int* rep_need, int nrep, int* need) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrep; r += gridDim.x * blockDim.x) rep_need[r] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *need = 0;
}

__global__ void k_ens_set_need(int* rep_need, int nrep, int* need) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrep; r += gridDim.x * blockDim.x) rep_need[r] = 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) *need = 1;
}

// Advances the noise call counter; optionally clears one counter.
__global__ void k_ens_advance(unsigned long long* call, int by, int* zero_me) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *call += (unsigned long long)by;
    if (zero_me) *zero_me = 0;
  }
}

__device__ __forceinline__ void flag_rebuild(const RepD& R, const NListD& nl, int s, int r, double x, double y,
                                             double z, int* rep_need, int* need) {
  double dx = nl.ox[s] - x, dy = nl.oy[s] - y, dz = nl.oz[s] - z;
  min_image(R.box, dx, dy, dz);
  if (!(dx * dx + dy * dy + dz * dz < R.pad2q)) {
    rep_need[r] = 1;
    *need = 1;
  }
}

// Soft force, its energy and (AL) the polar torque on slot s over its list.
// The soft sum follows pair order d = x_j - x_i, F_i -= ff d; the torque on i
// is J n_i x n_j per neighbour.
template <int BATCH, bool AL, bool EN>
__device__ __forceinline__ void ens_pair_sum(const EState& e, int b, const NListD& nl, const RepD& R, int s,
                                             const PS& p, double ri, double& fx, double& fy, double& fz,
                                             double& tx, double& ty, double& tz, double& en) {
  static_assert(BATCH % 4 == 0, "batches of whole four-entry groups");
  const int kn = nl.nnb[s];
  const uint4* G = reinterpret_cast<const uint4*>(nl.nbr);
  const D4* P = e.p[b];
  const D4* Dn = e.d[b];
  for (int k0 = 0; k0 < kn; k0 += BATCH) {
    int t[BATCH];
#pragma unroll
    for (int g = 0; g < BATCH / 4; ++g) {
      const uint4 q = G[(long)((k0 >> 2) + g) * nl.np + s];
      t[4 * g] = (int)q.x; t[4 * g + 1] = (int)q.y; t[4 * g + 2] = (int)q.z; t[4 * g + 3] = (int)q.w;
    }
#pragma unroll
    for (int u = 0; u < BATCH; ++u)
      if (k0 + u >= kn) t[u] = s;
    D4 pj[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; ++u) pj[u] = P[t[u]];
#pragma unroll
    for (int u = 0; u < BATCH; ++u) {
      if (k0 + u >= kn) break;
      double dx = pj[u].x - p.x, dy = pj[u].y - p.y, dz = pj[u].z - p.z;
      min_image(R.box, dx, dy, dz);
      const double r_sq = dx * dx + dy * dy + dz * dz;
      const double r = sqrt(r_sq);
      const double a = R.radii ? ri + pj[u].w : R.a;
      if (r < a) {
        const double diff = a - r;
        if (EN) en += 0.5 * R.k * diff * diff;
        const double ff = (r > 0.0) ? R.k * diff / r : R.k * diff;
        fx -= ff * dx; fy -= ff * dy; fz -= ff * dz;
      }
      if (AL && r_sq <= R.rc2_al) {
        const D4 dj = Dn[t[u]];
        tx += R.J * (p.ny * dj.z - p.nz * dj.y);
        ty += R.J * (p.nz * dj.x - p.nx * dj.z);
        tz += R.J * (p.nx * dj.y - p.ny * dj.x);
      }
    }
  }
}

__device__ __forceinline__ void load_ps(const EState& e, int b, int s, PS& p, double& rad) {
  const D4 q = e.p[b][s];
  const D4 d = e.d[b][s];
  p.x = q.x; p.y = q.y; p.z = q.z;
  p.nx = d.x; p.ny = d.y; p.nz = d.z;
  p.vx = p.vy = p.vz = 0.0;
  p.dix = p.diy = p.diz = 0;
  rad = q.w;
}

__device__ __forceinline__ void store_pd(const EState& e, int b, int s, const PS& p, double rad) {
  e.p[b][s] = D4{p.x, p.y, p.z, rad};
  e.d[b][s] = D4{p.nx, p.ny, p.nz, 0.0};
}

// One run-loop iteration of every replica (Brownian, T = 0): forces and
// torques at (x_t, n_t) from buffer b, the integrator update, the rebuild test
// on the integrated positions, then the next iteration's loop-head enforce
// (head != 0) and the store into buffer b ^ 1.  Velocities, forces, torques
// and omega are stored on the last iteration of a run only (last != 0): no
// step reads them back.
template <int BATCH>
__global__ void __launch_bounds__(kBlock) k_ens_brownian(EState e, int b, const RepD* __restrict__ reps, NListD nl,
                                                         int* rep_need, int* need, int n,
                                                         const unsigned long long* callp, int j, int head, int last) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int r = e.rep[s];
  const unsigned long long call = *callp + (unsigned long long)j;
  with_rep(reps, r, [&](const RepD& R) {
    PS p;
    double rad;
    load_ps(e, b, s, p, rad);
    double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, en = 0;
    const double ri = R.radii ? rad : 0.0;
    if (R.has_al) ens_pair_sum<BATCH, true, false>(e, b, nl, R, s, p, ri, fx, fy, fz, tx, ty, tz, en);
    else ens_pair_sum<BATCH, false, false>(e, b, nl, R, s, p, ri, fx, fy, fz, tx, ty, tz, en);
    const double dt = R.dt;
    const double fdx = R.v0 * p.nx + R.mu * fx, fdy = R.v0 * p.ny + R.mu * fy, fdz = R.v0 * p.nz + R.mu * fz;
    p.vx = fdx; p.vy = fdy; p.vz = fdz;
    p.x += dt * fdx; p.y += dt * fdy; p.z += dt * fdz;
    const bool fz0 = enforce<kSurfBasic>(R.box, R.con, p);
    const double om = R.mur * project_torque<kSurfBasic>(R.con, p, tx, ty, tz);
    double g, u;
    gauss1_u1(R.seed, TAG_BROWNIAN, (uint32_t)e.lid[s], call, g, u);
    rotate_director<kSurfBasic>(R.con, p, dt * om + R.stoch * g);
    flag_rebuild(R, nl, s, r, p.x, p.y, p.z, rep_need, need);
    if (last) {
      e.vx[s] = p.vx; e.vy[s] = p.vy; e.vz[s] = p.vz;
      e.fx[s] = fx; e.fy[s] = fy; e.fz[s] = fz0 ? 0.0 : fz;
      e.tx[s] = tx; e.ty[s] = ty; e.tz[s] = tz;
      e.omega[s] = om;
    }
    if (head) enforce<kSurfBasic>(R.box, R.con, p);
    store_pd(e, b ^ 1, s, p, rad);
  });
}

// age += dt once per iteration, repeated exactly as often as the run advanced.
__global__ void k_ens_age(EState e, const RepD* __restrict__ reps, int n, int iters) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double dt = reps[e.rep[s]].dt;
  double a = e.age[s];
  for (int i = 0; i < iters; ++i) a += dt;
  e.age[s] = a;
}

__global__ void k_ens_enforce(EState e, int b, const RepD* __restrict__ reps, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const RepD& R = reps[e.rep[s]];
  PS p;
  double rad;
  load_ps(e, b, s, p, rad);
  p.vx = e.vx[s]; p.vy = e.vy[s]; p.vz = e.vz[s];
  if (enforce<kSurfBasic>(R.box, R.con, p)) e.fz[s] = 0.0;
  store_pd(e, b, s, p, rad);
  e.vx[s] = p.vx; e.vy[s] = p.vy; e.vz[s] = p.vz;
}

// Forces, torques and per-slot energies (0.5 k (a - r)^2 per listed neighbour).
__global__ void k_ens_forces(EState e, int b, const RepD* __restrict__ reps, NListD nl, double* e_out, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const RepD& R = reps[e.rep[s]];
  PS p;
  double rad;
  load_ps(e, b, s, p, rad);
  double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, en = 0;
  const double ri = R.radii ? rad : 0.0;
  if (R.has_al) ens_pair_sum<4, true, true>(e, b, nl, R, s, p, ri, fx, fy, fz, tx, ty, tz, en);
  else ens_pair_sum<4, false, true>(e, b, nl, R, s, p, ri, fx, fy, fz, tx, ty, tz, en);
  e.fx[s] = fx; e.fy[s] = fy; e.fz[s] = fz;
  e.tx[s] = tx; e.ty[s] = ty; e.tz[s] = tz;
  if (e_out) e_out[s] = en;
}

// ---- FIRE ----------------------------------------------------------------
// Work-list slot v -> (replica, slot).
__device__ __forceinline__ int work_slot(const RepD* __restrict__ reps, const Work& w, int v, int& r) {
  const int i = seg_of(w.poff, w.cnt[0], v);
  r = w.rep[i];
  return reps[r].n0 + v - w.poff[i];
}

// The run loop's head enforce, then IntegratorFIRE::integrate's first loop:
// half kick with the current forces, drift, enforce
// (src/integrators/integrator_fire.cpp:50-70).  No torques act (soft
// potential only), so the director rotation is the identity.
__global__ void k_ens_fire_a(EState e, int b, const RepD* __restrict__ reps, Work w) {
  const int total = w.cnt[1];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    int r;
    const int s = work_slot(reps, w, v, r);
    const RepD& R = reps[r];
    const double dt = R.fdt, dt_2 = 0.5 * R.fdt;
    PS p;
    double rad;
    load_ps(e, b, s, p, rad);
    p.vx = e.vx[s]; p.vy = e.vy[s]; p.vz = e.vz[s];
    double fz = e.fz[s];
    if (enforce<kSurfBasic>(R.box, R.con, p)) fz = 0.0;
    p.vx += dt_2 * e.fx[s];
    p.vy += dt_2 * e.fy[s];
    p.vz += dt_2 * fz;
    p.x += dt * p.vx; p.y += dt * p.vy; p.z += dt * p.vz;
    enforce<kSurfBasic>(R.box, R.con, p);
    store_pd(e, b, s, p, rad);
    e.vx[s] = p.vx; e.vy[s] = p.vy; e.vz[s] = p.vz;
  }
}

// Forces at the drifted positions, second half kick, the per-slot terms of
// P = v.F, |F|^2, |v|^2, the energy and |F|, enforce of the velocities, and
// the rebuild test (src/integrators/integrator_fire.cpp:72-96).  Columns: 0 P,
// 1 F2, 2 V2, 3 E, 4 |F|.  Other threads read this slot's position as a
// neighbour, so positions and directors are not stored here: k_ens_fire_mix
// applies the same projection after every force has been formed.
template <int BATCH>
__global__ void __launch_bounds__(kBlock) k_ens_fire_b(EState e, int b, const RepD* __restrict__ reps, Work w,
                                                       NListD nl, double* cols, long ld, int* rep_need, int* need) {
  const int total = w.cnt[1];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    int r;
    const int s = work_slot(reps, w, v, r);
    with_rep(reps, r, [&](const RepD& R) {
      PS p;
      double rad;
      load_ps(e, b, s, p, rad);
      double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, en = 0;
      ens_pair_sum<BATCH, false, true>(e, b, nl, R, s, p, R.radii ? rad : 0.0, fx, fy, fz, tx, ty, tz, en);
      const double dt_2 = 0.5 * R.fdt;
      p.vx = e.vx[s] + dt_2 * fx;
      p.vy = e.vy[s] + dt_2 * fy;
      p.vz = e.vz[s] + dt_2 * fz;
      const double f2 = fx * fx + fy * fy + fz * fz;
      cols[s] = p.vx * fx + p.vy * fy + p.vz * fz;
      cols[ld + s] = f2;
      cols[2 * ld + s] = p.vx * p.vx + p.vy * p.vy + p.vz * p.vz;
      cols[3 * ld + s] = en;
      cols[4 * ld + s] = sqrt(f2);
      if (enforce<kSurfBasic>(R.box, R.con, p)) fz = 0.0;
      e.fx[s] = fx; e.fy[s] = fy; e.fz[s] = fz;
      e.vx[s] = p.vx; e.vy[s] = p.vy; e.vz[s] = p.vz;
      flag_rebuild(R, nl, s, r, p.x, p.y, p.z, rep_need, need);
    });
  }
}

// Per-replica sums of columns 0-3 and the max of column 4 for each listed
// replica: one workgroup per replica, a fixed summation tree.
__device__ void fire_update_one(RepD& R, int* remaining);

__global__ void __launch_bounds__(kBlock) k_ens_fire_reduce(const double* cols, long ld, RepD* reps, Work w,
                                                            int* remaining) {
  if ((int)blockIdx.x >= w.cnt[0]) return;
  RepD& R = reps[w.rep[blockIdx.x]];
  __shared__ double lds[5][kBlock / 64];
  double acc[5] = {0, 0, 0, 0, 0};
  for (int s = R.n0 + threadIdx.x; s < R.n0 + R.n; s += blockDim.x) {
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += cols[q * ld + s];
    acc[4] = fmax(acc[4], cols[4 * ld + s]);
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    double v = acc[q];
    for (int o = 32; o > 0; o >>= 1) {
      const double x = __shfl_xor(v, o, 64);
      v = q < 4 ? v + x : fmax(v, x);
    }
    if ((threadIdx.x & 63) == 0) lds[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[5] = {0, 0, 0, 0, 0};
    for (int wv = 0; wv < kBlock / 64; ++wv)
      for (int q = 0; q < 5; ++q) t[q] = q < 4 ? t[q] + lds[q][wv] : fmax(t[q], lds[q][wv]);
    R.P = t[0]; R.F2 = t[1]; R.V2 = t[2]; R.E = 0.5 * t[3]; R.Fmax = t[4];
    // every listed replica is running (not frozen): its FIRE update follows
    fire_update_one(R, remaining);
  }
}

// The scalar part of IntegratorFIRE::integrate per replica
// (src/integrators/integrator_fire.cpp:97-147), preceded by the stop rule
// max |F| < stop_fmax, which stops a converged replica.  *remaining counts the
// replicas still running.
__device__ void fire_update_one(RepD& R, int* remaining) {
  if (R.frozen) return;
  R.iters++;
  if (R.Fmax < R.stop_fmax) {
    R.frozen = 1;
    R.f1 = 1.0; R.f2 = 0.0; R.zero_v = 0;
    return;
  }
  atomicAdd(remaining, 1);
  const double Fn = sqrt(R.F2), Vn = sqrt(R.V2);
  const double sqrt_ndof = sqrt((double)(3 * R.n));
  if (Fn / sqrt_ndof < R.f_tol && fabs(R.E - R.E_old) < R.e_tol) {
    R.f1 = 1.0; R.f2 = 0.0; R.zero_v = 0;
    return;
  }
  const double inv_F = 1.0 / Fn;
  R.f1 = 1.0 - R.alpha;
  R.f2 = R.alpha * inv_F * Vn;
  R.zero_v = 0;
  if (R.P > 0.0) {
    R.last_neg++;
    if (R.last_neg > R.n_min) { R.fdt = fmin(R.fdt * R.f_inc, R.dt_max); R.alpha *= R.f_alpha; }
  } else {
    R.fdt *= R.f_dec;
    R.alpha = R.alpha0;
    R.last_neg = 0;
    R.zero_v = 1;
  }
  R.E_old = R.E;
}

// The position and director part of the second enforce (it depends on them
// only, so it equals what k_ens_fire_b computed), then, unless the replica
// stopped in this iteration, v = (1 - alpha) v + alpha |v| F / |F| with the
// coefficients from before the adaptation of dt and alpha; v = 0 after
// uphill motion.
__global__ void k_ens_fire_mix(EState e, int b, const RepD* __restrict__ reps, Work w) {
  const int total = w.cnt[1];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total; v += gridDim.x * blockDim.x) {
    int r;
    const int s = work_slot(reps, w, v, r);
    const RepD& R = reps[r];
    PS q;
    double rad;
    load_ps(e, b, s, q, rad);
    enforce<kSurfBasic>(R.box, R.con, q);
    store_pd(e, b, s, q, rad);
    if (R.frozen) continue;
    double ux = R.f1 * e.vx[s] + R.f2 * e.fx[s];
    double uy = R.f1 * e.vy[s] + R.f2 * e.fy[s];
    double uz = R.f1 * e.vz[s] + R.f2 * e.fz[s];
    if (R.zero_v) { ux = 0.0; uy = 0.0; uz = 0.0; }
    e.vx[s] = ux; e.vy[s] = uy; e.vz[s] = uz;
  }
}

__global__ void k_ens_morton(EState e, int b, const RepD* __restrict__ reps, unsigned long long* keys, int* slots,
                             int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int r = e.rep[s];
    const D4 q = e.p[b][s];
    int i, j, k;
    cell_ijk(reps[r], q.x, q.y, q.z, i, j, k);
    unsigned long long m = 0;
    for (int bit = 0; bit < 10; ++bit)
      m |= ((unsigned long long)((i >> bit) & 1) << (3 * bit)) |
           ((unsigned long long)((j >> bit) & 1) << (3 * bit + 1)) |
           ((unsigned long long)((k >> bit) & 1) << (3 * bit + 2));
    keys[s] = ((unsigned long long)r << 32) | m;
    slots[s] = s;
  }
}

__global__ void k_ens_gather4(D4* dst, const D4* src, const int* perm, int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) dst[s] = src[perm[s]];
}

// One component of a record column from / into a plain column.
__global__ void k_ens_put(D4* a, int comp, const double* src, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double* q = reinterpret_cast<double*>(a + s);
  q[comp] = src[s];
}

__global__ void k_ens_take(const D4* a, int comp, double* dst, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[s] = reinterpret_cast<const double*>(a + s)[comp];
}

int grid_cap(long n) {
  const int g = div_up(n, kBlock);
  return g < 1 ? 1 : (g > 2048 ? 2048 : g);
}

}  // namespace
}  // namespace samos

using namespace samos;

// ===========================================================================
// C-ABI
// ===========================================================================

namespace {

thread_local std::string g_ens_err;

#define ECHK(expr)                                                                \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      E->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
      return -2;                                                                  \
    }                                                                             \
  } while (0)

// Ints of a list of the given capacity: groups of four entries per slot, one
// spare group (a batch of eight may load the group after the last written).
size_t ens_nbr_ints(int stride, int np) { return ((size_t)(stride + 3) / 4 + 1) * 4 * (size_t)np; }

template <class T>
hipError_t ens_alloc(samos_ens* E, T** p, size_t count) {
  void* q = nullptr;
  hipError_t r = hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1));
  if (r == hipSuccess) {
    E->owned.push_back(q);
    *p = static_cast<T*>(q);
  }
  return r;
}

void ens_free(samos_ens* E, void* p) {
  if (!p) return;
  auto it = std::find(E->owned.begin(), E->owned.end(), p);
  if (it != E->owned.end()) E->owned.erase(it);
  hipFree(p);
}

void drop_graphs(samos_ens* E) {
  if (E->g_brown) { hipGraphExecDestroy(E->g_brown); E->g_brown = nullptr; }
  if (E->g_fire) { hipGraphExecDestroy(E->g_fire); E->g_fire = nullptr; }
}

int ens_upload(samos_ens* E) {
  ECHK(hipMemcpyAsync(E->d_rep, E->h_rep.data(), sizeof(RepD) * E->nrep, hipMemcpyHostToDevice, E->st));
  ECHK(hipStreamSynchronize(E->st));
  return 0;
}

int ens_download(samos_ens* E) {
  ECHK(hipMemcpyAsync(E->h_rep.data(), E->d_rep, sizeof(RepD) * E->nrep, hipMemcpyDeviceToHost, E->st));
  ECHK(hipStreamSynchronize(E->st));
  return 0;
}

int rep_grid(const samos_ens* E) { return div_up(std::max(E->nrep, 1), kBlock); }

std::pair<hipEvent_t, hipEvent_t> prof_begin(samos_ens* E) {
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (!E->prof) return ev;
  hipEventCreate(&ev.first);
  hipEventCreate(&ev.second);
  hipEventRecord(ev.first, E->st);
  return ev;
}

void prof_end(samos_ens* E, std::pair<hipEvent_t, hipEvent_t> ev, std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
  if (!ev.first) return;
  hipEventRecord(ev.second, E->st);
  v.push_back(ev);
}

// Rebuild of the replicas flagged in rep_need, every stage gated by *need.
void ens_rebuild(samos_ens* E, const int* need) {
  const int b = E->cur;
  const int g = grid_cap(E->n);
  const int gt = (int)std::min<long long>(E->ntiles, 1024);
  hipLaunchKernelGGL(k_ens_select<0>, dim3(1), dim3(kSelThreads), 0, E->st, E->d_rep, E->nrep, E->rep_need, E->wr,
                     need, (int*)nullptr, (int*)nullptr, (int*)nullptr);
  hipLaunchKernelGGL(k_ens_bin, dim3(g), dim3(kBlock), 0, E->st, E->e, b, E->d_rep, E->wr, E->nl, need);
  hipLaunchKernelGGL(k_ens_scan_tiles, dim3(gt), dim3(kTileThreads), 0, E->st, E->d_rep, E->wr, E->nl, need);
  hipLaunchKernelGGL(k_ens_scan_add, dim3(gt), dim3(kTileThreads), 0, E->st, E->d_rep, E->wr, E->nl, need);
  hipLaunchKernelGGL(k_ens_scatter, dim3(g), dim3(kBlock), 0, E->st, E->d_rep, E->wr, E->nl, need);
  hipLaunchKernelGGL(k_ens_sort, dim3(grid_cap(E->nl.ncell)), dim3(kBlock), 0, E->st, E->d_rep, E->wr, E->nl, need);
  hipLaunchKernelGGL(k_ens_build, dim3(g), dim3(kBlock), 0, E->st, E->e, b, E->d_rep, E->wr, E->nl, need);
}

}  // namespace

void samos::ens_enforce_async(samos_ens* E) {
  hipLaunchKernelGGL(k_ens_enforce, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->n);
}

int samos::ens_build_all(samos_ens* E) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    hipLaunchKernelGGL(k_ens_set_need, dim3(rep_grid(E)), dim3(kBlock), 0, E->st, E->rep_need, E->nrep, E->d_flags);
    ECHK(hipMemsetAsync(E->d_flags + 2, 0, 2 * sizeof(int), E->st));
    ECHK(hipMemsetAsync(E->d_flags + 5, 0, sizeof(int), E->st));
    ens_rebuild(E, E->d_flags);
    ECHK(hipGetLastError());
    int h[8];
    ECHK(hipMemcpyAsync(h, E->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, E->st));
    ECHK(hipStreamSynchronize(E->st));
    if (h[2] == 0 && E->nl.stride >= (int)(h[5] * 1.25) + 8) {
      hipLaunchKernelGGL(k_ens_clear_need, dim3(rep_grid(E)), dim3(kBlock), 0, E->st, E->rep_need, E->nrep,
                         E->d_flags);
      ECHK(hipStreamSynchronize(E->st));
      return 0;
    }
    drop_graphs(E);
    ens_free(E, E->nl.nbr);
    E->nl.nbr = nullptr;
    E->nl.stride = (int)std::ceil(h[5] * 1.5) + 16;
    ECHK(ens_alloc(E, &E->nl.nbr, ens_nbr_ints(E->nl.stride, E->np)));
  }
  E->err = "neighbour list capacity could not be established";
  return -1;
}

namespace {

int check_run_flags(samos_ens* E) {
  int h[8];
  ECHK(hipMemcpyAsync(h, E->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, E->st));
  ECHK(hipStreamSynchronize(E->st));
  if (h[2]) {
    // a list overflowed during the run: grow it for the next run and report
    ens_build_all(E);
    E->err = "neighbour list overflow during the run (capacity grown; repeat the run from a saved state)";
    return -1;
  }
  return 0;
}

// Where a field lives: a record column and component, or a plain column.
struct Col {
  D4* rec[2];
  int comp;
  double* plain;
};

bool ens_column(samos_ens* E, int field, Col& c) {
  EState& e = E->e;
  const int b = E->cur;
  c = Col{{nullptr, nullptr}, 0, nullptr};
  switch (field) {
    case SAMOS_X: case SAMOS_Y: case SAMOS_Z:
      c.rec[0] = e.p[b]; c.comp = field - SAMOS_X; return true;
    case SAMOS_RADIUS:
      c.rec[0] = e.p[b]; c.rec[1] = e.p[b ^ 1]; c.comp = 3; return true;
    case SAMOS_NX: case SAMOS_NY: case SAMOS_NZ:
      c.rec[0] = e.d[b]; c.comp = field - SAMOS_NX; return true;
    case SAMOS_VX: c.plain = e.vx; return true;
    case SAMOS_VY: c.plain = e.vy; return true;
    case SAMOS_VZ: c.plain = e.vz; return true;
    case SAMOS_FX: c.plain = e.fx; return true;
    case SAMOS_FY: c.plain = e.fy; return true;
    case SAMOS_FZ: c.plain = e.fz; return true;
    case SAMOS_TAUX: c.plain = e.tx; return true;
    case SAMOS_TAUY: c.plain = e.ty; return true;
    case SAMOS_TAUZ: c.plain = e.tz; return true;
    case SAMOS_OMEGA: c.plain = e.omega; return true;
    case SAMOS_AGE: c.plain = e.age; return true;
  }
  return false;
}

// One Brownian iteration: j is the iteration's offset from the device call
// counter, which advances at the end of each chunk.
void brownian_iter(samos_ens* E, int j, int head, int last) {
  hipLaunchKernelGGL(k_ens_clear_need, dim3(rep_grid(E)), dim3(kBlock), 0, E->st, E->rep_need, E->nrep, E->d_flags);
  const dim3 g(grid1(E->n)), bl(kBlock);
  auto ev = prof_begin(E);
  if (E->batch == 8)
    hipLaunchKernelGGL(k_ens_brownian<8>, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->nl, E->rep_need, E->d_flags,
                       E->n, E->d_call, j, head, last);
  else
    hipLaunchKernelGGL(k_ens_brownian<4>, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->nl, E->rep_need, E->d_flags,
                       E->n, E->d_call, j, head, last);
  prof_end(E, ev, E->ev_step);
  E->cur ^= 1;
  ev = prof_begin(E);
  ens_rebuild(E, E->d_flags);
  prof_end(E, ev, E->ev_rebuild);
}

// One FIRE iteration over the replicas still running; count_remaining clears
// the running-replica counter first so that it holds this iteration's count.
void fire_iter(samos_ens* E, bool count_remaining) {
  const dim3 g(grid_cap(E->n)), bl(kBlock);
  // the running-replica work list; clears the rebuild requests (and the
  // running-replica counter when counting)
  hipLaunchKernelGGL(k_ens_select<1>, dim3(1), dim3(kSelThreads), 0, E->st, E->d_rep, E->nrep, E->rep_need, E->wa,
                     (const int*)nullptr, E->rep_need, E->d_flags, count_remaining ? E->d_flags + 8 : (int*)nullptr);
  hipLaunchKernelGGL(k_ens_fire_a, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->wa);
  auto ev = prof_begin(E);
  if (E->batch == 8)
    hipLaunchKernelGGL(k_ens_fire_b<8>, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->wa, E->nl, E->cols, (long)E->np,
                       E->rep_need, E->d_flags);
  else
    hipLaunchKernelGGL(k_ens_fire_b<4>, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->wa, E->nl, E->cols, (long)E->np,
                       E->rep_need, E->d_flags);
  prof_end(E, ev, E->ev_step);
  // per-replica sums, then IntegratorFIRE's scalar update of that replica
  hipLaunchKernelGGL(k_ens_fire_reduce, dim3(E->nrep), bl, 0, E->st, E->cols, (long)E->np, E->d_rep, E->wa,
                     E->d_flags + 8);
  hipLaunchKernelGGL(k_ens_fire_mix, g, bl, 0, E->st, E->e, E->cur, E->d_rep, E->wa);
  ev = prof_begin(E);
  ens_rebuild(E, E->d_flags);
  prof_end(E, ev, E->ev_rebuild);
}

// Cell grid of every replica for cells of at least rcut + pad
// (CellList::CellList, src/system/cell_list.cpp:37-68) and the bucket layout
// (2^h >= n buckets and one end entry per replica).  Host table only.
int ens_grid(samos_ens* E, double rcut, double pad, long long* nb_total_out) {
  if (!(rcut > 0.0) || pad < 0.0) { E->err = "nlist needs rcut > 0 and pad >= 0"; return -1; }
  const double cut = rcut + pad;
  for (int r = 0; r < E->nrep; ++r) {
    const RepD& R = E->h_rep[r];
    const int ncx = (int)(R.box.Lx / cut), ncy = (int)(R.box.Ly / cut), ncz = (int)(R.box.Lz / cut);
    if (ncx < 3 || ncy < 3 || ncz < 3) {
      E->err = "replica " + std::to_string(r) + ": every box side must hold at least three cells of rcut + pad";
      return -1;
    }
    if (ncx > 1024 || ncy > 1024 || ncz > 1024) {
      E->err = "replica " + std::to_string(r) + ": more than 1024 cells per side";
      return -1;
    }
  }
  long long nb_total = 0;
  for (int r = 0; r < E->nrep; ++r) {
    RepD& R = E->h_rep[r];
    R.ncx = (int)(R.box.Lx / cut); R.ncy = (int)(R.box.Ly / cut); R.ncz = (int)(R.box.Lz / cut);
    R.wx = R.box.Lx / R.ncx; R.wy = R.box.Ly / R.ncy; R.wz = R.box.Lz / R.ncz;
    R.cut2 = cut * cut;
    R.pad2q = 0.25 * pad * pad;
    int h = 1;
    while ((1LL << h) < R.n) ++h;
    R.hbits = h;
    R.nb = 1 << h;
    R.b0 = (int)nb_total;
    nb_total += (long long)R.nb + 1;
  }
  if (nb_total >= (1LL << 30)) { E->err = "bucket table too large"; return -1; }
  E->rcut = rcut; E->pad = pad;
  *nb_total_out = nb_total;
  return 0;
}

int alloc_work(samos_ens* E, Work& w) {
  ECHK(ens_alloc(E, &w.rep, E->nrep));
  ECHK(ens_alloc(E, &w.poff, E->nrep + 1));
  ECHK(ens_alloc(E, &w.boff, E->nrep + 1));
  ECHK(ens_alloc(E, &w.cnt, 4));
  ECHK(hipMemsetAsync(w.cnt, 0, 4 * sizeof(int), E->st));
  return 0;
}

}  // namespace

extern "C" {

const char* samos_ens_last_error(const samos_ens* E) { return E ? E->err.c_str() : g_ens_err.c_str(); }

int samos_ens_create(samos_ens** out, int device, int nrep, const int* counts, const double* boxes,
                     const int* periodic, int constraint, const double* con_param) {
  if (!out) { g_ens_err = "null output handle"; return -1; }
  *out = nullptr;
  if (nrep <= 0 || !counts || !boxes || !periodic) { g_ens_err = "no replicas"; return -1; }
  if (constraint < 0 || constraint > 2) { g_ens_err = "unknown constraint"; return -1; }
  if (constraint == 2 && !con_param) { g_ens_err = "sphere radii missing"; return -1; }
  long long total = 0;
  for (int r = 0; r < nrep; ++r) {
    if (counts[r] <= 0) { g_ens_err = "replica " + std::to_string(r) + " has no particles"; return -1; }
    total += counts[r];
  }
  if (total >= (1LL << 30)) { g_ens_err = "too many particles"; return -1; }
  if (hipSetDevice(device) != hipSuccess) { g_ens_err = "hipSetDevice failed"; return -2; }
  samos_ens* E = new samos_ens();
  E->nrep = nrep;
  E->n = (int)total;
  E->np = ((E->n + 63) / 64) * 64;
  E->h_rep.assign(nrep, RepD{});
  long long n0 = 0;
  for (int r = 0; r < nrep; ++r) {
    RepD& R = E->h_rep[r];
    R.n0 = (int)n0; R.n = counts[r];
    n0 += counts[r];
    BoxD& B = R.box;
    B.Lx = boxes[3 * r]; B.Ly = boxes[3 * r + 1]; B.Lz = boxes[3 * r + 2];
    B.xlo = -0.5 * B.Lx; B.xhi = 0.5 * B.Lx; B.ylo = -0.5 * B.Ly; B.yhi = 0.5 * B.Ly;
    B.zlo = -0.5 * B.Lz; B.zhi = 0.5 * B.Lz;
    B.periodic = periodic[r] ? 1 : 0;
    R.con.kind = constraint;
    R.con.r = constraint == 2 ? con_param[r] : 0.0;
    R.con.zpos = (constraint == 1 && con_param) ? con_param[r] : 0.0;
    R.con.unlimited = 0;
    R.k = 1.0; R.a = 2.0; R.mu = 1.0; R.mur = 1.0; R.dt = 0.01;
  }
  const size_t np = E->np;
  auto fail = [&](const char* what) {
    g_ens_err = std::string("device allocation failed: ") + what;
    samos_ens_destroy(E);
    return -2;
  };
  if (hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking) != hipSuccess) return fail("stream");
  EState& e = E->e;
  for (int b = 0; b < 2; ++b) {
    if (ens_alloc(E, &e.p[b], np) != hipSuccess || ens_alloc(E, &e.d[b], np) != hipSuccess)
      return fail("position records");
    hipMemsetAsync(e.p[b], 0, sizeof(D4) * np, E->st);
    hipMemsetAsync(e.d[b], 0, sizeof(D4) * np, E->st);
  }
  double** dcols[] = {&e.vx, &e.vy, &e.vz, &e.fx, &e.fy, &e.fz, &e.tx, &e.ty, &e.tz, &e.omega, &e.age};
  for (double** c : dcols) {
    if (ens_alloc(E, c, np) != hipSuccess) return fail("columns");
    hipMemsetAsync(*c, 0, sizeof(double) * np, E->st);
  }
  if (ens_alloc(E, &e.rep, np) != hipSuccess || ens_alloc(E, &e.lid, np) != hipSuccess ||
      ens_alloc(E, &e.cellc, np) != hipSuccess)
    return fail("index columns");
  if (ens_alloc(E, &E->d_rep, nrep) != hipSuccess || ens_alloc(E, &E->rep_need, nrep) != hipSuccess ||
      ens_alloc(E, &E->d_flags, 16) != hipSuccess || ens_alloc(E, &E->d_call, 1) != hipSuccess ||
      ens_alloc(E, &E->cols, 5 * np) != hipSuccess)
    return fail("replica state");
  if (alloc_work(E, E->wr) || alloc_work(E, E->wa)) return fail("work lists");
  hipMemsetAsync(E->d_flags, 0, 16 * sizeof(int), E->st);
  hipMemsetAsync(E->rep_need, 0, sizeof(int) * nrep, E->st);
  hipMemsetAsync(E->d_call, 0, sizeof(unsigned long long), E->st);
  std::vector<int> repof(E->n), lid(E->n);
  E->h_gid.resize(E->n);
  for (int r = 0; r < nrep; ++r)
    for (int i = 0; i < counts[r]; ++i) {
      const int s = E->h_rep[r].n0 + i;
      repof[s] = r; lid[s] = i; E->h_gid[s] = s;
    }
  std::vector<double> ones(E->n, 1.0);
  hipMemcpyAsync(e.rep, repof.data(), sizeof(int) * E->n, hipMemcpyHostToDevice, E->st);
  hipMemcpyAsync(e.lid, lid.data(), sizeof(int) * E->n, hipMemcpyHostToDevice, E->st);
  hipMemcpyAsync(E->cols, ones.data(), sizeof(double) * E->n, hipMemcpyHostToDevice, E->st);
  for (int b = 0; b < 2; ++b)
    hipLaunchKernelGGL(k_ens_put, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, e.p[b], 3, E->cols, E->n);
  if (hipStreamSynchronize(E->st) != hipSuccess) return fail("initialisation");
  if (ens_upload(E)) return fail("replica table");
  *out = E;
  return 0;
}

int samos_ens_destroy(samos_ens* E) {
  if (!E) return 0;
  if (E->st) hipStreamSynchronize(E->st);
  drop_graphs(E);
  for (auto* v : {&E->ev_step, &E->ev_rebuild})
    for (auto& p : *v) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  for (void* p : E->owned) hipFree(p);
  if (E->st) hipStreamDestroy(E->st);
  delete E;
  return 0;
}

int samos_ens_num_particles(const samos_ens* E) { return E ? E->n : -1; }

int samos_ens_set_field(samos_ens* E, int field, const double* host) {
  Col c;
  if (!ens_column(E, field, c)) { E->err = "field " + std::to_string(field) + " is not held by an ensemble"; return -1; }
  std::vector<double> tmp(E->n);
  for (int s = 0; s < E->n; ++s) tmp[s] = host[E->h_gid[s]];
  double* dst = c.plain ? c.plain : E->cols;
  ECHK(hipMemcpyAsync(dst, tmp.data(), sizeof(double) * E->n, hipMemcpyHostToDevice, E->st));
  for (int k = 0; k < 2; ++k)
    if (c.rec[k])
      hipLaunchKernelGGL(k_ens_put, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, c.rec[k], c.comp, E->cols, E->n);
  ECHK(hipGetLastError());
  ECHK(hipStreamSynchronize(E->st));
  return 0;
}

int samos_ens_get_field(samos_ens* E, int field, double* host) {
  Col c;
  if (!ens_column(E, field, c)) { E->err = "field " + std::to_string(field) + " is not held by an ensemble"; return -1; }
  const double* src = c.plain;
  if (!src) {
    hipLaunchKernelGGL(k_ens_take, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, c.rec[0], c.comp, E->cols, E->n);
    src = E->cols;
  }
  std::vector<double> tmp(E->n);
  ECHK(hipMemcpyAsync(tmp.data(), src, sizeof(double) * E->n, hipMemcpyDeviceToHost, E->st));
  ECHK(hipStreamSynchronize(E->st));
  for (int s = 0; s < E->n; ++s) host[E->h_gid[s]] = tmp[s];
  return 0;
}

int samos_ens_set_params(samos_ens* E, const double* p) {
  for (int r = 0; r < E->nrep; ++r) {
    const double* q = p + 12 * r;
    if (!(q[10] > 0.0)) { E->err = "replica " + std::to_string(r) + ": timestep must be positive"; return -1; }
    if (q[8] < 0.0) { E->err = "replica " + std::to_string(r) + ": nu must be non-negative"; return -1; }
    if (q[0] < 0.0 || (q[2] == 0.0 && q[1] <= 0.0)) {
      E->err = "replica " + std::to_string(r) + ": soft potential needs k >= 0 and a > 0";
      return -1;
    }
  }
  for (int r = 0; r < E->nrep; ++r) {
    const double* q = p + 12 * r;
    RepD& R = E->h_rep[r];
    R.k = q[0]; R.a = q[1]; R.radii = q[2] != 0.0; R.J = q[3];
    R.has_al = (q[4] > 0.0 && q[3] != 0.0) ? 1 : 0;
    R.rc2_al = q[4] * q[4];
    R.v0 = q[5]; R.mu = q[6]; R.mur = q[7];
    R.seed = (uint32_t)(unsigned long long)q[9];
    R.dt = q[10];
    R.stoch = std::sqrt(q[8] * q[10]);
  }
  drop_graphs(E);
  return ens_upload(E);
}

int samos_ens_nlist(samos_ens* E, double rcut, double pad) {
  if (E->has_nlist) { E->err = "the ensemble already has a neighbour list"; return -1; }
  long long nb_total = 0;
  if (int rc = ens_grid(E, rcut, pad, &nb_total)) return rc;
  const double cut = rcut + pad;
  NListD& nl = E->nl;
  nl = NListD{};
  nl.ncell = (int)nb_total;
  nl.use_cell = 1;
  nl.np = E->np;
  nl.flags = E->d_flags;
  nl.cut2 = cut * cut;
  nl.pad2q = 0.25 * pad * pad;
  E->ntiles = (nb_total + kTile - 1) / kTile;
  const size_t np = E->np;
  ECHK(ens_alloc(E, &nl.count, nb_total));
  ECHK(ens_alloc(E, &nl.start, nb_total + 1));
  ECHK(ens_alloc(E, &nl.bsum, E->ntiles + 1));
  ECHK(ens_alloc(E, &nl.cell_of, np));
  ECHK(ens_alloc(E, &nl.members, np));
  ECHK(ens_alloc(E, &nl.nnb, np));
  ECHK(ens_alloc(E, &nl.ox, np));
  ECHK(ens_alloc(E, &nl.oy, np));
  ECHK(ens_alloc(E, &nl.oz, np));
  ECHK(hipMemsetAsync(nl.count, 0, sizeof(int) * nb_total, E->st));
  ECHK(hipMemsetAsync(nl.nnb, 0, sizeof(int) * np, E->st));
  nl.stride = 32;
  ECHK(ens_alloc(E, &nl.nbr, ens_nbr_ints(nl.stride, np)));
  if (int rc = ens_upload(E)) return rc;
  // Morton re-sort inside each replica: key = (replica, Morton code of the cell)
  {
    unsigned long long *k1 = nullptr, *k2 = nullptr;
    int *p1 = nullptr, *p2 = nullptr, *itmp = nullptr;
    D4* rtmp = nullptr;
    double* dtmp = nullptr;
    char* tmp = nullptr;
    const size_t tb = radix_sort_temp_bytes(E->n);
    ECHK(ens_alloc(E, &k1, np));
    ECHK(ens_alloc(E, &k2, np));
    ECHK(ens_alloc(E, &p1, np));
    ECHK(ens_alloc(E, &p2, np));
    ECHK(ens_alloc(E, &itmp, np));
    ECHK(ens_alloc(E, &dtmp, np));
    ECHK(ens_alloc(E, &rtmp, np));
    ECHK(ens_alloc(E, &tmp, tb + 256));
    EState& e = E->e;
    const int b = E->cur;
    hipLaunchKernelGGL(k_ens_morton, dim3(grid_cap(E->n)), dim3(kBlock), 0, E->st, e, b, E->d_rep, k1, p1, E->n);
    radix_sort_pairs(E->st, tmp, tb, k1, k2, p1, p2, E->n);
    for (D4* c : {e.p[b], e.d[b]}) {
      hipLaunchKernelGGL(k_ens_gather4, dim3(grid_cap(E->n)), dim3(kBlock), 0, E->st, rtmp, c, p2, E->n);
      ECHK(hipMemcpyAsync(c, rtmp, sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
    }
    ECHK(hipMemcpyAsync(e.p[b ^ 1], e.p[b], sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
    for (double* c : {e.vx, e.vy, e.vz, e.fx, e.fy, e.fz, e.tx, e.ty, e.tz, e.omega, e.age}) {
      launch_gather_d(E->st, dtmp, c, p2, E->n);
      ECHK(hipMemcpyAsync(c, dtmp, sizeof(double) * E->n, hipMemcpyDeviceToDevice, E->st));
    }
    launch_gather_i(E->st, itmp, e.lid, p2, E->n);
    ECHK(hipMemcpyAsync(e.lid, itmp, sizeof(int) * E->n, hipMemcpyDeviceToDevice, E->st));
    std::vector<int> perm(E->n);
    ECHK(hipMemcpyAsync(perm.data(), p2, sizeof(int) * E->n, hipMemcpyDeviceToHost, E->st));
    ECHK(hipStreamSynchronize(E->st));
    const std::vector<int> old = E->h_gid;
    for (int s = 0; s < E->n; ++s) E->h_gid[s] = old[perm[s]];
    for (void* q : {(void*)k1, (void*)k2, (void*)p1, (void*)p2, (void*)itmp, (void*)dtmp, (void*)rtmp, (void*)tmp})
      ens_free(E, q);
  }
  E->has_nlist = true;
  return ens_build_all(E);
}

int samos_ens_nlist_info(samos_ens* E, long long* out) {
  int h[8];
  ECHK(hipMemcpyAsync(h, E->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, E->st));
  ECHK(hipStreamSynchronize(E->st));
  out[0] = h[4]; out[1] = E->nl.stride; out[2] = E->nl.ncell; out[3] = h[3]; out[4] = h[5];
  out[5] = E->n; out[6] = E->nrep;
  return 0;
}

int samos_ens_pairs(samos_ens* E, int rep, int* out, long long cap, long long* count) {
  if (!E->has_nlist) { E->err = "no neighbour list"; return -1; }
  if (rep < 0 || rep >= E->nrep) { E->err = "replica out of range"; return -1; }
  const RepD& R = E->h_rep[rep];
  std::vector<int> nnb(R.n);
  ECHK(hipMemcpy(nnb.data(), E->nl.nnb + R.n0, sizeof(int) * R.n, hipMemcpyDeviceToHost));
  // the replica's rows of each four-entry group
  const int groups = (E->nl.stride + 3) / 4;
  std::vector<int> nbr((size_t)groups * 4 * R.n);
  for (int g = 0; g < groups; ++g)
    ECHK(hipMemcpy(nbr.data() + (size_t)g * 4 * R.n, E->nl.nbr + ((size_t)g * E->np + R.n0) * 4,
                   sizeof(int) * 4 * R.n, hipMemcpyDeviceToHost));
  std::vector<std::pair<int, int>> pr;
  for (int i = 0; i < R.n; ++i) {
    const int gi = E->h_gid[R.n0 + i] - R.n0;
    for (int k = 0; k < nnb[i]; ++k) {
      const int t = nbr[((size_t)(k >> 2) * R.n + i) * 4 + (k & 3)];
      const int gt = E->h_gid[t] - R.n0;
      if (gt > gi) pr.push_back({gi, gt});
    }
  }
  std::sort(pr.begin(), pr.end());
  *count = (long long)pr.size();
  if (out) {
    if ((long long)pr.size() > cap) { E->err = "output capacity too small"; return -1; }
    for (size_t k = 0; k < pr.size(); ++k) { out[2 * k] = pr[k].first; out[2 * k + 1] = pr[k].second; }
  }
  return 0;
}

int samos_ens_set_option(samos_ens* E, const char* name, double value) {
  const std::string k = name ? name : "";
  if (k == "gather_batch") {
    if (value != 4.0 && value != 8.0) { E->err = "gather_batch must be 4 or 8"; return -1; }
    E->batch = (int)value;
    drop_graphs(E);
    return 0;
  }
  E->err = "unknown option '" + k + "'";
  return -1;
}

int samos_ens_enforce(samos_ens* E) {
  hipLaunchKernelGGL(k_ens_enforce, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->n);
  ECHK(hipGetLastError());
  ECHK(hipStreamSynchronize(E->st));
  return 0;
}

int samos_ens_forces(samos_ens* E, double* energies) {
  if (!E->has_nlist) { E->err = "no neighbour list"; return -1; }
  hipLaunchKernelGGL(k_ens_forces, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->nl, E->cols,
                     E->n);
  ECHK(hipGetLastError());
  std::vector<double> en(E->n);
  ECHK(hipMemcpyAsync(en.data(), E->cols, sizeof(double) * E->n, hipMemcpyDeviceToHost, E->st));
  ECHK(hipStreamSynchronize(E->st));
  if (energies)
    for (int r = 0; r < E->nrep; ++r) {
      double t = 0.0;
      for (int s = E->h_rep[r].n0; s < E->h_rep[r].n0 + E->h_rep[r].n; ++s) t += en[s];
      energies[r] = 0.5 * t;
    }
  return 0;
}

int samos_ens_run_brownian(samos_ens* E, int steps, int flags) {
  if (!E->has_nlist) { E->err = "no neighbour list"; return -1; }
  if (steps < 0) { E->err = "negative step count"; return -1; }
  E->kind = 0;
  ECHK(hipMemsetAsync(E->d_flags + 2, 0, sizeof(int), E->st));
  hipLaunchKernelGGL(k_ens_enforce, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->n);
  const int iters = steps + 1;
  const int K = 32;  // even: a chunk returns the buffers to their parity
  int t = 0;
  if (!(flags & SAMOS_RUN_NO_GRAPH) && !E->prof && iters > K) {
    if (E->g_brown && (E->g_brown_cur != E->cur || E->g_brown_batch != E->batch)) {
      hipGraphExecDestroy(E->g_brown);
      E->g_brown = nullptr;
    }
    if (!E->g_brown) {
      const int cur0 = E->cur;
      hipGraph_t g = nullptr;
      ECHK(hipStreamBeginCapture(E->st, hipStreamCaptureModeThreadLocal));
      for (int j = 0; j < K; ++j) brownian_iter(E, j, 1, 0);
      hipLaunchKernelGGL(k_ens_advance, dim3(1), dim3(64), 0, E->st, E->d_call, K, (int*)nullptr);
      ECHK(hipStreamEndCapture(E->st, &g));
      const hipError_t ir = hipGraphInstantiate(&E->g_brown, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      E->cur = cur0;
      ECHK(ir);
      E->g_brown_cur = cur0;
      E->g_brown_batch = E->batch;
    }
    while (iters - t > K) {
      ECHK(hipGraphLaunch(E->g_brown, E->st));
      t += K;
    }
  }
  for (; t < iters; ++t) {
    brownian_iter(E, 0, t + 1 < iters, t + 1 == iters);
    hipLaunchKernelGGL(k_ens_advance, dim3(1), dim3(64), 0, E->st, E->d_call, 1, (int*)nullptr);
  }
  hipLaunchKernelGGL(k_ens_age, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->d_rep, E->n, iters);
  ECHK(hipGetLastError());
  ECHK(hipStreamSynchronize(E->st));
  return check_run_flags(E);
}

int samos_ens_autotune(samos_ens* E, int trial_steps, const double* pads, int npads, double* out) {
  if (!E->has_nlist) { E->err = "no neighbour list"; return -1; }
  if (trial_steps < 1 || npads < 1 || !pads) { E->err = "autotune needs trial steps and candidate pads"; return -1; }
  for (int i = 0; i < npads; ++i)
    if (!(pads[i] >= 0.0)) { E->err = "autotune: pads must be non-negative"; return -1; }
  // snapshot of the state (current buffers) and the noise counter
  EState& e = E->e;
  const int b = E->cur;
  double* snap = nullptr;
  D4* rsnap = nullptr;
  unsigned long long* call0 = nullptr;
  double* live[] = {e.vx, e.vy, e.vz, e.fx, e.fy, e.fz, e.tx, e.ty, e.tz, e.omega, e.age};
  const int ncol = (int)(sizeof(live) / sizeof(live[0]));
  ECHK(ens_alloc(E, &snap, (size_t)ncol * E->np));
  ECHK(ens_alloc(E, &rsnap, (size_t)2 * E->np));
  ECHK(ens_alloc(E, &call0, 1));
  for (int k = 0; k < ncol; ++k)
    ECHK(hipMemcpyAsync(snap + (size_t)k * E->np, live[k], sizeof(double) * E->n, hipMemcpyDeviceToDevice, E->st));
  ECHK(hipMemcpyAsync(rsnap, e.p[b], sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
  ECHK(hipMemcpyAsync(rsnap + E->np, e.d[b], sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
  ECHK(hipMemcpyAsync(call0, E->d_call, sizeof(unsigned long long), hipMemcpyDeviceToDevice, E->st));
  auto restore = [&]() -> int {
    const int c = E->cur;
    for (int k = 0; k < ncol; ++k)
      ECHK(hipMemcpyAsync(live[k], snap + (size_t)k * E->np, sizeof(double) * E->n, hipMemcpyDeviceToDevice, E->st));
    ECHK(hipMemcpyAsync(e.p[c], rsnap, sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
    ECHK(hipMemcpyAsync(e.d[c], rsnap + E->np, sizeof(D4) * E->n, hipMemcpyDeviceToDevice, E->st));
    ECHK(hipMemcpyAsync(E->d_call, call0, sizeof(unsigned long long), hipMemcpyDeviceToDevice, E->st));
    return 0;
  };
  const double rcut = E->rcut, pad0 = E->pad;
  const int batch0 = E->batch;
  double best_ms = 1e300, best_pad = pad0;
  int best_batch = batch0;
  hipEvent_t ev0, ev1;
  ECHK(hipEventCreate(&ev0));
  ECHK(hipEventCreate(&ev1));
  // whole graph chunks dominate the timed trial
  const int trial = std::max(trial_steps, 96);
  int rc = 0;
  for (int i = 0; i < npads && rc == 0; ++i) {
    long long nb = 0;
    if (ens_grid(E, rcut, pads[i], &nb)) continue;  // a pad the boxes cannot hold is skipped
    for (int batch : {4, 8}) {
      if ((rc = restore())) break;
      if ((rc = ens_upload(E))) break;
      if ((rc = ens_build_all(E))) break;
      E->batch = batch;
      // one untimed run captures the chunk graph for this variant
      if ((rc = samos_ens_run_brownian(E, 33, 0))) break;
      if ((rc = restore())) break;
      if ((rc = ens_build_all(E))) break;
      hipEventRecord(ev0, E->st);
      if ((rc = samos_ens_run_brownian(E, trial - 1, 0))) break;
      hipEventRecord(ev1, E->st);
      hipEventSynchronize(ev1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev0, ev1);
      const double per = ms / trial;
      if (per < best_ms) { best_ms = per; best_pad = pads[i]; best_batch = batch; }
    }
  }
  hipEventDestroy(ev0);
  hipEventDestroy(ev1);
  // leave the state as it was, with the chosen configuration
  long long nb = 0;
  if (rc == 0) rc = restore();
  if (rc == 0) rc = ens_grid(E, rcut, best_pad, &nb);
  if (rc == 0) rc = ens_upload(E);
  E->batch = best_batch;
  drop_graphs(E);
  if (rc == 0) rc = ens_build_all(E);
  ens_free(E, snap);
  ens_free(E, rsnap);
  ens_free(E, call0);
  if (rc) return rc;
  if (out) { out[0] = best_pad; out[1] = best_batch; out[2] = best_ms; }
  return 0;
}

int samos_ens_fire_setup(samos_ens* E, double dt0, double alpha, double alpha_scale, double f_tol, double e_tol,
                         double dt_inc, double dt_dec, double dt_max, int min_neg, double stop_fmax) {
  if (!E->has_nlist) { E->err = "no neighbour list"; return -1; }
  if (!(dt0 > 0.0)) { E->err = "FIRE timestep must be positive"; return -1; }
  if (int rc = ens_download(E)) return rc;
  for (const auto& R : E->h_rep)
    if (R.has_al) { E->err = "FIRE ensembles take the soft potential only (no aligner)"; return -1; }
  for (auto& R : E->h_rep) {
    R.fdt = dt0; R.alpha = alpha; R.alpha0 = alpha; R.f_alpha = alpha_scale; R.f_tol = f_tol; R.e_tol = e_tol;
    R.f_inc = dt_inc; R.f_dec = dt_dec; R.dt_max = dt_max; R.n_min = min_neg; R.last_neg = 0; R.frozen = 0;
    R.E_old = 1e15; R.iters = 0; R.stop_fmax = stop_fmax; R.f1 = 1.0; R.f2 = 0.0; R.zero_v = 0;
    R.P = R.F2 = R.V2 = R.E = 0.0; R.Fmax = INFINITY;
  }
  E->kind = 1;
  if (int rc = ens_upload(E)) return rc;
  // the run loop's first head: enforce, then forces at the enforced positions
  hipLaunchKernelGGL(k_ens_enforce, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->n);
  hipLaunchKernelGGL(k_ens_forces, dim3(grid1(E->n)), dim3(kBlock), 0, E->st, E->e, E->cur, E->d_rep, E->nl,
                     (double*)nullptr, E->n);
  ECHK(hipGetLastError());
  ECHK(hipStreamSynchronize(E->st));
  return 0;
}

int samos_ens_fire_run(samos_ens* E, int max_iter, int check, int* remaining, int* iters) {
  if (E->kind != 1) { E->err = "call samos_ens_fire_setup first"; return -1; }
  if (check < 1) check = 1;
  ECHK(hipMemsetAsync(E->d_flags + 2, 0, sizeof(int), E->st));
  int done = 0, rem = E->nrep;
  const bool graph = check >= 4 && !E->prof;
  if (graph && (!E->g_fire || E->g_fire_k != check || E->g_fire_cur != E->cur || E->g_fire_batch != E->batch)) {
    if (E->g_fire) { hipGraphExecDestroy(E->g_fire); E->g_fire = nullptr; }
    hipGraph_t g = nullptr;
    ECHK(hipStreamBeginCapture(E->st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < check; ++j) fire_iter(E, j == check - 1);
    ECHK(hipStreamEndCapture(E->st, &g));
    const hipError_t ir = hipGraphInstantiate(&E->g_fire, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    ECHK(ir);
    E->g_fire_k = check;
    E->g_fire_cur = E->cur;
    E->g_fire_batch = E->batch;
  }
  while (done < max_iter) {
    const int chunk = std::min(check, max_iter - done);
    if (graph && chunk == check) {
      ECHK(hipGraphLaunch(E->g_fire, E->st));
    } else {
      for (int j = 0; j < chunk; ++j) fire_iter(E, j == chunk - 1);
    }
    done += chunk;
    ECHK(hipMemcpyAsync(&rem, E->d_flags + 8, sizeof(int), hipMemcpyDeviceToHost, E->st));
    ECHK(hipStreamSynchronize(E->st));
    if (rem == 0) break;
  }
  ECHK(hipGetLastError());
  if (remaining) *remaining = rem;
  if (iters) *iters = done;
  return check_run_flags(E);
}

int samos_ens_fire_state(samos_ens* E, double* out) {
  if (int rc = ens_download(E)) return rc;
  for (int r = 0; r < E->nrep; ++r) {
    const RepD& R = E->h_rep[r];
    double* o = out + 8 * r;
    o[0] = (double)R.iters; o[1] = R.frozen; o[2] = R.E; o[3] = R.Fmax;
    o[4] = std::sqrt(R.F2); o[5] = R.fdt; o[6] = R.alpha; o[7] = R.P;
  }
  return 0;
}

int samos_ens_sync(samos_ens* E) {
  ECHK(hipStreamSynchronize(E->st));
  ECHK(hipGetLastError());
  return 0;
}

static void clear_events(std::vector<std::pair<hipEvent_t, hipEvent_t>>& v) {
  for (auto& p : v) { hipEventDestroy(p.first); hipEventDestroy(p.second); }
  v.clear();
}

int samos_ens_profile(samos_ens* E, int enable) {
  ECHK(hipStreamSynchronize(E->st));
  clear_events(E->ev_step);
  clear_events(E->ev_rebuild);
  E->prof = enable != 0;
  return 0;
}

int samos_ens_profile_read(samos_ens* E, const char* cls, double* ms_total, long long* launches) {
  const std::string c = cls ? cls : "";
  std::vector<std::pair<hipEvent_t, hipEvent_t>>* v =
      c == "step" ? &E->ev_step : (c == "rebuild" ? &E->ev_rebuild : nullptr);
  if (!v) { E->err = "unknown timing class '" + c + "' (step, rebuild)"; return -1; }
  ECHK(hipStreamSynchronize(E->st));
  double t = 0.0;
  for (auto& p : *v) {
    float ms = 0.f;
    ECHK(hipEventElapsedTime(&ms, p.first, p.second));
    t += ms;
  }
  *ms_total = t;
  *launches = (long long)v->size();
  return 0;
}

}  // extern "C"
