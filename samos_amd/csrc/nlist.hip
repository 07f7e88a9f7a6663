// Cell binning and Verlet neighbour-list build on the device.
//
// Follows the reference's cell list and build (src/system/cell_list.cpp:37-104,
// src/system/neighbour_list.cpp:117-302) but produces a FULL list (both
// directions of every pair) so that force kernels need no atomics.  The pair
// SET is identical to the reference's half list: pair (i, j) is kept when
// d2 = dx*dx + dy*dy + dz*dz < (rcut + pad)^2, with the same single-wrap
// minimum image, the same 27-cell stencil (duplicates included when a grid side
// has two cells) and the same truncating cell index.  This file is compiled
// with -ffp-contract=off so d2 rounds exactly as the reference's expression.
//
// A rebuild has five grid-wide phases: bin (cell index + count), scan of the
// counts, scatter into per-cell member ranges, per-cell ascending sort of the
// members (determinism), and the list build.  They exist both as separate
// kernels and fused into one cooperative kernel with grid barriers, which
// returns at once when the rebuild flag is clear.
//
// PLANAR: positions come from the interleaved (x, y) column; all particles sit
// on one plane, so dz = 0 and d2 is bit-identical to the 3-D expression.
#include <climits>

#include "kernels.hpp"
#include <hip/hip_cooperative_groups.h>
#include <hipcub/hipcub.hpp>

namespace cg = cooperative_groups;

namespace samos {

namespace {

__device__ __forceinline__ bool req_off(const int* req) { return req != nullptr && *req == 0; }

__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int gsize() { return gridDim.x * blockDim.x; }

// CellList::populate wrap + get_cell_idx (src/system/cell_list.cpp:72-104).
// The index is clamped to the grid: the reference indexes one past the grid for
// a particle exactly on the high face.
// Adds 1 to counter[key] for every active lane, one atomic per distinct key
// in the wavefront (slots in curve order share cells in runs).  Every lane
// of the wavefront must call it.
__device__ __forceinline__ void wave_count(int* counter, int key, bool active) {
  unsigned long long todo = __ballot(active);
  const int lane = threadIdx.x & 63;
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const int k0 = __shfl(key, leader, 64);
    const unsigned long long same = todo & __ballot(key == k0);
    if (lane == leader) atomicAdd(&counter[k0], __popcll(same));
    todo &= ~same;
  }
}

// Takes one from counter[key] per active lane and returns each lane a
// distinct index in [new value, old value): one atomic per distinct key.  The
// groups of equal keys are found first (register work only), then every
// group's first lane issues its atomic at once, so the wavefront waits for one
// round trip instead of one per distinct key.
__device__ __forceinline__ int wave_take(int* counter, int key, bool active) {
  unsigned long long todo = __ballot(active);
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int my_leader = lane;
  unsigned long long my_same = 0;
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const int k0 = __shfl(key, leader, 64);
    const unsigned long long same = todo & __ballot(key == k0);
    if (active && key == k0) { my_leader = leader; my_same = same; }
    todo &= ~same;
  }
  int old = 0;
  if (active && lane == my_leader) old = atomicSub(&counter[key], __popcll(my_same));
  old = __shfl(old, my_leader, 64);
  return active ? old - 1 - __popcll(my_same & below) : 0;
}

template <bool PLANAR>
__device__ void bin_phase(Fields& f, NListD& nl, const BoxD& b, int n) {
  // whole wavefronts step together so the per-cell counts can be aggregated
  const int lane = threadIdx.x & 63;
  for (int s0 = gtid() - lane; s0 < n; s0 += gsize()) {
    const int s = s0 + lane;
    const bool active = s < n;
    int c = 0;
    if (active) {
      double x, y, z;
      if (PLANAR) { const double2 p = f.xy[s]; x = p.x; y = p.y; z = f.z[s]; }
      else { x = f.x[s]; y = f.y[s]; z = f.z[s]; }
      if (b.periodic) {
        bool w = false;
        if (x < b.xlo) { x += b.Lx; w = true; } else if (x > b.xhi) { x -= b.Lx; w = true; }
        if (y < b.ylo) { y += b.Ly; w = true; } else if (y > b.yhi) { y -= b.Ly; w = true; }
        if (z < b.zlo) { z += b.Lz; w = true; } else if (z > b.zhi) { z -= b.Lz; w = true; }
        if (w) {
          f.x[s] = x; f.y[s] = y; f.z[s] = z;
          if (f.xy) f.xy[s] = make_double2(x, y);
        }
      }
      int i = (int)((x - b.xlo) / nl.wx);
      int j = (int)((y - b.ylo) / nl.wy);
      int k = (int)((z - b.zlo) / nl.wz);
      i = min(max(i, 0), nl.ncx - 1);
      j = min(max(j, 0), nl.ncy - 1);
      k = min(max(k, 0), nl.ncz - 1);
      c = nl.ncy * nl.ncz * i + nl.ncz * j + k;
      nl.cell_of[s] = c;
    }
    wave_count(nl.count, c, active);
  }
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread across the workgroup.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  int wave_off = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    int ws = lds[w];
    if (w < wave) wave_off += ws;
    tot += ws;
  }
  __syncthreads();
  total = tot;
  return wave_off + incl - v;
}

// Tile-local exclusive scan of the counts; tile sums into bsum.  Whole-block.
__device__ void scan_tile(NListD& nl, int tile, int* lds) {
  const int base = tile * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int idx = base + q;
    v[q] = idx < nl.ncell ? nl.count[idx] : 0;
    sum += v[q];
  }
  int total;
  int off = block_exclusive_scan(sum, lds, total);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int idx = base + q;
    if (idx < nl.ncell) nl.start[idx] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) nl.bsum[tile] = total;
}

// Adds the sum of all preceding tiles (a few hundred at most) to one tile.
__device__ void scan_add_tile(NListD& nl, int tile, int ntiles, int* lds) {
  int part = 0;
  for (int i = threadIdx.x; i < tile; i += blockDim.x) part += nl.bsum[i];
  int add;
  block_exclusive_scan(part, lds, add);
  const int base = tile * kScanTile;
  for (int q = threadIdx.x; q < kScanTile; q += blockDim.x) {
    const int idx = base + q;
    if (idx < nl.ncell) nl.start[idx] += add;
  }
  if (tile == ntiles - 1 && threadIdx.x == 0) nl.start[nl.ncell] = add + nl.bsum[tile];
}

// Counts return to zero as members are placed, so the next build starts clean.
__device__ void scatter_phase(NListD& nl, int n) {
  const int lane = threadIdx.x & 63;
  for (int s0 = gtid() - lane; s0 < n; s0 += gsize()) {
    const int s = s0 + lane;
    const bool active = s < n;
    const int c = active ? nl.cell_of[s] : 0;
    const int r = wave_take(nl.count, c, active);
    if (active) nl.members[nl.start[c] + r] = s;
  }
}

// Bitonic sorting network over N values held in registers (N a power of
// two; every index is a compile-time constant, so nothing spills).
template <int N>
__device__ __forceinline__ void bitonic_sort(int (&v)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const int lo = min(v[i], v[l]), hi = max(v[i], v[l]);
          v[i] = up ? lo : hi;
          v[l] = up ? hi : lo;
        }
      }
    }
  }
}

// One cell's members (m <= N) sorted in registers: all loads are issued
// together instead of one dependent global access per insertion step.
template <int N>
__device__ __forceinline__ void sort_cell_registers(int* mem, int m) {
  int v[N];
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = q < m ? mem[q] : INT_MAX;
  bitonic_sort<N>(v);
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (q < m) mem[q] = v[q];
}

// Ascending slot order inside each cell makes the list deterministic.
__device__ void sort_phase(NListD& nl) {
  int local_max = 0;
  for (int c = gtid(); c < nl.ncell; c += gsize()) {
    const int a = nl.start[c], e = nl.start[c + 1];
    const int m = e - a;
    local_max = max(local_max, m);
    if (m <= 1) continue;
    if (m <= 8) { sort_cell_registers<8>(nl.members + a, m); continue; }
    if (m <= 16) { sort_cell_registers<16>(nl.members + a, m); continue; }
    for (int i = a + 1; i < e; ++i) {
      const int key = nl.members[i];
      int j = i - 1;
      while (j >= a && nl.members[j] > key) { nl.members[j + 1] = nl.members[j]; --j; }
      nl.members[j + 1] = key;
    }
  }
  for (int o = 32; o > 0; o >>= 1) local_max = max(local_max, __shfl_xor(local_max, o, 64));
  if ((threadIdx.x & 63) == 0 && local_max > 0) atomic_max_if_above(&nl.flags[3], local_max);
}

// After a slot's last entry (k of them found): the last group completed with
// the own index, and one more group of it when the list would end inside an
// 8-entry batch (the pair loops load whole batches and skip the padding).
__device__ __forceinline__ void list_finish(const NListD& nl, int s, int k, uint4& qb) {
  list_close(nl, s, k, qb);
  const int groups = (min(k, nl.stride) + 3) >> 2;
  if (groups & 1) group_store(nl, groups, s, uint4{(unsigned)s, (unsigned)s, (unsigned)s, (unsigned)s});
}

// Flags: [2] overflow, [4] builds, [5] largest neighbour count.  One atomic per
// wavefront.
// ovf2 (optional): a second overflow flag (the captured chunk's).
__device__ __forceinline__ void build_epilogue(NListD& nl, bool overflow, int kmax, int* ovf2 = nullptr) {
  for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
  const bool any_over = __ballot(overflow) != 0;
  if ((threadIdx.x & 63) == 0) {
    if (any_over) {
      atomicOr(&nl.flags[2], 1);
      if (ovf2) atomicOr(ovf2, 1);
    }
    if (kmax > 0) atomic_max_if_above(&nl.flags[5], kmax);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&nl.flags[4], 1);
}

// NeighbourList::build_cell (src/system/neighbour_list.cpp:216-273), full list:
// candidate t of slot s is kept when d2 < (rcut + pad)^2; the coordination
// counts the kept candidates closer than r_s + r_t.
template <bool PLANAR>
__device__ __forceinline__ void build_candidate(const NListD& nl, const BoxD& b, int s, double xi, double yi,
                                                double zi, double ri, bool wrap, int t, double xj, double yj,
                                                double zj, double rj, int& k, int& coord, uint4& qb) {
  if (t == s) return;
  double dx = xi - xj, dy = yi - yj, dz = PLANAR ? 0.0 : zi - zj;
  if (wrap) {
    if (PLANAR) min_image_xy(b, dx, dy);
    else min_image(b, dx, dy, dz);
  }
  // planar: dz = 0, and adding +0 to a sum of squares changes no bit
  const double d2 = PLANAR ? dx * dx + dy * dy : dx * dx + dy * dy + dz * dz;
  if (d2 < nl.cut2) {
    list_push(nl, s, k, (unsigned)t, qb);
    ++k;
    const double r = ri + rj;
    if (d2 < r * r) ++coord;
  }
}

// Candidates from the 3 x 3 (x 3) cells around the particle's own, read from
// global memory: cell members are consumed in batches so several candidate
// loads are in flight per thread.
template <bool PLANAR>
__device__ __forceinline__ void build_cells_global(const Fields& f, const NListD& nl, const BoxD& b, int s, int c,
                                                   double xi, double yi, double zi, double ri, bool runi, bool wrap,
                                                   int& k, int& coord, uint4& qb) {
  constexpr int U = 8;
  const int ci = c / (nl.ncy * nl.ncz), cj = (c / nl.ncz) % nl.ncy, ck = c % nl.ncz;
  // Planar: every particle is in the same z-layer ck, and a used cell grid
  // has ncz >= 2, so the stencil's other z-layers hold nobody and the own
  // layer appears exactly once; only it is visited.
  const int iz_lo = PLANAR ? 0 : -1, iz_hi = PLANAR ? 0 : 1;
  for (int ix = -1; ix <= 1; ix++) {
    int iix = ci + ix;
    if (iix < 0) iix = nl.ncx - 1; else if (iix == nl.ncx) iix = 0;
    for (int iy = -1; iy <= 1; iy++) {
      int iiy = cj + iy;
      if (iiy < 0) iiy = nl.ncy - 1; else if (iiy == nl.ncy) iiy = 0;
      for (int iz = iz_lo; iz <= iz_hi; iz++) {
        int iiz = ck + iz;
        if (iiz < 0) iiz = nl.ncz - 1; else if (iiz == nl.ncz) iiz = 0;
        const int cc = nl.ncy * nl.ncz * iix + nl.ncz * iiy + iiz;
        const int a = nl.start[cc], e = nl.start[cc + 1];
        for (int m0 = a; m0 < e; m0 += U) {
          int t[U];
          double xj[U], yj[U], zj[U], rj[U];
#pragma unroll
          for (int u = 0; u < U; ++u) t[u] = (m0 + u < e) ? nl.members[m0 + u] : s;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (PLANAR) { const double2 p = f.xy[t[u]]; xj[u] = p.x; yj[u] = p.y; zj[u] = zi; }
            else { xj[u] = f.x[t[u]]; yj[u] = f.y[t[u]]; zj[u] = f.z[t[u]]; }
            rj[u] = runi ? nl.r_uni : f.radius[t[u]];
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            if (m0 + u >= e) break;
            build_candidate<PLANAR>(nl, b, s, xi, yi, zi, ri, wrap, t[u], xj[u], yj[u], zj[u], rj[u], k, coord,
                                    qb);
          }
        }
      }
    }
  }
}

// Planar builds stage candidates per wavefront: the 64 slots of a wavefront
// (consecutive along the storage curve) lie in a small block of cells, so the
// members and positions of that block plus one cell around it are loaded into
// LDS once, coalesced, and every lane reads its 3 x 3 stencil from there in
// the same order as the global path (the same list, entry for entry).  A
// wavefront whose block touches the grid's edge (stencil cells that wrap) or
// does not fit the stage takes the global path.
constexpr int kStageCells = 64;   // cells of a wavefront's block, halo included (one lane each)
constexpr int kStageCand = 448;   // staged candidates per wavefront
static_assert(kStageCells <= 64, "the cell table is filled and scanned by the lanes of one wavefront");
struct WaveStage {
  int cstart[kStageCells + 1];    // first staged entry of each block cell; [nq] = total
  int gstart[kStageCells];        // its first index in members[]
  int mem[kStageCand];            // staged slots
  double2 pos[kStageCand];        // their (x, y)
};

// LDS written by some lanes of a wavefront and read by others of the same one.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Every lane of the wavefront calls it (lanes past n with active false).
// Returns false (nothing staged) when the wavefront takes the global path.
__device__ bool stage_wave_block(const Fields& f, const NListD& nl, WaveStage& S, bool active, int c, int& bx0,
                                 int& by0, int& bw) {
  const int lane = threadIdx.x & 63;
  const int ci = c / (nl.ncy * nl.ncz), cj = (c / nl.ncz) % nl.ncy, ck = c % nl.ncz;
  const int imin = wave_min_i(active ? ci : INT_MAX), imax = wave_max_i(active ? ci : INT_MIN);
  const int jmin = wave_min_i(active ? cj : INT_MAX), jmax = wave_max_i(active ? cj : INT_MIN);
  const int kmin = wave_min_i(active ? ck : INT_MAX), kmax = wave_max_i(active ? ck : INT_MIN);
  if (imin > imax) return false;  // no active lane
  bx0 = imin - 1; by0 = jmin - 1;
  bw = imax - imin + 3;
  const int bh = jmax - jmin + 3, nq = bw * bh;
  if (kmin != kmax || bx0 < 0 || imax + 1 >= nl.ncx || by0 < 0 || jmax + 1 >= nl.ncy || nq > kStageCells)
    return false;
  int cnt = 0, gs = 0;
  if (lane < nq) {
    const int cc = nl.ncy * nl.ncz * (bx0 + lane % bw) + nl.ncz * (by0 + lane / bw) + kmin;
    gs = nl.start[cc];
    cnt = nl.start[cc + 1] - gs;
  }
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const int total = __shfl(incl, 63, 64);
  if (total > kStageCand) return false;
  if (lane < nq) { S.cstart[lane] = incl - cnt; S.gstart[lane] = gs; }
  if (lane == 0) S.cstart[nq] = total;
  wave_lds_sync();
  for (int q = lane; q < total; q += 64) {
    int lo = 0, hi = nq - 1;  // the last block cell starting at or before q
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (S.cstart[mid] <= q) lo = mid; else hi = mid - 1;
    }
    const int t = nl.members[S.gstart[lo] + (q - S.cstart[lo])];
    S.mem[q] = t;
    S.pos[q] = f.xy[t];
  }
  wave_lds_sync();
  return true;
}

// The same stencil walk as build_cells_global, from the wavefront's stage.
__device__ __forceinline__ void build_cells_staged(const NListD& nl, const BoxD& b, const WaveStage& S, int bx0,
                                                   int by0, int bw, int s, int c, double xi, double yi, double ri,
                                                   double rj_u, bool wrap, int& k, int& coord, uint4& qb) {
  const int ci = c / (nl.ncy * nl.ncz), cj = (c / nl.ncz) % nl.ncy;
  for (int ix = -1; ix <= 1; ix++) {
    for (int iy = -1; iy <= 1; iy++) {
      const int q = (ci + ix - bx0) + (cj + iy - by0) * bw;
      const int a = S.cstart[q], e = S.cstart[q + 1];
      for (int m = a; m < e; ++m) {
        // both LDS reads issued together; the own slot (at distance 0) is
        // excluded after the distance test, as build_candidate does first
        const double2 p = S.pos[m];
        const int t = S.mem[m];
        double dx = xi - p.x, dy = yi - p.y;
        if (wrap) min_image_xy(b, dx, dy);
        const double d2 = dx * dx + dy * dy;
        if (d2 < nl.cut2 && t != s) {
          list_push(nl, s, k, (unsigned)t, qb);
          ++k;
          const double r = ri + rj_u;
          if (d2 < r * r) ++coord;
        }
      }
    }
  }
}

template <bool PLANAR>
__device__ void build_phase(Fields& f, NListD& nl, const BoxD& b, int n, int* ovf2 = nullptr) {
  bool overflow = false;
  int kmax = 0;
  __shared__ WaveStage stage[PLANAR ? kBlock / 64 : 1];
  const int lane = threadIdx.x & 63;
  // whole wavefronts step together (the planar stage is per wavefront)
  for (int s0 = xcd_block() * blockDim.x + threadIdx.x - lane; s0 < n; s0 += gsize()) {
    const int s = s0 + lane;
    const bool active = s < n;
    const int c = active ? nl.cell_of[s] : 0;
    const bool runi = nl.r_uni >= 0.0;
    int bx0 = 0, by0 = 0, bw = 0;
    // planar with one common radius: candidates from the wavefront's stage
    const bool staged = PLANAR && runi && stage_wave_block(f, nl, stage[PLANAR ? threadIdx.x >> 6 : 0], active, c,
                                                           bx0, by0, bw);
    if (!active) continue;
    double xi, yi, zi;
    if (PLANAR) { const double2 p = f.xy[s]; xi = p.x; yi = p.y; zi = f.z[s]; }
    else { xi = f.x[s]; yi = f.y[s]; zi = f.z[s]; }
    // one common radius (the usual case): the coordination test needs no
    // per-candidate radius gather
    const double ri = runi ? nl.r_uni : f.radius[s];
    int k = 0, coord = 0;
    uint4 qb{0u, 0u, 0u, 0u};
    // Candidates come from the 3 x 3 (x 3) cells around the particle's own,
    // all within two cell widths of it per axis.  Farther than that from every
    // periodic face, none of them is across a face, the raw difference is the
    // minimum image and the image step is skipped (an exact no-op there).
    bool wrap = b.periodic != 0;
    if (wrap) {
      const double Rx = 2.0 * nl.wx, Ry = 2.0 * nl.wy, Rz = 2.0 * nl.wz;
      bool inner = xi - b.xlo > Rx && b.xhi - xi > Rx && yi - b.ylo > Ry && b.yhi - yi > Ry;
      if (!PLANAR) inner = inner && zi - b.zlo > Rz && b.zhi - zi > Rz;
      wrap = !inner;
    }
    if (PLANAR && staged)
      build_cells_staged(nl, b, stage[PLANAR ? threadIdx.x >> 6 : 0], bx0, by0, bw, s, c, xi, yi, ri, nl.r_uni, wrap,
                         k, coord, qb);
    else
      build_cells_global<PLANAR>(f, nl, b, s, c, xi, yi, zi, ri, runi, wrap, k, coord, qb);
    if (k > nl.stride) overflow = true;
    kmax = max(kmax, k);
    nl.nnb[s] = min(k, nl.stride);
    list_finish(nl, s, k, qb);
    f.coord[s] = coord;
    nl.ox[s] = xi; nl.oy[s] = yi; nl.oz[s] = zi;
  }
  build_epilogue(nl, overflow, kmax, ovf2);
}

// ---- standalone stage kernels ---------------------------------------------
template <bool PLANAR>
__global__ void k_bin(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  bin_phase<PLANAR>(f, nl, b, n);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(NListD nl, const int* req) {
  if (req_off(req)) return;
  __shared__ int lds[kScanThreads / 64];
  scan_tile(nl, blockIdx.x, lds);
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(NListD nl, int nb, const int* req) {
  if (req_off(req)) return;
  __shared__ int lds[kScanThreads / 64];
  scan_add_tile(nl, blockIdx.x, nb, lds);
}

__global__ void k_scatter(NListD nl, int n, const int* req) {
  if (req_off(req)) return;
  scatter_phase(nl, n);
}

__global__ void k_cell_sort(NListD nl, const int* req) {
  if (req_off(req)) return;
  sort_phase(nl);
}

template <bool PLANAR>
__global__ void k_build(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  build_phase<PLANAR>(f, nl, b, n);
}

// ---- the end of a captured step chunk ----------------------------------------
// First kernel: the chunk's positions (and directors) settle into buffer 0
// when its last slot left them in buffer 1, the rebuild request of its slots
// is recorded, and when there is one the slots are binned, each by the thread
// that has just copied it.
template <bool PLANAR>
__global__ void k_settle_bin(ChunkCtl* ck, Fields f0, Fields f1, NListD nl, BoxD b, int n, int directors) {
  const bool odd = (ck->last & 1) != 0;
  const int rb = ck->halt | ck->req[0] | ck->req[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) ck->rebuild = rb;
  if (!odd && !rb) return;
  if (odd) {
    for (int i = gtid(); i < n; i += gsize()) {
      if (PLANAR) {
        f0.xy[i] = f1.xy[i];
      } else {
        f0.x[i] = f1.x[i]; f0.y[i] = f1.y[i]; f0.z[i] = f1.z[i];
      }
      if (directors) {
        f0.nx[i] = f1.nx[i]; f0.ny[i] = f1.ny[i];
        if (!PLANAR) f0.nz[i] = f1.nz[i];
      }
    }
  }
  if (rb) bin_phase<PLANAR>(f0, nl, b, n);
}

// Last kernel: the chunk's bookkeeping (run step and noise calls advanced by
// the slots that ran, requests cleared), then the list build when requested;
// an overflow of it is recorded in the chunk's stop word as well.
template <bool PLANAR>
__global__ void k_build_finish(Fields f, NListD nl, BoxD b, int n, ChunkCtl* ck, ClockD* clk) {
  if (blockIdx.x == 0) {
    const int last = ck->last;
    __syncthreads();
    if (threadIdx.x < 8) clk->calls[threadIdx.x] += (unsigned long long)last;
    if (threadIdx.x == 0) {
      clk->run_step += last;
      ck->base += last;
      ck->halt = 0; ck->req[0] = 0; ck->req[1] = 0; ck->last = 0;
    }
  }
  if (ck->rebuild == 0) return;
  build_phase<PLANAR>(f, nl, b, n, &ck->stop);
}

// ---- the whole rebuild in one cooperative launch ----------------------------
// Every thread reads the same flag, so either the whole grid returns at once or
// the whole grid reaches every barrier.
template <bool PLANAR>
__global__ void __launch_bounds__(kScanThreads) k_rebuild_coop(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  cg::grid_group grid = cg::this_grid();
  __shared__ int lds[kScanThreads / 64];
  const int ntiles = (nl.ncell + kScanTile - 1) / kScanTile;
  bin_phase<PLANAR>(f, nl, b, n);
  grid.sync();
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) scan_tile(nl, tile, lds);
  grid.sync();
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) scan_add_tile(nl, tile, ntiles, lds);
  grid.sync();
  scatter_phase(nl, n);
  grid.sync();
  sort_phase(nl);
  grid.sync();
  build_phase<PLANAR>(f, nl, b, n);
}

// NeighbourList::build_nsq (src/system/neighbour_list.cpp:178-213), full list.
__global__ void k_build_nsq(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  bool overflow = false;
  int kmax = 0;
  for (int s = gtid(); s < n; s += gsize()) {
    const double xi = f.x[s], yi = f.y[s], zi = f.z[s], ri = f.radius[s];
    int k = 0, coord = 0;
    uint4 qb{0u, 0u, 0u, 0u};
    for (int t = 0; t < n; ++t) {
      if (t == s) continue;
      double dx = xi - f.x[t], dy = yi - f.y[t], dz = zi - f.z[t];
      min_image(b, dx, dy, dz);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < nl.cut2) {
        list_push(nl, s, k, (unsigned)t, qb);
        ++k;
        const double r = ri + f.radius[t];
        if (d2 < r * r) ++coord;
      }
    }
    if (k > nl.stride) overflow = true;
    kmax = max(kmax, k);
    nl.nnb[s] = min(k, nl.stride);
    list_finish(nl, s, k, qb);
    f.coord[s] = coord;
    nl.ox[s] = xi; nl.oy[s] = yi; nl.oz[s] = zi;
  }
  build_epilogue(nl, overflow, kmax);
}

// NeighbourList::need_update (src/system/neighbour_list.cpp:278-302) for every
// particle, reduced with a wavefront ballot: one atomic per wavefront that saw
// a particle move by pad/2 or more.
__global__ void k_need_update(Fields f, NListD nl, BoxD b, int n, int* req_out) {
  bool need = false;
  for (int s = gtid(); s < n; s += gsize()) {
    double dx = nl.ox[s] - f.x[s], dy = nl.oy[s] - f.y[s], dz = nl.oz[s] - f.z[s];
    min_image(b, dx, dy, dz);
    need |= !(dx * dx + dy * dy + dz * dz < nl.pad2q);
  }
  if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomic_set_once(req_out);
}

__device__ __forceinline__ unsigned long long spread3(unsigned long long v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// Position of cell (x, y) along the Hilbert curve over a 2^order x 2^order
// grid: the iterative form that reads one bit of each coordinate per level,
// from the top, adds the quadrant's offset and rotates or reflects the
// coordinates into the quadrant's frame (Hilbert curve, "xy2d").
__device__ __forceinline__ unsigned long long hilbert2(unsigned x, unsigned y, int order) {
  const unsigned side = 1u << order;
  unsigned long long d = 0;
  for (unsigned h = side >> 1; h > 0; h >>= 1) {
    const unsigned rx = (x & h) ? 1u : 0u, ry = (y & h) ? 1u : 0u;
    d += (unsigned long long)h * h * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) { x = side - 1 - x; y = side - 1 - y; }
      const unsigned t = x; x = y; y = t;
    }
  }
  return d;
}

// Storage keys.  Plane-constrained systems (every particle in one z layer of
// cells) are ordered along a 2-D Hilbert curve per layer, whose consecutive
// cells are always adjacent, so the 64 slots of a wavefront cover a compact
// block of cells (the planar build stages such blocks); other systems along
// a 3-D Morton curve of the grid.
__global__ void k_morton(Fields f, NListD nl, BoxD b, int n, unsigned long long* keys, int* slots, int plane) {
  const int m = max(nl.ncx, nl.ncy);
  const int order = m > 1 ? 32 - __clz(m - 1) : 0;
  for (int s = gtid(); s < n; s += gsize()) {
    int i = (int)((f.x[s] - b.xlo) / nl.wx);
    int j = (int)((f.y[s] - b.ylo) / nl.wy);
    int k = (int)((f.z[s] - b.zlo) / nl.wz);
    i = min(max(i, 0), nl.ncx - 1);
    j = min(max(j, 0), nl.ncy - 1);
    k = min(max(k, 0), nl.ncz - 1);
    keys[s] = plane ? ((unsigned long long)k << 42) | hilbert2((unsigned)i, (unsigned)j, order)
                    : spread3(i) | (spread3(j) << 1) | (spread3(k) << 2);
    slots[s] = s;
  }
}

__global__ void k_gather_d(double* dst, const double* src, const int* perm, int n) {
  for (int s = gtid(); s < n; s += gsize()) dst[s] = src[perm[s]];
}
__global__ void k_gather_i(int* dst, const int* src, const int* perm, int n) {
  for (int s = gtid(); s < n; s += gsize()) dst[s] = src[perm[s]];
}

__global__ void k_export_pairs(Fields f, NListD nl, int n, int* out, unsigned long long* counter, long cap) {
  for (int s = gtid(); s < n; s += gsize()) {
    const int is = f.id[s];
    const int kn = nl.nnb[s];
    for (int k = 0; k < kn; ++k) {
      const int t = nl.nbr[nbr_at(nl, k, s)];
      const int it = f.id[t];
      if (it > is) {
        unsigned long long p = atomicAdd(counter, 1ULL);
        if ((long)p < cap) { out[2 * p] = is; out[2 * p + 1] = it; }
      }
    }
  }
}

// Grid-stride launches are capped at 1024 workgroups (4 per CU): the stages of
// a device-conditional rebuild then cost one small dispatch when skipped.
int grid_for(int n) {
  int g = div_up(n, kBlock);
  return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

int coop_grid_cache[2] = {-1, -1};

}  // namespace

void launch_bin(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  hipLaunchKernelGGL(k_bin<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_scan(hipStream_t st, NListD nl, const int* req) {
  const int nb = div_up(nl.ncell, kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kScanThreads), 0, st, nl, req);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, st, nl, nb, req);
}

void launch_scatter(hipStream_t st, NListD nl, int n, const int* req) {
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, st, nl, n, req);
}

void launch_cell_sort(hipStream_t st, NListD nl, const int* req) {
  hipLaunchKernelGGL(k_cell_sort, dim3(grid_for(nl.ncell)), dim3(kBlock), 0, st, nl, req);
}

void launch_build(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar) {
  if (planar) hipLaunchKernelGGL(k_build<true>, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
  else hipLaunchKernelGGL(k_build<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_build_nsq(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  hipLaunchKernelGGL(k_build_nsq, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_rebuild(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar) {
  if (!nl.use_cell) { launch_build_nsq(st, f, nl, box, n, req); return; }
  if (planar) hipLaunchKernelGGL(k_bin<true>, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
  else launch_bin(st, f, nl, box, n, req);
  launch_scan(st, nl, req);
  launch_scatter(st, nl, n, req);
  launch_cell_sort(st, nl, req);
  launch_build(st, f, nl, box, n, req, planar);
}

void launch_chunk_rebuild(hipStream_t st, ChunkCtl* ck, ClockD* clk, Fields f0, Fields f1, NListD nl, BoxD box,
                          int n, int planar, int directors) {
  if (!nl.use_cell) {
    launch_chunk_settle(st, ck, f0, f1, n, planar, directors);
    launch_build_nsq(st, f0, nl, box, n, &ck->rebuild);
    launch_chunk_finish(st, ck, clk, nl.flags);
    return;
  }
  const int* req = &ck->rebuild;
  if (planar) hipLaunchKernelGGL(k_settle_bin<true>, dim3(grid_for(n)), dim3(kBlock), 0, st, ck, f0, f1, nl, box, n, directors);
  else hipLaunchKernelGGL(k_settle_bin<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, ck, f0, f1, nl, box, n, directors);
  launch_scan(st, nl, req);
  launch_scatter(st, nl, n, req);
  launch_cell_sort(st, nl, req);
  if (planar) hipLaunchKernelGGL(k_build_finish<true>, dim3(grid_for(n)), dim3(kBlock), 0, st, f0, nl, box, n, ck, clk);
  else hipLaunchKernelGGL(k_build_finish<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, f0, nl, box, n, ck, clk);
}

hipError_t launch_rebuild_coop(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar) {
  if (!nl.use_cell) { launch_build_nsq(st, f, nl, box, n, req); return hipGetLastError(); }
  int& grid = coop_grid_cache[planar ? 1 : 0];
  if (grid < 0) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = planar ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rebuild_coop<true>, kScanThreads, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rebuild_coop<false>, kScanThreads, 0);
    if (e != hipSuccess) return e;
    grid = cus * per_cu;
    if (grid < 1) return hipErrorInvalidConfiguration;
  }
  void* args[] = {&f, &nl, &box, &n, &req};
  return planar ? hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_rebuild_coop<true>), dim3(grid),
                                             dim3(kScanThreads), args, 0, st)
                : hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_rebuild_coop<false>), dim3(grid),
                                             dim3(kScanThreads), args, 0, st);
}

void launch_need_update(hipStream_t st, Fields f, NListD nl, BoxD box, int n, int* req_out) {
  hipLaunchKernelGGL(k_need_update, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req_out);
}

void launch_morton_keys(hipStream_t st, Fields f, NListD nl, BoxD box, int n, unsigned long long* keys, int* slots,
                        int plane) {
  hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, keys, slots, plane);
}

void launch_gather_d(hipStream_t st, double* dst, const double* src, const int* perm, int n) {
  hipLaunchKernelGGL(k_gather_d, dim3(grid_for(n)), dim3(kBlock), 0, st, dst, src, perm, n);
}

void launch_gather_i(hipStream_t st, int* dst, const int* src, const int* perm, int n) {
  hipLaunchKernelGGL(k_gather_i, dim3(grid_for(n)), dim3(kBlock), 0, st, dst, src, perm, n);
}

void launch_export_pairs(hipStream_t st, Fields f, NListD nl, int n, int* out, unsigned long long* counter, long cap) {
  hipLaunchKernelGGL(k_export_pairs, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, n, out, counter, cap);
}

size_t radix_sort_temp_bytes(int n) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                     (const int*)nullptr, (int*)nullptr, n);
  return bytes;
}

void radix_sort_pairs(hipStream_t st, void* temp, size_t temp_bytes, const unsigned long long* kin,
                      unsigned long long* kout, const int* vin, int* vout, int n) {
  hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, n, 0, 64, st);
}

}  // namespace samos
