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
#include "kernels.hpp"
#include <hipcub/hipcub.hpp>

namespace samos {

namespace {

__device__ __forceinline__ bool req_off(const int* req) { return req != nullptr && *req == 0; }

// CellList::populate wrap + get_cell_idx (src/system/cell_list.cpp:72-104).
// The index is clamped to the grid: the reference indexes one past the grid for
// a particle exactly on the high face.
__global__ void k_bin(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    double x = f.x[s], y = f.y[s], z = f.z[s];
    if (b.periodic) {
      bool w = false;
      if (x < b.xlo) { x += b.Lx; w = true; } else if (x > b.xhi) { x -= b.Lx; w = true; }
      if (y < b.ylo) { y += b.Ly; w = true; } else if (y > b.yhi) { y -= b.Ly; w = true; }
      if (z < b.zlo) { z += b.Lz; w = true; } else if (z > b.zhi) { z -= b.Lz; w = true; }
      if (w) { f.x[s] = x; f.y[s] = y; f.z[s] = z; }
    }
    int i = (int)((x - b.xlo) / nl.wx);
    int j = (int)((y - b.ylo) / nl.wy);
    int k = (int)((z - b.zlo) / nl.wz);
    i = min(max(i, 0), nl.ncx - 1);
    j = min(max(j, 0), nl.ncy - 1);
    k = min(max(k, 0), nl.ncz - 1);
    int c = nl.ncy * nl.ncz * i + nl.ncz * j + k;
    nl.cell_of[s] = c;
    atomicAdd(&nl.count[c], 1);
  }
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Inclusive scan of one value per thread across the workgroup (LDS, Hillis-Steele
// over wavefront sums).
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

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(NListD nl, const int* req) {
  if (req_off(req)) return;
  __shared__ int lds[kScanThreads / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int idx = base + q;
    v[q] = idx < nl.ncell ? nl.count[idx] : 0;
    sum += v[q];
  }
  int total;
  int off = block_exclusive_scan(sum, lds, total);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    int idx = base + q;
    if (idx < nl.ncell) nl.start[idx] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) nl.bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_bsum(NListD nl, int nb, const int* req) {
  if (req_off(req)) return;
  __shared__ int lds[16];
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    int idx = base + threadIdx.x;
    int v = idx < nb ? nl.bsum[idx] : 0;
    int total;
    int ex = block_exclusive_scan(v, lds, total);
    if (idx < nb) nl.bsum[idx] = carry + ex;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) nl.start[nl.ncell] = carry;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(NListD nl, const int* req) {
  if (req_off(req)) return;
  const int add = nl.bsum[blockIdx.x];
  const int base = blockIdx.x * kScanTile;
  for (int q = threadIdx.x; q < kScanTile; q += kScanThreads) {
    int idx = base + q;
    if (idx < nl.ncell) nl.start[idx] += add;
  }
}

// Counts return to zero as members are placed, so the next build starts clean.
__global__ void k_scatter(NListD nl, int n, const int* req) {
  if (req_off(req)) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    int c = nl.cell_of[s];
    int pos = nl.start[c] + atomicSub(&nl.count[c], 1) - 1;
    nl.members[pos] = s;
  }
}

// Ascending slot order inside each cell makes the list deterministic.
__global__ void k_cell_sort(NListD nl, const int* req) {
  if (req_off(req)) return;
  int local_max = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nl.ncell; c += gridDim.x * blockDim.x) {
    int a = nl.start[c], e = nl.start[c + 1];
    local_max = max(local_max, e - a);
    for (int i = a + 1; i < e; ++i) {
      int key = nl.members[i];
      int j = i - 1;
      while (j >= a && nl.members[j] > key) { nl.members[j + 1] = nl.members[j]; --j; }
      nl.members[j + 1] = key;
    }
  }
  // wavefront max, one atomic per wavefront
  for (int o = 32; o > 0; o >>= 1) local_max = max(local_max, __shfl_xor(local_max, o, 64));
  if ((threadIdx.x & 63) == 0 && local_max > 0) atomicMax(&nl.flags[3], local_max);
}

// Flags: [2] overflow, [4] builds, [5] largest neighbour count.  One atomic per
// wavefront.
__device__ __forceinline__ void build_epilogue(NListD& nl, bool overflow, int kmax) {
  for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
  const bool any_over = __ballot(overflow) != 0;
  if ((threadIdx.x & 63) == 0) {
    if (any_over) atomicOr(&nl.flags[2], 1);
    if (kmax > 0) atomicMax(&nl.flags[5], kmax);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&nl.flags[4], 1);
}

// NeighbourList::build_cell (src/system/neighbour_list.cpp:216-273), full list.
__global__ void k_build(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  bool overflow = false;
  int kmax = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const double xi = f.x[s], yi = f.y[s], zi = f.z[s], ri = f.radius[s];
    const int c = nl.cell_of[s];
    const int ci = c / (nl.ncy * nl.ncz), cj = (c / nl.ncz) % nl.ncy, ck = c % nl.ncz;
    int k = 0, coord = 0;
    for (int ix = -1; ix <= 1; ix++) {
      int iix = ci + ix;
      if (iix < 0) iix = nl.ncx - 1; else if (iix == nl.ncx) iix = 0;
      for (int iy = -1; iy <= 1; iy++) {
        int iiy = cj + iy;
        if (iiy < 0) iiy = nl.ncy - 1; else if (iiy == nl.ncy) iiy = 0;
        for (int iz = -1; iz <= 1; iz++) {
          int iiz = ck + iz;
          if (iiz < 0) iiz = nl.ncz - 1; else if (iiz == nl.ncz) iiz = 0;
          const int cc = nl.ncy * nl.ncz * iix + nl.ncz * iiy + iiz;
          const int a = nl.start[cc], e = nl.start[cc + 1];
          for (int m = a; m < e; ++m) {
            const int t = nl.members[m];
            if (t == s) continue;
            double dx = xi - f.x[t], dy = yi - f.y[t], dz = zi - f.z[t];
            min_image(b, dx, dy, dz);
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < nl.cut2) {
              if (k < nl.stride) nl.nbr[(long)k * nl.np + s] = t;
              ++k;
              const double r = ri + f.radius[t];
              if (d2 < r * r) ++coord;
            }
          }
        }
      }
    }
    if (k > nl.stride) overflow = true;
    kmax = max(kmax, k);
    nl.nnb[s] = min(k, nl.stride);
    f.coord[s] = coord;
    nl.ox[s] = xi; nl.oy[s] = yi; nl.oz[s] = zi;
  }
  build_epilogue(nl, overflow, kmax);
}

// NeighbourList::build_nsq (src/system/neighbour_list.cpp:178-213), full list.
__global__ void k_build_nsq(Fields f, NListD nl, BoxD b, int n, const int* req) {
  if (req_off(req)) return;
  bool overflow = false;
  int kmax = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const double xi = f.x[s], yi = f.y[s], zi = f.z[s], ri = f.radius[s];
    int k = 0, coord = 0;
    for (int t = 0; t < n; ++t) {
      if (t == s) continue;
      double dx = xi - f.x[t], dy = yi - f.y[t], dz = zi - f.z[t];
      min_image(b, dx, dy, dz);
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < nl.cut2) {
        if (k < nl.stride) nl.nbr[(long)k * nl.np + s] = t;
        ++k;
        const double r = ri + f.radius[t];
        if (d2 < r * r) ++coord;
      }
    }
    if (k > nl.stride) overflow = true;
    kmax = max(kmax, k);
    nl.nnb[s] = min(k, nl.stride);
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
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    double dx = nl.ox[s] - f.x[s], dy = nl.oy[s] - f.y[s], dz = nl.oz[s] - f.z[s];
    min_image(b, dx, dy, dz);
    need |= !(dx * dx + dy * dy + dz * dz < nl.pad2q);
  }
  if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomicOr(req_out, 1);
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

__global__ void k_morton(Fields f, NListD nl, BoxD b, int n, unsigned long long* keys, int* slots) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    int i = (int)((f.x[s] - b.xlo) / nl.wx);
    int j = (int)((f.y[s] - b.ylo) / nl.wy);
    int k = (int)((f.z[s] - b.zlo) / nl.wz);
    i = min(max(i, 0), nl.ncx - 1);
    j = min(max(j, 0), nl.ncy - 1);
    k = min(max(k, 0), nl.ncz - 1);
    keys[s] = spread3(i) | (spread3(j) << 1) | (spread3(k) << 2);
    slots[s] = s;
  }
}

__global__ void k_gather_d(double* dst, const double* src, const int* perm, int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) dst[s] = src[perm[s]];
}
__global__ void k_gather_i(int* dst, const int* src, const int* perm, int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) dst[s] = src[perm[s]];
}

__global__ void k_export_pairs(Fields f, NListD nl, int n, int* out, unsigned long long* counter, long cap) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int is = f.id[s];
    const int kn = nl.nnb[s];
    for (int k = 0; k < kn; ++k) {
      const int t = nl.nbr[(long)k * nl.np + s];
      const int it = f.id[t];
      if (it > is) {
        unsigned long long p = atomicAdd(counter, 1ULL);
        if ((long)p < cap) { out[2 * p] = is; out[2 * p + 1] = it; }
      }
    }
  }
}

int grid_for(int n) {
  int g = div_up(n, kBlock);
  return g < 1 ? 1 : (g > 8192 ? 8192 : g);
}

}  // namespace

void launch_bin(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  hipLaunchKernelGGL(k_bin, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_scan(hipStream_t st, NListD nl, const int* req) {
  int nb = div_up(nl.ncell, kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kScanThreads), 0, st, nl, req);
  hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, st, nl, nb, req);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanThreads), 0, st, nl, req);
}

void launch_scatter(hipStream_t st, NListD nl, int n, const int* req) {
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(kBlock), 0, st, nl, n, req);
}

void launch_cell_sort(hipStream_t st, NListD nl, const int* req) {
  hipLaunchKernelGGL(k_cell_sort, dim3(grid_for(nl.ncell)), dim3(kBlock), 0, st, nl, req);
}

void launch_build(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  hipLaunchKernelGGL(k_build, dim3(grid1(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_build_nsq(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  hipLaunchKernelGGL(k_build_nsq, dim3(grid1(n)), dim3(kBlock), 0, st, f, nl, box, n, req);
}

void launch_rebuild(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req) {
  if (!nl.use_cell) { launch_build_nsq(st, f, nl, box, n, req); return; }
  launch_bin(st, f, nl, box, n, req);
  launch_scan(st, nl, req);
  launch_scatter(st, nl, n, req);
  launch_cell_sort(st, nl, req);
  launch_build(st, f, nl, box, n, req);
}

void launch_need_update(hipStream_t st, Fields f, NListD nl, BoxD box, int n, int* req_out) {
  hipLaunchKernelGGL(k_need_update, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, req_out);
}

void launch_morton_keys(hipStream_t st, Fields f, NListD nl, BoxD box, int n, unsigned long long* keys, int* slots) {
  hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBlock), 0, st, f, nl, box, n, keys, slots);
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
