// Cell-tiled fused Langevin step for planar single systems.
//
// Same physics and per-pair arithmetic as k_baoab_fused (integrate.hip) and
// the same pair set as the neighbour-list build (nlist.hip), restating
//   cell list / build     src/system/cell_list.cpp:37-104, src/system/neighbour_list.cpp:216-302
//   soft, LJ pair terms   src/potentials/pair/pair_soft_potential.cpp:66-109,
//                         src/potentials/pair/pair_lj_potential.cpp:48-121
//   Langevin BAOAB        src/integrators/integrator_langevin.cpp:54-118
//   plane constraint      src/constraints/constraint_plane.cpp:36-88
// with a different data layout:
//   * the state of a run lives in cell order (row-major cells of the 2-D grid,
//     members of a cell contiguous), in two sets that a build swaps between;
//   * a workgroup owns T consecutive cells of one row and stages the positions
//     of the 3 x (T + 2) cells around them in LDS with coalesced loads;
//   * a particle's Verlet list holds the 16-bit stage indices of the
//     candidates of its 3 x 3 stencil that were within rcut + pad at the last
//     build (the stage's layout follows the cell starts, fixed between
//     builds), so a step reads 2 bytes of list per neighbour and takes each
//     neighbour position from LDS.
// Launch order and control: the position buffer a step reads alternates with
// the step and is a launch argument; which state set is current and whether a
// build is requested live in one control word changed only by atomics (see
// TileCtl), so a captured sequence of step and build kernels needs no host
// decisions and no device-wide fences.
#include "tiles.hpp"
#include "constraint.hpp"

namespace samos {

namespace {

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int gsize() { return gridDim.x * blockDim.x; }


__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ int block_exscan(int v, int* lds4, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) lds4[wave] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kTileThreads / 64; ++w) {
    const int ws = lds4[w];
    if (w < wave) off += ws;
    tot += ws;
  }
  __syncthreads();
  total = tot;
  return off + incl - v;
}

__device__ __forceinline__ void wave_max_to(int v, int* dst) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0 && v > 0) atomic_max_if_above(dst, v);
}

// counter[key] += 1 for every active lane with one atomic per distinct key
// in the wavefront (lanes of a cell-ordered set share keys in runs).  Every
// lane of the wavefront must call it.
__device__ __forceinline__ void wave_add1(int* counter, int key, bool active) {
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

__device__ __forceinline__ int ctl_word(const TileCtl* ctl) {
  return __hip_atomic_load(&ctl->word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tile of the workgroup: XCD-aware, so the eight XCDs (workgroup i goes to
// XCD i mod 8) each sweep one contiguous band of rows and keep the halo rows
// they share in their own L2.
struct TileGeom {
  bool active;
  int ty, i0, nc, nh;
};

__device__ __forceinline__ int xcd_tile(const TileD& td) {
  const int per = (td.ntiles + 7) >> 3;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

__device__ __forceinline__ TileGeom tile_geom(const TileD& td, int t) {
  TileGeom g;
  g.active = t < td.ntiles;
  const int tt = g.active ? t : 0;
  g.ty = tt / td.ntx;
  const int tx = tt - g.ty * td.ntx;
  g.i0 = tx * td.T;
  g.nc = min(td.T, td.ncx - g.i0);
  g.nh = 3 * (g.nc + 2);
  return g;
}

// Halo cell h (row r = h / (nc + 2), column k = h % (nc + 2)) is cell
// (i0 - 1 + k, ty - 1 + r), both wrapped.  s_hcs[h] is its first staged
// index, s_hgs[h] its first index in the set's cell order; s_hcs[nh] = total.
__device__ __forceinline__ int stage_index(const TileD& td, const TileGeom& g, const int* cs, int* s_hcs, int* s_hgs,
                                           int* s_scan) {
  // thread t takes halo cells 2t and 2t + 1 (at most 2 * kTileThreads)
  int cnt[2] = {0, 0};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int h = 2 * (int)threadIdx.x + u;
    if (g.active && h < g.nh) {
      const int r = h / (g.nc + 2), k = h - r * (g.nc + 2);
      const int ci = wrapi(g.i0 - 1 + k, td.ncx), cj = wrapi(g.ty - 1 + r, td.ncy);
      const int c = cj * td.ncx + ci;
      const int a = cs[c];
      cnt[u] = cs[c + 1] - a;
      s_hgs[h] = a;
    }
  }
  int total;
  const int ex = block_exscan(cnt[0] + cnt[1], s_scan, total);
  const int h = 2 * (int)threadIdx.x;
  if (h < g.nh) s_hcs[h] = ex;
  if (h + 1 < g.nh) s_hcs[h + 1] = ex + cnt[0];
  if (threadIdx.x == 0) s_hcs[g.nh] = total;
  __syncthreads();
  return total;
}

// Global (set-order) index of staged index k: each row is three contiguous
// pieces (left halo cell, the tile's own cells, right halo cell).
__device__ __forceinline__ int staged_global(const TileGeom& g, const int* s_hcs, const int* s_hgs, int k) {
  int h = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int base = r * (g.nc + 2);
    if (k >= s_hcs[base]) h = base;
    if (k >= s_hcs[base + 1]) h = base + 1;
    if (k >= s_hcs[base + g.nc + 1]) h = base + g.nc + 1;
  }
  return s_hgs[h] + (k - s_hcs[h]);
}

// The same for a tile whose halo rows are contiguous in the set's cell order
// (no periodic wrap along x inside the halo): row r is one run.
__device__ __forceinline__ int staged_global_rows(const TileGeom& g, const int* s_hcs, const int* s_hgs, int k) {
  const int row1 = g.nc + 2;
  const int r = (k >= s_hcs[row1] ? 1 : 0) + (k >= s_hcs[2 * row1] ? 1 : 0);
  return s_hgs[r * row1] + (k - s_hcs[r * row1]);
}

// Loads and stores through the global address space (the set's arrays are
// reached through pointers read from memory, which the compiler would
// otherwise access with flat instructions).
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)(p);
}
typedef double gdouble2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gld(const double2* p) {
  const gdouble2 v = *(const __attribute__((address_space(1))) gdouble2*)(p);
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gst(double2* p, const double2& a) {
  gdouble2 v;
  v.x = a.x;
  v.y = a.y;
  *(__attribute__((address_space(1))) gdouble2*)(p) = v;
}

// A tile's stage descriptor, read with scalar loads (uniform over the
// workgroup, written by an earlier kernel).
typedef const __attribute__((address_space(4))) int* DescPtr;
__device__ __forceinline__ DescPtr tile_desc(const TileD& td, int t) {
  return (DescPtr)(td.desc + (long)t * kTileDesc);
}

// Set index of stage entry k: the last run starting at or before k.  The
// runs past the last hold INT_MAX as their first stage entry (the builds
// write every slot), so a fixed, unrolled search needs no run count.
__device__ __forceinline__ int desc_global(DescPtr D, int k) {
  int so = D[5], gs = D[6];
#pragma unroll
  for (int r = 1; r < kTileRuns; ++r) {
    const int o = D[5 + 2 * r];
    if (k >= o) { so = o; gs = D[6 + 2 * r]; }
  }
  return gs + (k - so);
}

// The same for a stage of at most three runs (a tile whose halo does not wrap
// along x: one run per row).
__device__ __forceinline__ int desc_global3(DescPtr D, int k) {
  int so = D[5], gs = D[6];
  if (k >= D[7]) { so = D[7]; gs = D[8]; }
  if (k >= D[9]) { so = D[9]; gs = D[10]; }
  return gs + (k - so);
}

// ---------------------------------------------------------------------------
// Build, kernel 1: bin every particle by its snapshot position
// (CellList::get_cell_idx, src/system/cell_list.cpp:72-79, clamped) and count
// per cell and per tile.  par: the position buffer holding the current
// positions.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kTileThreads) k_tile_bin(TileD td, int par) {
  const TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  if (!(w & (4 << par)) || ctl->far) return;
  const int S = w & 1, init = ctl->init;
  const TileSet& A = td.sets[S];
  bool far = false;
  for (int base = blockIdx.x * blockDim.x; base < td.n; base += gsize()) {
    const int p = base + threadIdx.x;
    const bool active = p < td.n;
    double2 q = make_double2(0.0, 0.0);
    if (active) q = init ? A.o[p] : A.xy[par ^ 1][p];
    int i = (int)((q.x - td.xlo) / td.wx);
    int j = (int)((q.y - td.ylo) / td.wy);
    i = min(max(i, 0), td.ncx - 1);
    j = min(max(j, 0), td.ncy - 1);
    const int c = j * td.ncx + i;
    if (active) td.cell_of[p] = c;
    wave_add1(td.count, c, active);
    wave_add1(td.tcount, j * td.ntx + i / td.T, active);
    if (active && !init) {
      const double2 o = A.o[p];
      int oi = (int)((o.x - td.xlo) / td.wx);
      int oj = (int)((o.y - td.ylo) / td.wy);
      oi = min(max(oi, 0), td.ncx - 1);
      oj = min(max(oj, 0), td.ncy - 1);
      int di = abs(i - oi), dj = abs(j - oj);
      di = min(di, td.ncx - di);
      dj = min(dj, td.ncy - dj);
      far |= (di > 1 || dj > 1);
    }
  }
  if (__ballot(far) != 0 && (threadIdx.x & 63) == 0) atomic_set_once(&td.ctl->far);
}

// Build, kernel 2 (a range of kTilesPerBlock tiles per workgroup): the tile
// starts (the sum of all earlier tiles' counts, then a scan over the range),
// each cell's start (tile start + counts of the tile's earlier cells) and,
// after the first build, each cell's member list gathered from the 3 x 3 old
// cells around it (a particle moves by at most one cell between builds, which
// kernel 1 checks), sorted by slot.
constexpr int kTilesPerBlock = 8;

__global__ void __launch_bounds__(kTileThreads) k_tile_cells(TileD td, int par) {
  const TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  if (!(w & (4 << par)) || ctl->far) return;
  const int S = w & 1, init = ctl->init, N = S ^ 1;
  const int* cso = td.sets[S].cstart;
  int* csn = td.sets[N].cstart;
  __shared__ int s_scan[kTileThreads / 64];
  __shared__ int s_tstart[kTilesPerBlock + 1];
  const int tb = blockIdx.x * kTilesPerBlock, te = min(tb + kTilesPerBlock, td.ntiles);
  int part = 0;
  for (int t = threadIdx.x; t < tb; t += kTileThreads) part += td.tcount[t];
  int before;
  block_exscan(part, s_scan, before);
  const int own = (int)threadIdx.x < te - tb ? td.tcount[tb + threadIdx.x] : 0;
  int total;
  const int ex = block_exscan(own, s_scan, total);
  if ((int)threadIdx.x < te - tb) s_tstart[threadIdx.x] = before + ex;
  if (threadIdx.x == 0 && te == td.ntiles) csn[td.ncell] = before + total;
  __syncthreads();
  // cells of the range: tile t covers cells [j*ncx + tx*T, j*ncx + min(tx*T + T, ncx))
  bool far = false;
  for (int u = threadIdx.x; u < (te - tb) * td.T; u += kTileThreads) {
    const int t = tb + u / td.T, l = u % td.T;
    const int j = t / td.ntx, i0 = (t - j * td.ntx) * td.T;
    const int i = i0 + l;
    if (i >= td.ncx) continue;
    const int cb = j * td.ncx + i0, c = cb + l;
    int start = s_tstart[t - tb];
    for (int k = cb; k < c; ++k) start += td.count[k];
    csn[c] = start;
    if (!init) {
      const int cnt = td.count[c];
      int m = 0;
      for (int dj = -1; dj <= 1; ++dj) {
        const int jj = wrapi(j + dj, td.ncy);
        for (int di = -1; di <= 1; ++di) {
          const int oc = jj * td.ncx + wrapi(i + di, td.ncx);
          const int e = cso[oc + 1];
          for (int p = cso[oc]; p < e; ++p) {
            if (td.cell_of[p] == c) {
              if (m < cnt) td.members[start + m] = p;
              ++m;
            }
          }
        }
      }
      far |= (m != cnt);
      // canonical order inside a cell: ascending slot, as the first build of a
      // run sorts it, so the order depends only on the snapshot
      const int* key = td.sets[S].slot;
      const int e = start + min(m, cnt);
      for (int x = start + 1; x < e; ++x) {
        const int mx = td.members[x];
        const int kx = key[mx];
        int y = x - 1;
        while (y >= start && key[td.members[y]] > kx) {
          td.members[y + 1] = td.members[y];
          --y;
        }
        td.members[y + 1] = mx;
      }
    }
  }
  if (__ballot(far) != 0 && (threadIdx.x & 63) == 0) atomic_set_once(&td.ctl->far);
}

// First build of a run only: members by atomic scatter (counts return to
// zero), then each cell sorted by slot, which makes the order deterministic.
__global__ void __launch_bounds__(kTileThreads) k_tile_scatter(TileD td, int par) {
  const TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  if (!(w & (4 << par)) || !ctl->init) return;
  const int* csn = td.sets[(w & 1) ^ 1].cstart;
  for (int p = gtid(); p < td.n; p += gsize()) {
    const int c = td.cell_of[p];
    const int r = atomicSub(&td.count[c], 1) - 1;
    td.members[csn[c] + r] = p;
  }
}

__global__ void __launch_bounds__(kTileThreads) k_tile_sort(TileD td, int par) {
  const TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  if (!(w & (4 << par)) || !ctl->init) return;
  const int S = w & 1;
  const int* csn = td.sets[S ^ 1].cstart;
  const int* key = td.sets[S].slot;
  for (int c = gtid(); c < td.ncell; c += gsize()) {
    const int a = csn[c], e = csn[c + 1];
    for (int i = a + 1; i < e; ++i) {
      const int m = td.members[i];
      const int km = key[m];
      int j = i - 1;
      while (j >= a && key[td.members[j]] > km) {
        td.members[j + 1] = td.members[j];
        --j;
      }
      td.members[j + 1] = m;
    }
  }
}

// Build, kernel 3 (per tile): stage the snapshot positions of the halo in new
// order, write the new set for the tile's own particles and their lists of
// stage indices (NeighbourList::build_cell's test, src/system/neighbour_list.cpp:
// 240-262: dx = x_i - x_j, single-wrap image, d2 = dx*dx + dy*dy + dz*dz,
// kept when d2 < (rcut + pad)^2).  The current positions are in buffer
// par ^ 1, and the new set's go there too (the next step reads it).  Tiles
// are visited grid-stride.  Marks the build finished (bit 1 of the word).
__global__ void __launch_bounds__(kTileThreads) k_tile_build(TileD td, BoxD b, int par_req) {
  TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  if (!(w & (4 << par_req)) || ctl->far) return;
  const int S = w & 1, init = ctl->init, N = S ^ 1, par = par_req ^ 1;
  __shared__ double2 s_snap[kTileHaloMax];
  __shared__ int s_hcs[3 * (kTileMaxCells + 2) + 1];
  __shared__ int s_hgs[3 * (kTileMaxCells + 2)];
  __shared__ int s_scan[kTileThreads / 64];
  bool over = false;
  int cmax = 0, hmax = 0;
  for (int t = blockIdx.x; t < td.ntiles; t += gridDim.x) {
  const TileGeom g = tile_geom(td, t);
  const int* csn = td.sets[N].cstart;
  const int total = stage_index(td, g, csn, s_hcs, s_hgs, s_scan);
  const bool staged = total <= kTileHaloMax;
  over |= g.active && total > kTileStageMax;
  if (g.active && threadIdx.x == 0) {
    // the stage as runs of consecutive set indices (empty cells skipped,
    // neighbouring cells of the set's order merged)
    int* D = td.desc + (long)t * kTileDesc;
    const int row1 = g.nc + 2;
    D[0] = total;
    D[1] = s_hcs[row1 + 1];
    D[2] = s_hcs[row1 + g.nc + 1];
    D[3] = s_hgs[row1 + 1];
    int nr = 0, end = -1;
    for (int h = 0; h < g.nh; ++h) {
      const int cnt = s_hcs[h + 1] - s_hcs[h];
      if (cnt == 0) continue;
      if (nr > 0 && s_hgs[h] == end) { end += cnt; continue; }
      D[5 + 2 * nr] = s_hcs[h];
      D[6 + 2 * nr] = s_hgs[h];
      end = s_hgs[h] + cnt;
      ++nr;
    }
    if (nr == 0) { D[5] = 0; D[6] = 0; nr = 1; }
    D[4] = nr;
    for (int r = nr; r < kTileRuns; ++r) { D[5 + 2 * r] = 0x7fffffff; D[6 + 2 * r] = 0; }
  }
  const double2* snap_src = init ? td.sets[S].o : td.sets[S].xy[par];
  if (g.active && staged) {
    for (int k = threadIdx.x; k < total; k += kTileThreads)
      s_snap[k] = snap_src[td.members[staged_global(g, s_hcs, s_hgs, k)]];
  }
  __syncthreads();
  hmax = max(hmax, total);
  if (g.active) {
    const int row1 = g.nc + 2;
    const int kb = s_hcs[row1 + 1], ke = s_hcs[row1 + g.nc + 1];
    const int qa = s_hgs[row1 + 1];
    const double zero = 0.0;
    over |= ke - kb > kTileMaxOwn;  // the step kernels take at most this many per tile
    for (int k = kb + (int)threadIdx.x; k < ke; k += kTileThreads) {
      const int q = qa + (k - kb);
      const int mem = td.members[q];
      const double2 si = staged ? s_snap[k] : snap_src[mem];
      td.sets[N].xy[par][q] = td.sets[S].xy[par][mem];
      td.sets[N].o[q] = si;
      td.sets[N].v[q] = td.sets[S].v[mem];
      const int dp = td.dirs2 ? par : 0;  // the buffer holding the current directors
      td.sets[N].nd[dp][q] = td.sets[S].nd[dp][mem];
      td.sets[N].om[q] = td.sets[S].om[mem];
      td.sets[N].id[q] = td.sets[S].id[mem];
      td.sets[N].slot[q] = td.sets[S].slot[mem];
      const int lc = (td.cell_of[mem] % td.ncx) - g.i0 + 1;
      // entries in ascending stage order, four per 8-byte row
      unsigned long long* L = reinterpret_cast<unsigned long long*>(td.lst) + q;
      unsigned long long acc = 0ULL;
      int cnt = 0;
      for (int r = 0; r < 3; ++r) {
        const int a = s_hcs[r * row1 + lc - 1], e = s_hcs[r * row1 + lc + 2];
        for (int kk = a; kk < e; ++kk) {
          if (kk == k) continue;
          const double2 sj = staged ? s_snap[kk] : snap_src[td.members[staged_global(g, s_hcs, s_hgs, kk)]];
          double dx = si.x - sj.x, dy = si.y - sj.y, dz = zero;
          min_image(b, dx, dy, dz);
          const double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < td.cut2) {
            if (cnt < td.K) {
              acc |= (unsigned long long)kk << (16 * (cnt & 3));
              if ((cnt & 3) == 3) {
                L[(long)(cnt >> 2) * td.n] = acc;
                acc = 0ULL;
              }
            }
            ++cnt;
          }
        }
      }
      const int c = min(cnt, td.K);
      if (c & 3) {
        // pad the last row with the sentinel (the stage's slot after the halo)
        for (int u = c & 3; u < 4; ++u) acc |= (unsigned long long)total << (16 * u);
        L[(long)(c >> 2) * td.n] = acc;
      }
      td.len[q] = c;
      over |= cnt > td.K;
      cmax = max(cmax, cnt);
    }
    if ((int)threadIdx.x < g.nc) td.count[g.ty * td.ncx + g.i0 + threadIdx.x] = 0;
    if (threadIdx.x == 0) td.tcount[t] = 0;
  }
  __syncthreads();  // the stage is reused by the next tile
  }
  if (__ballot(over) != 0 && (threadIdx.x & 63) == 0) atomicOr(&ctl->overflow, 1);
  wave_max_to(cmax, &ctl->maxcand);
  if (threadIdx.x == 0 && hmax > 0) atomicMax(&ctl->maxhalo, hmax);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicOr(&ctl->word, 2);
    if (!init) atomicAdd(&ctl->rebuilds, 1);
  }
}

// ---------------------------------------------------------------------------
// Per-pair force terms of a planar system (dz = 0), as pair_force_sum in
// pair_sum.hpp with one particle type: d = x_j - x_i.
// ---------------------------------------------------------------------------
// Constants of the pair loop: the pair parameters and the box's image
// thresholds.  list_forces keeps them in vector registers (an empty asm with
// a "+v" operand): they are uniform, but held in scalar registers across the
// pair loop they would be spilled and reloaded lane by lane.
struct PairK {
  double soft_k, soft_a, lj_rc2, lj_sig2, lj_e48;
  double xlo, xhi, lx, ylo, yhi, ly;
  double al_J, al_rc2;  // aligner coupling and cutoff^2 (Brownian runs)
};

__device__ __forceinline__ PairK pair_k(const PairConst& pc, const BoxD& b) {
  PairK k;
  k.soft_k = pc.soft_k; k.soft_a = pc.soft_a;
  k.lj_rc2 = pc.lj_rc2; k.lj_sig2 = pc.lj_sig2; k.lj_e48 = 48.0 * pc.lj_eps;
  k.xlo = b.xlo; k.xhi = b.xhi; k.lx = b.Lx; k.ylo = b.ylo; k.yhi = b.yhi; k.ly = b.Ly;
  k.al_J = 0.0; k.al_rc2 = 0.0;
  return k;
}

// r_sq = dx * dx + dy * dy (dz = 0: adding +0 to a sum of squares changes no
// bit).  The arithmetic of pair_force_loop (pair_sum.hpp) for one type: each
// product accumulated with one rounding (fused multiply-add); FP32 factors
// from the hardware square root and reciprocal, accumulated in FP64.
template <int POT, bool F32>
__device__ __forceinline__ void pair_term(const PairK& pk, double dx, double dy, double r_sq, double& fx,
                                          double& fy) {
  if (POT & POT_SOFT) {
    if (F32) {
      const float r = __builtin_amdgcn_sqrtf((float)r_sq);
      if (r < (float)pk.soft_a) {
        const float diff = (float)pk.soft_a - r;
        const float ff = (r > 0.0f) ? (float)pk.soft_k * diff * __builtin_amdgcn_rcpf(r) : (float)pk.soft_k * diff;
        fx = fma(-(double)ff, dx, fx); fy = fma(-(double)ff, dy, fy);
      }
    } else {
      const double r = sqrt(r_sq);
      if (r < pk.soft_a) {
        const double diff = pk.soft_a - r;
        const double ff = (r > 0.0) ? pk.soft_k * diff / r : pk.soft_k * diff;
        fx = fma(-ff, dx, fx); fy = fma(-ff, dy, fy);
      }
    }
  }
  if (POT & POT_LJ) {
    if (r_sq <= pk.lj_rc2) {
      if (F32) {
        const float s2 = (float)pk.lj_sig2 * __builtin_amdgcn_rcpf((float)r_sq);
        const float s6 = s2 * s2 * s2;
        const float ff = (float)pk.lj_e48 * s6 * (s6 - 0.5f) * s2;
        fx = fma(-(double)ff, dx, fx); fy = fma(-(double)ff, dy, fy);
      } else {
        const double inv_r_sq = lj_inv_r_sq(pk.lj_sig2, r_sq);
        const double inv_r_6 = inv_r_sq * inv_r_sq * inv_r_sq;
        const double ff = pk.lj_e48 * inv_r_6 * (inv_r_6 - 0.5) * inv_r_sq;
        fx = fma(-ff, dx, fx); fy = fma(-ff, dy, fy);
      }
    }
  }
}

// Single-wrap minimum image of a planar pair (src/system/system.cpp:1083-1095).
__device__ __forceinline__ void min_image_k(const PairK& pk, double& dx, double& dy) {
  if (dx > pk.xhi) dx -= pk.lx; else if (dx < pk.xlo) dx += pk.lx;
  if (dy > pk.yhi) dy -= pk.ly; else if (dy < pk.ylo) dy += pk.ly;
}

// Forces on particle q from its list: one 8-byte load brings four stage
// indices (the next row's load is issued before the current row's reads),
// then four staged positions are read from LDS together.  The padding of the
// last row is the stage's sentinel slot, a position so far away that no pair
// term is formed with it.  wrap: the tile's halo crosses the periodic
// boundary, so the single-wrap image is applied (inside the box it never
// changes a difference and is skipped); uniform over the workgroup.
template <int POT, bool F32>
__device__ __forceinline__ void list_forces(PairK pk, const TileD& td, int q, int kn, const double2* s_pos, bool wrap,
                                            double xi, double yi, double& fx, double& fy) {
  asm volatile("" : "+v"(pk.lj_rc2), "+v"(pk.lj_sig2), "+v"(pk.lj_e48), "+v"(pk.soft_k), "+v"(pk.soft_a));
  asm volatile("" : "+v"(pk.xlo), "+v"(pk.xhi), "+v"(pk.lx), "+v"(pk.ylo), "+v"(pk.yhi), "+v"(pk.ly));
  const unsigned long long* L = reinterpret_cast<const unsigned long long*>(td.lst) + q;
  const long n = td.n;
  const int rows = (kn + 3) >> 2;
  unsigned long long nxt = rows > 0 ? gld(L) : 0ULL;
  for (int r = 0; r < rows; ++r) {
    const unsigned long long cur = nxt;
    if (r + 1 < rows) nxt = gld(L + (long)(r + 1) * n);
    double2 pj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) pj[u] = s_pos[(int)((cur >> (16 * u)) & 0xffffULL)];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double dx = pj[u].x - xi, dy = pj[u].y - yi;
      if (wrap) min_image_k(pk, dx, dy);
      pair_term<POT, F32>(pk, dx, dy, dx * dx + dy * dy, fx, fy);
    }
  }
}

// The same for a tile whose halo is larger than the LDS stage: each listed
// position is read from the set through L2, one entry at a time.
template <int POT, bool F32>
__device__ __forceinline__ void list_forces_l2(const PairK& pk, const TileD& td, int q, int kn, const double2* cur,
                                               DescPtr D, int nr, double xi, double yi, double& fx, double& fy) {
  const unsigned short* L = reinterpret_cast<const unsigned short*>(td.lst);
#pragma unroll 1
  for (int k = 0; k < kn; ++k) {
    const int e = L[((long)(k >> 2) * td.n + q) * 4 + (k & 3)];
    const double2 pj = gld(cur + desc_global(D, e));
    double dx = pj.x - xi, dy = pj.y - yi;
    min_image_k(pk, dx, dy);
    pair_term<POT, F32>(pk, dx, dy, dx * dx + dy * dy, fx, fy);
  }
}

// One particle of a fused BAOAB step (phases as k_baoab_fused): k is its
// stage index, q its index in the set.  Returns the rebuild request.
template <int POT, int PHASE, bool F32>
__device__ __forceinline__ bool tile_particle(const TileD& td, const TileSet& A, const SlotOut& so, const BoxD& b,
                                              const ConstraintD& c, const PairConst& pc, const FusedArgs& fa, int par,
                                              const double2* cur, const double2* s_pos, bool staged, bool wrap,
                                              DescPtr D, int nr, int k, int q) {
  const double dt = fa.dt, dt2 = 0.5 * dt, m = fa.mass_u, fact = fa.fact_u;
  const int kn = gld(td.len + q);
  double2 pi;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (staged) {
    pi = s_pos[k];
    list_forces<POT, F32>(pair_k(pc, b), td, q, kn, s_pos, wrap, pi.x, pi.y, fx, fy);
  } else {
    pi = gld(cur + q);
    list_forces_l2<POT, F32>(pair_k(pc, b), td, q, kn, cur, D, nr, pi.x, pi.y, fx, fy);
  }
  // On a plane only the enforce's normalisation changes a director in a
  // Langevin run, and the first step of a run (phase 0) normalises them all:
  // later steps read a director only for self propulsion.
  constexpr bool kDirs = PHASE == 0;
  const double2 n0 = (kDirs || pc.has_sp) ? gld(A.nd[0] + q) : make_double2(0.0, 0.0);
  if (pc.has_sp) { fx += pc.sp_alpha * n0.x; fy += pc.sp_alpha * n0.y; fz += pc.sp_alpha * 0.0; }
  const double2 v0 = gld(A.v + q);
  double vx = v0.x, vy = v0.y, vz = 0.0;
  if (PHASE >= 1) {  // closing B of step t-1
    vx += fact * fx; vy += fact * fy; vz += fact * fz;
  }
  if (PHASE <= 1) {
    // the O-step noise; the z draw only reaches v_z, which the plane enforce
    // resets, so only the first pair of the block is formed
    double gn[2] = {0.0, 0.0};
    double B = fa.Bc;
    if (B < 0.0) {
      const double T = fa.temp.at(eff_step(fa.clock, fa.t_off));
      B = sqrt(T * (1.0 - exp(-2.0 * fa.gamma * dt)));
    }
    if (B != 0.0)
      gauss2(fa.seed, TAG_LANGEVIN, (uint32_t)gld(A.id + q),
             eff_call(fa.clock, fa.clock_slot, 0ULL) + (unsigned long long)fa.t_off, gn);
    PS p;
    p.x = pi.x; p.y = pi.y; p.z = c.zpos; p.vx = vx; p.vy = vy; p.vz = vz;
    p.nx = n0.x; p.ny = n0.y; p.nz = 0.0;
    p.dix = p.diy = p.diz = 0;
    p.vx += fact * fx; p.vy += fact * fy; p.vz += fact * fz;
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    enforce<kSurfPlane>(b, c, p, kDirs);
    p.vx *= fa.exp_dt; p.vy *= fa.exp_dt; p.vz *= fa.exp_dt;
    if (B != 0.0) {
      const double sfac = fa.sf_u >= 0.0 ? fa.sf_u : B / sqrt(m);
      p.vx += sfac * gn[0]; p.vy += sfac * gn[1]; p.vz += sfac * 0.0;
    }
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    // the director was normalised by the first enforce of the step
    enforce<kSurfPlane>(b, c, p, false);
    if (fa.head_enforce) enforce<kSurfPlane>(b, c, p, false);
    gst(A.xy[par ^ 1] + q, make_double2(p.x, p.y));
    gst(A.v + q, make_double2(p.vx, p.vy));
    if (kDirs && (p.nx != n0.x || p.ny != n0.y)) gst(A.nd[0] + q, make_double2(p.nx, p.ny));
    if (p.dix | p.diy | p.diz) {
      const int s = A.slot[q];
      if (p.dix) so.ix[s] += p.dix;
      if (p.diy) so.iy[s] += p.diy;
      if (p.diz) so.iz[s] += p.diz;
    }
    const double2 o = gld(A.o + q);
    double ddx = o.x - p.x, ddy = o.y - p.y, ddz = 0.0;
    min_image(b, ddx, ddy, ddz);
    return !(ddx * ddx + ddy * ddy + ddz * ddz < td.pad2q);
  }
  // closing phase: F(x_t) is the state the reference leaves behind
  A.v[q] = make_double2(vx, vy);
  const int s = A.slot[q];
  so.fx[s] = fx; so.fy[s] = fy; so.fz[s] = fz;
  return false;
}

// One fused BAOAB step on the current set: phases as k_baoab_fused.
// The run's constant arguments come from device memory, read with scalar
// loads where they are used: as kernel arguments they would occupy scalar
// registers for the whole kernel and be spilled and reloaded lane by lane.
template <int POT, int PHASE, bool F32>
__global__ void __launch_bounds__(kTileThreads, 8) k_tile_step(const TileStepArgs* __restrict__ P, int par,
                                                               int t_off, int head_enforce) {
  const TileD& td = P->td;
  const SlotOut& so = P->so;
  const BoxD& b = P->b;
  const ConstraintD& c = P->c;
  const PairConst& pc = P->pc;
  FusedArgs fa = P->fa;
  fa.t_off = t_off;
  fa.head_enforce = head_enforce;
  TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  const int S = (w ^ (w >> 1)) & 1;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    // commit a finished build, clear the request it answered
    if (w & 2) atomicXor(&ctl->word, 3);
    atomicAnd(&ctl->word, ~(4 << (par ^ 1)));
    if (ctl->init) ctl->init = 0;
  }
  const TileSet& A = td.sets[S];
  __shared__ double2 s_pos[kTileHaloMax + 1];
  const int t = xcd_tile(td);
  const TileGeom g = tile_geom(td, t);
  if (!g.active) return;
  const DescPtr D = tile_desc(td, t);
  const int total = D[0], kb = D[1], ke = D[2], qa = D[3], nr = D[4];
  // Workgroup (t, r) takes the tile's particles r * kTileThreads .. + kTileThreads - 1
  // (one per thread, no loop: a loop would keep the step's loop-invariant
  // values in registers across the pair loop).  A later-round workgroup
  // whose tile holds no particle of its round returns at once.
  const int round = (int)blockIdx.y;
  if (round > 0 && kb + round * kTileThreads >= ke) return;
  const double2* cur = A.xy[par];
  const bool staged = total <= kTileHaloMax;
  if (staged) {
    // every load of the stage issued before the first LDS store
    double2 v[kTileHaloMax / kTileThreads];
    if (nr <= 3) {
#pragma unroll
      for (int u = 0; u < kTileHaloMax / kTileThreads; ++u) {
        const int k = (int)threadIdx.x + u * kTileThreads;
        if (k < total) v[u] = gld(cur + desc_global3(D, k));
      }
    } else {
#pragma unroll
      for (int u = 0; u < kTileHaloMax / kTileThreads; ++u) {
        const int k = (int)threadIdx.x + u * kTileThreads;
        if (k < total) v[u] = gld(cur + desc_global(D, k));
      }
    }
#pragma unroll
    for (int u = 0; u < kTileHaloMax / kTileThreads; ++u) {
      const int k = (int)threadIdx.x + u * kTileThreads;
      if (k < total) s_pos[k] = v[u];
    }
    if (threadIdx.x == 0) s_pos[total] = make_double2(1.0e30, 1.0e30);  // the lists' padding
  }
  __syncthreads();
  const bool wrap = b.periodic && (g.i0 == 0 || g.i0 + g.nc >= td.ncx || g.ty == 0 || g.ty == td.ncy - 1);
  bool need = false;
  const int k = kb + round * kTileThreads + (int)threadIdx.x;
  if (k < ke)
    need = tile_particle<POT, PHASE, F32>(td, A, so, b, c, pc, fa, par, cur, s_pos, staged, wrap, D, nr, k,
                                          qa + (k - kb));
  if (PHASE != 2) {
    if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomicOr(&ctl->word, 4 << par);
  }
}

// ---------------------------------------------------------------------------
// Fused Brownian step on the tiles: loop-head forces and aligner torques of
// step t from the staged positions and directors, then
// IntegratorBrownian::integrate (src/integrators/integrator_brownian.cpp:
// 42-117) on the plane.  Pair terms as the list-based k_brownian_fused
// (pair_sum.hpp: soft, LJ; polar src/aligner/pair/pair_polar_aligner.cpp:
// 58-103, nematic src/aligner/pair/pair_nematic_aligner.cpp:53-92, with
// J the effective coupling).  Positions and directors are read from buffer
// par and written to par ^ 1.
// ---------------------------------------------------------------------------
template <int POT, int AL, bool F32>
__device__ __forceinline__ void align_term(const PairK& pk, double r_sq, double nxi, double nyi, double nxj,
                                           double nyj, double& tz) {
  // n_i x n_j: on a plane only its z component is non-zero
  const double cz = nxi * nyj - nyi * nxj;
  if (AL & AL_POLAR) {
    if (r_sq <= pk.al_rc2) tz += pk.al_J * cz;
  }
  if (AL & AL_NEMATIC) {
    if (r_sq <= pk.al_rc2) {
      const double nn = nxi * nxj + nyi * nyj;
      tz += pk.al_J * nn * cz;
    }
  }
}

template <int POT, int AL, bool F32>
__device__ __forceinline__ void list_forces_al(PairK pk, const TileD& td, int q, int kn, const double2* s_pos,
                                               const double2* s_dir, bool wrap, double xi, double yi, double nxi,
                                               double nyi, double& fx, double& fy, double& tz) {
  asm volatile("" : "+v"(pk.lj_rc2), "+v"(pk.lj_sig2), "+v"(pk.lj_e48), "+v"(pk.soft_k), "+v"(pk.soft_a));
  asm volatile("" : "+v"(pk.xlo), "+v"(pk.xhi), "+v"(pk.lx), "+v"(pk.ylo), "+v"(pk.yhi), "+v"(pk.ly));
  asm volatile("" : "+v"(pk.al_J), "+v"(pk.al_rc2));
  const unsigned long long* L = reinterpret_cast<const unsigned long long*>(td.lst) + q;
  const long n = td.n;
  const int rows = (kn + 3) >> 2;
  unsigned long long nxt = rows > 0 ? gld(L) : 0ULL;
  for (int r = 0; r < rows; ++r) {
    const unsigned long long cur = nxt;
    if (r + 1 < rows) nxt = gld(L + (long)(r + 1) * n);
    double2 pj[4], nj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = (int)((cur >> (16 * u)) & 0xffffULL);
      pj[u] = s_pos[e];
      if (AL) nj[u] = s_dir[e];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double dx = pj[u].x - xi, dy = pj[u].y - yi;
      if (wrap) min_image_k(pk, dx, dy);
      const double r_sq = dx * dx + dy * dy;
      pair_term<POT, F32>(pk, dx, dy, r_sq, fx, fy);
      if (AL) align_term<POT, AL, F32>(pk, r_sq, nxi, nyi, nj[u].x, nj[u].y, tz);
    }
  }
}

// The same through L2 for a tile whose halo is larger than the LDS stage.
template <int POT, int AL, bool F32>
__device__ __forceinline__ void list_forces_al_l2(const PairK& pk, const TileD& td, int q, int kn,
                                                  const double2* cur, const double2* dir, DescPtr D, double xi,
                                                  double yi, double nxi, double nyi, double& fx, double& fy,
                                                  double& tz) {
  const unsigned short* L = reinterpret_cast<const unsigned short*>(td.lst);
#pragma unroll 1
  for (int k = 0; k < kn; ++k) {
    const int e = L[((long)(k >> 2) * td.n + q) * 4 + (k & 3)];
    const int j = desc_global(D, e);
    const double2 pj = gld(cur + j);
    double dx = pj.x - xi, dy = pj.y - yi;
    min_image_k(pk, dx, dy);
    const double r_sq = dx * dx + dy * dy;
    pair_term<POT, F32>(pk, dx, dy, r_sq, fx, fy);
    if (AL) {
      const double2 nj = gld(dir + j);
      align_term<POT, AL, F32>(pk, r_sq, nxi, nyi, nj.x, nj.y, tz);
    }
  }
}

template <int POT, int AL, bool F32>
__device__ __forceinline__ bool tile_brownian_particle(const TileBrownArgs& P, const TileSet& A, int par, int t_off,
                                                       int head_enforce, int store, const double2* cur,
                                                       const double2* dir,
                                                       const double2* s_pos, const double2* s_dir, bool staged,
                                                       bool wrap, DescPtr D, int k, int q) {
  const TileD& td = P.td;
  const BoxD& b = P.b;
  const ConstraintD& c = P.c;
  const PairConst& pc = P.pc;
  const int kn = gld(td.len + q);
  PairK pk = pair_k(pc, b);
  pk.al_J = P.al_J;
  pk.al_rc2 = P.al_rc2;
  double2 pi, ni;
  double fx = 0.0, fy = 0.0, fz = 0.0, tz = 0.0;
  if (staged) {
    pi = s_pos[k];
    ni = AL ? s_dir[k] : gld(dir + q);  // directors are staged only for an aligner
    list_forces_al<POT, AL, F32>(pk, td, q, kn, s_pos, s_dir, wrap, pi.x, pi.y, ni.x, ni.y, fx, fy, tz);
  } else {
    pi = gld(cur + q);
    ni = gld(dir + q);
    list_forces_al_l2<POT, AL, F32>(pk, td, q, kn, cur, dir, D, pi.x, pi.y, ni.x, ni.y, fx, fy, tz);
  }
  if (pc.has_sp) { fx += pc.sp_alpha * ni.x; fy += pc.sp_alpha * ni.y; fz += pc.sp_alpha * 0.0; }
  const uint32_t pid = (uint32_t)gld(A.id + q);
  const double2 v0 = gld(A.v + q);
  PS p;
  p.x = pi.x; p.y = pi.y; p.z = c.zpos;
  p.vx = v0.x; p.vy = v0.y; p.vz = 0.0;
  p.nx = ni.x; p.ny = ni.y; p.nz = 0.0;
  p.dix = p.diy = p.diz = 0;
  // IntegratorBrownian::integrate for this particle (k_brownian_fused, mode 0)
  const unsigned long long call = eff_call(P.clock, P.clock_slot, 0ULL) + (unsigned long long)t_off;
  double g_rot = 0.0, u_flip = 1.0;
  gauss1_u1(P.seed, TAG_BROWNIAN, pid, call, g_rot, u_flip);
  if (P.nematic && u_flip < P.flip_p) {
    p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz;
    if (P.velocity) { p.vx = -p.vx; p.vy = -p.vy; p.vz = -p.vz; }
  }
  const double T = P.temp.at(eff_step(P.clock, t_off));
  const double B = sqrt(2.0 * P.mu * T);
  const double sqrt_dt = sqrt(P.dt);
  const double fdx = P.v0 * p.nx + P.mu * fx, fdy = P.v0 * p.ny + P.mu * fy, fdz = P.v0 * p.nz + P.mu * fz;
  p.vx = fdx; p.vy = fdy; p.vz = fdz;
  p.x += P.dt * fdx; p.y += P.dt * fdy; p.z += P.dt * fdz;
  if (T > 0.0) {
    double g[3];
    gauss3(P.seed, TAG_BROWNIAN, pid, call, g);
    const double frx = B * g[0], fry = B * g[1], frz = B * g[2];
    p.vx += frx; p.vy += fry; p.vz += frz;
    p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
  }
  const bool fz0 = enforce<kSurfPlane>(b, c, p);
  const double om = P.mur * project_torque<kSurfPlane>(c, p, 0.0, 0.0, tz);
  if (store) {
    const int s = A.slot[q];
    P.so.fx[s] = fx; P.so.fy[s] = fy; P.so.fz[s] = fz0 ? 0.0 : fz;
    if (AL) { P.so.tx[s] = 0.0; P.so.ty[s] = 0.0; P.so.tz[s] = tz; }
  }
  const double dtheta = P.dt * om + P.stoch * g_rot;
  rotate_director<kSurfPlane>(c, p, dtheta);
  if (P.velocity) rotate_velocity<kSurfPlane>(c, p, dtheta);
  if (head_enforce) enforce<kSurfPlane>(b, c, p);
  gst(A.xy[par ^ 1] + q, make_double2(p.x, p.y));
  gst(A.nd[par ^ 1] + q, make_double2(p.nx, p.ny));
  gst(A.v + q, make_double2(p.vx, p.vy));
  A.om[q] = om;
  if (p.dix | p.diy | p.diz) {
    const int s = A.slot[q];
    if (p.dix) P.so.ix[s] += p.dix;
    if (p.diy) P.so.iy[s] += p.diy;
    if (p.diz) P.so.iz[s] += p.diz;
  }
  const double2 o = gld(A.o + q);
  double ddx = o.x - p.x, ddy = o.y - p.y, ddz = 0.0;
  min_image(b, ddx, ddy, ddz);
  return !(ddx * ddx + ddy * ddy + ddz * ddz < td.pad2q);
}

// One fused Brownian step on the current set (launch layout as k_tile_step;
// the stage holds positions and, with an aligner, directors).
template <int POT, int AL, bool F32>
__global__ void __launch_bounds__(kTileThreads, 5) k_tile_brownian(const TileBrownArgs* __restrict__ P, int par,
                                                                   int t_off, int head_enforce, int store) {
  const TileD& td = P->td;
  TileCtl* ctl = td.ctl;
  const int w = ctl_word(ctl);
  const int S = (w ^ (w >> 1)) & 1;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    // commit a finished build, clear the request it answered
    if (w & 2) atomicXor(&ctl->word, 3);
    atomicAnd(&ctl->word, ~(4 << (par ^ 1)));
    if (ctl->init) ctl->init = 0;
  }
  const TileSet& A = td.sets[S];
  __shared__ double2 s_pos[kTileHaloMax + 1];
  __shared__ double2 s_dir[AL ? kTileHaloMax + 1 : 1];
  const int t = xcd_tile(td);
  const TileGeom g = tile_geom(td, t);
  if (!g.active) return;
  const DescPtr D = tile_desc(td, t);
  const int total = D[0], kb = D[1], ke = D[2], qa = D[3], nr = D[4];
  const int round = (int)blockIdx.y;
  if (round > 0 && kb + round * kTileThreads >= ke) return;
  const double2* cur = A.xy[par];
  const double2* dir = A.nd[par];
  const bool staged = total <= kTileHaloMax;
  if (staged) {
    double2 v[kTileHaloMax / kTileThreads], d[kTileHaloMax / kTileThreads];
#pragma unroll
    for (int u = 0; u < kTileHaloMax / kTileThreads; ++u) {
      const int k = (int)threadIdx.x + u * kTileThreads;
      if (k < total) {
        const int j = nr <= 3 ? desc_global3(D, k) : desc_global(D, k);
        v[u] = gld(cur + j);
        if (AL) d[u] = gld(dir + j);
      }
    }
#pragma unroll
    for (int u = 0; u < kTileHaloMax / kTileThreads; ++u) {
      const int k = (int)threadIdx.x + u * kTileThreads;
      if (k < total) {
        s_pos[k] = v[u];
        if (AL) s_dir[k] = d[u];
      }
    }
    if (threadIdx.x == 0) {
      s_pos[total] = make_double2(1.0e30, 1.0e30);  // the lists' padding
      if (AL) s_dir[total] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  const bool wrap = P->b.periodic && (g.i0 == 0 || g.i0 + g.nc >= td.ncx || g.ty == 0 || g.ty == td.ncy - 1);
  bool need = false;
  const int k = kb + round * kTileThreads + (int)threadIdx.x;
  if (k < ke)
    need = tile_brownian_particle<POT, AL, F32>(*P, A, par, t_off, head_enforce, store, cur, dir, s_pos, s_dir,
                                                staged, wrap, D, k, qa + (k - kb));
  if (__ballot(need) != 0 && (threadIdx.x & 63) == 0) atomicOr(&ctl->word, 4 << par);
}

__global__ void k_tile_load(TileD td, Fields f, NListD nl) {
  const int p = gtid();
  if (p == 0) {
    TileCtl z{};
    z.word = 1 | (4 << 1);  // set 1 holds the loaded state; the first build answers parity 1
    z.init = 1;
    *td.ctl = z;
  }
  if (p < td.n) {
    td.sets[1].xy[0][p] = make_double2(f.x[p], f.y[p]);
    td.sets[1].o[p] = make_double2(nl.ox[p], nl.oy[p]);
    td.sets[1].v[p] = make_double2(f.vx[p], f.vy[p]);
    td.sets[1].nd[0][p] = make_double2(f.nx[p], f.ny[p]);
    td.sets[1].om[p] = f.omega[p];
    td.sets[1].id[p] = f.id[p];
    td.sets[1].slot[p] = p;
  }
}

__global__ void k_tile_store(TileD td, Fields f, NListD nl, int par) {
  const int w = ctl_word(td.ctl);
  const int S = (w ^ (w >> 1)) & 1;
  const int q = gtid();
  if (q < td.n) {
    const int s = td.sets[S].slot[q];
    const double2 p = td.sets[S].xy[par][q];
    f.x[s] = p.x; f.y[s] = p.y;
    if (f.xy) f.xy[s] = p;
    const double2 v = td.sets[S].v[q];
    f.vx[s] = v.x; f.vy[s] = v.y;
    const double2 d = td.sets[S].nd[td.dirs2 ? par : 0][q];
    f.omega[s] = td.sets[S].om[q];
    f.nx[s] = d.x; f.ny[s] = d.y;
    const double2 o = td.sets[S].o[q];
    nl.ox[s] = o.x; nl.oy[s] = o.y;
  }
}

int capped_grid(long n) {
  long g = (n + kTileThreads - 1) / kTileThreads;
  return (int)std::max(1L, std::min(g, 2048L));
}

int tile_grid(const TileD& td) { return 8 * ((td.ntiles + 7) / 8); }

template <int POT, int PHASE>
void step_dispatch(hipStream_t st, int fp32, const TileStepArgs* args, int grid, int par, int t_off,
                   int head_enforce) {
  const dim3 g(grid, kTileMaxOwn / kTileThreads), blk(kTileThreads);
  if (fp32) hipLaunchKernelGGL((k_tile_step<POT, PHASE, true>), g, blk, 0, st, args, par, t_off, head_enforce);
  else hipLaunchKernelGGL((k_tile_step<POT, PHASE, false>), g, blk, 0, st, args, par, t_off, head_enforce);
}

template <int POT>
void step_phase(hipStream_t st, int phase, int fp32, const TileStepArgs* args, int grid, int par, int t_off,
                int head_enforce) {
  if (phase == 0) step_dispatch<POT, 0>(st, fp32, args, grid, par, t_off, head_enforce);
  else if (phase == 1) step_dispatch<POT, 1>(st, fp32, args, grid, par, t_off, head_enforce);
  else step_dispatch<POT, 2>(st, fp32, args, grid, par, t_off, head_enforce);
}

}  // namespace

void launch_tile_load(hipStream_t st, TileD td, Fields f, NListD nl) {
  hipLaunchKernelGGL(k_tile_load, dim3(grid1(td.n)), dim3(kBlock), 0, st, td, f, nl);
}

void launch_tile_build(hipStream_t st, TileD td, BoxD box, int init, int par) {
  const int g = std::max(1, std::min(td.ntiles, 2048));
  hipLaunchKernelGGL(k_tile_bin, dim3(capped_grid(td.n)), dim3(kTileThreads), 0, st, td, par);
  hipLaunchKernelGGL(k_tile_cells, dim3((td.ntiles + kTilesPerBlock - 1) / kTilesPerBlock), dim3(kTileThreads), 0,
                     st, td, par);
  if (init) {
    hipLaunchKernelGGL(k_tile_scatter, dim3(capped_grid(td.n)), dim3(kTileThreads), 0, st, td, par);
    hipLaunchKernelGGL(k_tile_sort, dim3(capped_grid(td.ncell)), dim3(kTileThreads), 0, st, td, par);
  }
  hipLaunchKernelGGL(k_tile_build, dim3(g), dim3(kTileThreads), 0, st, td, box, par);
}

void launch_tile_step(hipStream_t st, int pot_mask, int phase, int fp32, const TileD& td, const TileStepArgs* args,
                      int par, int t_off, int head_enforce) {
  const int grid = tile_grid(td);
  switch (pot_mask) {
    case 1: step_phase<1>(st, phase, fp32, args, grid, par, t_off, head_enforce); break;
    case 2: step_phase<2>(st, phase, fp32, args, grid, par, t_off, head_enforce); break;
    default: step_phase<3>(st, phase, fp32, args, grid, par, t_off, head_enforce); break;
  }
}

template <int POT, int AL>
void brownian_dispatch(hipStream_t st, int fp32, const TileBrownArgs* args, int grid, int par, int t_off,
                       int head_enforce, int store) {
  const dim3 g(grid, kTileMaxOwn / kTileThreads), blk(kTileThreads);
  if (fp32)
    hipLaunchKernelGGL((k_tile_brownian<POT, AL, true>), g, blk, 0, st, args, par, t_off, head_enforce, store);
  else
    hipLaunchKernelGGL((k_tile_brownian<POT, AL, false>), g, blk, 0, st, args, par, t_off, head_enforce, store);
}

template <int POT>
void brownian_al(hipStream_t st, int al_mask, int fp32, const TileBrownArgs* args, int grid, int par, int t_off,
                 int head_enforce, int store) {
  if (al_mask == AL_POLAR)
    brownian_dispatch<POT, AL_POLAR>(st, fp32, args, grid, par, t_off, head_enforce, store);
  else if (al_mask == AL_NEMATIC)
    brownian_dispatch<POT, AL_NEMATIC>(st, fp32, args, grid, par, t_off, head_enforce, store);
  else
    brownian_dispatch<POT, 0>(st, fp32, args, grid, par, t_off, head_enforce, store);
}

void launch_tile_brownian(hipStream_t st, int pot_mask, int al_mask, int fp32, const TileD& td,
                          const TileBrownArgs* args, int par, int t_off, int head_enforce, int store) {
  const int grid = tile_grid(td);
  switch (pot_mask) {
    case 1: brownian_al<1>(st, al_mask, fp32, args, grid, par, t_off, head_enforce, store); break;
    case 2: brownian_al<2>(st, al_mask, fp32, args, grid, par, t_off, head_enforce, store); break;
    default: brownian_al<3>(st, al_mask, fp32, args, grid, par, t_off, head_enforce, store); break;
  }
}

void launch_tile_store(hipStream_t st, TileD td, Fields f, NListD nl, int par) {
  hipLaunchKernelGGL(k_tile_store, dim3(grid1(td.n)), dim3(kBlock), 0, st, td, f, nl, par);
}

}  // namespace samos
