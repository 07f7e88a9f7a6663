// C-ABI of the device pipeline (include/samos_hip.h): the context object,
// state transfer in particle-id order, plugin registration with the
// reference's parameter semantics, the run loop and captured step chunks.
#include "samos_hip.h"
#include "kernels.hpp"
#include "tiles.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace samos;

namespace {

thread_local std::string g_create_error;

struct Integ {
  std::string name;     // brownian, brownian_pos, brownian_align, langevin, fire
  std::string group = "all";  // Integrator::m_group_name (src/integrators/integrator.hpp:98-108)
  unsigned gbit = 1u;         // that group's bit in the per-slot group column
  int slot = 0;       // clock slot
  double dt = 0.0;
  TempD temp{0, 0.0, 1.0, 1000};
  double v0 = 1.0, nu = 1.0, mu = 1.0, mur = 1.0, stoch = 0.0, flip_p = 0.0;
  int nematic = 0, velocity = 0;
  uint32_t seed = 0;
  unsigned long long calls = 0;
  double gamma = 1.0; int method = 0; double* R = nullptr;
  double alpha = 0.1, alpha_init = 0.1, f_alpha = 0.99, F_tol = 1e-4, E_tol = 1e-6, f_inc = 1.1, f_dec = 0.5,
         dt_max = 1.0;
  int N_min = 5, last_neg = 0; bool converged = false; double old_energy = 1e15;
  double last[4] = {0, 0, 0, 0};
  double limit = -1.0, theta_limit = -1.0;  // nve; negative: not given
  FireD* fire = nullptr;  // fire: scalar state on the device
  int fire_ndof3 = -1;    // 3 N of the group as last uploaded
};

struct ProfAcc {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
};

}  // namespace

struct samos_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  int n = 0, np = 0;
  BoxD box{};
  std::string err;
  double* dcol[kNumDoubleFields] = {};
  int* icol[kNumIntFields] = {};      // type, ix, iy, iz, coord, id, groups
  double* alt[3] = {};                // second position buffer for fused steps
  double* altn[3] = {};               // second director buffer for fused Brownian steps
  double2* xy[2] = {};                // interleaved (x, y) of the two position buffers
  std::vector<int> h_id;              // slot -> id
  // particle groups (System::make_group): bit index per name ("all" = 0) and
  // each particle's bits in id order; the device holds them per slot
  std::map<std::string, int> group_index{{"all", 0}};
  std::vector<unsigned> h_groups;
  std::string next_group = "all";     // group of the next integrator created
  std::string last_key;               // map key of the integrator created last
  double dt = 0.0;
  int ntypes = 1;
  // neighbour list
  bool has_nlist = false;
  double rcut = 3.0, pad = 0.5;
  NListD nl{};
  int* d_flags = nullptr;
  int* d_nnb = nullptr;               // neighbour counts, zero while no list exists
  // sort scratch
  void* sort_tmp = nullptr; size_t sort_tmp_bytes = 0;
  unsigned long long *keys = nullptr, *keys2 = nullptr; int *perm = nullptr, *perm2 = nullptr;
  double* dtmp = nullptr; int* itmp = nullptr;
  // interactions
  bool has_soft = false, has_lj = false, has_pol = false, has_nem = false, has_sp = false;
  double soft_k = 1, soft_a = 2; int soft_radii = 0;
  double lj_eps = 1, lj_sig = 1, lj_rc = 3; int lj_shifted = 0, lj_radii = 0;
  double pol_J = 1, pol_rc = 0; double nem_J = 1, nem_rc = 0; bool nem_pp = false;
  double sp_alpha = 1.0;
  std::vector<double> soft_tab, lj_tab, pol_tab, nem_tab;  // host copies [3][16][16]
  double* d_tab = nullptr;                                  // 4 tables x 3 x 256
  ConstraintD con{-1, 1.0, 0.0, 0, 0.0, 0.0, 0.0, 100, 1e-6};
  // sphere radius schedule (Constraint rescale parameters)
  double rescale = 1.0, rescale_scale = 1.0;
  int rescale_steps = 1000, rescale_freq = 10;
  std::map<std::string, Integ> integs;
  ClockD* d_clock = nullptr;
  ChunkCtl* d_chunk = nullptr;  // control words of the captured step chunks
  int chunk_slots = 32;         // fused steps per captured chunk (even)
  int time_step = 0;
  double* d_en = nullptr;       // 5 * np per-slot energies: soft, lj, polar, nematic, further pair potentials
  double* d_cols = nullptr;     // 3 * np FIRE columns
  double* d_partial = nullptr;
  double* d_out = nullptr;      // 8 reduced values
  double* h_out = nullptr;      // pinned
  unsigned long long* d_counter = nullptr;
  bool prof = false;
  std::map<std::string, ProfAcc> profs;
  // Cooperative rebuild: -1 untried, 1 works, 0 off.  Off by default: the five
  // stage kernels replayed from a captured graph measured faster on MI355X
  // (skipped stages cost little inside a graph; the cooperative kernel runs
  // its build phase at lower occupancy).
  int coop = 0;
  int gather_batch = 4;         // neighbour gathers in flight per thread in the fused step
  bool run_begun = false;       // loop head of t = 0 of the current run done
  bool mass_known = false;      // mass_u valid
  double mass_u = -1.0;         // common particle mass, or -1
  AgeCtl* d_age = nullptr;      // common-age state of the fused ranges
  bool age_checked = false;     // d_age describes the age column as it is now
  // Instantiated step-chunk graphs, reused across run ranges.  The key holds
  // every value the captured nodes bake in (pointers, sizes, parameters), so
  // an equal key means an identical graph.
  struct GraphEntry { std::string key; hipGraphExec_t exec; unsigned long long used; };
  std::vector<GraphEntry> graphs;
  unsigned long long graph_tick = 0;
  long long graph_stats[2] = {0, 0};  // instantiations, reuses
  long long overflow_recoveries = 0;  // list overflows recovered inside runs
  int resort_every = 200;             // list builds between storage re-sorts inside runs (0: never)
  int builds_at_sort = 0;             // the builds count after the last re-sort
  long long resorts = 0;              // re-sorts done inside runs
  // cell-tiled engine of planar fused runs (allocated on first use)
  TileD td{};
  void* tile_mem = nullptr;
  TileStepArgs* td_args = nullptr;  // the step kernels' constant arguments (inside tile_mem)
  TileStepArgs h_tile_args{};       // their host copy
  TileBrownArgs* tb_args = nullptr; // the same for tiled Brownian runs
  TileBrownArgs h_tb_args{};
  size_t tile_bytes = 0;
  int tile_key[5] = {-1, -1, -1, -1, -1};  // n, ncx, ncy, T, K of the allocation
  long long tile_stats[4] = {0, 0, 0, 0};  // tiled runs, rebuilds, largest candidate count, largest halo
  int tile_cells = 0;                      // cells per tile along x; 0: from the density
  int brownian_lanes = 0;                  // threads per particle in the fused Brownian step; 0: automatic
  // further pair potentials (morse, gaussian, yukawa, soft_attractive), by name
  std::map<std::string, std::pair<int, std::vector<double>>> extras;
  std::map<std::string, std::pair<int, int>> extra_flags;  // use_particle_radii, shifted
  ExtraD extra{};
};

namespace {

constexpr int kTab = 3 * kMaxTypes * kMaxTypes;

// Fills used inside captured step sequences (where a memset node is avoided).
__global__ void k_fill_d(double* p, double v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_i(int* p, int v, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}


void fill_d(hipStream_t st, double* p, double v, long n) {
  const int g = (int)std::max<long>(1, std::min<long>((n + kBlock - 1) / kBlock, 1024));
  hipLaunchKernelGGL(k_fill_d, dim3(g), dim3(kBlock), 0, st, p, v, n);
}

void fill_i(hipStream_t st, int* p, int v, long n) {
  const int g = (int)std::max<long>(1, std::min<long>((n + kBlock - 1) / kBlock, 1024));
  hipLaunchKernelGGL(k_fill_i, dim3(g), dim3(kBlock), 0, st, p, v, n);
}

#define HIPCHK(expr)                                                            \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ctx->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr; \
      return -2;                                                                \
    }                                                                           \
  } while (0)

int fail(samos_ctx* ctx, const std::string& m) { ctx->err = m; return -1; }

Fields fields(samos_ctx* ctx, int parity = 0) {
  Fields f;
  double** d = ctx->dcol;
  f.x = parity ? ctx->alt[0] : d[0]; f.y = parity ? ctx->alt[1] : d[1]; f.z = parity ? ctx->alt[2] : d[2];
  f.vx = d[3]; f.vy = d[4]; f.vz = d[5]; f.nx = d[6]; f.ny = d[7]; f.nz = d[8];
  f.fx = d[9]; f.fy = d[10]; f.fz = d[11]; f.tx = d[12]; f.ty = d[13]; f.tz = d[14];
  f.omega = d[15]; f.age = d[16]; f.radius = d[17]; f.mass = d[18];
  f.type = ctx->icol[0]; f.ix = ctx->icol[1]; f.iy = ctx->icol[2]; f.iz = ctx->icol[3];
  f.coord = ctx->icol[4]; f.id = ctx->icol[5]; f.groups = ctx->icol[6];
  f.xy = nullptr;
  return f;
}

int pot_mask(const samos_ctx* c) { return (c->has_soft ? POT_SOFT : 0) | (c->has_lj ? POT_LJ : 0); }
int al_mask(const samos_ctx* c) { return (c->has_pol ? AL_POLAR : 0) | (c->has_nem ? AL_NEMATIC : 0); }

InteractD interact_params(samos_ctx* ctx) {
  InteractD ip{};
  const int nt = kMaxTypes;
  ip.soft_tab = ctx->d_tab; ip.soft_nt = nt; ip.soft_radii = ctx->soft_radii;
  ip.lj_tab = ctx->d_tab + kTab; ip.lj_nt = nt; ip.lj_radii = ctx->lj_radii; ip.lj_shifted = ctx->lj_shifted;
  ip.pol_tab = ctx->d_tab + 2 * kTab; ip.pol_nt = nt;
  ip.nem_tab = ctx->d_tab + 3 * kTab; ip.nem_nt = nt;
  ip.has_sp = ctx->has_sp ? 1 : 0; ip.sp_alpha = ctx->sp_alpha;
  return ip;
}

int upload_tables(samos_ctx* ctx) {
  std::vector<double> all(4 * kTab, 0.0);
  auto put = [&](int k, const std::vector<double>& t) { if (!t.empty()) std::copy(t.begin(), t.end(), all.begin() + k * kTab); };
  put(0, ctx->soft_tab); put(1, ctx->lj_tab); put(2, ctx->pol_tab);
  if (!ctx->nem_tab.empty()) {
    std::vector<double> t = ctx->nem_tab;
    // PairNematicAlign::compute: global J is doubled, per-type J used as given
    // (src/aligner/pair/pair_nematic_aligner.cpp:37,57-61).
    if (!ctx->nem_pp) for (int q = 0; q < kMaxTypes * kMaxTypes; ++q) t[q] = 2.0 * ctx->nem_J;
    put(3, t);
  }
  HIPCHK(hipMemcpyAsync(ctx->d_tab, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

std::vector<double> make_tab(double a, double b, double c) {
  std::vector<double> t(kTab);
  const int nn = kMaxTypes * kMaxTypes;
  for (int q = 0; q < nn; ++q) { t[q] = a; t[nn + q] = b; t[2 * nn + q] = c; }
  return t;
}

void set_tab(std::vector<double>& t, int t1, int t2, double a, double b, double c) {
  const int nn = kMaxTypes * kMaxTypes;
  int i = (t1 - 1) * kMaxTypes + (t2 - 1), j = (t2 - 1) * kMaxTypes + (t1 - 1);
  t[i] = t[j] = a; t[nn + i] = t[nn + j] = b; t[2 * nn + i] = t[2 * nn + j] = c;
}

// ---- cached step-chunk graphs ----------------------------------------------
// Byte key of scalars and pointers, added one member at a time (no struct
// padding enters it).
struct KeyBuf {
  std::string s;
  template <class T> KeyBuf& operator<<(const T& v) {
    s.append(reinterpret_cast<const char*>(&v), sizeof(T));
    return *this;
  }
};

KeyBuf& operator<<(KeyBuf& k, const TempD& t) { return k << t.kind << t.tmin << t.tmax << t.steps; }

// Everything of the context a captured step sequence reads on the host side:
// column and scratch pointers, sizes, list geometry, box, constraint and the
// interaction switches (the parameter tables are read by the kernels from
// device memory).
void key_context(KeyBuf& k, const samos_ctx* c) {
  k << c->n << c->np << c->ntypes;
  for (int i = 0; i < kNumDoubleFields; ++i) k << c->dcol[i];
  for (int i = 0; i < kNumIntFields; ++i) k << c->icol[i];
  for (int i = 0; i < 3; ++i) k << c->alt[i] << c->altn[i];
  k << c->xy[0] << c->xy[1];
  const NListD& l = c->nl;
  k << c->has_nlist << l.ncx << l.ncy << l.ncz << l.ncell << l.wx << l.wy << l.wz << l.cut2 << l.pad2q << l.use_cell
    << l.np << l.stride << l.cell_of << l.count << l.start << l.members << l.bsum << l.nbr << l.nnb << l.ox << l.oy
    << l.oz << l.flags << l.r_uni;
  const BoxD& b = c->box;
  k << b.Lx << b.Ly << b.Lz << b.xlo << b.xhi << b.ylo << b.yhi << b.zlo << b.zhi << b.periodic;
  const ConstraintD& q = c->con;
  k << q.kind << q.r << q.zpos << q.unlimited << q.q0 << q.q1 << q.q2 << q.max_iter << q.tol;
  k << c->has_soft << c->has_lj << c->has_pol << c->has_nem << c->has_sp << c->soft_radii << c->lj_radii
    << c->lj_shifted << c->sp_alpha << c->d_tab;
  const ExtraD& e = c->extra;
  k << e.count;
  for (int i = 0; i < e.count; ++i) {
    k << e.kind[i] << e.radii[i] << e.shifted[i];
    for (int j = 0; j < 8; ++j) k << e.q[i][j];
  }
  k << c->d_flags << c->d_clock << c->d_chunk << c->d_age << c->d_en << c->gather_batch << c->coop;
}

void key_integ(KeyBuf& k, const Integ& I) {
  k << I.name.size();
  k.s += I.name;
  k << I.slot << I.gbit << I.dt << I.temp << I.v0 << I.nu << I.mu << I.mur << I.stoch << I.flip_p << I.nematic
    << I.velocity << I.seed << I.gamma << I.method << I.R << I.limit << I.theta_limit << I.fire;
}

// The instantiated graph of `key`, recorded by `record` on first use.  The
// least recently used entry is dropped beyond 24.
int cached_graph(samos_ctx* ctx, const std::string& key, const std::function<int()>& record, hipGraphExec_t* out) {
  for (auto& g : ctx->graphs)
    if (g.key == key) {
      g.used = ++ctx->graph_tick;
      ctx->graph_stats[1]++;
      *out = g.exec;
      return 0;
    }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
  const int rc = record();
  const hipError_t ec = hipStreamEndCapture(ctx->st, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  HIPCHK(ec);
  hipGraphExec_t ge = nullptr;
  const hipError_t ir = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  HIPCHK(ir);
  if (ctx->graphs.size() >= 24) {
    auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                [](const samos_ctx::GraphEntry& a, const samos_ctx::GraphEntry& b) { return a.used < b.used; });
    hipGraphExecDestroy(lru->exec);
    ctx->graphs.erase(lru);
  }
  ctx->graphs.push_back({key, ge, ++ctx->graph_tick});
  ctx->graph_stats[0]++;
  *out = ge;
  return 0;
}

// Instantiated chunk graphs of one step sequence, one per size, ascending:
// 2, 4, ... below the configured size, then the configured size.
using ChunkSet = std::vector<std::pair<int, hipGraphExec_t>>;

// Every size of the sequence keyed by `base`, instantiated together on first
// use, so that a range of any length finds its chunks ready.
int chunk_graphs(samos_ctx* ctx, const std::string& base, const std::function<int(int)>& record, ChunkSet* out) {
  out->clear();
  std::vector<int> sizes;
  for (int m = 2; m < ctx->chunk_slots; m <<= 1) sizes.push_back(m);
  sizes.push_back(ctx->chunk_slots);
  for (int m : sizes) {
    KeyBuf k;
    k << m;
    hipGraphExec_t ge = nullptr;
    if (int r = cached_graph(ctx, base + k.s, [&]() { return record(m); }, &ge)) return r;
    out->push_back({m, ge});
  }
  return 0;
}

// Launches chunks until the chunk base reaches `end` (each chunk runs at
// least its first slot, and a slot whose positions need a rebuild stops its
// chunk).  A stretch longer than the largest chunk is covered by a batch of
// largest chunks, as many as the slots per chunk seen so far suggest; a
// shorter remainder by the largest chunks that fit while more than 8 slots
// are left and then by one chunk that covers the rest.  The host reads the
// base after each batch.  Returns the base reached in *t, and 1
// when the batch stopped at a neighbour-list overflow.
// stall: rounds that may end without progress (a request left before the
// chunks stops their first slot; the chunk's rebuild serves it).  Returns 2
// when the slots stopped because a storage re-sort is due.
int replay_chunks(samos_ctx* ctx, const ChunkSet& set, int end, int* t, int stall = 0) {
  const int Mmax = set.back().first;
  double per = Mmax;
  int* h_base = reinterpret_cast<int*>(ctx->h_out);
  while (*t < end) {
    const int r = end - *t;
    int k = 0;
    if (r > Mmax) {
      k = std::max(1, std::min(256, (int)(r / per)));
      for (int i = 0; i < k; ++i) HIPCHK(hipGraphLaunch(set.back().second, ctx->st));
    } else {
      // largest chunks while more than 8 slots are left, then the smallest
      // chunk that covers the rest: a chunk boundary (5 dispatches) costs more
      // than the up to 7 slots that return at once past the range's end
      for (int left = r; left > 0;) {
        const std::pair<int, hipGraphExec_t>* g = nullptr;
        if (left <= 8) {
          for (const auto& e : set)
            if (e.first >= left) { g = &e; break; }
        }
        if (!g) {
          auto it = set.rbegin();
          while (it != set.rend() && it->first > left) ++it;
          g = it != set.rend() ? &*it : &set.front();
        }
        HIPCHK(hipGraphLaunch(g->second, ctx->st));
        left -= g->first;
      }
    }
    // the next run step and the stop word, adjacent in ChunkCtl
    HIPCHK(hipMemcpyAsync(h_base, &ctx->d_chunk->base, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    const int tn = h_base[0];
    if (h_base[1] & 1) {  // a rebuild overflowed: the slots after it stopped
      *t = tn;
      return 1;
    }
    if (h_base[1] & 2) {  // the slots stopped for a re-sort
      *t = tn;
      return 2;
    }
    if (tn == *t && stall > 0) { --stall; continue; }
    if (tn <= *t || tn > end) return fail(ctx, "captured step chunks made no progress");
    if (k) per = std::max(1.0, (double)(tn - *t) / k);
    *t = tn;
  }
  return 0;
}

// ---- profiling ----------------------------------------------------------
hipEvent_t prof_begin(samos_ctx* ctx, const char* cls) {
  if (!ctx->prof) return nullptr;
  ProfAcc& a = ctx->profs[cls];
  if (a.used == a.ev.size()) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return nullptr;
    a.ev.push_back({e0, e1});
  }
  hipEventRecord(a.ev[a.used].first, ctx->st);
  return a.ev[a.used].first;
}
void prof_end(samos_ctx* ctx, const char* cls, hipEvent_t b) {
  if (!ctx->prof || !b) return;
  ProfAcc& a = ctx->profs[cls];
  hipEventRecord(a.ev[a.used].second, ctx->st);
  a.used++;
}
struct ProfScope {
  samos_ctx* c; const char* cls; hipEvent_t b;
  ProfScope(samos_ctx* c_, const char* cls_) : c(c_), cls(cls_) { b = prof_begin(c, cls); }
  ~ProfScope() { prof_end(c, cls, b); }
};

// ---- neighbour list -------------------------------------------------------
void free_nlist(samos_ctx* ctx) {
  NListD& nl = ctx->nl;
  hipFree(nl.cell_of); hipFree(nl.count); hipFree(nl.start); hipFree(nl.members); hipFree(nl.bsum);
  hipFree(nl.nbr); hipFree(nl.ox); hipFree(nl.oy); hipFree(nl.oz);
  nl.cell_of = nl.count = nl.start = nl.members = nl.bsum = nl.nbr = nullptr;
  nl.ox = nl.oy = nl.oz = nullptr;
}

int alloc_nbr(samos_ctx* ctx, int stride) {
  NListD& nl = ctx->nl;
  hipFree(nl.nbr);
  nl.nbr = nullptr;
  nl.stride = stride;
  // 8 entries beyond the capacity, in groups of four per slot: the builds pad
  // each list to a whole number of 8-entry batches
  const size_t groups = (size_t)(stride + 8 + 3) / 4;
  HIPCHK(hipMalloc(&nl.nbr, sizeof(int) * 4 * groups * (size_t)ctx->np + 64));
  return 0;
}

int read_flags(samos_ctx* ctx, int* h) {
  HIPCHK(hipMemcpyAsync(h, ctx->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

// Synchronous build with neighbour capacity growth: rebuilds until no slot
// overflowed, then keeps 1.5x headroom for the device-side rebuilds of a run.
// `at` (optional) replaces the slot columns the build reads positions from
// (planar: from its interleaved (x, y) column).
int build_sync(samos_ctx* ctx, const Fields* at = nullptr, bool planar = false) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    HIPCHK(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->st));
    HIPCHK(hipMemsetAsync(ctx->d_flags + 5, 0, sizeof(int), ctx->st));
    {
      ProfScope ps(ctx, "rebuild");
      launch_rebuild(ctx->st, at ? *at : fields(ctx), ctx->nl, ctx->box, ctx->n, nullptr, planar ? 1 : 0);
    }
    HIPCHK(hipGetLastError());
    int h[8];
    if (int r = read_flags(ctx, h)) return r;
    const int need = h[5];
    if (h[2] == 0 && ctx->nl.stride >= (int)(need * 1.25) + 8) return 0;
    if (int r = alloc_nbr(ctx, (int)std::ceil(need * 1.5) + 16)) return r;
  }
  return fail(ctx, "neighbour list capacity could not be established");
}

// A device rebuild of this run found more neighbours than the list holds and
// stored a truncated list (flags[2]); no step has read it (chunk slots stop at
// the flag, direct steps are checked before the next one).  The capacity
// grows and the list is rebuilt from the same positions, which gives the
// rebuild's pair set, count and snapshot exactly; the run then continues.
// Returns 1 when a recovery happened.
// builds (optional): the builds count after the call.
int recover_overflow(samos_ctx* ctx, const Fields& at, bool planar, int* builds = nullptr) {
  int h[8];
  if (int r = read_flags(ctx, h)) return r;
  if (builds) *builds = h[4];
  if (!h[2]) return 0;
  const int counted = h[4];
  if (int r = alloc_nbr(ctx, (int)std::ceil(h[5] * 1.5) + 16)) return r;
  if (int r = build_sync(ctx, &at, planar)) return r;
  // the rebuild that overflowed has been counted; its repetition is not
  HIPCHK(hipMemcpyAsync(ctx->d_flags + 4, &counted, sizeof(int), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->overflow_recoveries++;
  return 1;
}

int resort(samos_ctx* ctx) {
  if (!ctx->has_nlist || !ctx->nl.use_cell || ctx->n < 2) return 0;
  const int n = ctx->n;
  Fields f = fields(ctx);
  launch_morton_keys(ctx->st, f, ctx->nl, ctx->box, n, ctx->keys, ctx->perm, ctx->con.kind == 1 ? 1 : 0);
  radix_sort_pairs(ctx->st, ctx->sort_tmp, ctx->sort_tmp_bytes, ctx->keys, ctx->keys2, ctx->perm, ctx->perm2, n);
  for (int k = 0; k < kNumDoubleFields; ++k) {
    launch_gather_d(ctx->st, ctx->dtmp, ctx->dcol[k], ctx->perm2, n);
    HIPCHK(hipMemcpyAsync(ctx->dcol[k], ctx->dtmp, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  for (int k = 0; k < kNumIntFields; ++k) {
    launch_gather_i(ctx->st, ctx->itmp, ctx->icol[k], ctx->perm2, n);
    HIPCHK(hipMemcpyAsync(ctx->icol[k], ctx->itmp, sizeof(int) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  // the list's snapshot follows its slots
  double* snap[3] = {ctx->nl.ox, ctx->nl.oy, ctx->nl.oz};
  for (double* col : snap) {
    launch_gather_d(ctx->st, ctx->dtmp, col, ctx->perm2, n);
    HIPCHK(hipMemcpyAsync(col, ctx->dtmp, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  HIPCHK(hipMemcpyAsync(ctx->h_id.data(), ctx->icol[5], sizeof(int) * n, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipGetLastError());
  return 0;
}

// Re-sorts the storage and rebuilds the list from its permuted snapshot: the
// same pair set as the last build, which stays the one counted.
int resort_at_snapshot(samos_ctx* ctx) {
  int h[8];
  if (int r = read_flags(ctx, h)) return r;
  if (int r = resort(ctx)) return r;
  Fields at = fields(ctx);
  at.x = ctx->nl.ox; at.y = ctx->nl.oy; at.z = ctx->nl.oz;
  if (int r = build_sync(ctx, &at, false)) return r;
  HIPCHK(hipMemcpyAsync(ctx->d_flags + 4, &h[4], sizeof(int), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->builds_at_sort = h[4];
  return 0;
}

// Storage re-sort between captured chunks of a run (particles drift away from
// their storage neighbours over long runs, and the gathers lose locality).
// The positions are in buffer 0 then; every per-slot column, the list's
// snapshot among them, is permuted, and the list is rebuilt from the
// permuted snapshot: the same pair set as the last build, which stays the
// one counted.  Only storage order changes (and with it the order of the
// force sums).
// par: the buffer holding the positions (and, with directors, the directors
// of a schedule that double-buffers them); they are in buffer 0 afterwards.
int resort_in_run(samos_ctx* ctx, bool planar, int par, bool directors) {
  const size_t bytes = sizeof(double) * ctx->n;
  if (par) {
    if (planar) {
      HIPCHK(hipMemcpyAsync(ctx->xy[0], ctx->xy[1], 2 * bytes, hipMemcpyDeviceToDevice, ctx->st));
    } else {
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(ctx->dcol[k], ctx->alt[k], bytes, hipMemcpyDeviceToDevice, ctx->st));
    }
    if (directors)
      for (int k = 0; k < (planar ? 2 : 3); ++k)
        HIPCHK(hipMemcpyAsync(ctx->dcol[6 + k], ctx->altn[k], bytes, hipMemcpyDeviceToDevice, ctx->st));
  }
  if (planar) {
    Fields f0 = fields(ctx, 0);
    f0.xy = ctx->xy[0];
    launch_unpack_xy(ctx->st, f0, ctx->n);
  }
  if (int r = resort_at_snapshot(ctx)) return r;
  if (planar) {
    Fields f0 = fields(ctx, 0);
    f0.xy = ctx->xy[0];
    launch_pack_xy(ctx->st, f0, ctx->n);
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->resorts++;
  return 0;
}

// The builds count at which a run re-sorts next.
int resort_at(const samos_ctx* ctx) {
  return ctx->resort_every > 0 ? ctx->builds_at_sort + ctx->resort_every : INT_MAX;
}

// ---- forces ---------------------------------------------------------------
int energies_to_host(samos_ctx* ctx, double* e4) {
  launch_reduce_sum(ctx->st, ctx->d_en, 4, ctx->np, ctx->n, ctx->d_partial, ctx->d_out);
  HIPCHK(hipMemcpyAsync(ctx->h_out, ctx->d_out, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  for (int k = 0; k < 4; ++k) e4[k] = 0.5 * ctx->h_out[k];
  return 0;
}

// Soft / LJ / aligners in one pass, then the further pair potentials, which
// add to the forces (mode 0 overwrites first, mode 2 leaves forces alone).
void interact(samos_ctx* ctx, int pm, int am, int mode, double* e_out, bool f32 = false) {
  const bool ext = ctx->has_nlist && ctx->extra.count > 0;
  if (!ctx->has_nlist) { pm = 0; am = 0; }
  if (pm == 0 && am == 0 && !ctx->has_sp) {
    if (e_out) fill_d(ctx->st, e_out, 0.0, 4 * ctx->np);
    if (ext && mode == 0)
      for (int k = 9; k <= 11; ++k) fill_d(ctx->st, ctx->dcol[k], 0.0, ctx->n);
  } else {
    ProfScope ps(ctx, "interact");
    if (f32) launch_interact_f32(ctx->st, pm, am, mode, fields(ctx), ctx->nl, ctx->box, interact_params(ctx), ctx->n, e_out);
    else launch_interact(ctx->st, pm, am, mode, fields(ctx), ctx->nl, ctx->box, interact_params(ctx), ctx->n, e_out);
  }
  if (ext) {
    ProfScope ps(ctx, "interact");
    launch_interact_extra(ctx->st, fields(ctx), ctx->nl, ctx->box, ctx->extra, ctx->n,
                          e_out ? e_out + 4L * ctx->np : nullptr, mode != 2 ? 1 : 0);
  }
}

// Loop head: reset f and tau (System::reset_forces / reset_torques), then the
// pair potentials and aligners in one pass.  Columns the pass does not write
// are cleared here.
void loop_head_forces(samos_ctx* ctx, double* e_out, bool f32 = false) {
  const int pm = ctx->has_nlist ? pot_mask(ctx) : 0;
  const int am = ctx->has_nlist ? al_mask(ctx) : 0;
  if (pm == 0 && !ctx->has_sp)
    for (int k = 9; k <= 11; ++k) fill_d(ctx->st, ctx->dcol[k], 0.0, ctx->n);
  if (am == 0)
    for (int k = 12; k <= 14; ++k) fill_d(ctx->st, ctx->dcol[k], 0.0, ctx->n);
  interact(ctx, pm, am, 0, e_out, f32);
}

// The general pair-sum form (per-type tables, per-particle radii) instead of
// the single-type form with loop-constant parameters.
bool general_pairs(const samos_ctx* c) { return c->ntypes > 1 || c->soft_radii || c->lj_radii; }

bool needs_nlist(const samos_ctx* c) {
  return c->has_nlist && (c->has_soft || c->has_lj || c->has_pol || c->has_nem || c->extra.count > 0);
}

// Device-side conditional rebuild after a step: flags[0] = any(need_update).
void check_and_rebuild(samos_ctx* ctx) {
  if (!needs_nlist(ctx)) return;
  fill_i(ctx->st, ctx->d_flags, 0, 1);
  Fields f = fields(ctx);
  launch_need_update(ctx->st, f, ctx->nl, ctx->box, ctx->n, ctx->d_flags);
  ProfScope ps(ctx, "rebuild");
  launch_rebuild(ctx->st, f, ctx->nl, ctx->box, ctx->n, ctx->d_flags);
}

// ---- integrators ----------------------------------------------------------
int group_size(const samos_ctx* ctx, unsigned gbit) {
  if (gbit == 1u) return ctx->n;
  int c = 0;
  for (unsigned g : ctx->h_groups) c += (g & gbit) ? 1 : 0;
  return c;
}

// True when no particle belongs to two of the integrators' groups.
bool groups_disjoint(const samos_ctx* ctx, const std::vector<const Integ*>& v) {
  for (size_t a = 0; a < v.size(); ++a)
    for (size_t b = a + 1; b < v.size(); ++b) {
      if (v[a]->gbit == 1u || v[b]->gbit == 1u) return false;
      const unsigned both = v[a]->gbit | v[b]->gbit;
      for (unsigned g : ctx->h_groups)
        if ((g & both) == both) return false;
    }
  return true;
}

// j >= 0: the step is the j-th of a captured chunk; its noise call and run
// step are read from the device clock plus j instead of being baked in.
int integrate_one(samos_ctx* ctx, Integ& I, int run_step, int j = -1) {
  ctx->age_checked = false;  // these kernels advance each slot's age
  Fields f = fields(ctx);
  const int n = ctx->n;
  const bool clocked = j >= 0;
  const unsigned long long call = clocked ? (unsigned long long)j : I.calls++;
  const ClockD* clk = clocked ? ctx->d_clock : nullptr;
  const int t_off = clocked ? j : run_step;
  if (I.name == "brownian" || I.name == "brownian_pos" || I.name == "brownian_align") {
    int which = I.name == "brownian" ? 0 : (I.name == "brownian_pos" ? 1 : 2);
    ProfScope ps(ctx, "integrate");
    launch_brownian(ctx->st, which, f, ctx->box, ctx->con, n, I.dt, I.v0, I.mu, I.mur, I.stoch, I.flip_p, I.nematic,
                    I.velocity, I.temp, I.seed, call, clk, I.slot, t_off, I.gbit);
  } else if (I.name == "langevin") {
    {
      ProfScope ps(ctx, "integrate");
      launch_langevin_a(ctx->st, I.method, f, ctx->box, ctx->con, n, I.dt, I.gamma, I.temp, I.seed, call, clk,
                        I.slot, t_off, I.R, I.gbit);
    }
    interact(ctx, pot_mask(ctx), 0, 0, nullptr);
    ProfScope ps(ctx, "integrate");
    launch_langevin_b(ctx->st, I.method, f, n, I.dt, I.gamma, I.temp, I.seed, call, clk, I.slot, t_off, I.R,
                      ctx->box, ctx->con, I.gbit);
  } else if (I.name == "nve") {
    {
      ProfScope ps(ctx, "integrate");
      launch_nve_a(ctx->st, f, ctx->box, ctx->con, n, I.dt, I.limit, I.theta_limit, I.gbit);
    }
    // reset forces and torques, Potential::compute, Aligner::compute
    loop_head_forces(ctx, nullptr);
    ProfScope ps(ctx, "integrate");
    launch_nve_b(ctx->st, f, ctx->box, ctx->con, n, I.dt, I.limit, I.theta_limit, I.gbit);
  } else if (I.name == "fire") {
    {
      ProfScope ps(ctx, "integrate");
      launch_fire_a(ctx->st, f, ctx->box, ctx->con, n, I.fire, I.gbit);
    }
    // reset forces and torques, Potential::compute, Aligner::compute
    loop_head_forces(ctx, ctx->d_en);
    {
      ProfScope ps(ctx, "integrate");
      launch_fire_b(ctx->st, f, ctx->box, ctx->con, n, I.fire, ctx->d_cols, ctx->np, I.gbit);
    }
    // P, |F|^2, |v|^2 and the pair energies in fixed-order sums, then the
    // scalar update (src/integrators/integrator_fire.cpp:97-147) on the device
    launch_reduce_sum(ctx->st, ctx->d_cols, 3, ctx->np, n, ctx->d_partial, ctx->d_out);
    launch_reduce_sum(ctx->st, ctx->d_en, 2, ctx->np, n, ctx->d_partial, ctx->d_out + 3);
    if (ctx->extra.count > 0)
      launch_reduce_sum(ctx->st, ctx->d_en + 4L * ctx->np, 1, ctx->np, n, ctx->d_partial, ctx->d_out + 5);
    launch_fire_update(ctx->st, I.fire, ctx->d_out, ctx->extra.count > 0 ? 1 : 0);
    launch_fire_mix(ctx->st, f, n, I.fire, I.gbit);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// The group's size enters FIRE's sqrt(3 N); groups may gain members after the
// integrator exists, so it is refreshed before each run range (outside any
// capture).
int sync_fire_groups(samos_ctx* ctx) {
  for (auto& kv : ctx->integs) {
    Integ& I = kv.second;
    if (I.name != "fire") continue;
    const int ndof3 = 3 * group_size(ctx, I.gbit);
    if (ndof3 == I.fire_ndof3) continue;
    HIPCHK(hipMemcpy(&I.fire->ndof3, &ndof3, sizeof(int), hipMemcpyHostToDevice));
    I.fire_ndof3 = ndof3;
  }
  return 0;
}

// Clears a few flag words on the run paths: one small kernel launch costs the
// host less time than hipMemsetAsync, and the run's first kernel waits on it.
void clear_flags(samos_ctx* ctx, int first, int count) {
  hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, ctx->st, ctx->d_flags + first, 0, (long)count);
}

// ---- fused Langevin run ---------------------------------------------------
int upload_clock(samos_ctx* ctx, int run_step) {
  ClockD c{};
  for (auto& kv : ctx->integs) c.calls[kv.second.slot] = kv.second.calls;
  c.run_step = run_step;
  HIPCHK(hipMemcpyAsync(ctx->d_clock, &c, sizeof(ClockD), hipMemcpyHostToDevice, ctx->st));
  return 0;
}

// One fused step t of the Langevin BAOAB schedule with positions in buffer
// (rel % 2), followed (outside captured chunks) by the device-conditional
// rebuild on the new positions: one cooperative launch when available
// (coop), else the stage kernels (bin, scan, scatter, sort, build).  In a
// chunk (chunk != null) the step is slot `slot` and the chunk's own rebuild
// follows the slots.
// rebuild = false leaves the step's rebuild request in the parity flag for the
// captured chunks that follow to serve.
// call0 >= 0: a launch outside the chunks, which takes noise call call0 and
// run step t directly instead of the device clock.
int fused_step(samos_ctx* ctx, Integ& I, int phase, int t, int rel, int t_off, int steps, bool f32, bool planar,
               double mass_u, bool coop, ChunkCtl* chunk = nullptr, int slot = 0, bool rebuild = true,
               long long call0 = -1) {
  const int par = rel & 1;
  Fields cur = fields(ctx, par), nxt = fields(ctx, par ^ 1);
  if (planar) { cur.xy = ctx->xy[par]; nxt.xy = ctx->xy[par ^ 1]; }
  {
    ProfScope ps(ctx, "fused");
    FusedArgs fa;
    fa.dt = I.dt; fa.gamma = I.gamma; fa.temp = I.temp; fa.seed = I.seed;
    fa.clock = ctx->d_clock; fa.clock_slot = I.slot; fa.t_off = t_off; fa.parity = par;
    if (call0 >= 0) { fa.clock = nullptr; fa.call0 = (unsigned long long)call0; fa.t_off = t; }
    fa.head_enforce = t + 1 <= steps ? 1 : 0;
    fa.mass_u = mass_u;
    // BAOAB O-step constants (src/integrators/integrator_langevin.cpp:57-58)
    fa.exp_dt = std::exp(-I.gamma * I.dt);
    fa.Bc = I.temp.kind == SAMOS_VALUE_CONSTANT ? std::sqrt(I.temp.tmin * (1.0 - std::exp(-2.0 * I.gamma * I.dt))) : -1.0;
    // the per-particle quotients of a common mass, formed once (IEEE division
    // and square root round the same on both sides)
    fa.fact_u = mass_u > 0.0 ? (0.5 * I.dt) / mass_u : 0.0;
    fa.sf_u = (mass_u > 0.0 && fa.Bc >= 0.0) ? fa.Bc / std::sqrt(mass_u) : -1.0;
    fa.batch = ctx->gather_batch;
    // with a group other than "all" only members age: no common age then
    fa.age = I.gbit == 1u ? ctx->d_age : nullptr;
    fa.gbit = I.gbit;
    fa.chunk = chunk;
    fa.slot = slot;
    if (phase == 2 && planar) { fa.out_x = ctx->dcol[0]; fa.out_y = ctx->dcol[1]; }
    launch_baoab_fused(ctx->st, pot_mask(ctx), phase, cur, nxt, ctx->nl, ctx->box, ctx->con, interact_params(ctx),
                       ctx->n, fa, f32 ? 1 : 0, planar ? 1 : 0, general_pairs(ctx) ? 1 : 0);
  }
  if (phase != 2 && !chunk && rebuild) {
    ProfScope ps(ctx, "rebuild");
    if (coop && ctx->coop != 0) {
      hipError_t e = launch_rebuild_coop(ctx->st, nxt, ctx->nl, ctx->box, ctx->n, ctx->d_flags + par, planar ? 1 : 0);
      if (e == hipSuccess) { ctx->coop = 1; return 0; }
      if (ctx->coop == 1) { ctx->err = std::string("cooperative rebuild failed: ") + hipGetErrorString(e); return -2; }
      (void)hipGetLastError();
      ctx->coop = 0;  // not available here: stage kernels from now on
    }
    launch_rebuild(ctx->st, nxt, ctx->nl, ctx->box, ctx->n, ctx->d_flags + par, planar ? 1 : 0);
  }
  return 0;
}

// The common mass when every particle has the same one, else -1.
int uniform_mass(samos_ctx* ctx, double* out) {
  if (!ctx->mass_known) {
    std::vector<double> m(ctx->n);
    if (ctx->n) HIPCHK(hipMemcpy(m.data(), ctx->dcol[18], sizeof(double) * ctx->n, hipMemcpyDeviceToHost));
    ctx->mass_u = -1.0;
    if (ctx->n && m[0] > 0.0 && std::all_of(m.begin(), m.end(), [&](double v) { return v == m[0]; })) ctx->mass_u = m[0];
    ctx->mass_known = true;
  }
  *out = ctx->mass_u;
  return 0;
}

// Iterations a..b-1 of a run whose last index is `steps`, fused: each launch
// computes F(x_t) once and applies the closing half-kick of step t-1 and the
// first part of step t.  Positions of iteration t live in buffer (t - a) % 2;
// x_a is in the primary columns on entry and x_b is back there on return,
// with f = F(x_b): the state the reference has at the loop head of b.
bool tile_eligible(samos_ctx* ctx, int flags, double mass_u);
int run_tiled_langevin(samos_ctx* ctx, Integ& I, int a, int b, int steps, int flags, double mass_u);

// packed: the interleaved (x, y) column of buffer 1 already holds the
// positions (written by the run's first enforce).
int run_fused_langevin(samos_ctx* ctx, Integ& I, int a, int b, int steps, int flags, bool packed) {
  const bool f32 = (flags & SAMOS_RUN_FP32_FORCE) != 0;
  double mass_u = -1.0;
  if (int r = uniform_mass(ctx, &mass_u)) return r;
  if (I.gbit != 1u) {
    ctx->age_checked = false;  // each member's age advances in the step kernel
  } else if (!ctx->age_checked) {
    launch_age_check(ctx->st, ctx->d_age, ctx->dcol[16], ctx->n);
    ctx->age_checked = true;
  }
  if (I.gbit == 1u && tile_eligible(ctx, flags, mass_u)) {
    const int r = run_tiled_langevin(ctx, I, a, b, steps, flags, mass_u);
    if (r <= 0) return r;
  }
  // (the rebuild-request flags were cleared at the range's start)
  // Planar layout: the plane constraint puts every particle on z = zpos, so
  // neighbour gathers read the interleaved (x, y) column.  The N^2 build reads
  // the x, y, z columns, so lists without a cell grid keep the 3-D layout.
  const bool planar = ctx->con.kind == 1 && ctx->nl.use_cell;
  const bool graph = ctx->coop == 0 && !(flags & SAMOS_RUN_NO_GRAPH) && !ctx->prof;
  const bool chunks = graph && b - (a + 1) >= 2;
  // x_a is read from buffer r0: planar runs keep it in the interleaved
  // column of buffer 1, so that x_{a+1} lands in buffer 0, where captured
  // chunks start
  const int r0 = planar ? 1 : 0;
  // the loop-head enforce of iteration a was applied by the run's begin or by
  // the last kernel of the previous range
  if (planar) {
    Fields f1 = fields(ctx, 0);
    f1.xy = ctx->xy[1];
    if (!packed) launch_pack_xy(ctx->st, f1, ctx->n);
    // the direct steps' rebuilds bin either buffer and read z (= zpos after
    // the enforce); chunks bin buffer 0 only
    if (!chunks)
      HIPCHK(hipMemcpyAsync(ctx->alt[2], ctx->dcol[2], sizeof(double) * ctx->n, hipMemcpyDeviceToDevice, ctx->st));
  }
  auto at_buffer = [&](int par) {
    Fields f = fields(ctx, par & 1);
    if (planar) f.xy = ctx->xy[par & 1];
    return f;
  };
  // Launches outside the chunks take their noise call and run step from the
  // host: iteration t draws call c0 + (t - a).
  const unsigned long long c0 = I.calls;
  auto call_of = [&](int t) { return (long long)(c0 + (unsigned long long)(t - a)); };
  // With chunks to follow, this step's rebuild request is served by the first
  // chunk (its slot 0 stops at once, the chunk's rebuild runs); otherwise the
  // rebuild runs here and an overflow of it is caught before the next step.
  if (int r = fused_step(ctx, I, 0, a, r0, 0, steps, f32, planar, mass_u, true, nullptr, 0, !chunks, call_of(a)))
    return r;
  int t = a + 1;
  int buf = r0 ^ 1;  // buffer holding x_t
  // steps outside the chunks: an overflow of the step's rebuild is recovered,
  // and a due re-sort done, before the next step reads the list
  auto after_direct = [&](int par) -> int {
    int builds = 0;
    if (int r = recover_overflow(ctx, at_buffer(par), planar, &builds); r < 0) return r;
    if (builds < resort_at(ctx)) return 0;
    if (int r = resort_in_run(ctx, planar, par, false)) return r;
    return 1;
  };
  if (!chunks) {
    const int r = after_direct(buf);
    if (r < 0) return r;
    if (r == 1) buf = 0;
  }
  // Captured chunks: M fused steps ("slots") followed by one device-conditional
  // rebuild.  A slot whose new positions need a rebuild stops the chunk: the
  // later slots return at once, the positions settle into buffer 0, the
  // rebuild runs, and the next chunk resumes at the following step.  Slots
  // past the chunk's limit return at once as well, so one chunk serves any
  // range; short ranges use a smaller power-of-two chunk.  The rebuild
  // decisions and the steps are exactly those of one launch per step; only
  // the skipped-rebuild dispatches between steps are gone.  Slots cover the
  // iterations up to b - 1 and read the noise call and run step from the
  // device clock, which each chunk's bookkeeping advances.  Instantiated
  // chunks are kept per context.
  if (chunks) {
    Fields f0 = at_buffer(0), f1 = at_buffer(1);
    auto record = [&](int m) -> int {
      // the loop-head enforce of each slot's next iteration is decided on the
      // device from the run's last iteration
      for (int j = 0; j < m; ++j)
        if (int r = fused_step(ctx, I, 1, b - 2, j, j, steps, f32, planar, mass_u, false, ctx->d_chunk, j)) return r;
      launch_chunk_rebuild(ctx->st, ctx->d_chunk, ctx->d_clock, f0, f1, ctx->nl, ctx->box, ctx->n, planar ? 1 : 0, 0);
      return 0;
    };
    // slots start at buffer parity 0
    if (buf == 1)
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(ctx->dcol[k], ctx->alt[k], sizeof(double) * ctx->n, hipMemcpyDeviceToDevice, ctx->st));
    buf = 0;
    bool pending = true;
    for (;;) {
      // the list's arrays (and so the chunks' key) change when it grows
      KeyBuf base;
      base << 'L' << f32 << planar << mass_u << pot_mask(ctx);
      key_context(base, ctx);
      key_integ(base, I);
      ChunkSet set;
      if (int r = chunk_graphs(ctx, base.s, record, &set)) return r;
      // the first chunks take over the first step's request (its parity
      // flag); the device clock starts at iteration t
      launch_chunk_init(ctx->st, ctx->d_chunk, t, b, steps, resort_at(ctx), pending ? ctx->d_flags + r0 : nullptr,
                        ctx->d_clock, I.slot, (unsigned long long)call_of(t));
      const int r = replay_chunks(ctx, set, b, &t, pending ? 1 : 0);
      pending = false;
      if (r < 0) return r;
      if (r == 0) break;
      if (r == 2) {
        if (int rc = resort_in_run(ctx, planar, 0, false)) return rc;
        continue;
      }
      // stopped at an overflowed rebuild of the positions now in buffer 0
      if (int rc = recover_overflow(ctx, f0, planar); rc < 0) return rc;
    }
    // the last chunk served any request; direct steps use the parity flags
    if (t < b) clear_flags(ctx, 0, 2);
  }
  for (; t < b; ++t) {
    if (int r = fused_step(ctx, I, 1, t, buf, 0, steps, f32, planar, mass_u, true, nullptr, 0, true, call_of(t)))
      return r;
    buf ^= 1;
    const int r = after_direct(buf);
    if (r < 0) return r;
    if (r == 1) buf = 0;
  }
  if (int r = fused_step(ctx, I, 2, b, buf, 0, steps, f32, planar, mass_u, true, nullptr, 0, true, call_of(b)))
    return r;
  const int pf = buf;  // buffer holding x_b
  // planar: the closing step wrote x_b back to the x, y columns
  if (!planar && pf) {
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync(ctx->dcol[k], ctx->alt[k], sizeof(double) * ctx->n, hipMemcpyDeviceToDevice, ctx->st));
  }
  // one closing half-kick, and one age step, per iteration
  if (I.gbit == 1u) launch_age_finish(ctx->st, ctx->d_age, ctx->dcol[16], ctx->n, I.dt, b - a);
  HIPCHK(hipGetLastError());
  I.calls += (unsigned long long)(b - a);
  return 0;
}


// ---- fused Brownian run ------------------------------------------------------
// One IntegratorBrownian ("brownian") with pair potentials and at most one
// aligner: each launch evaluates the loop-head forces and torques of step t
// and integrates it (k_brownian_fused), followed by the device-conditional
// rebuild; chunks are captured as for the fused Langevin run.  On entry the
// primary columns hold the loop-head state of a; on return that of b.
// Either one "brownian" integrator, or exactly "brownian_align" and
// "brownian_pos" (both over all particles, the reference's example confs).
// Or up to four "brownian" integrators on pairwise disjoint groups (each
// particle is integrated by at most one of them, so their map order does not
// matter), such as the active and passive groups of a two-component conf.
bool tile_brownian_eligible(samos_ctx* ctx, int flags);
int run_tiled_brownian(samos_ctx* ctx, Integ& I, int a, int b, int steps, int flags);

bool brownian_groups(const samos_ctx* ctx) {
  if (ctx->integs.empty() || ctx->integs.size() > 4) return false;
  std::vector<const Integ*> v;
  for (auto& kv : ctx->integs) {
    if (kv.second.name != "brownian") return false;
    v.push_back(&kv.second);
  }
  return v.size() == 1 || groups_disjoint(ctx, v);
}

// The fused step kernels gather with 32-bit byte offsets (16 np < 2^32).
bool fused_size_ok(const samos_ctx* ctx) { return ctx->np < (1 << 28); }

bool fused_brownian_eligible(samos_ctx* ctx, int flags) {
  if ((flags & SAMOS_RUN_FAITHFUL) || ctx->extra.count != 0 || !fused_size_ok(ctx)) return false;
  if (!needs_nlist(ctx) || !brownian_fused_supported(pot_mask(ctx), al_mask(ctx))) return false;
  if (brownian_groups(ctx)) return true;
  return ctx->integs.size() == 2 && ctx->integs.count("brownian_align_all") && ctx->integs.count("brownian_pos_all");
}

int run_fused_brownian(samos_ctx* ctx, int a, int b, int steps, int flags) {
  const bool pair = !brownian_groups(ctx);
  Integ& I = pair ? ctx->integs["brownian_align_all"] : ctx->integs.begin()->second;
  Integ* P = pair ? &ctx->integs["brownian_pos_all"] : nullptr;
  // integrators on groups other than "all": per-slot parameters (multi-type kernels)
  std::vector<Integ*> grouped;
  if (!pair && !(ctx->integs.size() == 1 && I.gbit == 1u))
    for (auto& kv : ctx->integs) grouped.push_back(&kv.second);
  const bool multi = general_pairs(ctx) || !grouped.empty();
  const bool f32 = (flags & SAMOS_RUN_FP32_FORCE) != 0;
  const bool planar = ctx->con.kind == 1 && ctx->nl.use_cell;
  const int n = ctx->n;
  auto buffer = [&](int par) {
    Fields f = fields(ctx, par);
    if (par) { f.nx = ctx->altn[0]; f.ny = ctx->altn[1]; f.nz = ctx->altn[2]; }
    if (planar) f.xy = ctx->xy[par];
    return f;
  };
  if (!grouped.empty()) {
    ctx->age_checked = false;  // only members age
  } else if (!ctx->age_checked) {
    launch_age_check(ctx->st, ctx->d_age, ctx->dcol[16], n);
    ctx->age_checked = true;
  }
  if (!pair && grouped.empty() && tile_brownian_eligible(ctx, flags)) {
    const int r = run_tiled_brownian(ctx, I, a, b, steps, flags);
    if (r <= 0) return r;
  }
  if (int r = upload_clock(ctx, a)) return r;
  clear_flags(ctx, 0, 2);
  if (planar) {
    Fields f0 = buffer(0);
    launch_pack_xy(ctx->st, f0, n);
    // the binning of either buffer reads z (= zpos after the enforce)
    HIPCHK(hipMemcpyAsync(ctx->alt[2], ctx->dcol[2], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  auto step = [&](int t, int par, int t_off, ChunkCtl* chunk, int slot) {
    Fields cur = buffer(par), nxt = buffer(par ^ 1);
    BrownArgs ba;
    ba.dt = I.dt; ba.v0 = I.v0; ba.mu = I.mu; ba.mur = I.mur; ba.stoch = I.stoch; ba.flip_p = I.flip_p;
    ba.nematic = I.nematic ? 1 : 0; ba.velocity = I.velocity ? 1 : 0;
    ba.temp = I.temp; ba.seed = I.seed; ba.clock = ctx->d_clock; ba.clock_slot = I.slot; ba.t_off = t_off;
    ba.parity = par; ba.head_enforce = t + 1 <= steps ? 1 : 0;
    ba.chunk = chunk; ba.slot = slot;
    ba.age = grouped.empty() ? ctx->d_age : nullptr;
    ba.ngroups = (int)grouped.size();
    for (int k = 0; k < ba.ngroups; ++k) {
      const Integ& G = *grouped[k];
      BrownArgs::Group& q = ba.grp[k];
      q.dt = G.dt; q.v0 = G.v0; q.mu = G.mu; q.mur = G.mur; q.stoch = G.stoch; q.flip_p = G.flip_p;
      q.nematic = G.nematic ? 1 : 0; q.velocity = G.velocity ? 1 : 0; q.temp = G.temp; q.seed = G.seed;
      q.clock_slot = G.slot; q.gbit = G.gbit;
    }
    if (pair) {
      ba.mode = 1;
      ba.pos_dt = P->dt; ba.pos_mu = P->mu; ba.pos_temp = P->temp; ba.pos_seed = P->seed; ba.pos_slot = P->slot;
    }
    // small systems (the step is then latency-bound): several threads per
    // slot, up to 16, while the launch stays under 16 Ki threads; larger
    // systems measured faster with one thread per slot (a sweep of 1-16
    // threads on MI355X: 4096 particles fastest with 2, 100,000 with 1)
    ba.lanes = 1; ba.lane_shift = 0;
    if (ctx->brownian_lanes > 0) {
      while (ba.lanes < ctx->brownian_lanes) { ba.lanes <<= 1; ++ba.lane_shift; }
    } else {
      while (ba.lanes < 16 && (long)n * ba.lanes * 2 <= 16384L) { ba.lanes <<= 1; ++ba.lane_shift; }
    }
    {
      ProfScope ps(ctx, "fused");
      launch_brownian_fused(ctx->st, pot_mask(ctx), al_mask(ctx), cur, nxt, ctx->nl, ctx->box, ctx->con,
                            interact_params(ctx), n, ba, f32 ? 1 : 0, planar ? 1 : 0, multi ? 1 : 0);
    }
    if (!chunk) {
      ProfScope ps(ctx, "rebuild");
      launch_rebuild(ctx->st, nxt, ctx->nl, ctx->box, n, ctx->d_flags + par, planar ? 1 : 0);
    }
  };
  int t = a, buf = 0;
  const bool graph = ctx->coop == 0 && !(flags & SAMOS_RUN_NO_GRAPH) && !ctx->prof;
  if (graph && b - t >= 2) {
    Fields f0 = buffer(0), f1 = buffer(1);
    auto record = [&](int m) -> int {
      for (int j = 0; j < m; ++j) step(b - 2, j & 1, j, ctx->d_chunk, j);
      launch_chunk_rebuild(ctx->st, ctx->d_chunk, ctx->d_clock, f0, f1, ctx->nl, ctx->box, n, planar ? 1 : 0, 1);
      return 0;
    };
    for (;;) {
      // the list's arrays (and so the chunks' key) change when it grows
      KeyBuf base;
      base << 'B' << f32 << planar << pot_mask(ctx) << al_mask(ctx) << ctx->brownian_lanes;
      key_context(base, ctx);
      key_integ(base, I);
      if (P) key_integ(base, *P);
      for (const Integ* G : grouped) key_integ(base, *G);
      ChunkSet set;
      if (int r = chunk_graphs(ctx, base.s, record, &set)) return r;
      launch_chunk_init(ctx->st, ctx->d_chunk, t, b, steps, resort_at(ctx));
      const int r = replay_chunks(ctx, set, b, &t);
      if (r < 0) return r;
      if (r == 0) break;
      if (r == 2) {
        if (int rc = resort_in_run(ctx, planar, 0, true)) return rc;
        continue;
      }
      // stopped at an overflowed rebuild of the positions now in buffer 0
      if (int rc = recover_overflow(ctx, f0, planar); rc < 0) return rc;
    }
    clear_flags(ctx, 0, 2);
  }
  const int t0 = t;
  for (; t < b; ++t) {
    step(t, buf, t - t0, nullptr, 0);
    buf ^= 1;
    // the next step reads the list this step's rebuild may have overflowed;
    // a due re-sort happens here as well
    int builds = 0;
    if (int r = recover_overflow(ctx, buffer(buf), planar, &builds); r < 0) return r;
    if (builds >= resort_at(ctx)) {
      if (int r = resort_in_run(ctx, planar, buf, true)) return r;
      buf = 0;
    }
  }
  if (b > t0) launch_clock_advance(ctx->st, ctx->d_clock, b - t0);
  const int pf = buf;  // buffer holding x_b and n_b
  if (planar) {
    Fields f0 = fields(ctx, 0);
    f0.xy = ctx->xy[pf];
    launch_unpack_xy(ctx->st, f0, n);
  } else if (pf) {
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemcpyAsync(ctx->dcol[k], ctx->alt[k], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  if (pf)
    for (int k = 0; k < (planar ? 2 : 3); ++k)
      HIPCHK(hipMemcpyAsync(ctx->dcol[6 + k], ctx->altn[k], sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  if (grouped.empty()) launch_age_finish(ctx->st, ctx->d_age, ctx->dcol[16], n, P ? P->dt : I.dt, b - a);
  if (grouped.empty()) I.calls += (unsigned long long)(b - a);
  for (Integ* G : grouped) G->calls += (unsigned long long)(b - a);
  if (P) P->calls += (unsigned long long)(b - a);
  // loop head of b: its enforce was applied with the positions; forces and torques
  if (b <= steps) loop_head_forces(ctx, nullptr, f32);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- cell-tiled fused Langevin run (planar systems) ------------------------
// Used when the run asks for it (SAMOS_RUN_TILES) and the system fits: plane
// constraint, cell grid with at least three cells per side, one particle type,
// no per-particle radii, one common mass and one common age (the fused run's
// uniform-age path).
bool tile_eligible(samos_ctx* ctx, int flags, double mass_u) {
  const NListD& nl = ctx->nl;
  if (!((flags & SAMOS_RUN_TILES) && ctx->con.kind == 1 && nl.use_cell && nl.ncx >= 3 && nl.ncy >= 3 &&
        ctx->ntypes == 1 && !ctx->soft_radii && !ctx->lj_radii && mass_u > 0.0 && ctx->n > 0))
    return false;
  // the engine keeps no age column: it needs the common age decided at the
  // start of this range
  AgeCtl h{};
  if (hipMemcpyAsync(&h, ctx->d_age, sizeof(AgeCtl), hipMemcpyDeviceToHost, ctx->st) != hipSuccess ||
      hipStreamSynchronize(ctx->st) != hipSuccess)
    return false;
  return h.uniform != 0;
}

// Carves the tiled engine's arrays out of one allocation, reused while the
// particle count, grid and tile width stay the same.
int tile_prepare(samos_ctx* ctx) {
  const NListD& nl = ctx->nl;
  const int n = ctx->n, ncx = nl.ncx, ncy = nl.ncy;
  const double mean = (double)n / ((double)ncx * ncy);
  // about 224 particles per tile, so most tiles fill one 256-thread workgroup
  int T = ctx->tile_cells > 0 ? ctx->tile_cells : (int)std::lround(224.0 / std::max(mean, 1e-9));
  T = std::max(1, std::min(T, std::min(kTileMaxCells, ncx)));
  // automatic width: the fewest tiles per row of at most T cells, of even width
  if (ctx->tile_cells <= 0) T = (ncx + (ncx + T - 1) / T - 1) / ((ncx + T - 1) / T);
  // list capacity: twice the mean count within rcut + pad, plus 8, in rows of 4
  const double area = ctx->box.Lx * ctx->box.Ly;
  const double mean_nb = area > 0.0 ? M_PI * nl.cut2 * (double)n / area : 0.0;
  const int K = 4 * std::max(2, (int)std::ceil((2.0 * mean_nb + 8.0) / 4.0));
  TileD& td = ctx->td;
  const int key[5] = {n, ncx, ncy, T, K};
  if (!(ctx->tile_mem && std::equal(key, key + 5, ctx->tile_key))) {
    hipFree(ctx->tile_mem);
    ctx->tile_mem = nullptr;
    const int ntx = (ncx + T - 1) / T;
    const long ncell = (long)ncx * ncy, ntiles = (long)ntx * ncy;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t d2 = al(sizeof(double2) * n), i4 = al(sizeof(int) * n), d8 = al(sizeof(double) * n);
    const size_t bytes = 2 * (2 * d2 + 4 * d2 + d8 + 2 * i4) + 2 * al(sizeof(int) * (ncell + 1)) +
                         al(sizeof(ushort4) * (size_t)(K / 4) * (size_t)n) + i4 +
                         al(sizeof(int) * kTileDesc * (size_t)ntiles) + al(sizeof(int) * ncell) +
                         al(sizeof(int) * ntiles) + al(sizeof(int) * (ntiles + 1)) + 2 * i4 + al(sizeof(TileCtl)) +
                         al(2 * sizeof(TileSet)) + al(sizeof(TileStepArgs)) + al(sizeof(TileBrownArgs));
    HIPCHK(hipMalloc(&ctx->tile_mem, bytes));
    ctx->tile_bytes = bytes;
    char* p = static_cast<char*>(ctx->tile_mem);
    auto take = [&](size_t b) { char* q = p; p += b; return q; };
    td = TileD{};
    TileSet sets[2];
    for (int s = 0; s < 2; ++s) {
      sets[s].xy[0] = (double2*)take(d2); sets[s].xy[1] = (double2*)take(d2);
      sets[s].v = (double2*)take(d2); sets[s].nd[0] = (double2*)take(d2); sets[s].nd[1] = (double2*)take(d2);
      sets[s].o = (double2*)take(d2); sets[s].om = (double*)take(d8);
      sets[s].id = (int*)take(i4); sets[s].slot = (int*)take(i4);
      sets[s].cstart = (int*)take(al(sizeof(int) * (ncell + 1)));
    }
    TileSet* d_sets = (TileSet*)take(al(2 * sizeof(TileSet)));
    HIPCHK(hipMemcpy(d_sets, sets, 2 * sizeof(TileSet), hipMemcpyHostToDevice));
    td.sets = d_sets;
    td.lst = (ushort4*)take(al(sizeof(ushort4) * (size_t)(K / 4) * (size_t)n));
    td.len = (int*)take(i4);
    td.desc = (int*)take(al(sizeof(int) * kTileDesc * (size_t)ntiles));
    td.count = (int*)take(al(sizeof(int) * ncell));
    td.tcount = (int*)take(al(sizeof(int) * ntiles));
    td.tstart = (int*)take(al(sizeof(int) * (ntiles + 1)));
    td.cell_of = (int*)take(i4);
    td.members = (int*)take(i4);
    td.ctl = (TileCtl*)take(al(sizeof(TileCtl)));
    ctx->td_args = (TileStepArgs*)take(al(sizeof(TileStepArgs)));
    ctx->tb_args = (TileBrownArgs*)take(al(sizeof(TileBrownArgs)));
    td.n = n; td.ncx = ncx; td.ncy = ncy; td.ncell = (int)ncell; td.T = T; td.ntx = ntx; td.ntiles = (int)ntiles;
    td.K = K;
    std::copy(key, key + 5, ctx->tile_key);
  }
  HIPCHK(hipMemsetAsync(td.count, 0, sizeof(int) * (size_t)td.ncell, ctx->st));
  HIPCHK(hipMemsetAsync(td.tcount, 0, sizeof(int) * (size_t)td.ntiles, ctx->st));
  td.wx = nl.wx; td.wy = nl.wy; td.xlo = ctx->box.xlo; td.ylo = ctx->box.ylo;
  td.cut2 = nl.cut2; td.pad2q = nl.pad2q;
  td.dirs2 = 0;
  return 0;
}

int read_tile_ctl(samos_ctx* ctx, TileCtl* h) {
  HIPCHK(hipMemcpyAsync(h, ctx->td.ctl, sizeof(TileCtl), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

// The constant arguments of a tiled run's step kernels, uploaded to the
// context's device copy before the run's first launch.
int tile_upload_args(samos_ctx* ctx, Integ& I, double mass_u) {
  // the host copy lives in the context, so the asynchronous upload never reads
  // a released stack frame
  TileStepArgs& A = ctx->h_tile_args;
  A = TileStepArgs{};
  A.td = ctx->td;
  const Fields f = fields(ctx);
  A.so = SlotOut{f.fx, f.fy, f.fz, f.ix, f.iy, f.iz, f.tx, f.ty, f.tz};
  A.b = ctx->box;
  A.c = ctx->con;
  FusedArgs& fa = A.fa;
  fa.dt = I.dt; fa.gamma = I.gamma; fa.temp = I.temp; fa.seed = I.seed;
  fa.clock = ctx->d_clock; fa.clock_slot = I.slot; fa.t_off = 0; fa.parity = 0;
  fa.head_enforce = 0;
  fa.mass_u = mass_u;
  // BAOAB O-step constants (src/integrators/integrator_langevin.cpp:57-58)
  fa.exp_dt = std::exp(-I.gamma * I.dt);
  fa.Bc = I.temp.kind == SAMOS_VALUE_CONSTANT ? std::sqrt(I.temp.tmin * (1.0 - std::exp(-2.0 * I.gamma * I.dt))) : -1.0;
  fa.fact_u = (0.5 * I.dt) / mass_u;  // the tiled engine runs with a common mass only
  fa.sf_u = fa.Bc >= 0.0 ? fa.Bc / std::sqrt(mass_u) : -1.0;
  fa.batch = 4;
  fa.age = ctx->d_age;
  PairConst& pc = A.pc;
  const int nn = kMaxTypes * kMaxTypes;
  if (ctx->has_soft) { pc.soft_k = ctx->soft_tab[0]; pc.soft_a = ctx->soft_tab[nn]; }
  if (ctx->has_lj) {
    pc.lj_eps = ctx->lj_tab[0];
    pc.lj_sig2 = ctx->lj_tab[nn] * ctx->lj_tab[nn];
    pc.lj_rc2 = ctx->lj_tab[2 * nn] * ctx->lj_tab[2 * nn];
  }
  pc.has_sp = ctx->has_sp ? 1 : 0;
  pc.sp_alpha = ctx->sp_alpha;
  HIPCHK(hipMemcpyAsync(ctx->td_args, &A, sizeof(A), hipMemcpyHostToDevice, ctx->st));
  return 0;
}

// par: the position buffer step t reads (its offset from the range start, mod 2).
void tile_fused_step(samos_ctx* ctx, int phase, int t, int t_off, int steps, bool f32, int par) {
  {
    ProfScope ps(ctx, "fused");
    launch_tile_step(ctx->st, pot_mask(ctx), phase, f32 ? 1 : 0, ctx->td, ctx->td_args, par, t_off,
                     t + 1 <= steps ? 1 : 0);
  }
  if (phase != 2) {
    ProfScope ps(ctx, "rebuild");
    launch_tile_build(ctx->st, ctx->td, ctx->box, 0, par);
  }
}

// Iterations a..b-1 of a run on the tiled engine; same contract as
// run_fused_langevin.  Returns 1 (nothing changed) when the system does not fit
// the engine, so the caller runs the list-based path instead.
int run_tiled_langevin(samos_ctx* ctx, Integ& I, int a, int b, int steps, int flags, double mass_u) {
  const bool f32 = (flags & SAMOS_RUN_FP32_FORCE) != 0;
  if (int r = tile_prepare(ctx)) return r;
  Fields f = fields(ctx);
  f.xy = ctx->xy[0];
  launch_tile_load(ctx->st, ctx->td, f, ctx->nl);
  launch_tile_build(ctx->st, ctx->td, ctx->box, 1, 1);
  TileCtl h;
  if (int r = read_tile_ctl(ctx, &h)) return r;
  if (h.overflow || !(h.word & 2)) return 1;  // too dense for the lists: list path
  if (int r = upload_clock(ctx, a)) return r;
  if (int r = tile_upload_args(ctx, I, mass_u)) return r;
  tile_fused_step(ctx, 0, a, 0, steps, f32, 0);
  launch_clock_advance(ctx->st, ctx->d_clock, 1);
  int t = a + 1;
  const int K = 64;  // steps per captured chunk
  const bool graph = !(flags & SAMOS_RUN_NO_GRAPH) && !ctx->prof;
  if (graph && b - t >= 2 * K) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < K; ++j) tile_fused_step(ctx, 1, t + j, j, steps, f32, (t + j - a) & 1);
    launch_clock_advance(ctx->st, ctx->d_clock, K);
    HIPCHK(hipStreamEndCapture(ctx->st, &g));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    while (b - t > K) {
      HIPCHK(hipGraphLaunch(ge, ctx->st));
      t += K;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipGraphExecDestroy(ge));
    HIPCHK(hipGraphDestroy(g));
  }
  const int t0 = t;
  for (; t < b; ++t) tile_fused_step(ctx, 1, t, t - t0, steps, f32, (t - a) & 1);
  launch_clock_advance(ctx->st, ctx->d_clock, b - t0);
  tile_fused_step(ctx, 2, b, 0, steps, f32, (b - a) & 1);
  launch_tile_store(ctx->st, ctx->td, f, ctx->nl, (b - a) & 1);
  HIPCHK(hipGetLastError());
  if (int r = read_tile_ctl(ctx, &h)) return r;
  ctx->tile_stats[0] += 1;
  ctx->tile_stats[1] += h.rebuilds;
  ctx->tile_stats[2] = std::max<long long>(ctx->tile_stats[2], h.maxcand);
  ctx->tile_stats[3] = std::max<long long>(ctx->tile_stats[3], h.maxhalo);
  if (h.far) return fail(ctx, "tiled run: a particle moved more than one cell width between neighbour-list builds");
  if (h.overflow)
    return fail(ctx, "tiled run: a particle had more than " + std::to_string(ctx->td.K) +
                         " neighbours (largest " + std::to_string(h.maxcand) + ")");
  if (h.rebuilds > 0) {
    // The slot-ordered list is rebuilt at the last build's snapshot, so it
    // holds exactly the pair set (and coordination) of that build.
    int before[8], after[8];
    if (int r = read_flags(ctx, before)) return r;
    Fields at = fields(ctx);
    at.x = ctx->nl.ox; at.y = ctx->nl.oy;
    if (int r = build_sync(ctx, &at)) return r;
    if (int r = read_flags(ctx, after)) return r;
    after[4] = before[4] + h.rebuilds;
    HIPCHK(hipMemcpy(ctx->d_flags + 4, &after[4], sizeof(int), hipMemcpyHostToDevice));
  }
  clear_flags(ctx, 0, 2);
  I.calls += (unsigned long long)(b - a);
  // one closing half-kick, and one age step, per iteration
  launch_age_finish(ctx->st, ctx->d_age, ctx->dcol[16], ctx->n, I.dt, b - a);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- cell-tiled fused Brownian run (planar systems) ------------------------
// One "brownian" integrator on every particle of a planar single-type system
// without per-particle radii, asked for with SAMOS_RUN_TILES: the fused
// Brownian step (forces, aligner torques and IntegratorBrownian) on the tiled
// engine, with positions and directors staged in LDS.  Same contract as
// run_fused_brownian; returns 1 (nothing changed) when the system does not fit.
bool tile_brownian_eligible(samos_ctx* ctx, int flags) {
  const NListD& nl = ctx->nl;
  if (!((flags & SAMOS_RUN_TILES) && ctx->integs.size() == 1 && ctx->con.kind == 1 && nl.use_cell &&
        nl.ncx >= 3 && nl.ncy >= 3 && ctx->ntypes == 1 && !ctx->soft_radii && !ctx->lj_radii && ctx->n > 0))
    return false;
  const Integ& I = ctx->integs.begin()->second;
  if (I.name != "brownian" || I.gbit != 1u) return false;
  AgeCtl h{};
  if (hipMemcpyAsync(&h, ctx->d_age, sizeof(AgeCtl), hipMemcpyDeviceToHost, ctx->st) != hipSuccess ||
      hipStreamSynchronize(ctx->st) != hipSuccess)
    return false;
  return h.uniform != 0;
}

int tile_upload_brownian_args(samos_ctx* ctx, const Integ& I) {
  TileBrownArgs& A = ctx->h_tb_args;
  A = TileBrownArgs{};
  A.td = ctx->td;
  const Fields f = fields(ctx);
  A.so = SlotOut{f.fx, f.fy, f.fz, f.ix, f.iy, f.iz, f.tx, f.ty, f.tz};
  A.b = ctx->box;
  A.c = ctx->con;
  PairConst& pc = A.pc;
  const int nn = kMaxTypes * kMaxTypes;
  if (ctx->has_soft) { pc.soft_k = ctx->soft_tab[0]; pc.soft_a = ctx->soft_tab[nn]; }
  if (ctx->has_lj) {
    pc.lj_eps = ctx->lj_tab[0];
    pc.lj_sig2 = ctx->lj_tab[nn] * ctx->lj_tab[nn];
    pc.lj_rc2 = ctx->lj_tab[2 * nn] * ctx->lj_tab[2 * nn];
  }
  pc.has_sp = ctx->has_sp ? 1 : 0;
  pc.sp_alpha = ctx->sp_alpha;
  A.dt = I.dt; A.v0 = I.v0; A.mu = I.mu; A.mur = I.mur; A.stoch = I.stoch; A.flip_p = I.flip_p;
  A.nematic = I.nematic ? 1 : 0; A.velocity = I.velocity ? 1 : 0;
  A.temp = I.temp; A.seed = I.seed; A.clock = ctx->d_clock; A.clock_slot = I.slot;
  // the aligner's coupling as the pair loop applies it: the global nematic J
  // is doubled, a per-type J is used as given
  // (src/aligner/pair/pair_nematic_aligner.cpp:37,57-61)
  if (ctx->has_pol) { A.al_J = ctx->pol_tab[0]; A.al_rc2 = ctx->pol_tab[nn] * ctx->pol_tab[nn]; }
  if (ctx->has_nem) {
    A.al_J = ctx->nem_pp ? ctx->nem_tab[0] : 2.0 * ctx->nem_J;
    A.al_rc2 = ctx->nem_tab[nn] * ctx->nem_tab[nn];
  }
  HIPCHK(hipMemcpyAsync(ctx->tb_args, &A, sizeof(A), hipMemcpyHostToDevice, ctx->st));
  return 0;
}

// store: the range's last step, whose loop-head forces and torques are the
// state left behind
void tile_brownian_step(samos_ctx* ctx, int t, int t_off, int steps, bool f32, int par, bool store = false) {
  {
    ProfScope ps(ctx, "fused");
    launch_tile_brownian(ctx->st, pot_mask(ctx), al_mask(ctx), f32 ? 1 : 0, ctx->td, ctx->tb_args, par, t_off,
                         t + 1 <= steps ? 1 : 0, store ? 1 : 0);
  }
  ProfScope ps(ctx, "rebuild");
  launch_tile_build(ctx->st, ctx->td, ctx->box, 0, par);
}

int run_tiled_brownian(samos_ctx* ctx, Integ& I, int a, int b, int steps, int flags) {
  const bool f32 = (flags & SAMOS_RUN_FP32_FORCE) != 0;
  if (int r = tile_prepare(ctx)) return r;
  ctx->td.dirs2 = 1;
  Fields f = fields(ctx);
  f.xy = ctx->xy[0];
  launch_tile_load(ctx->st, ctx->td, f, ctx->nl);
  launch_tile_build(ctx->st, ctx->td, ctx->box, 1, 1);
  TileCtl h;
  if (int r = read_tile_ctl(ctx, &h)) return r;
  if (h.overflow || !(h.word & 2)) return 1;  // too dense for the lists: list path
  if (int r = upload_clock(ctx, a)) return r;
  if (int r = tile_upload_brownian_args(ctx, I)) return r;
  int t = a;
  const int K = 64;  // steps per captured chunk
  const bool graph = !(flags & SAMOS_RUN_NO_GRAPH) && !ctx->prof;
  if (graph && b - t >= 2 * K) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < K; ++j) tile_brownian_step(ctx, t + j, j, steps, f32, (t + j - a) & 1);
    launch_clock_advance(ctx->st, ctx->d_clock, K);
    HIPCHK(hipStreamEndCapture(ctx->st, &g));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    while (b - t > K) {
      HIPCHK(hipGraphLaunch(ge, ctx->st));
      t += K;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipGraphExecDestroy(ge));
    HIPCHK(hipGraphDestroy(g));
  }
  const int t0 = t;
  for (; t < b; ++t) tile_brownian_step(ctx, t, t - t0, steps, f32, (t - a) & 1, t == b - 1);
  launch_clock_advance(ctx->st, ctx->d_clock, b - t0);
  launch_tile_store(ctx->st, ctx->td, f, ctx->nl, (b - a) & 1);
  HIPCHK(hipGetLastError());
  if (int r = read_tile_ctl(ctx, &h)) return r;
  ctx->tile_stats[0] += 1;
  ctx->tile_stats[1] += h.rebuilds;
  ctx->tile_stats[2] = std::max<long long>(ctx->tile_stats[2], h.maxcand);
  ctx->tile_stats[3] = std::max<long long>(ctx->tile_stats[3], h.maxhalo);
  if (h.far) return fail(ctx, "tiled run: a particle moved more than one cell width between neighbour-list builds");
  if (h.overflow)
    return fail(ctx, "tiled run: a particle had more than " + std::to_string(ctx->td.K) +
                         " neighbours (largest " + std::to_string(h.maxcand) + ")");
  if (h.rebuilds > 0) {
    // the slot-ordered list is rebuilt at the last build's snapshot, so it
    // holds exactly the pair set (and coordination) of that build
    int before[8], after[8];
    if (int r = read_flags(ctx, before)) return r;
    Fields at = fields(ctx);
    at.x = ctx->nl.ox; at.y = ctx->nl.oy;
    if (int r = build_sync(ctx, &at)) return r;
    if (int r = read_flags(ctx, after)) return r;
    after[4] = before[4] + h.rebuilds;
    HIPCHK(hipMemcpy(ctx->d_flags + 4, &after[4], sizeof(int), hipMemcpyHostToDevice));
  }
  clear_flags(ctx, 0, 2);
  launch_age_finish(ctx->st, ctx->d_age, ctx->dcol[16], ctx->n, I.dt, b - a);
  I.calls += (unsigned long long)(b - a);
  // loop head of b: its enforce was applied with the positions; forces and torques
  if (b <= steps) loop_head_forces(ctx, nullptr, f32);
  HIPCHK(hipGetLastError());
  return 0;
}

int check_overflow(samos_ctx* ctx) {
  int h[8];
  if (int r = read_flags(ctx, h)) return r;
  if (h[2]) {
    return fail(ctx, "Neighbour list overflow during the run: a particle had " + std::to_string(h[5]) +
                         " neighbours, capacity " + std::to_string(ctx->nl.stride) + ".");
  }
  return 0;
}

// ---- scratch of the allocation-free pair export ----------------------------
// Staging of samos_nlist_pairs_noalloc: a counter in the first 8 bytes, the
// exported rows from byte kScratchHead on.  One copy per GPU, held until the
// process ends.  Host code uses only the per-device address resolved below.
constexpr long kScratchPairs = 1L << 21;  // half pairs: 2,097,152 rows of two ints (16 MiB)
constexpr size_t kScratchHead = 256;
static __device__ __align__(256) unsigned char scratch[kScratchHead + 8 * (size_t)kScratchPairs];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device `device`, resolved
// once per device.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

__global__ void k_clear_counter(unsigned long long* c) {
  if (threadIdx.x == 0) *c = 0ULL;
}

// Rows (id_lo, id_hi) as stored in the caller's int array.
struct PairRow { int a, b; };

}  // namespace

extern "C" {

int samos_create(samos_ctx** out, int device, int n, double lx, double ly, double lz, int periodic) {
  *out = nullptr;
  if (n < 0) { g_create_error = "negative particle count"; return -1; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { g_create_error = "no HIP device"; return -2; }
  if (device < 0 || device >= ndev) { g_create_error = "device index out of range"; return -1; }
  if (hipSetDevice(device) != hipSuccess) { g_create_error = "hipSetDevice failed"; return -2; }
  {
    // set-up of the allocation-free entry points on this device
    void* sp = nullptr;
    const hipError_t e = scratch_address(device, &sp);
    if (e != hipSuccess) { g_create_error = std::string("scratch address: ") + hipGetErrorString(e); return -2; }
  }
  samos_ctx* ctx = new samos_ctx();
  ctx->device = device;
  ctx->n = n;
  ctx->np = ((n + 63) / 64) * 64;
  if (ctx->np == 0) ctx->np = 64;
  BoxD& b = ctx->box;
  b.Lx = lx; b.Ly = ly; b.Lz = lz;
  b.xlo = -0.5 * lx; b.xhi = 0.5 * lx; b.ylo = -0.5 * ly; b.yhi = 0.5 * ly; b.zlo = -0.5 * lz; b.zhi = 0.5 * lz;
  b.periodic = periodic ? 1 : 0;
  auto bad = [&](const char* what) { g_create_error = std::string("allocation failed: ") + what; samos_destroy(ctx); return -2; };
  if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess) return bad("stream");
  const size_t np = ctx->np;
  for (int k = 0; k < kNumDoubleFields; ++k)
    if (hipMalloc(&ctx->dcol[k], sizeof(double) * np) != hipSuccess) return bad("columns");
  for (int k = 0; k < kNumIntFields; ++k)
    if (hipMalloc(&ctx->icol[k], sizeof(int) * np) != hipSuccess) return bad("columns");
  for (int k = 0; k < 3; ++k)
    if (hipMalloc(&ctx->alt[k], sizeof(double) * np) != hipSuccess || hipMalloc(&ctx->altn[k], sizeof(double) * np) != hipSuccess)
      return bad("columns");
  for (int k = 0; k < 2; ++k)
    if (hipMalloc(&ctx->xy[k], sizeof(double2) * np) != hipSuccess) return bad("columns");
  if (hipMalloc(&ctx->d_flags, 16 * sizeof(int)) != hipSuccess) return bad("flags");
  if (hipMalloc(&ctx->d_nnb, sizeof(int) * np) != hipSuccess) return bad("neighbour counts");
  hipMemsetAsync(ctx->d_nnb, 0, sizeof(int) * np, ctx->st);
  ctx->nl.nnb = ctx->d_nnb;
  ctx->nl.np = ctx->np;
  ctx->nl.flags = ctx->d_flags;
  if (hipMalloc(&ctx->d_tab, 4 * kTab * sizeof(double)) != hipSuccess) return bad("tables");
  if (hipMalloc(&ctx->d_clock, sizeof(ClockD)) != hipSuccess) return bad("clock");
  if (hipMalloc(&ctx->d_chunk, sizeof(ChunkCtl)) != hipSuccess) return bad("chunk control");
  if (hipMalloc(&ctx->d_age, sizeof(AgeCtl)) != hipSuccess) return bad("age control");
  if (hipMalloc(&ctx->d_en, 5 * np * sizeof(double)) != hipSuccess) return bad("energies");
  if (hipMalloc(&ctx->d_cols, 3 * np * sizeof(double)) != hipSuccess) return bad("fire columns");
  if (hipMalloc(&ctx->d_partial, 8 * (size_t)reduce_blocks(ctx->np) * sizeof(double)) != hipSuccess) return bad("partials");
  if (hipMalloc(&ctx->d_out, 16 * sizeof(double)) != hipSuccess) return bad("out");
  if (hipHostMalloc(&ctx->h_out, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess) return bad("pinned out");
  if (hipMalloc(&ctx->d_counter, sizeof(unsigned long long)) != hipSuccess) return bad("counter");
  if (hipMalloc(&ctx->keys, sizeof(unsigned long long) * np) != hipSuccess) return bad("sort");
  if (hipMalloc(&ctx->keys2, sizeof(unsigned long long) * np) != hipSuccess) return bad("sort");
  if (hipMalloc(&ctx->perm, sizeof(int) * np) != hipSuccess) return bad("sort");
  if (hipMalloc(&ctx->perm2, sizeof(int) * np) != hipSuccess) return bad("sort");
  if (hipMalloc(&ctx->dtmp, sizeof(double) * np) != hipSuccess) return bad("sort");
  if (hipMalloc(&ctx->itmp, sizeof(int) * np) != hipSuccess) return bad("sort");
  ctx->sort_tmp_bytes = radix_sort_temp_bytes(ctx->n > 0 ? ctx->n : 1);
  if (hipMalloc(&ctx->sort_tmp, ctx->sort_tmp_bytes + 256) != hipSuccess) return bad("sort");
  // defaults: type 1, radius 1, mass 1, everything else zero; id = slot
  hipMemsetAsync(ctx->d_flags, 0, 16 * sizeof(int), ctx->st);
  hipMemsetAsync(ctx->d_chunk, 0, sizeof(ChunkCtl), ctx->st);
  hipMemsetAsync(ctx->d_tab, 0, 4 * kTab * sizeof(double), ctx->st);
  for (int k = 0; k < kNumDoubleFields; ++k) hipMemsetAsync(ctx->dcol[k], 0, sizeof(double) * np, ctx->st);
  for (int k = 0; k < kNumIntFields; ++k) hipMemsetAsync(ctx->icol[k], 0, sizeof(int) * np, ctx->st);
  ctx->h_id.resize(n);
  for (int i = 0; i < n; ++i) ctx->h_id[i] = i;
  ctx->h_groups.assign(n, 1u);  // every particle is in group "all"
  std::vector<double> ones(n, 1.0);
  std::vector<int> ione(n, 1);
  if (n > 0) {
    hipMemcpyAsync(ctx->dcol[17], ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st);
    ctx->nl.r_uni = 1.0;  // every radius is the default 1
    hipMemcpyAsync(ctx->dcol[18], ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st);
    hipMemcpyAsync(ctx->icol[0], ione.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->st);
    hipMemcpyAsync(ctx->icol[5], ctx->h_id.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->st);
    hipMemcpyAsync(ctx->icol[6], ione.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->st);
  }
  if (hipStreamSynchronize(ctx->st) != hipSuccess) return bad("initialisation");
  *out = ctx;
  return 0;
}

int samos_destroy(samos_ctx* ctx) {
  if (!ctx) return 0;
  if (ctx->st) hipStreamSynchronize(ctx->st);
  for (auto& p : ctx->profs)
    for (auto& e : p.second.ev) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
  for (auto& g : ctx->graphs) hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
  free_nlist(ctx);
  for (auto& kv : ctx->integs) { hipFree(kv.second.R); hipFree(kv.second.fire); }
  for (int k = 0; k < kNumDoubleFields; ++k) hipFree(ctx->dcol[k]);
  for (int k = 0; k < kNumIntFields; ++k) hipFree(ctx->icol[k]);
  for (int k = 0; k < 3; ++k) { hipFree(ctx->alt[k]); hipFree(ctx->altn[k]); }
  for (int k = 0; k < 2; ++k) hipFree(ctx->xy[k]);
  hipFree(ctx->d_flags); hipFree(ctx->d_nnb); hipFree(ctx->d_tab); hipFree(ctx->d_clock); hipFree(ctx->d_chunk); hipFree(ctx->d_age); hipFree(ctx->d_en); hipFree(ctx->d_cols);
  hipFree(ctx->d_partial); hipFree(ctx->d_out); hipFree(ctx->d_counter);
  if (ctx->h_out) hipHostFree(ctx->h_out);
  hipFree(ctx->keys); hipFree(ctx->keys2); hipFree(ctx->perm); hipFree(ctx->perm2);
  hipFree(ctx->dtmp); hipFree(ctx->itmp); hipFree(ctx->sort_tmp);
  hipFree(ctx->tile_mem);
  if (ctx->st) hipStreamDestroy(ctx->st);
  delete ctx;
  return 0;
}

const char* samos_last_error(const samos_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int samos_num_particles(const samos_ctx* ctx) { return ctx ? ctx->n : -1; }

int samos_set_field(samos_ctx* ctx, int field, const void* host) {
  const int n = ctx->n;
  if (n == 0) return 0;
  if (field >= 0 && field < kNumDoubleFields) {
    const double* h = static_cast<const double*>(host);
    std::vector<double> tmp(n);
    for (int s = 0; s < n; ++s) tmp[s] = h[ctx->h_id[s]];
    HIPCHK(hipMemcpyAsync(ctx->dcol[field], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
    if (field == SAMOS_MASS) ctx->mass_known = false;
    if (field == SAMOS_AGE) ctx->age_checked = false;
    if (field == SAMOS_RADIUS) {
      bool uni = true;
      for (int s = 1; s < n && uni; ++s) uni = tmp[s] == tmp[0];
      ctx->nl.r_uni = uni ? tmp[0] : -1.0;
    }
    if (field <= 2) HIPCHK(hipMemcpyAsync(ctx->alt[field], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
  } else if (field >= SAMOS_TYPE && field < SAMOS_TYPE + 5) {
    const int* h = static_cast<const int*>(host);
    std::vector<int> tmp(n);
    for (int s = 0; s < n; ++s) tmp[s] = h[ctx->h_id[s]];
    if (field == SAMOS_TYPE) {
      int mx = 1;
      for (int v : tmp) {
        if (v < 1 || v > kMaxTypes)
          return fail(ctx, "particle type " + std::to_string(v) + " outside 1.." + std::to_string(kMaxTypes));
        mx = std::max(mx, v);
      }
      ctx->ntypes = mx;
    }
    HIPCHK(hipMemcpyAsync(ctx->icol[field - SAMOS_TYPE], tmp.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->st));
  } else {
    return fail(ctx, "unknown field code " + std::to_string(field));
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_get_field(samos_ctx* ctx, int field, void* host) {
  const int n = ctx->n;
  if (n == 0) return 0;
  if (field >= 0 && field < kNumDoubleFields) {
    std::vector<double> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), ctx->dcol[field], sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    double* h = static_cast<double*>(host);
    for (int s = 0; s < n; ++s) h[ctx->h_id[s]] = tmp[s];
  } else if (field >= SAMOS_TYPE && field < SAMOS_TYPE + 5) {
    std::vector<int> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), ctx->icol[field - SAMOS_TYPE], sizeof(int) * n, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    int* h = static_cast<int*>(host);
    for (int s = 0; s < n; ++s) h[ctx->h_id[s]] = tmp[s];
  } else {
    return fail(ctx, "unknown field code " + std::to_string(field));
  }
  return 0;
}

int samos_set_timestep(samos_ctx* ctx, double dt) {
  ctx->dt = dt;
  return 0;
}

int samos_nlist_create(samos_ctx* ctx, double rcut, double pad) {
  free_nlist(ctx);
  ctx->has_nlist = true;
  ctx->rcut = rcut;
  ctx->pad = pad;
  NListD& nl = ctx->nl;
  const double r_uni = nl.r_uni;  // a property of the particles, kept across lists
  nl = NListD{};
  nl.r_uni = r_uni;
  const double cut = rcut + pad;
  nl.cut2 = cut * cut;
  nl.pad2q = 0.25 * pad * pad;
  nl.np = ctx->np;
  nl.flags = ctx->d_flags;
  const BoxD& b = ctx->box;
  nl.use_cell = (b.Lx > 2.0 * cut && b.Ly > 2.0 * cut && b.Lz > 2.0 * cut) ? 1 : 0;
  if (nl.use_cell) {
    // CellList::CellList (src/system/cell_list.cpp:37-68)
    nl.ncx = (int)(b.Lx / cut); nl.ncy = (int)(b.Ly / cut); nl.ncz = (int)(b.Lz / cut);
    const long ncell = (long)nl.ncx * nl.ncy * nl.ncz;
    if (ncell > 400000000L) return fail(ctx, "cell grid too large");
    nl.ncell = (int)ncell;
    nl.wx = b.Lx / nl.ncx; nl.wy = b.Ly / nl.ncy; nl.wz = b.Lz / nl.ncz;
    HIPCHK(hipMalloc(&nl.count, sizeof(int) * (size_t)nl.ncell));
    HIPCHK(hipMalloc(&nl.start, sizeof(int) * ((size_t)nl.ncell + 1)));
    HIPCHK(hipMalloc(&nl.bsum, sizeof(int) * ((size_t)nl.ncell / 4096 + 2)));
    HIPCHK(hipMemsetAsync(nl.count, 0, sizeof(int) * (size_t)nl.ncell, ctx->st));
  } else {
    nl.ncx = nl.ncy = nl.ncz = 0;
    nl.ncell = 0;
  }
  HIPCHK(hipMalloc(&nl.cell_of, sizeof(int) * ctx->np));
  HIPCHK(hipMalloc(&nl.members, sizeof(int) * ctx->np));
  nl.nnb = ctx->d_nnb;
  HIPCHK(hipMalloc(&nl.ox, sizeof(double) * ctx->np));
  HIPCHK(hipMalloc(&nl.oy, sizeof(double) * ctx->np));
  HIPCHK(hipMalloc(&nl.oz, sizeof(double) * ctx->np));
  HIPCHK(hipMemsetAsync(nl.nnb, 0, sizeof(int) * ctx->np, ctx->st));
  // initial capacity from the mean density, grown by build_sync when needed
  const double vol = b.Lx * b.Ly * b.Lz;
  const double dens = vol > 0 ? ctx->n / vol : 0.0;
  int guess = (int)(dens * 4.0 / 3.0 * M_PI * cut * cut * cut * 2.0) + 32;
  guess = std::min(std::max(guess, 32), std::max(ctx->n, 32));
  if (int r = alloc_nbr(ctx, guess)) return r;
  // the reference's initial build works on the input order; re-sorting first
  // changes only the storage order
  if (nl.use_cell) {
    if (int r = build_sync(ctx)) return r;
    if (int r = resort(ctx)) return r;
  }
  return build_sync(ctx);
}

int samos_nlist_build(samos_ctx* ctx) {
  if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
  return build_sync(ctx);
}

int samos_resort(samos_ctx* ctx) {
  if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
  if (int r = resort(ctx)) return r;
  return build_sync(ctx);
}

int samos_nlist_need_update(samos_ctx* ctx, int* out) {
  if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
  HIPCHK(hipMemsetAsync(ctx->d_flags + 6, 0, sizeof(int), ctx->st));
  launch_need_update(ctx->st, fields(ctx), ctx->nl, ctx->box, ctx->n, ctx->d_flags + 6);
  int h[8];
  if (int r = read_flags(ctx, h)) return r;
  *out = h[6] ? 1 : 0;
  return 0;
}

int samos_nlist_info(samos_ctx* ctx, long long* info) {
  int h[8];
  if (int r = read_flags(ctx, h)) return r;
  info[0] = ctx->nl.use_cell; info[1] = ctx->nl.ncx; info[2] = ctx->nl.ncy; info[3] = ctx->nl.ncz;
  info[4] = ctx->nl.stride; info[5] = h[4]; info[6] = h[3]; info[7] = h[2];
  return 0;
}

int samos_nlist_pairs(samos_ctx* ctx, int* out, long long cap, long long* count) {
  if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
  // total directed entries bound the half-pair count
  HIPCHK(hipMemsetAsync(ctx->d_counter, 0, sizeof(unsigned long long), ctx->st));
  std::vector<int> nnb(ctx->n);
  HIPCHK(hipMemcpyAsync(nnb.data(), ctx->nl.nnb, sizeof(int) * ctx->n, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  long long total = 0;
  for (int v : nnb) total += v;
  const long long half_cap = total / 2 + 1;
  int* d_out = nullptr;
  HIPCHK(hipMalloc(&d_out, sizeof(int) * 2 * (size_t)half_cap));
  launch_export_pairs(ctx->st, fields(ctx), ctx->nl, ctx->n, d_out, ctx->d_counter, half_cap);
  unsigned long long cnt = 0;
  HIPCHK(hipMemcpyAsync(&cnt, ctx->d_counter, sizeof(cnt), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if ((long long)cnt > half_cap) { hipFree(d_out); return fail(ctx, "pair export overflow"); }
  std::vector<std::pair<int, int>> pr(cnt);
  std::vector<int> raw(2 * cnt);
  if (cnt) HIPCHK(hipMemcpy(raw.data(), d_out, sizeof(int) * 2 * cnt, hipMemcpyDeviceToHost));
  hipFree(d_out);
  for (size_t k = 0; k < cnt; ++k) pr[k] = {raw[2 * k], raw[2 * k + 1]};
  std::sort(pr.begin(), pr.end());
  *count = (long long)cnt;
  if (out) {
    if ((long long)cnt > cap) return fail(ctx, "output capacity too small");
    for (size_t k = 0; k < cnt; ++k) { out[2 * k] = pr[k].first; out[2 * k + 1] = pr[k].second; }
  }
  return 0;
}

int samos_nlist_pairs_noalloc(samos_ctx* ctx, int* out, long long cap, long long* count) {
  *count = 0;
  if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
  void* base = nullptr;
  {
    const hipError_t e = scratch_address(ctx->device, &base);
    if (e != hipSuccess) return fail(ctx, std::string("scratch address: ") + hipGetErrorString(e));
  }
  auto* counter = static_cast<unsigned long long*>(base);
  int* rows = reinterpret_cast<int*>(static_cast<unsigned char*>(base) + kScratchHead);
  hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(64), 0, ctx->st, counter);
  launch_export_pairs(ctx->st, fields(ctx), ctx->nl, ctx->n, rows, counter, kScratchPairs);
  HIPCHK(hipGetLastError());
  unsigned long long cnt = 0;
  HIPCHK(hipMemcpyAsync(&cnt, counter, sizeof(cnt), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (cnt > (unsigned long long)kScratchPairs)
    return fail(ctx, "samos_nlist_pairs_noalloc: the list holds " + std::to_string(cnt) +
                         " half pairs, more than its scratch maximum of " + std::to_string(kScratchPairs) +
                         "; use samos_nlist_pairs");
  *count = (long long)cnt;
  if (!out) return 0;
  if ((long long)cnt > cap) return fail(ctx, "output capacity too small");
  if (cnt) HIPCHK(hipMemcpyAsync(out, rows, sizeof(int) * 2 * cnt, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  // ascending (id_lo, id_hi), sorted in place in the caller's array
  PairRow* r = reinterpret_cast<PairRow*>(out);
  std::sort(r, r + cnt, [](const PairRow& x, const PairRow& y) { return x.a < y.a || (x.a == y.a && x.b < y.b); });
  return 0;
}

int samos_pair_soft(samos_ctx* ctx, double k, double a, int use_radii) {
  ctx->has_soft = true; ctx->soft_k = k; ctx->soft_a = a; ctx->soft_radii = use_radii ? 1 : 0;
  ctx->soft_tab = make_tab(k, a, 0.0);
  return upload_tables(ctx);
}

int samos_pair_lj(samos_ctx* ctx, double eps, double sigma, double rcut, int shifted, int use_radii) {
  ctx->has_lj = true; ctx->lj_eps = eps; ctx->lj_sig = sigma; ctx->lj_rc = rcut;
  ctx->lj_shifted = shifted ? 1 : 0; ctx->lj_radii = use_radii ? 1 : 0;
  ctx->lj_tab = make_tab(eps, sigma, rcut);
  return upload_tables(ctx);
}

int samos_pair_type_param(samos_ctx* ctx, int kind, int t1, int t2, double p0, double p1, double p2, int shifted) {
  if (t1 < 1 || t2 < 1 || t1 > kMaxTypes || t2 > kMaxTypes) return fail(ctx, "type out of range");
  if (kind == SAMOS_PAIR_SOFT) {
    if (!ctx->has_soft) return fail(ctx, "soft potential not defined");
    set_tab(ctx->soft_tab, t1, t2, p0, p1, 0.0);
  } else if (kind == SAMOS_PAIR_LJ) {
    if (!ctx->has_lj) return fail(ctx, "lj potential not defined");
    set_tab(ctx->lj_tab, t1, t2, p0, p1, p2);
    if (shifted) ctx->lj_shifted = 1;
  } else {
    return fail(ctx, "unknown pair potential kind");
  }
  return upload_tables(ctx);
}

int samos_align_polar(samos_ctx* ctx, double J, double rcut) {
  if (rcut <= 0.0) {
    if (!ctx->has_nlist) return fail(ctx, "polar aligner needs a neighbour list");
    rcut = ctx->rcut;
  }
  ctx->has_pol = true; ctx->pol_J = J; ctx->pol_rc = rcut;
  ctx->pol_tab = make_tab(J, rcut, 0.0);
  return upload_tables(ctx);
}

int samos_align_nematic(samos_ctx* ctx, double J, double rcut) {
  if (rcut <= 0.0) {
    if (!ctx->has_nlist) return fail(ctx, "nematic aligner needs a neighbour list");
    rcut = ctx->rcut;
  }
  ctx->has_nem = true; ctx->nem_J = J; ctx->nem_rc = rcut; ctx->nem_pp = false;
  ctx->nem_tab = make_tab(J, rcut, 0.0);
  return upload_tables(ctx);
}

int samos_align_type_param(samos_ctx* ctx, int kind, int t1, int t2, double J, double rcut) {
  if (t1 < 1 || t2 < 1 || t1 > kMaxTypes || t2 > kMaxTypes) return fail(ctx, "type out of range");
  if (kind == SAMOS_ALIGN_POLAR) {
    if (!ctx->has_pol) return fail(ctx, "polar aligner not defined");
    set_tab(ctx->pol_tab, t1, t2, J, rcut, 0.0);
  } else if (kind == SAMOS_ALIGN_NEMATIC) {
    if (!ctx->has_nem) return fail(ctx, "nematic aligner not defined");
    set_tab(ctx->nem_tab, t1, t2, J, rcut, 0.0);
    ctx->nem_pp = true;
  } else {
    return fail(ctx, "unknown aligner kind");
  }
  return upload_tables(ctx);
}

int samos_external_self_propulsion(samos_ctx* ctx, double alpha) {
  ctx->has_sp = true;
  ctx->sp_alpha = alpha;
  return 0;
}

int samos_constraint(samos_ctx* ctx, int kind, double param, int flag) {
  if (kind < 0 || kind > 2) return fail(ctx, "unknown constraint kind");
  const int was = ctx->con.kind;
  ctx->con.kind = kind;
  if (kind == 1) { ctx->con.zpos = param; ctx->con.unlimited = flag ? 1 : 0; }
  if (kind == 2) ctx->con.r = param;
  launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipGetLastError());
  // a plane keeps its slots in Hilbert order (resort); a list made before
  // the constraint was set is re-sorted into it (the same pair set)
  if (kind != was && (kind == 1 || was == 1) && ctx->has_nlist && ctx->nl.use_cell && ctx->n >= 2)
    return resort_at_snapshot(ctx);
  return 0;
}

int samos_constraint_surface(samos_ctx* ctx, int kind, const double* params, int nparams, int max_iter, double tol) {
  static const int need[] = {0, 0, 0, 1, 2, 3};
  if (kind < SAMOS_CONSTRAINT_CYLINDER || kind > SAMOS_CONSTRAINT_ELLIPSOID) return fail(ctx, "unknown constraint kind");
  if (nparams != need[kind]) return fail(ctx, "wrong parameter count for the constraint");
  if (max_iter < 0) return fail(ctx, "Constraint. Maximum number of iterations has to be non-negative.");
  ConstraintD c = ctx->con;
  c.kind = kind;
  c.max_iter = max_iter;
  c.tol = tol;
  if (kind == SAMOS_CONSTRAINT_CYLINDER) c.r = params[0];
  if (kind == SAMOS_CONSTRAINT_TORUS) {
    c.q0 = params[0]; c.q1 = params[1];
    if (c.q1 >= c.q0) return fail(ctx, "Toroidal constraint. Tube radius has to be smaller than torus radius.");
  }
  if (kind == SAMOS_CONSTRAINT_ELLIPSOID) { c.q0 = params[0]; c.q1 = params[1]; c.q2 = params[2]; }
  ctx->con = c;
  launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipGetLastError());
  return 0;
}

int samos_enforce_all(samos_ctx* ctx) {
  launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

static int add_integ(samos_ctx* ctx, Integ& I, const std::string& name) {
  const std::string group = ctx->next_group;
  ctx->next_group = "all";
  if (ctx->dt <= 0.0)
    return fail(ctx, "Integrator step size has to be larger than 0. Did you forget to use \"timestep\" command?");
  const std::string key = name + "_" + group;
  if (ctx->integs.count(key))
    return fail(ctx, "Integrator of type " + name + " (for group " + group +
                         ") already exists. You need to disable the old one first.");
  if (ctx->integs.size() >= 8) return fail(ctx, "too many integrators");
  I.name = name;
  I.group = group;
  I.gbit = 1u << ctx->group_index.at(group);
  I.dt = ctx->dt;
  int slot = 0;
  for (;; ++slot) {
    bool used = false;
    for (auto& kv : ctx->integs) used |= kv.second.slot == slot;
    if (!used) break;
  }
  I.slot = slot;
  ctx->integs[key] = I;
  ctx->last_key = key;
  return 0;
}

// The integrator that add_integ registered last (its map key holds the group).
static Integ& last_integ(samos_ctx* ctx, const char*) { return ctx->integs.at(ctx->last_key); }

int samos_integrator_brownian(samos_ctx* ctx, int which, double v0, double nu, double mu, double mur, int seed,
                              int nematic, double tau, int velocity_align, int tkind, double tmin, double tmax,
                              int tsteps) {
  if (which < 0 || which > 2) return fail(ctx, "unknown brownian integrator kind");
  Integ I;
  I.v0 = v0; I.nu = nu; I.mu = mu; I.mur = mur; I.seed = (uint32_t)seed;
  I.nematic = nematic ? 1 : 0; I.velocity = velocity_align ? 1 : 0;
  I.temp = TempD{tkind, tmin, tmax, tsteps};
  const char* nm = which == 0 ? "brownian" : (which == 1 ? "brownian_pos" : "brownian_align");
  if (int r = add_integ(ctx, I, nm)) return r;
  Integ& J = last_integ(ctx, nm);
  J.flip_p = J.dt / tau;
  J.stoch = std::sqrt(J.nu * J.dt);
  return 0;
}

int samos_integrator_langevin(samos_ctx* ctx, double gamma, int seed, int method, int tkind, double tmin,
                              double tmax, int tsteps) {
  if (method < 0 || method > 2) return fail(ctx, "Unknown integration method");
  Integ I;
  I.gamma = gamma; I.seed = (uint32_t)seed; I.method = method;
  I.temp = TempD{tkind, tmin, tmax, tsteps};
  if (int r = add_integ(ctx, I, "langevin")) return r;
  Integ& J = last_integ(ctx, "langevin");
  if (method == 2) {
    HIPCHK(hipMalloc(&J.R, sizeof(double) * 3 * (size_t)std::max(ctx->n, 1)));
    launch_bbk_init(ctx->st, fields(ctx), ctx->n, J.seed, J.R);
    HIPCHK(hipStreamSynchronize(ctx->st));
  }
  return 0;
}

int samos_integrator_fire(samos_ctx* ctx, double alpha, double alpha_scale, double f_tol, double e_tol,
                          double dt_inc, double dt_dec, double dt_max, int min_neg_steps) {
  Integ I;
  I.alpha = alpha; I.alpha_init = alpha; I.f_alpha = alpha_scale; I.F_tol = f_tol; I.E_tol = e_tol;
  I.f_inc = dt_inc; I.f_dec = dt_dec; I.dt_max = dt_max; I.N_min = min_neg_steps;
  if (int r = add_integ(ctx, I, "fire")) return r;
  Integ& J = last_integ(ctx, "fire");
  FireD q{};
  q.dt = J.dt; q.alpha = J.alpha; q.alpha_init = J.alpha_init; q.f_alpha = J.f_alpha; q.F_tol = J.F_tol;
  q.E_tol = J.E_tol; q.f_inc = J.f_inc; q.f_dec = J.f_dec; q.dt_max = J.dt_max; q.old_energy = 1e15;
  q.N_min = J.N_min;
  q.ndof3 = 3 * group_size(ctx, J.gbit);
  J.fire_ndof3 = q.ndof3;
  HIPCHK(hipMalloc(&J.fire, sizeof(FireD)));
  HIPCHK(hipMemcpy(J.fire, &q, sizeof(FireD), hipMemcpyHostToDevice));
  return 0;
}

int samos_integrator_nve(samos_ctx* ctx, double limit, double angle_limit) {
  Integ I;
  I.limit = limit < 0.0 ? -1.0 : limit;
  I.theta_limit = angle_limit < 0.0 ? -1.0 : angle_limit;
  return add_integ(ctx, I, "nve");
}

// Rebuilds the device table of further pair potentials from the by-name map
// (std::map order = the reference's registry order).
void upload_extras(samos_ctx* ctx) {
  ExtraD& ex = ctx->extra;
  ex = ExtraD{};
  for (auto& kv : ctx->extras) {
    const int e = ex.count++;
    ex.kind[e] = kv.second.first;
    for (int i = 0; i < 8; ++i) ex.q[e][i] = kv.second.second[i];
    ex.radii[e] = ctx->extra_flags[kv.first].first;
    ex.shifted[e] = ctx->extra_flags[kv.first].second;
  }
}

int samos_pair_extra(samos_ctx* ctx, int kind, const double* params, int nparams, int use_radii, int shifted) {
  static const char* names[] = {"", "", "morse", "gaussian", "yukawa", "soft_attractive"};
  static const int need[] = {0, 0, 4, 7, 3, 3};
  if (kind < SAMOS_PAIR_MORSE || kind > SAMOS_PAIR_SOFT_ATTRACTIVE) return fail(ctx, "unknown pair potential kind");
  if (nparams != need[kind]) return fail(ctx, std::string("wrong parameter count for ") + names[kind]);
  std::vector<double> q(8, 0.0);
  for (int i = 0; i < nparams; ++i) q[i] = params[i];
  if (!ctx->extras.count(names[kind]) && ctx->extras.size() >= (size_t)kMaxExtra)
    return fail(ctx, "too many further pair potentials");
  ctx->extras[names[kind]] = {kind, q};
  ctx->extra_flags[names[kind]] = {use_radii ? 1 : 0, shifted ? 1 : 0};
  upload_extras(ctx);
  return 0;
}

int samos_pair_extra_energy(samos_ctx* ctx, double* energy) {
  *energy = 0.0;
  if (ctx->extra.count == 0 || !ctx->has_nlist) return 0;
  {
    ProfScope ps(ctx, "interact");
    launch_interact_extra(ctx->st, fields(ctx), ctx->nl, ctx->box, ctx->extra, ctx->n, ctx->d_en + 4L * ctx->np, 0);
  }
  launch_reduce_sum(ctx->st, ctx->d_en + 4L * ctx->np, 1, ctx->np, ctx->n, ctx->d_partial, ctx->d_out);
  HIPCHK(hipMemcpyAsync(ctx->h_out, ctx->d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  *energy = 0.5 * ctx->h_out[0];
  return 0;
}

int samos_fire_state(samos_ctx* ctx, double* out) {
  auto it = ctx->integs.begin();
  while (it != ctx->integs.end() && it->second.name != "fire") ++it;
  if (it == ctx->integs.end()) return fail(ctx, "no FIRE integrator");
  FireD q{};
  HIPCHK(hipMemcpyAsync(&q, it->second.fire, sizeof(FireD), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  out[0] = q.dt; out[1] = q.alpha; out[2] = q.last[0]; out[3] = q.last[1]; out[4] = q.last[2];
  out[5] = q.last[3]; out[6] = q.converged ? 1.0 : 0.0; out[7] = q.last_neg;
  return 0;
}

int samos_compute_forces(samos_ctx* ctx, double* energies) {
  loop_head_forces(ctx, energies ? ctx->d_en : nullptr);
  HIPCHK(hipGetLastError());
  if (energies) return energies_to_host(ctx, energies);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_pair_compute(samos_ctx* ctx, int kind, double* energy) {
  int pm = kind == SAMOS_PAIR_SOFT ? POT_SOFT : POT_LJ;
  if ((pm == POT_SOFT && !ctx->has_soft) || (pm == POT_LJ && !ctx->has_lj)) return fail(ctx, "pair potential not defined");
  if (!ctx->has_nlist) return fail(ctx, "pair potential needs a neighbour list");
  const bool sp = ctx->has_sp;
  ctx->has_sp = false;
  {
    ProfScope ps(ctx, "interact");
    launch_interact(ctx->st, pm, 0, 1, fields(ctx), ctx->nl, ctx->box, interact_params(ctx), ctx->n,
                    energy ? ctx->d_en : nullptr);
  }
  ctx->has_sp = sp;
  HIPCHK(hipGetLastError());
  if (energy) {
    double e4[4];
    if (int r = energies_to_host(ctx, e4)) return r;
    *energy = pm == POT_SOFT ? e4[0] : e4[1];
    return 0;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_align_compute(samos_ctx* ctx, int kind, double* energy) {
  int am = kind == SAMOS_ALIGN_POLAR ? AL_POLAR : AL_NEMATIC;
  if ((am == AL_POLAR && !ctx->has_pol) || (am == AL_NEMATIC && !ctx->has_nem)) return fail(ctx, "aligner not defined");
  if (!ctx->has_nlist) return fail(ctx, "aligner needs a neighbour list");
  const bool sp = ctx->has_sp;
  ctx->has_sp = false;
  {
    ProfScope ps(ctx, "interact");
    launch_interact(ctx->st, 0, am, 1, fields(ctx), ctx->nl, ctx->box, interact_params(ctx), ctx->n,
                    energy ? ctx->d_en : nullptr);
  }
  ctx->has_sp = sp;
  HIPCHK(hipGetLastError());
  if (energy) {
    double e4[4];
    if (int r = energies_to_host(ctx, e4)) return r;
    *energy = am == AL_POLAR ? e4[2] : e4[3];
    return 0;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_integrate(samos_ctx* ctx) {
  if (ctx->con.kind < 0) return fail(ctx, "Constraint has not been defined. Please define it using \"constraint\" command before running simulation.");
  if (int r = sync_fire_groups(ctx)) return r;
  for (auto& kv : ctx->integs)
    if (int r = integrate_one(ctx, kv.second, 0)) return r;
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

}  // extern "C"

namespace {

// ConstraintSphere::rescale (src/constraints/constraint_sphere.cpp:87-104): due
// at loop heads whose time step is a multiple of rescale_freq below rescale_steps.
bool rescale_due(const samos_ctx* ctx, int time_step) {
  return ctx->con.kind == 2 && ctx->rescale != 1.0 && time_step % ctx->rescale_freq == 0 &&
         time_step < ctx->rescale_steps;
}

// The radius grows by its factor, every particle is enforced on the new sphere,
// and the neighbour list is rebuilt (src/samos.cpp:882-886).
int apply_rescale(samos_ctx* ctx) {
  ctx->con.r *= ctx->rescale_scale;
  launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
  HIPCHK(hipGetLastError());
  if (ctx->has_nlist) return build_sync(ctx);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int run_range_core(samos_ctx* ctx, int t_begin, int t_end, int steps, int flags) {
  if (int r = sync_fire_groups(ctx)) return r;
  const bool f32 = (flags & SAMOS_RUN_FP32_FORCE) != 0;
  // rebuild requests of both step parities and the overflow flag
  clear_flags(ctx, 0, 3);
  // Fused schedule: one Langevin BAOAB integrator, pair potentials only.  The
  // reference's loop-head force evaluation repeats the evaluation that ended
  // the previous step on the same positions, so it is done once.
  Integ* only = ctx->integs.size() == 1 ? &ctx->integs.begin()->second : nullptr;
  const bool fused = !(flags & SAMOS_RUN_FAITHFUL) && ctx->extra.count == 0 && only && only->name == "langevin" &&
                     only->method == 0 && al_mask(ctx) == 0 && pot_mask(ctx) != 0 && needs_nlist(ctx) &&
                     fused_size_ok(ctx);
  const bool fused_b = !fused && fused_brownian_eligible(ctx, flags);
  bool packed = false;
  if (t_begin == 0 && !ctx->run_begun) {
    // Loop head of t = 0 (src/samos.cpp:887-892): enforce, reset, forces and
    // torques.  The pre-run evaluation before it (:872-875) is overwritten by
    // this reset and leaves no trace, so it is not repeated here.  The fused
    // schedules evaluate the loop-head forces of every step they integrate,
    // so they need the evaluation only when the range stops at t = 0.
    Fields f = fields(ctx);
    // the fused planar Langevin step reads positions from the interleaved
    // column, which this enforce writes as it goes
    if (fused && t_end > t_begin && ctx->con.kind == 1 && ctx->nl.use_cell) {
      f.xy = ctx->xy[1];
      packed = true;
    }
    launch_enforce(ctx->st, f, ctx->box, ctx->con, ctx->n);
    if (!(fused || fused_b) || t_end == t_begin) loop_head_forces(ctx, nullptr, f32);
    ctx->run_begun = true;
  }
  if (t_end == t_begin) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
  }
  if (t_end == steps + 1) ctx->run_begun = false;
  if (fused) {
    // every rebuild of the fused schedule is checked for overflow as the run
    // goes (and a truncated list is rebuilt before any step reads it)
    if (int r = run_fused_langevin(ctx, *only, t_begin, t_end, steps, flags, packed)) return r;
    ctx->time_step += t_end - t_begin;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return 0;
  }
  if (fused_b) {
    if (int r = run_fused_brownian(ctx, t_begin, t_end, steps, flags)) return r;
    ctx->time_step += t_end - t_begin;
    return check_overflow(ctx);
  }
  int t = t_begin;
  // The step sequence is launch-only (FIRE's scalar update included): chunks of
  // K iterations are captured once, with the noise call and run step read from
  // the device clock plus the iteration's offset, and replayed; K is the
  // largest power of two up to 32 that fits what is left.
  const bool graphable = !(flags & SAMOS_RUN_NO_GRAPH) && !ctx->prof && ctx->coop == 0;
  if (graphable && needs_nlist(ctx)) {
    // Captured iterations cannot stop at a list overflow: keep headroom (a
    // quarter of the capacity) before replaying them.
    int h[8];
    if (int r = read_flags(ctx, h)) return r;
    if (4 * h[5] > 3 * ctx->nl.stride) {
      if (int r = alloc_nbr(ctx, (int)std::ceil(h[5] * 1.5) + 16)) return r;
      Fields at = fields(ctx);
      at.x = ctx->nl.ox; at.y = ctx->nl.oy; at.z = ctx->nl.oz;
      const int builds = h[4];
      if (int r = build_sync(ctx, &at)) return r;  // the same list, from the last build's snapshot
      HIPCHK(hipMemcpyAsync(ctx->d_flags + 4, &builds, sizeof(int), hipMemcpyHostToDevice, ctx->st));
    }
  }
  if (graphable) {
    ctx->age_checked = false;
    // every captured iteration ends with the next loop head: t + K <= steps
    for (int K = 32; K >= 4; K >>= 1) {
      if (!(t_end - t >= K && t + K <= steps)) continue;
      if (int r = upload_clock(ctx, t)) return r;
      KeyBuf key;
      key << 'G' << K << f32 << pot_mask(ctx) << al_mask(ctx);
      key_context(key, ctx);
      for (auto& kv : ctx->integs) key_integ(key, kv.second);
      hipGraphExec_t ge = nullptr;
      auto record = [&]() -> int {
        for (int j = 0; j < K; ++j) {
          for (auto& kv : ctx->integs)
            if (int r = integrate_one(ctx, kv.second, 0, j)) return r;
          check_and_rebuild(ctx);
          launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
          loop_head_forces(ctx, nullptr, f32);
        }
        launch_clock_advance(ctx->st, ctx->d_clock, K);
        return 0;
      };
      if (int r = cached_graph(ctx, key.s, record, &ge)) return r;
      while (t_end - t >= K && t + K <= steps) {
        HIPCHK(hipGraphLaunch(ge, ctx->st));
        t += K;
        ctx->time_step += K;
        for (auto& kv : ctx->integs) kv.second.calls += K;
        int h[8];
        if (int r = read_flags(ctx, h)) return r;
        if (h[2]) {
          // Iterations after the overflowed rebuild read a truncated list:
          // stop here, with the capacity grown for a rerun.
          alloc_nbr(ctx, (int)std::ceil(h[5] * 1.5) + 16);
          build_sync(ctx);
          return fail(ctx, "Neighbour list overflow inside the captured step sequence ending at run iteration " +
                               std::to_string(t) + " (a particle had " + std::to_string(h[5]) +
                               " neighbours); the capacity has grown, rerun from a saved state or use "
                               "SAMOS_RUN_NO_GRAPH, whose steps recover from an overflow.");
        }
      }
    }
  }
  for (; t < t_end; ++t) {
    for (auto& kv : ctx->integs)
      if (int r = integrate_one(ctx, kv.second, t)) return r;
    check_and_rebuild(ctx);
    // the loop-head forces below read the list this rebuild may have overflowed
    if (needs_nlist(ctx))
      if (int r = recover_overflow(ctx, fields(ctx), false); r < 0) return r;
    ctx->time_step++;
    if (t + 1 <= steps) {  // loop head of t + 1: enforce, reset, forces, torques
      launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
      loop_head_forces(ctx, nullptr, f32);
    }
  }
  HIPCHK(hipGetLastError());
  return check_overflow(ctx);
}

}  // namespace

extern "C" {

int samos_run_range(samos_ctx* ctx, int t_begin, int t_end, int steps, int flags) {
  if (ctx->con.kind < 0)
    return fail(ctx, "Constraint has not been defined. Please define it using \"constraint\" command before running simulation.");
  if (ctx->integs.empty())
    return fail(ctx, "Integrator has not been defined. Please define it using \"integrator\" command before running simulation.");
  if (steps < 0 || t_begin < 0 || t_end > steps + 1 || t_begin > t_end || (t_begin == t_end && t_begin != 0))
    return fail(ctx, "invalid run range");
  if (t_begin > 0 && !ctx->run_begun) return fail(ctx, "run range does not continue a started run");
  const int ts0 = ctx->time_step - t_begin;  // time step of the run's iteration 0
  bool rescaling = false;
  for (int t = t_begin; t <= std::min(t_end, steps) && !rescaling; ++t) rescaling = rescale_due(ctx, ts0 + t);
  if (!rescaling) return run_range_core(ctx, t_begin, t_end, steps, flags);
  // The range is split at the loop heads where the sphere rescales: each piece
  // ends in the loop-head state of its last iteration, the rescale follows
  // there, and the loop-head enforce and forces are redone on the new sphere.
  // The radius changes between pieces, so they launch step by step.
  flags |= SAMOS_RUN_NO_GRAPH;
  if (t_begin == 0 && !ctx->run_begun && rescale_due(ctx, ts0))
    if (int r = apply_rescale(ctx)) return r;
  int a = t_begin;
  if (a == t_end) return run_range_core(ctx, a, t_end, steps, flags);
  while (a < t_end) {
    int r = -1;
    for (int t = a + 1; t <= t_end && t <= steps && r < 0; ++t)
      if (rescale_due(ctx, ts0 + t)) r = t;
    if (r < 0) return run_range_core(ctx, a, t_end, steps, flags);
    if (int rc = run_range_core(ctx, a, r, steps, flags)) return rc;
    if (int rc = apply_rescale(ctx)) return rc;
    launch_enforce(ctx->st, fields(ctx), ctx->box, ctx->con, ctx->n);
    loop_head_forces(ctx, nullptr, (flags & SAMOS_RUN_FP32_FORCE) != 0);
    HIPCHK(hipGetLastError());
    a = r;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_constraint_rescale(samos_ctx* ctx, double rescale, int rescale_steps, int rescale_freq) {
  if (ctx->con.kind != 2) return fail(ctx, "only the sphere constraint rescales");
  if (rescale_steps <= 0 || rescale_freq <= 0) return fail(ctx, "rescale_steps and rescale_freq must be positive");
  ctx->rescale = rescale;
  ctx->rescale_steps = rescale_steps;
  ctx->rescale_freq = rescale_freq;
  // Constraint constructor (src/constraints/constraint.hpp:103)
  ctx->rescale_scale = std::pow(rescale, static_cast<double>(rescale_freq) / static_cast<double>(rescale_steps));
  return 0;
}

int samos_constraint_radius(samos_ctx* ctx, double* r) {
  *r = ctx->con.r;
  return 0;
}

int samos_integrator_disable(samos_ctx* ctx, const char* type, const char* group) {
  const std::string key = std::string(type) + "_" + group;
  auto it = ctx->integs.find(key);
  if (it == ctx->integs.end())
    return fail(ctx, std::string("Integrator ") + type + " (for group " + group + ") does not exist.");
  hipFree(it->second.R);
  hipFree(it->second.fire);
  ctx->integs.erase(it);
  return 0;
}

int samos_group_define(samos_ctx* ctx, const char* name, const int* member) {
  const std::string nm = name ? name : "";
  if (nm.empty() || nm == "all") return fail(ctx, "group all always holds every particle");
  if (nm == "boundary" || nm == "tissue")
    return fail(ctx, "Groups \"tissue\" and \"boundary\" are reserved and cannot be used.");
  auto it = ctx->group_index.find(nm);
  int idx;
  if (it == ctx->group_index.end()) {
    idx = (int)ctx->group_index.size();
    if (idx >= 32) return fail(ctx, "too many particle groups (at most 31 besides all)");
    ctx->group_index[nm] = idx;
  } else {
    idx = it->second;
  }
  const int n = ctx->n;
  for (int i = 0; i < n; ++i)
    if (member[i]) ctx->h_groups[i] |= 1u << idx;
  std::vector<int> tmp(n);
  for (int s = 0; s < n; ++s) tmp[s] = (int)ctx->h_groups[ctx->h_id[s]];
  if (n) HIPCHK(hipMemcpyAsync(ctx->icol[6], tmp.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return 0;
}

int samos_group_size(samos_ctx* ctx, const char* name, int* size) {
  auto it = ctx->group_index.find(name ? name : "");
  if (it == ctx->group_index.end()) return fail(ctx, std::string("Group ") + (name ? name : "") + " does not exist.");
  *size = group_size(ctx, 1u << it->second);
  return 0;
}

int samos_set_integrator_group(samos_ctx* ctx, const char* group) {
  const std::string g = group ? group : "all";
  if (!ctx->group_index.count(g)) return fail(ctx, "Group " + g + " does not exist.");
  ctx->next_group = g;
  return 0;
}

int samos_integrator_remove(samos_ctx* ctx, const char* type) {
  auto it = ctx->integs.find(std::string(type) + "_all");
  if (it == ctx->integs.end()) return fail(ctx, std::string("Integrator ") + type + " (for group all) does not exist.");
  hipFree(it->second.R);
  hipFree(it->second.fire);
  ctx->integs.erase(it);
  return 0;
}

int samos_run(samos_ctx* ctx, int steps, int flags) { return samos_run_range(ctx, 0, steps + 1, steps, flags); }

int samos_energies(samos_ctx* ctx, double* energies) {
  interact(ctx, pot_mask(ctx), al_mask(ctx), 2, ctx->d_en);
  HIPCHK(hipGetLastError());
  return energies_to_host(ctx, energies);
}

}  // extern "C"

namespace {

__global__ void k_observe(Fields f, int n, int what, double* col, long ld, unsigned long long* maxbits) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (what == 0) {
    // |F| >= 0 (or NaN), so the largest bit pattern is the largest value and a
    // NaN wins as it would in a scalar max; one atomic per wavefront
    unsigned long long u = 0;
    if (s < n) u = (unsigned long long)__double_as_longlong(sqrt(f.fx[s] * f.fx[s] + f.fy[s] * f.fy[s] + f.fz[s] * f.fz[s]));
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long t = __shfl_xor(u, o, 64);
      u = t > u ? t : u;
    }
    if ((threadIdx.x & 63) == 0 && u != 0) atomicMax(maxbits, u);
    return;
  }
  if (s >= n) return;
  switch (what) {
    case 1: col[s] = f.fx[s] * f.fx[s] + f.fy[s] * f.fy[s] + f.fz[s] * f.fz[s]; break;
    // LogKineticEng sums 0.5 v^2 / m (src/log/log_kinetic_eng.hpp:58-64)
    case 2: col[s] = 0.5 * (f.vx[s] * f.vx[s] + f.vy[s] * f.vy[s] + f.vz[s] * f.vz[s]) / f.mass[s]; break;
    case 3: col[s] = f.vx[s]; col[ld + s] = f.vy[s]; col[2 * ld + s] = f.vz[s]; break;
    case 4: col[s] = sqrt(f.vx[s] * f.vx[s] + f.vy[s] * f.vy[s] + f.vz[s] * f.vz[s]); break;
    case 5: col[s] = f.nx[s]; col[ld + s] = f.ny[s]; col[2 * ld + s] = f.nz[s]; break;
  }
}

}  // namespace

extern "C" {

int samos_reduce(samos_ctx* ctx, int what, double* out) {
  if (what < 0 || what > 5) return fail(ctx, "unknown observable");
  const int n = ctx->n;
  if (n == 0) { *out = 0.0; return 0; }
  HIPCHK(hipMemsetAsync(ctx->d_counter, 0, sizeof(unsigned long long), ctx->st));
  hipLaunchKernelGGL(k_observe, dim3(grid1(n)), dim3(kBlock), 0, ctx->st, fields(ctx), n, what, ctx->d_cols,
                     (long)ctx->np, ctx->d_counter);
  HIPCHK(hipGetLastError());
  if (what == 0) {
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, ctx->d_counter, sizeof(bits), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *out = v;
    return 0;
  }
  const int nc = (what == 3 || what == 5) ? 3 : 1;
  launch_reduce_sum(ctx->st, ctx->d_cols, nc, ctx->np, n, ctx->d_partial, ctx->d_out);
  HIPCHK(hipMemcpyAsync(ctx->h_out, ctx->d_out, nc * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (nc == 3) {
    const double* v = ctx->h_out;
    *out = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / n;
  } else if (what == 4) {
    *out = ctx->h_out[0] / n;
  } else {
    *out = ctx->h_out[0];
  }
  return 0;
}

int samos_profile(samos_ctx* ctx, int enable) {
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->prof = enable != 0;
  for (auto& p : ctx->profs) p.second.used = 0;
  return 0;
}

int samos_profile_read(samos_ctx* ctx, const char* cls, double* ms_total, long long* launches) {
  HIPCHK(hipStreamSynchronize(ctx->st));
  *ms_total = 0.0;
  *launches = 0;
  auto it = ctx->profs.find(cls);
  if (it == ctx->profs.end()) return 0;
  double tot = 0.0;
  for (size_t k = 0; k < it->second.used; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, it->second.ev[k].first, it->second.ev[k].second));
    tot += ms;
  }
  *ms_total = tot;
  *launches = (long long)it->second.used;
  return 0;
}

int samos_run_stats(samos_ctx* ctx, long long* out) {
  out[0] = ctx->graph_stats[0];
  out[1] = ctx->graph_stats[1];
  out[2] = ctx->overflow_recoveries;
  out[3] = ctx->nl.stride;
  out[4] = ctx->resorts;
  return 0;
}

int samos_sync(samos_ctx* ctx) {
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipGetLastError());
  return 0;
}

int samos_tile_info(samos_ctx* ctx, long long* info) {
  for (int k = 0; k < 4; ++k) info[k] = ctx->tile_stats[k];
  info[4] = ctx->tile_mem ? ctx->td.T : 0;
  info[5] = ctx->tile_mem ? ctx->td.K : 0;
  return 0;
}

int samos_set_option(samos_ctx* ctx, const char* name, double value) {
  const std::string k = name ? name : "";
  if (k == "rebuild_coop") {
    ctx->coop = value != 0.0 ? -1 : 0;
  } else if (k == "gather_batch") {
    if (value != 4.0 && value != 8.0) return fail(ctx, "gather_batch must be 4 or 8");
    ctx->gather_batch = (int)value;
  } else if (k == "chunk_slots") {
    if (value < 2.0 || value > 256.0 || value != std::floor(value) || ((int)value & 1))
      return fail(ctx, "chunk_slots must be an even integer in 2..256");
    ctx->chunk_slots = (int)value;
  } else if (k == "neighbour_capacity") {
    // capacity per slot: max(value, the current largest count); the list is
    // rebuilt from the last build's snapshot, so it holds the same pairs
    if (!ctx->has_nlist) return fail(ctx, "no neighbour list");
    if (value < 0.0 || value != std::floor(value)) return fail(ctx, "neighbour_capacity must be a non-negative integer");
    int h[8];
    if (int r = read_flags(ctx, h)) return r;
    if (int r = alloc_nbr(ctx, std::max((int)value, h[5]))) return r;
    Fields at = fields(ctx);
    at.x = ctx->nl.ox; at.y = ctx->nl.oy; at.z = ctx->nl.oz;
    HIPCHK(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->st));
    launch_rebuild(ctx->st, at, ctx->nl, ctx->box, ctx->n, nullptr);
    HIPCHK(hipMemcpyAsync(ctx->d_flags + 4, &h[4], sizeof(int), hipMemcpyHostToDevice, ctx->st));
    if (int r = read_flags(ctx, h)) return r;
    if (h[2]) return fail(ctx, "neighbour_capacity below the current largest count");
  } else if (k == "resort_every") {
    if (value < 0.0 || value != std::floor(value)) return fail(ctx, "resort_every must be a non-negative integer");
    ctx->resort_every = (int)value;
  } else if (k == "brownian_lanes") {
    const int v = (int)value;
    if (value != v || !(v == 0 || v == 1 || v == 2 || v == 4 || v == 8 || v == 16))
      return fail(ctx, "brownian_lanes must be 0 (automatic), 1, 2, 4, 8 or 16");
    ctx->brownian_lanes = v;
  } else if (k == "tile_cells") {
    if (value < 0.0 || value > kTileMaxCells || value != std::floor(value))
      return fail(ctx, "tile_cells must be 0 (automatic) or an integer in 1.." + std::to_string(kTileMaxCells));
    ctx->tile_cells = (int)value;
  } else {
    return fail(ctx, "unknown option " + k);
  }
  return 0;
}

}  // extern "C"
