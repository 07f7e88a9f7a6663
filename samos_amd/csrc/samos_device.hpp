// Device-side data model and shared helpers for the SAMoS particle pipeline on
// CDNA4 (gfx950).  All particle state lives in HBM as structure-of-arrays FP64
// columns in "slot" order: slots follow a space-filling curve over the cell
// grid (Morton / Z-order in 3-D, Hilbert per cell layer on a plane), so
// particles that are close in space are close in memory.  The
// original particle id of each slot is kept in `id`, so every export to the
// host is in the reference's id order.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace samos {

constexpr int kBlock = 256;          // threads per workgroup for per-particle kernels
constexpr int kMaxTypes = 16;        // particle types supported by the per-type tables

// Box centred at the origin (src/system/box.hpp:56-62).
struct BoxD {
  double Lx, Ly, Lz, xlo, xhi, ylo, yhi, zlo, zhi;
  int periodic;
};

// Single-wrap minimum image (src/system/system.cpp:1083-1095).
__device__ __forceinline__ void min_image(const BoxD& b, double& dx, double& dy, double& dz) {
  if (b.periodic) {
    if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
    if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
    if (dz > b.zhi) dz -= b.Lz; else if (dz < b.zlo) dz += b.Lz;
  }
}

// sig2 / r_sq for the LJ force factor without the FP64 division sequence:
// the hardware reciprocal refined by two Newton-Raphson steps (each doubles
// the correct bits; the result is within one rounding of 1 / r_sq), then one
// multiply.  Differs from the correctly rounded quotient by at most about one
// unit in the last place.
__device__ __forceinline__ double lj_inv_r_sq(double sig2, double r_sq) {
  double y = __builtin_amdgcn_rcp(r_sq);
  double e = fma(-r_sq, y, 1.0);
  y = fma(y, e, y);
  e = fma(-r_sq, y, 1.0);
  y = fma(y, e, y);
  return sig2 * y;
}

// Workgroup b of a launch runs on XCD b % 8, and each XCD has its own L2.
// This maps the launch's workgroups so that each XCD takes one contiguous
// range of logical blocks (a bijection on [0, gridDim.x)): slots that are
// neighbours in memory, and mostly in space, then share one L2.
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x, q = G >> 3, r = G & 7, x = blockIdx.x & 7;
  return x * q + min(x, r) + (blockIdx.x >> 3);
}

// Device-scope atomics on one address from thousands of wavefronts are
// serialised where the address lives.  These flags only grow within a launch
// (a running maximum; a 0 -> 1 request), so a plain read that already shows
// the value the atomic would leave makes the atomic a no-op and it is skipped.
// A stale read can only be smaller, which costs the atomic, never a result.
__device__ __forceinline__ void atomic_max_if_above(int* p, int v) {
  if (v > *(volatile int*)p) atomicMax(p, v);
}
__device__ __forceinline__ void atomic_set_once(int* p) {
  if (*(volatile int*)p == 0) atomicOr(p, 1);
}

// The x and y steps of min_image, for pairs with dz = 0.
__device__ __forceinline__ void min_image_xy(const BoxD& b, double& dx, double& dy) {
  if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
  if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
}

// Particle columns (slot order).  Passed by value into every kernel.
struct Fields {
  double *x, *y, *z, *vx, *vy, *vz, *nx, *ny, *nz;
  double *fx, *fy, *fz, *tx, *ty, *tz, *omega, *age, *radius, *mass;
  int *type, *ix, *iy, *iz, *coord, *id;
  int* groups;        // bit g set: the slot belongs to particle group g (bit 0: "all")
  double2* xy;        // interleaved (x, y) for neighbour gathers on planar systems; may be null
};
constexpr int kNumDoubleFields = 19;
constexpr int kNumIntFields = 7;

// Membership of slot s in the group whose bit mask is gbit (System::make_group,
// src/system/system.cpp:462-570): every particle is in "all" (bit 0), and for
// that group the column is not read.
__device__ __forceinline__ bool in_group(const Fields& f, int s, unsigned gbit) {
  return gbit == 1u || ((unsigned)f.groups[s] & gbit) != 0u;
}

// Neighbour-list state.  The list is a full (both directions) Verlet list in a
// transposed layout of groups of four: entries 4g .. 4g + 3 of slot s are the
// 16 bytes at nbr[(g * np + s) * 4] (nbr_at), so one 16-byte load brings a
// slot four entries, the g-th loads of consecutive slots are contiguous, and
// a build stores each group whole (single systems and replica ensembles).
struct NListD {
  int ncx, ncy, ncz, ncell;
  double wx, wy, wz;
  double cut2;        // (rcut + pad)^2
  double pad2q;       // 0.25 * pad^2
  int use_cell;
  int np;             // slot-padded particle count (multiple of 64)
  int stride;         // neighbour capacity per slot
  int* cell_of;       // [n] linear cell index per slot
  int* count;         // [ncell] members per cell; zero between builds
  int* start;         // [ncell + 1] exclusive scan of count
  int* members;       // [n] slots grouped by cell, ascending slot within a cell
  int* bsum;          // [scan blocks] per-block sums for the scan
  int* nbr;           // [ceil((stride + 8) / 4) * np * 4]: entry k of slot s at nbr_at(k, s)
  int* nnb;           // [np]
  double *ox, *oy, *oz;  // positions at the last build
  int* flags;         // [0],[1] rebuild request by step parity, [2] overflow, [3] largest cell
                      // count, [4] builds, [5] largest neighbour count
  double r_uni = -1.0;  // the common particle radius, or -1: the build then reads each radius
};

// Index of entry k of slot s in NListD::nbr.
__device__ __forceinline__ long nbr_at(const NListD& nl, int k, int s) {
  return (((long)(k >> 2) * nl.np + s) << 2) + (k & 3);
}

// A list being built is written four entries at a time: the group being
// filled is a shift register (the newest entry enters at w), stored whole,
// 16 bytes, when its fourth entry arrives.
__device__ __forceinline__ void group_push(uint4& qb, unsigned t) {
  qb.x = qb.y; qb.y = qb.z; qb.z = qb.w; qb.w = t;
}
__device__ __forceinline__ void group_store(const NListD& nl, int g, int s, const uint4& qb) {
  *reinterpret_cast<uint4*>(nl.nbr + nbr_at(nl, 4 * g, s)) = qb;
}
// Entry t of slot s at position k of its list (nothing past the capacity).
__device__ __forceinline__ void list_push(const NListD& nl, int s, int k, unsigned t, uint4& qb) {
  if (k < nl.stride) {
    group_push(qb, t);
    if ((k & 3) == 3) group_store(nl, k >> 2, s, qb);
  }
}
// The last, partly filled group of a list of k entries, completed with the
// own index.
__device__ __forceinline__ void list_close(const NListD& nl, int s, int k, uint4& qb) {
  const int kk = min(k, nl.stride);
  if (kk & 3) {
    for (int q = kk & 3; q < 4; ++q) group_push(qb, (unsigned)s);
    group_store(nl, kk >> 2, s, qb);
  }
}

// Pair interactions evaluated in one pass over the neighbour list.
struct InteractD {
  // soft (src/potentials/pair/pair_soft_potential.cpp:33-119): table k, a
  const double* soft_tab; int soft_nt; int soft_radii;
  // lj (src/potentials/pair/pair_lj_potential.cpp:32-124): table eps, sigma, rcut
  const double* lj_tab; int lj_nt; int lj_radii; int lj_shifted;
  // polar (src/aligner/pair/pair_polar_aligner.cpp:32-107): table J, rcut
  const double* pol_tab; int pol_nt;
  // nematic (src/aligner/pair/pair_nematic_aligner.cpp:32-96): table J (already doubled
  // when no per-type parameter was set), rcut
  const double* nem_tab; int nem_nt;
  // external self propulsion (src/potentials/external/external_self_propulsion.cpp:33-50)
  int has_sp; double sp_alpha;
};

enum PotMask { POT_SOFT = 1, POT_LJ = 2 };

// Further pair potentials with global parameters, evaluated in their own pass
// after soft / LJ, in the reference's registry-name order.
enum ExtraKind { EXTRA_MORSE = 2, EXTRA_GAUSSIAN = 3, EXTRA_YUKAWA = 4, EXTRA_SOFT_ATTRACTIVE = 5 };
constexpr int kMaxExtra = 4;
struct ExtraD {
  int count;
  int kind[kMaxExtra];
  int radii[kMaxExtra];
  int shifted[kMaxExtra];
  double q[kMaxExtra][8];
};
enum AlMask { AL_POLAR = 1, AL_NEMATIC = 2 };

// Constraint applied per particle (src/constraints/*).
struct ConstraintD {
  int kind;           // 0 none, 1 plane, 2 sphere, 3 cylinder, 4 torus, 5 ellipsoid, -1 unset
  double r;           // sphere and cylinder radius
  double zpos;        // plane height
  int unlimited;      // plane, non-periodic: no reflective walls
  double q0, q1, q2;  // torus: c (ring radius), a (tube radius); ellipsoid: a, b, c
  int max_iter;       // SHAKE projection (torus, ellipsoid): iteration cap and
  double tol;         //   tolerance on |g(x)|
};

// Temperature schedule (src/system/value.hpp:115-170).
struct TempD {
  int kind; double tmin, tmax; int steps;
  __device__ __host__ double at(int run_step) const {
    if (kind == 0) return tmin;
    if (run_step > steps) return tmax;
    return tmin + (double)run_step * ((tmax - tmin) / (double)steps);
  }
};

// Per-step clock in device memory, advanced after each captured chunk of steps
// so that replayed graphs draw fresh noise.
// Control words of a captured chunk of fused Langevin steps ("slots").  A slot
// whose new positions need a rebuild sets req[slot & 1]; the next slot sees it
// and every later slot of the chunk then does nothing.  After the slots, the
// chunk settles the positions into buffer 0, rebuilds if requested and
// advances the clock by the number of slots that ran.
struct ChunkCtl {
  int halt;      // set once a slot has seen its predecessor's request
  int req[2];    // rebuild request of the slot with this parity
  int last;      // slots executed in this chunk
  int base;      // run step of slot 0
  int stop;      // why the chunk's slots stopped early (next to base: one read-back):
                 // bit 0 the list overflowed at a rebuild, bit 1 a storage re-sort is due
  int resort_at; // builds count (NListD::flags[4]) at which slots stop for a re-sort
  int limit;     // slot j runs only while base + j < limit
  int rebuild;   // rebuild request read by the chunk's rebuild stages
  int steps;     // last iteration of the run: slot j applies the next loop
                 // head's enforce while base + j + 1 <= steps
};

// The tests a captured slot j makes before its step, from words all loaded
// up front and combined before the first branch (one memory latency instead
// of one per test), so a slot that returns at once costs little more than
// its launch.  Bits: kGateOverflow the list overflowed at a rebuild,
// kGateResort a storage re-sort is due, kGateHalt an earlier slot asked for a
// rebuild, kGateEnd the slot lies past the range; head: the slot applies the
// next loop head's enforce.
constexpr int kGateOverflow = 1, kGateResort = 2, kGateHalt = 4, kGateEnd = 8;
struct SlotGate {
  int stop;  // kGate* bits
  int head;
};
__device__ __forceinline__ SlotGate slot_gate(const ChunkCtl* ck, const int* flags, int j) {
  const int overflow = flags[2], builds = flags[4];
  const int halt = ck->halt, req = ck->req[(j + 1) & 1];
  const int base = ck->base, resort_at = ck->resort_at, limit = ck->limit, steps = ck->steps;
  SlotGate g;
  g.stop = (overflow ? kGateOverflow : 0) | (builds >= resort_at ? kGateResort : 0) | ((halt | req) ? kGateHalt : 0) |
           (base + j >= limit ? kGateEnd : 0);
  g.head = base + j + 1 <= steps ? 1 : 0;
  return g;
}

struct ClockD {
  unsigned long long calls[8];   // per-integrator call counters
  int run_step;                  // run-loop index t
  int pad_;
};

// IntegratorFIRE's scalar state (src/integrators/integrator_fire.hpp:58-154),
// kept on the device so that an iteration is launches only: the step size
// and mixing factors the particle kernels use, the reduced P, |F|, |v|, E of
// the last iteration and the adaptation counters.
struct FireD {
  double dt, alpha, alpha_init, f_alpha, F_tol, E_tol, f_inc, f_dec, dt_max, old_energy;
  double last[4];   // P, |F|, |v|, E of the last iteration
  double f1, f2;    // velocity mixing factors of the last iteration
  int N_min, last_neg, converged, zero_v, mix;  // mix: the last iteration mixed velocities
  int ndof3;        // 3 N of the integrator's group
};

// Particle ages only grow by the same dt per step for every integrated
// particle (src/integrators/integrator_langevin.cpp:115,
// src/integrators/integrator_brownian.cpp:118).  While every age is equal the
// fused steps leave the column alone and the common value is advanced once
// per range by the exact sequence of additions; `uniform` is decided on the
// device at the start of a range.
struct AgeCtl {
  int uniform;
  int pad_;
  double age;
};

// Noise call and run step of a launch: read from the device clock plus an
// offset inside captured step sequences, or given directly (clk null).
__device__ __forceinline__ unsigned long long eff_call(const ClockD* clk, int slot, unsigned long long call) {
  return clk ? clk->calls[slot] + call : call;
}
__device__ __forceinline__ int eff_step(const ClockD* clk, int t_off) { return clk ? clk->run_step + t_off : t_off; }

// ---------------------------------------------------------------------------
// Counter-based noise: Philox4x32-10 (Salmon et al., SC'11), keyed by
// (seed, integrator tag) with counter (particle id, call lo, call hi, purpose).
// ---------------------------------------------------------------------------
struct U4 { uint32_t a, b, c, d; };

__device__ __forceinline__ U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  // the round's two 32x32 -> 64-bit products, each one v_mad_u64_u32
  const uint64_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    const uint64_t p0 = M0 * ctr.a, p1 = M1 * ctr.c;
    U4 n;
    n.a = (uint32_t)(p1 >> 32) ^ ctr.b ^ k0;
    n.b = (uint32_t)p1;
    n.c = (uint32_t)(p0 >> 32) ^ ctr.d ^ k1;
    n.d = (uint32_t)p0;
    ctr = n;
  }
  return ctr;
}

__device__ __forceinline__ double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

// Natural log of u in (0, 1) (normal, non-zero), within about 1.5 ulp:
// u = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 artanh(s) with
// s = (m - 1) / (m + 1), |s| <= 0.1716, summed as the series
// 2 s (1 + s^2/3 + s^4/5 + ... + s^20/21) (Abramowitz & Stegun 4.1.27;
// the next term is below 2^-55 relative), and e ln 2 with ln 2 split into a
// 33-bit head (e ln2_hi exact) and a tail.  Every operation is IEEE (division
// included) or an explicit fma, so a C compilation of the same expression
// without contraction returns the same bits.
__device__ __forceinline__ double bm_log(double u) {
  int e;
  double m = frexp(u, &e);
  if (m < 0x1.6a09e667f3bcdp-1) { m = 2.0 * m; e -= 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 0x1.8618618618618p-5;            // 1/21
  p = fma(p, z, 0x1.af286bca1af28p-5);        // 1/19
  p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5);        // 1/17
  p = fma(p, z, 0x1.1111111111111p-4);        // 1/15
  p = fma(p, z, 0x1.3b13b13b13b14p-4);        // 1/13
  p = fma(p, z, 0x1.745d1745d1746p-4);        // 1/11
  p = fma(p, z, 0x1.c71c71c71c71cp-4);        // 1/9
  p = fma(p, z, 0x1.2492492492492p-3);        // 1/7
  p = fma(p, z, 0x1.999999999999ap-3);        // 1/5
  p = fma(p, z, 0x1.5555555555555p-2);        // 1/3
  const double s2 = 2.0 * s;
  const double lm = fma(s2, z * p, s2);
  const double de = (double)e;
  const double lo = fma(de, 0x1.473de6af278edp-34, lm);  // ln2 - ln2_hi
  return de * 0x1.62e42fef00000p-1 + lo;                // e ln2_hi is exact
}

// sin and cos of 2 pi u for u in (0, 1), within about 1.5 ulp: t = 4u,
// q = rint(t), f = t - q in [-1/2, 1/2] (both exact), then the Taylor
// polynomials of sin(pi f / 2) (to f^17) and cos(pi f / 2) (to f^16), whose
// coefficients (pi/2)^k / k! are rounded from exact values; truncation is
// below 2^-60.  The quadrant q mod 4 swaps and negates them.
__device__ __forceinline__ void bm_sincos_2pi(double u, double& sn, double& cs) {
  const double t = 4.0 * u;
  const double q = rint(t);
  const double f = t - q;
  const double z = f * f;
  double ps = 0x1.aaec32af93359p-38;
  ps = fma(ps, z, -0x1.6fadb9f155744p-31);
  ps = fma(ps, z, 0x1.e8f434d018d63p-25);
  ps = fma(ps, z, -0x1.e3074fde8871fp-19);
  ps = fma(ps, z, 0x1.50783487ee782p-13);
  ps = fma(ps, z, -0x1.32d2cce62bd86p-8);
  ps = fma(ps, z, 0x1.466bc6775aae2p-4);
  ps = fma(ps, z, -0x1.4abbce625be53p-1);
  ps = fma(ps, z, 0x1.921fb54442d18p+0);
  double pc = 0x1.20c62c2f2d7f5p-34;
  pc = fma(pc, z, -0x1.b6e24f44b128fp-28);
  pc = fma(pc, z, 0x1.f9d38a3763cc3p-22);
  pc = fma(pc, z, -0x1.a6d1f2a204a8cp-16);
  pc = fma(pc, z, 0x1.e1f506891babbp-11);
  pc = fma(pc, z, -0x1.55d3c7e3cbffap-6);
  pc = fma(pc, z, 0x1.03c1f081b5ac4p-2);
  pc = fma(pc, z, -0x1.3bd3cc9be45dep+0);
  pc = fma(pc, z, 1.0);
  const double S = f * ps, C = pc;
  const int k = (int)q & 3;
  sn = (k == 0) ? S : (k == 1) ? C : (k == 2) ? -S : -C;
  cs = (k == 0) ? C : (k == 1) ? -S : (k == 2) ? -C : S;
}

// Box-Muller pair from two 32-bit words (both uniforms in (0, 1)).
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& g0, double& g1) {
  const double u1 = u01(a), u2 = u01(b);
  const double r = sqrt(-2.0 * bm_log(u1));
  double s, c;
  bm_sincos_2pi(u2, s, c);
  g0 = r * c;
  g1 = r * s;
}

enum { TAG_BROWNIAN = 1, TAG_LANGEVIN = 2, TAG_BROWNIAN_POS = 3, TAG_BROWNIAN_ALIGN = 4, TAG_BBK_INIT = 5 };

__device__ __forceinline__ void gauss3(uint32_t seed, uint32_t tag, uint32_t pid, unsigned long long call, double g[3]) {
  U4 c{pid, (uint32_t)call, (uint32_t)(call >> 32), 0u};
  U4 r = philox4x32_10(c, seed, tag);
  double g3;
  box_muller(r.a, r.b, g[0], g[1]);
  box_muller(r.c, r.d, g[2], g3);
}

// The first two Gaussians of gauss3 (the same draw block and pair): enough for
// a planar system, whose z velocity every plane enforce resets.
__device__ __forceinline__ void gauss2(uint32_t seed, uint32_t tag, uint32_t pid, unsigned long long call, double g[2]) {
  U4 c{pid, (uint32_t)call, (uint32_t)(call >> 32), 0u};
  U4 r = philox4x32_10(c, seed, tag);
  box_muller(r.a, r.b, g[0], g[1]);
}

__device__ __forceinline__ void gauss1_u1(uint32_t seed, uint32_t tag, uint32_t pid, unsigned long long call, double& g, double& u) {
  U4 c{pid, (uint32_t)call, (uint32_t)(call >> 32), 1u};
  U4 r = philox4x32_10(c, seed, tag);
  double g1;
  box_muller(r.a, r.b, g, g1);
  u = u01(r.c) - 1.1641532182693481e-10;
}

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }
// Grid size for one thread per slot; never zero, so empty systems launch validly.
inline int grid1(long n) { int g = div_up(n, kBlock); return g < 1 ? 1 : g; }

}  // namespace samos
