// SAMoS hot-path ORACLE -- test infrastructure only.
//
// A strict FP64 CPU restatement of the reference's per-step particle pipeline.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only as a checker or as the timed CPU baseline; the product
// path (samos_amd/) never links or calls it.
//
// Built twice (see oracle/Makefile):
//   liboracle.so       -O2 -ffp-contract=off, no fast-math: the arithmetic order of
//                      every expression follows the reference source as written.
//   liboracle_fast.so  -O3 -funroll-loops -ffast-math -march=x86-64-v3: the
//                      reference's own Release flags (CMakeCFlagsSetup.txt:46),
//                      used only to time the single-threaded CPU baseline.
//
// Layout follows the reference: an array-of-structures particle record and a
// half neighbour list stored as vector<vector<int>> (src/system/system.hpp:101,
// src/system/neighbour_list.hpp).  Every function cites the reference lines it
// restates (paths relative to the reference repository root).
//
// Parity status of each part (what pins it) is stated in the test suite header
// (tests/test_oracle.py) and in the design notes:
//   * cell list / nlist build: brute-force O(N^2) pair sets + example inputs;
//   * soft / LJ / polar / nematic: closed forms, Newton's third law,
//     finite-difference gradients;
//   * constraints: invariants (|x| = R, n.N = 0, |n| = 1, rotation inverses);
//     cylinder, torus and ellipsoid (generic SHAKE) against closed-form
//     projections with tol = 0 (tests/test_surfaces.py);
//   * morse / gaussian / yukawa / soft_attractive: two-particle closed forms,
//     the rcut shift, Newton's third law, finite-difference gradients, the
//     soft_attractive coordination against a brute-force count
//     (tests/test_pair_extra.py);
//   * integrators: single-particle closed forms (free ABP drift, O-step decay,
//     FIRE two-particle relaxation, NVE free flight, limits and LJ-pair energy
//     conservation), equipartition with noise on, the linear temperature ramp;
//   * the default noise stream is the project's counter-based Philox4x32-10
//     (checked against the Random123 known-answer vectors), the one the device
//     library draws.  A second stream restates the reference's own noise
//     source (GSL MT19937 + gsl_ran_gaussian, drawn sequentially in the
//     reference's per-particle order); it is used to check noisy runs
//     statistically.  Bit-level parity with GSL draws is unpinned (no GSL here).

#include <cmath>
#include <random>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace {

// ---------------------------------------------------------------------------
// Counter-based RNG (project-defined noise stream, shared definition with the
// device library; checked against Random123 philox4x32-10 KAT vectors).
// ---------------------------------------------------------------------------
struct U4 { uint32_t v[4]; };

inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, ctr.v[0], hi0, lo0);
    mulhilo(M1, ctr.v[2], hi1, lo1);
    U4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
  }
  return ctr;
}

// Uniform in (0,1): (u + 0.5) * 2^-32.
inline double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

// Natural log of u in (0, 1): u = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 artanh(s), s = (m - 1) / (m + 1), summed as the series
// 2 s (1 + s^2/3 + ... + s^20/21) (Abramowitz & Stegun 4.1.27), plus e ln 2
// with ln 2 split into a 33-bit head and a tail.  Within about 1.5 ulp of
// log(u) (pinned against the libm log in the tests); written with explicit
// fma and no contraction so that the device's evaluation gives the same bits.
inline double bm_log(double u) {
  int e;
  double m = std::frexp(u, &e);
  if (m < 0x1.6a09e667f3bcdp-1) { m = 2.0 * m; e -= 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  static const double c[10] = {0x1.8618618618618p-5, 0x1.af286bca1af28p-5, 0x1.e1e1e1e1e1e1ep-5,
                               0x1.1111111111111p-4, 0x1.3b13b13b13b14p-4, 0x1.745d1745d1746p-4,
                               0x1.c71c71c71c71cp-4, 0x1.2492492492492p-3, 0x1.999999999999ap-3,
                               0x1.5555555555555p-2};  // 1/21, 1/19, ..., 1/3
  double p = c[0];
  for (int k = 1; k < 10; ++k) p = std::fma(p, z, c[k]);
  const double s2 = 2.0 * s;
  const double lm = std::fma(s2, z * p, s2);
  const double de = (double)e;
  const double lo = std::fma(de, 0x1.473de6af278edp-34, lm);
  return de * 0x1.62e42fef00000p-1 + lo;
}

// sin and cos of 2 pi u, u in (0, 1): t = 4u, q = rint(t), f = t - q, the
// Taylor polynomials of sin(pi f / 2) and cos(pi f / 2) in f with
// coefficients (pi/2)^k / k!, and the quadrant q mod 4.
inline void bm_sincos_2pi(double u, double& sn, double& cs) {
  const double t = 4.0 * u;
  const double q = std::rint(t);
  const double f = t - q;
  const double z = f * f;
  static const double cs_s[9] = {0x1.aaec32af93359p-38, -0x1.6fadb9f155744p-31, 0x1.e8f434d018d63p-25,
                                 -0x1.e3074fde8871fp-19, 0x1.50783487ee782p-13, -0x1.32d2cce62bd86p-8,
                                 0x1.466bc6775aae2p-4, -0x1.4abbce625be53p-1, 0x1.921fb54442d18p+0};
  static const double cs_c[9] = {0x1.20c62c2f2d7f5p-34, -0x1.b6e24f44b128fp-28, 0x1.f9d38a3763cc3p-22,
                                 -0x1.a6d1f2a204a8cp-16, 0x1.e1f506891babbp-11, -0x1.55d3c7e3cbffap-6,
                                 0x1.03c1f081b5ac4p-2, -0x1.3bd3cc9be45dep+0, 1.0};
  double ps = cs_s[0], pc = cs_c[0];
  for (int k = 1; k < 9; ++k) { ps = std::fma(ps, z, cs_s[k]); pc = std::fma(pc, z, cs_c[k]); }
  const double S = f * ps, C = pc;
  const int k = (int)q & 3;
  sn = (k == 0) ? S : (k == 1) ? C : (k == 2) ? -S : -C;
  cs = (k == 0) ? C : (k == 1) ? -S : (k == 2) ? -C : S;
}

// Box-Muller pair from two 32-bit words.
inline void box_muller(uint32_t a, uint32_t b, double& g0, double& g1) {
  double u1 = u01(a), u2 = u01(b);
  double r = std::sqrt(-2.0 * bm_log(u1));
  double s, c;
  bm_sincos_2pi(u2, s, c);
  g0 = r * c;
  g1 = r * s;
}

// Stream tags: one per integrator family, so that two integrators with the
// same seed do not share draws.
enum { TAG_BROWNIAN = 1, TAG_LANGEVIN = 2, TAG_BROWNIAN_POS = 3, TAG_BROWNIAN_ALIGN = 4, TAG_BBK_INIT = 5 };

// Draw block for (seed, tag, particle id, call counter, purpose).
U4 draw(uint32_t seed, uint32_t tag, uint32_t pid, uint64_t call, uint32_t purpose) {
  U4 c; c.v[0] = pid; c.v[1] = (uint32_t)call; c.v[2] = (uint32_t)(call >> 32); c.v[3] = purpose;
  return philox4x32_10(c, seed, tag);
}

// Three Gaussians (translational noise) for one particle and call.
void gauss3(uint32_t seed, uint32_t tag, uint32_t pid, uint64_t call, double g[3]) {
  U4 a = draw(seed, tag, pid, call, 0);
  double g3;
  box_muller(a.v[0], a.v[1], g[0], g[1]);
  box_muller(a.v[2], a.v[3], g[2], g3);
}

// One Gaussian (rotational noise) plus one uniform (nematic flip) per call.
void gauss1_u1(uint32_t seed, uint32_t tag, uint32_t pid, uint64_t call, double& g, double& u) {
  U4 a = draw(seed, tag, pid, call, 1);
  double g1;
  box_muller(a.v[0], a.v[1], g, g1);
  u = u01(a.v[2]) - 1.1641532182693481e-10;  // in [0,1)
}

// ---------------------------------------------------------------------------
// Data model (src/system/particle.hpp:257-293, src/system/box.hpp:56-62)
// ---------------------------------------------------------------------------
struct Particle {
  int id, type;
  double radius, mass;
  double x, y, z, vx, vy, vz, nx, ny, nz;
  double fx, fy, fz, tau_x, tau_y, tau_z, omega, age;
  double Nx, Ny, Nz;
  double s[9];            // stress, zeroed by reset_forces
  int ix, iy, iz, coordination, boundary;
};

struct Box {
  double Lx, Ly, Lz, xlo, xhi, ylo, yhi, zlo, zhi;
  void set(double lx, double ly, double lz) {
    Lx = lx; Ly = ly; Lz = lz;
    xlo = -0.5 * Lx; xhi = 0.5 * Lx;
    ylo = -0.5 * Ly; yhi = 0.5 * Ly;
    zlo = -0.5 * Lz; zhi = 0.5 * Lz;
  }
};

struct PartPos { double x, y, z; };

struct TypeTable {
  int ntypes = 1;
  std::vector<double> p0, p1, p2;  // ntypes*ntypes each
  void init(int nt, double a, double b, double c) {
    ntypes = nt; p0.assign(nt * nt, a); p1.assign(nt * nt, b); p2.assign(nt * nt, c);
  }
};

struct PairPot {        // soft (kind 0), lj (1), morse (2), gaussian (3), yukawa (4), soft_attractive (5)
  int kind;
  double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // kinds 2-5: the global parameters (see extra_compute)
  double k, a;                         // soft (pair_soft_potential.hpp:63-180)
  double eps, sigma, rcut;             // lj (pair_lj_potential.hpp:61-248)
  bool shifted = false, use_radii = false, has_pair_params = false;
  TypeTable tab;                       // soft: k,a ; lj: eps,sigma,rcut
  double energy = 0.0;
};

struct PairAl {         // polar (kind 0) or nematic (kind 1)
  int kind;
  double J, rcut;
  bool has_pair_params = false, exclude_boundary = false;
  TypeTable tab;                       // J, rcut
  double energy = 0.0;
};

// Reference noise source (src/utils/rng.cpp:34-61): one sequential stream per
// integrator.  GSL's default generator is MT19937; gsl_rng_set(r, 0) seeds it
// with 4357 and otherwise with the seed, using the Matsumoto-Nishimura (2002)
// initialisation, which is std::mt19937's seeding.  gsl_rng_uniform returns
// word / 2^32; gsl_rng_uniform_pos rejects 0; gsl_ran_gaussian is the
// Marsaglia polar method (two uniform_pos draws mapped to the square
// (-1,1)^2, rejected outside the unit disc or at the origin) returning
// sigma * y * sqrt(-2 ln r2 / r2), one value per accepted pair.  Written from
// those published algorithms; GSL itself is absent, so bit parity is unpinned.
struct GslStream {
  std::mt19937 mt;
  void set(uint32_t seed) { mt.seed(seed == 0 ? 4357u : seed); }
  double uniform() { return (double)mt() / 4294967296.0; }
  double uniform_pos() {
    double u;
    do { u = uniform(); } while (u == 0.0);
    return u;
  }
  double gauss(double sigma) {
    double x, y, r2;
    do {
      x = -1.0 + 2.0 * uniform_pos();
      y = -1.0 + 2.0 * uniform_pos();
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    return sigma * y * std::sqrt(-2.0 * std::log(r2) / r2);
  }
};

struct Constraint {
  int kind = -1;        // 0 none, 1 plane, 2 sphere, 3 cylinder, 4 torus, 5 ellipsoid
  double r = 1.0, zpos = 0.0;
  bool unlimited = false;
  double tc = 1.0, ta = 0.5;                 // torus: ring radius c, tube radius a
  double ea = 1.0, eb = 1.0, ec = 1.0;       // ellipsoid semi-axes
  int max_iter = 100; double tol = 1e-6;     // SHAKE iteration (constraint.hpp:62-84)
  // radius schedule of the sphere (constraint.hpp:85-108): total factor,
  // steps it is spread over, steps between rescales, factor per rescale
  double rescale = 1.0; int rescale_steps = 1000, rescale_freq = 10; double scale = 1.0;
};

struct Integ {
  std::string name;     // "brownian", "brownian_pos", "brownian_align", "langevin", "fire"
  std::string group = "all";  // Integrator::m_group_name (src/integrators/integrator.hpp:98-108)
  double dt = 0.0;
  // temperature control (src/system/value.hpp:115-170): kind 0 constant, 1 linear
  int tkind = 0; double tmin = 0.0, tmax = 1.0; int tsteps = 1000;
  // brownian family (integrator_brownian.hpp:55-160)
  double v0 = 1.0, nu = 1.0, mu = 1.0, mur = 1.0, tau = 0.0, stoch = 0.0;
  bool nematic = false, velocity = false;
  uint32_t seed = 0; uint64_t calls = 0;
  int stream = 0;       // 0: counter-based Philox (device stream); 1: GslStream (reference order)
  GslStream gsl;
  // langevin (integrator_langevin.hpp:64-121)
  double gamma = 1.0; int method = 0;  // 0 baoab, 1 spv, 2 bbk
  std::vector<double> Rx, Ry, Rz;
  // fire (integrator_fire.hpp:58-154)
  double alpha = 0.1, alpha_init = 0.1, f_alpha = 0.99, F_tol = 1e-4, E_tol = 1e-6,
         f_inc = 1.1, f_dec = 0.5, dt_max = 1.0;
  int N_min = 5, last_neg = 0; bool converged = false; double old_energy = 1e15;
  // nve (integrator_nve.hpp:59-86)
  bool has_limit = false, has_theta_limit = false; double limit = 0.0, theta_limit = 0.0;
  double last_P = 0, last_Fnorm = 0, last_Vnorm = 0, last_E = 0;
  double Tval(int step) const {
    if (tkind == 0) return tmin;
    if (step > tsteps) return tmax;
    return tmin + (double)step * ((tmax - tmin) / (double)tsteps);
  }
};

struct Sim {
  std::vector<Particle> p;
  Box box;
  bool periodic = true;
  double dt = 0.0;
  int time_step = 0, run_step = 0;
  // neighbour list (src/system/neighbour_list.hpp:124-217)
  bool has_nlist = false;
  double cut = 3.0, pad = 0.5;
  bool use_cell = false;
  int cnx = 0, cny = 0, cnz = 0; double cwx = 0, cwy = 0, cwz = 0;
  std::vector<std::vector<int>> stencil, cells;
  std::vector<std::vector<int>> list;
  std::vector<PartPos> old;
  long builds = 0;
  // plugins; std::map order fixes evaluation order (src/potentials/potential.cpp:36-52)
  std::map<std::string, PairPot> pots;
  std::map<std::string, PairAl> als;
  std::vector<double> self_prop;   // external self_propulsion alpha (empty: none)
  bool has_self_prop = false; double sp_alpha = 1.0;
  Constraint con;
  std::map<std::string, Integ> integs;   // key: type + "_" + group (src/samos.cpp:666-690)
  // Particle groups (System::make_group, src/system/system.cpp:462-570):
  // member indices in the order the reference adds them; "all" holds everyone.
  std::map<std::string, std::vector<int>> groups;
  std::string next_group = "all";
  int ntypes = 1;
  int rng_stream = 0;   // noise stream of integrators added from now on
};

// System::apply_periodic (src/system/system.cpp:1083-1095)
inline void apply_periodic(const Sim& s, double& dx, double& dy, double& dz) {
  if (s.periodic) {
    const Box& b = s.box;
    if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
    if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
    if (dz > b.zhi) dz -= b.Lz; else if (dz < b.zlo) dz += b.Lz;
  }
}

// System::enforce_periodic (src/system/system.cpp:1119-1155)
inline void enforce_periodic(Sim& s, Particle& p) {
  if (s.periodic) {
    const Box& b = s.box;
    if (p.x <= b.xlo) { p.x += b.Lx; p.ix--; } else if (p.x > b.xhi) { p.x -= b.Lx; p.ix++; }
    if (p.y <= b.ylo) { p.y += b.Ly; p.iy--; } else if (p.y > b.yhi) { p.y -= b.Ly; p.iy++; }
    if (p.z <= b.zlo) { p.z += b.Lz; p.iz--; } else if (p.z > b.zhi) { p.z -= b.Lz; p.iz++; }
  }
}

// System::reset_forces / reset_torques (src/system/system.hpp:136-157)
void reset_forces(Sim& s) {
  for (auto& q : s.p) { q.fx = q.fy = q.fz = 0.0; for (double& v : q.s) v = 0.0; }
}
void reset_torques(Sim& s) {
  for (auto& q : s.p) { q.tau_x = q.tau_y = q.tau_z = 0.0; }
}

// Group::get_particles of an integrator's group.
const std::vector<int>& members(const Sim& s, const Integ& I) { return s.groups.at(I.group); }

// ---------------------------------------------------------------------------
// Cell list (src/system/cell_list.cpp:37-104)
// ---------------------------------------------------------------------------
void cell_list_create(Sim& s, double cutoff) {
  s.cnx = (int)(s.box.Lx / cutoff);
  s.cny = (int)(s.box.Ly / cutoff);
  s.cnz = (int)(s.box.Lz / cutoff);
  int size = s.cnx * s.cny * s.cnz;
  s.cwx = s.box.Lx / s.cnx; s.cwy = s.box.Ly / s.cny; s.cwz = s.box.Lz / s.cnz;
  s.stencil.assign(size, {});
  s.cells.assign(size, {});
  for (int i = 0; i < s.cnx; i++)
    for (int j = 0; j < s.cny; j++)
      for (int k = 0; k < s.cnz; k++) {
        int idx = s.cny * s.cnz * i + s.cnz * j + k;
        for (int ix = -1; ix <= 1; ix++)
          for (int iy = -1; iy <= 1; iy++)
            for (int iz = -1; iz <= 1; iz++) {
              int iix = i + ix, iiy = j + iy, iiz = k + iz;
              if (iix < 0) iix = s.cnx - 1; else if (iix == s.cnx) iix = 0;
              if (iiy < 0) iiy = s.cny - 1; else if (iiy == s.cny) iiy = 0;
              if (iiz < 0) iiz = s.cnz - 1; else if (iiz == s.cnz) iiz = 0;
              s.stencil[idx].push_back(s.cny * s.cnz * iix + s.cnz * iiy + iiz);
            }
      }
}

// CellList::get_cell_idx (src/system/cell_list.cpp:72-79).  The reference does
// not clamp; a particle exactly on the high face indexes one past the grid.
// Clamping to the last cell is the defined behaviour here (and on the device).
int cell_idx(const Sim& s, const Particle& p) {
  int i = (int)((p.x - s.box.xlo) / s.cwx);
  int j = (int)((p.y - s.box.ylo) / s.cwy);
  int k = (int)((p.z - s.box.zlo) / s.cwz);
  if (i >= s.cnx) i = s.cnx - 1; if (i < 0) i = 0;
  if (j >= s.cny) j = s.cny - 1; if (j < 0) j = 0;
  if (k >= s.cnz) k = s.cnz - 1; if (k < 0) k = 0;
  return s.cny * s.cnz * i + s.cnz * j + k;
}

// CellList::populate (src/system/cell_list.cpp:82-104)
void cell_populate(Sim& s) {
  for (auto& c : s.cells) c.clear();
  const Box& b = s.box;
  for (int i = 0; i < (int)s.p.size(); i++) {
    Particle& q = s.p[i];
    if (s.periodic) {
      if (q.x < b.xlo) q.x += b.Lx; else if (q.x > b.xhi) q.x -= b.Lx;
      if (q.y < b.ylo) q.y += b.Ly; else if (q.y > b.yhi) q.y -= b.Ly;
      if (q.z < b.zlo) q.z += b.Lz; else if (q.z > b.zhi) q.z -= b.Lz;
    }
    s.cells[cell_idx(s, q)].push_back(i);
  }
}

// NeighbourList::build_cell (src/system/neighbour_list.cpp:216-273)
void build_cell(Sim& s) {
  int N = (int)s.p.size();
  double cut = s.cut + s.pad;
  double cut2 = cut * cut;
  s.old.clear();
  cell_populate(s);
  for (int i = 0; i < N; i++) s.p[i].coordination = 0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    const std::vector<int>& nc = s.stencil[cell_idx(s, pi)];
    for (int c : nc) {
      const std::vector<int>& ids = s.cells[c];
      for (int jj : ids) {
        Particle& pj = s.p[jj];
        if (pj.id > pi.id) {
          double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          apply_periodic(s, dx, dy, dz);
          double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 < cut2) {
            s.list[i].push_back(pj.id);
            double r = pi.radius + pj.radius;
            if (d2 < r * r) { pi.coordination++; pj.coordination++; }
          }
        }
      }
    }
    s.old.push_back({pi.x, pi.y, pi.z});
  }
}

// NeighbourList::build_nsq (src/system/neighbour_list.cpp:178-213)
void build_nsq(Sim& s, int i) {
  double cut = s.cut + s.pad;
  double cut2 = cut * cut;
  Particle& pi = s.p[i];
  pi.coordination = 0;
  for (int j = 0; j < pi.id; j++) {
    Particle& pj = s.p[j];
    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    apply_periodic(s, dx, dy, dz);
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < cut2) {
      s.list[j].push_back(i);
      double r = pi.radius + pj.radius;
      if (d2 < r * r) { pi.coordination++; pj.coordination++; }
    }
  }
  s.old.push_back({pi.x, pi.y, pi.z});
}

// NeighbourList::build (src/system/neighbour_list.cpp:117-149)
void nlist_build(Sim& s) {
  s.list.assign(s.p.size(), {});
  if (s.use_cell) build_cell(s);
  else { s.old.clear(); for (int i = 0; i < (int)s.p.size(); i++) build_nsq(s, i); }
  s.builds++;
}

// NeighbourList constructor selector (src/system/neighbour_list.hpp:138-150)
void nlist_create(Sim& s, double rcut, double pad) {
  s.has_nlist = true; s.cut = rcut; s.pad = pad;
  if (s.box.Lx > 2.0 * (rcut + pad) && s.box.Ly > 2.0 * (rcut + pad) && s.box.Lz > 2.0 * (rcut + pad)) {
    s.use_cell = true;
    cell_list_create(s, rcut + pad);
  } else {
    s.use_cell = false;
  }
  nlist_build(s);
}

// NeighbourList::need_update (src/system/neighbour_list.cpp:278-302)
bool need_update(const Sim& s, const Particle& p) {
  const Box& b = s.box;
  double dx = s.old[p.id].x - p.x, dy = s.old[p.id].y - p.y, dz = s.old[p.id].z - p.z;
  if (s.periodic) {
    if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
    if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
    if (dz > b.zhi) dz -= b.Lz; else if (dz < b.zlo) dz += b.Lz;
  }
  return !(dx * dx + dy * dy + dz * dz < 0.25 * s.pad * s.pad);
}

// ---------------------------------------------------------------------------
// Pair potentials
// ---------------------------------------------------------------------------
// PairSoftPotential::compute (src/potentials/pair/pair_soft_potential.cpp:33-119);
// phase_in is not supported (alpha = 1).
void soft_compute(Sim& s, PairPot& P) {
  int N = (int)s.p.size();
  int nt = P.tab.ntypes;
  double tot = 0.0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    double ai = pi.radius;
    for (int j : s.list[i]) {
      Particle& pj = s.p[j];
      int tt = (pi.type - 1) * nt + (pj.type - 1);
      double k = P.tab.p0[tt];
      double aj = pj.radius;
      double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      apply_periodic(s, dx, dy, dz);
      double r_sq = dx * dx + dy * dy + dz * dz;
      double r = std::sqrt(r_sq);
      double a = P.use_radii ? ai + aj : P.tab.p1[tt];
      if (r < a) {
        double diff = a - r;
        double alpha = 1.0;
        double e = 0.5 * k * alpha * diff * diff;
        tot += e;
        double ff = (r > 0.0) ? k * alpha * diff / r : k * diff;
        pi.fx -= ff * dx; pi.fy -= ff * dy; pi.fz -= ff * dz;
        pj.fx += ff * dx; pj.fy += ff * dy; pj.fz += ff * dz;
      }
    }
  }
  P.energy = tot;
}

// PairLJPotential::compute (src/potentials/pair/pair_lj_potential.cpp:32-124)
void lj_compute(Sim& s, PairPot& P) {
  int N = (int)s.p.size();
  int nt = P.tab.ntypes;
  double sigma = P.sigma, eps = P.eps, rcut = P.rcut;
  double sigma_sq = sigma * sigma, rcut_sq = rcut * rcut;
  P.energy = 0.0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    double ai = pi.radius;
    for (int j : s.list[i]) {
      Particle& pj = s.p[j];
      int tt = (pi.type - 1) * nt + (pj.type - 1);
      if (P.has_pair_params) { rcut = P.tab.p2[tt]; rcut_sq = rcut * rcut; }
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      apply_periodic(s, dx, dy, dz);
      double r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq <= rcut_sq) {
        if (P.has_pair_params) { sigma = P.tab.p1[tt]; eps = P.tab.p0[tt]; sigma_sq = sigma * sigma; }
        if (P.use_radii) { sigma = ai + pj.radius; sigma_sq = sigma * sigma; }
        double alpha = 1.0;
        double inv_r_sq = sigma_sq / r_sq;
        double inv_r_6 = inv_r_sq * inv_r_sq * inv_r_sq;
        double e = 4.0 * eps * alpha * inv_r_6 * (inv_r_6 - 1.0);
        if (P.shifted) {
          double inv_rc_sq = sigma_sq / rcut_sq;
          double inv_rc_6 = inv_rc_sq * inv_rc_sq * inv_rc_sq;
          e -= 4.0 * eps * alpha * inv_rc_6 * (inv_rc_6 - 1.0);
        }
        P.energy += e;
        double ff = 48.0 * eps * alpha * inv_r_6 * (inv_r_6 - 0.5) * inv_r_sq;
        pi.fx += ff * dx; pi.fy += ff * dy; pi.fz += ff * dz;
        pj.fx -= ff * dx; pj.fy -= ff * dy; pj.fz -= ff * dz;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Pair aligners
// ---------------------------------------------------------------------------
// PairPolarAlign::compute (src/aligner/pair/pair_polar_aligner.cpp:32-107)
void polar_compute(Sim& s, PairAl& A) {
  int N = (int)s.p.size();
  int nt = A.tab.ntypes;
  const Box& b = s.box;
  double J = A.J, rcut = A.rcut;
  double tot = 0.0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    for (int j : s.list[i]) {
      Particle& pj = s.p[j];
      if (A.has_pair_params) { int tt = (pi.type - 1) * nt + (pj.type - 1); J = A.tab.p0[tt]; rcut = A.tab.p1[tt]; }
      double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      if (s.periodic) {
        if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
        if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
        if (dz > b.zhi) dz -= b.Lz; else if (dz < b.zlo) dz += b.Lz;
      }
      double r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq <= rcut * rcut) {
        double tx = pi.ny * pj.nz - pi.nz * pj.ny;
        double ty = pi.nz * pj.nx - pi.nx * pj.nz;
        double tz = pi.nx * pj.ny - pi.ny * pj.nx;
        if (!(A.exclude_boundary && pi.boundary)) { pi.tau_x += J * tx; pi.tau_y += J * ty; pi.tau_z += J * tz; }
        if (!(A.exclude_boundary && pj.boundary)) { pj.tau_x += -J * tx; pj.tau_y += -J * ty; pj.tau_z += -J * tz; }
        tot += -2.0 * J * (pi.nx * pj.nx + pi.ny * pj.ny + pi.nz * pj.nz);
      }
    }
  }
  A.energy = tot;
}

// PairNematicAlign::compute (src/aligner/pair/pair_nematic_aligner.cpp:32-96).
// The global coupling is doubled; per-type couplings are used as given (:37,:57-61).
void nematic_compute(Sim& s, PairAl& A) {
  int N = (int)s.p.size();
  int nt = A.tab.ntypes;
  const Box& b = s.box;
  double J = 2.0 * A.J, rcut = A.rcut;
  A.energy = 0.0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    for (int j : s.list[i]) {
      Particle& pj = s.p[j];
      if (A.has_pair_params) { int tt = (pi.type - 1) * nt + (pj.type - 1); J = A.tab.p0[tt]; rcut = A.tab.p1[tt]; }
      double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
      if (s.periodic) {
        if (dx > b.xhi) dx -= b.Lx; else if (dx < b.xlo) dx += b.Lx;
        if (dy > b.yhi) dy -= b.Ly; else if (dy < b.ylo) dy += b.Ly;
        if (dz > b.zhi) dz -= b.Lz; else if (dz < b.zlo) dz += b.Lz;
      }
      double r_sq = dx * dx + dy * dy + dz * dz;
      if (r_sq <= rcut * rcut) {
        double nn = pi.nx * pj.nx + pi.ny * pj.ny + pi.nz * pj.nz;
        double tx = pi.ny * pj.nz - pi.nz * pj.ny;
        double ty = pi.nz * pj.nx - pi.nx * pj.nz;
        double tz = pi.nx * pj.ny - pi.ny * pj.nx;
        pi.tau_x += J * nn * tx; pi.tau_y += J * nn * ty; pi.tau_z += J * nn * tz;
        pj.tau_x += -J * nn * tx; pj.tau_y += -J * nn * ty; pj.tau_z += -J * nn * tz;
        A.energy += -2.0 * J * (2.0 * nn * nn - 1.0);
      }
    }
  }
}

// ExternalSelfPropulsion::compute (src/potentials/external/external_self_propulsion.cpp:33-50)
void self_prop_compute(Sim& s) {
  for (auto& q : s.p) { q.fx += s.sp_alpha * q.nx; q.fy += s.sp_alpha * q.ny; q.fz += s.sp_alpha * q.nz; }
}

// Potential::compute (src/potentials/potential.cpp:36-52): pair (map order) then external.
// Global-parameter forms of four further pair potentials; phase_in is not
// supported (alpha = 1) and the per-type tables hold the globals.
//   morse            q = {D, a, re, rcut}        pair_morse_potential.cpp:32-107
//   gaussian         q = {A, B, alpha, beta, rA, rB, rcut}  pair_gaussian_potential.cpp:32-111
//   yukawa           q = {g, kappa, rcut}        pair_yukawa_potential.cpp:32-80
//   soft_attractive  q = {k, a, re_fact}         pair_soft_attractive_potential.cpp:32-118
// (morse and gaussian subtract the value at rcut when shifted; morse and
// gaussian with use_particle_radii take re / rB = a_i + a_j, soft_attractive
// takes a = a_i + a_j; soft_attractive also recounts the coordination.)
void extra_compute(Sim& s, PairPot& P) {
  int N = (int)s.p.size();
  const double* q = P.q;
  P.energy = 0.0;
  if (P.kind == 5)
    for (auto& p : s.p) p.coordination = 0;
  for (int i = 0; i < N; i++) {
    Particle& pi = s.p[i];
    double ai = pi.radius;
    for (int j : s.list[i]) {
      Particle& pj = s.p[j];
      if (P.kind == 5) {
        // PairSoftAttractivePotential: d = x_j - x_i, forces with the minus sign on i
        double dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        apply_periodic(s, dx, dy, dz);
        double r_sq = dx * dx + dy * dy + dz * dz;
        double r = std::sqrt(r_sq);
        double k = q[0], alpha = 1.0;
        double ai_p_aj = P.use_radii ? ai + pj.radius : q[1];
        double r_turn = q[2] * ai_p_aj;
        double r_f = 2.0 * r_turn - ai_p_aj;
        double ff = 0.0, e = 0.0;
        if (r < r_f) {
          pi.coordination += 1;
          pj.coordination += 1;
          if (r <= r_turn) {
            double diff = ai_p_aj - r;
            e = 0.5 * k * alpha * diff * diff;
            if (r > 0.0) ff = k * alpha * diff / r; else ff = k * diff;
          } else if (r > r_turn && r < r_f) {
            double diff = r_f - r;
            e = -0.5 * k * alpha * diff * diff;
            ff = -k * alpha * diff / r;
          }
          P.energy += e;
          pi.fx -= ff * dx; pi.fy -= ff * dy; pi.fz -= ff * dz;
          pj.fx += ff * dx; pj.fy += ff * dy; pj.fz += ff * dz;
        }
        continue;
      }
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      apply_periodic(s, dx, dy, dz);
      double r_sq = dx * dx + dy * dy + dz * dz;
      double ff = 0.0, e = 0.0;
      if (P.kind == 2) {                       // morse
        double rcut = q[3];
        if (!(r_sq <= rcut * rcut)) continue;
        double D = q[0], a = q[1], re = q[2], alpha = 1.0;
        if (P.use_radii) re = ai + pj.radius;
        double r = std::sqrt(r_sq);
        double exp_fact = std::exp(-a * (r - re));
        double pot_fact = exp_fact - 1.0;
        e = D * alpha * (pot_fact * pot_fact - 1.0);
        if (P.shifted) {
          double shift_fact = std::exp(-a * (rcut - re)) - 1.0;
          e -= D * alpha * (shift_fact * shift_fact - 1.0);
        }
        ff = 2.0 * D * a * alpha * exp_fact * pot_fact / r;
      } else if (P.kind == 3) {                // gaussian
        double rcut = q[6];
        if (!(r_sq <= rcut * rcut)) continue;
        double A = q[0], B = q[1], al = q[2], be = q[3], rA = q[4], rB = q[5], pf = 1.0;
        if (P.use_radii) rB = ai + pj.radius;
        double r = std::sqrt(r_sq);
        double r_m_rA = r - rA, r_m_rB = r - rB;
        e = A * pf * std::exp(-al * r_m_rA * r_m_rA) + B * std::exp(-be * r_m_rB * r_m_rB);
        if (P.shifted) {
          double rcut_m_rA = rcut - rA, rcut_m_rB = rcut - rB;
          e -= A * pf * std::exp(-al * rcut_m_rA * rcut_m_rA) + B * std::exp(-be * rcut_m_rB * rcut_m_rB);
        }
        ff = (2.0 / r) * pf * (A * al * r_m_rA * std::exp(-al * r_m_rA * r_m_rA) +
                               B * be * r_m_rB * std::exp(-be * r_m_rB * r_m_rB));
      } else if (P.kind == 4) {                // yukawa
        double rcut = q[2];
        if (!(r_sq <= rcut * rcut)) continue;
        double g = q[0], kappa = q[1];
        double r = std::sqrt(r_sq);
        double exp_fact = g * std::exp(-kappa * r);
        e = exp_fact / r;
        ff = exp_fact * (1 + kappa * r) / (r_sq * r);
      }
      P.energy += e;
      pi.fx += ff * dx; pi.fy += ff * dy; pi.fz += ff * dz;
      pj.fx -= ff * dx; pj.fy -= ff * dy; pj.fz -= ff * dz;
    }
  }
}

void pot_compute(Sim& s) {
  for (auto& kv : s.pots) {
    if (kv.second.kind == 0) soft_compute(s, kv.second);
    else if (kv.second.kind == 1) lj_compute(s, kv.second);
    else extra_compute(s, kv.second);
  }
  if (s.has_self_prop) self_prop_compute(s);
}
// Potential::compute_potential_energy (src/potentials/potential.cpp:57-75)
double pot_energy(const Sim& s) {
  double e = 0.0;
  for (auto& kv : s.pots) e += kv.second.energy;
  return e;
}
// Aligner::compute (src/aligner/aligner.cpp:35-45)
void align_compute(Sim& s) {
  for (auto& kv : s.als) {
    if (kv.second.kind == 0) polar_compute(s, kv.second); else nematic_compute(s, kv.second);
  }
}

// ---------------------------------------------------------------------------
// Constraints
// ---------------------------------------------------------------------------
// ConstraintSphere::enforce (src/constraints/constraint_sphere.cpp:38-81)
void sphere_enforce(Sim& s, Particle& p) {
  double R0 = s.con.r;
  double x = p.x, y = p.y, z = p.z;
  double R = std::sqrt(x * x + y * y + z * z);
  double sc = R0 / R;
  p.x *= sc; p.y *= sc; p.z *= sc;
  double Nx = p.x / R0, Ny = p.y / R0, Nz = p.z / R0;
  double vN = p.vx * Nx + p.vy * Ny + p.vz * Nz;
  double nN = p.nx * Nx + p.ny * Ny + p.nz * Nz;
  p.vx -= vN * Nx; p.vy -= vN * Ny; p.vz -= vN * Nz;
  p.nx -= nN * Nx; p.ny -= nN * Ny; p.nz -= nN * Nz;
  double inv_len = 1.0 / std::sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
  p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
  p.Nx = Nx; p.Ny = Ny; p.Nz = Nz;
}

// ConstraintPlane::enforce (src/constraints/constraint_plane.cpp:36-88)
void plane_enforce(Sim& s, Particle& p) {
  double Lx = s.box.Lx, Ly = s.box.Ly;
  double xlo = -0.5 * Lx, xhi = 0.5 * Lx, ylo = -0.5 * Ly, yhi = 0.5 * Ly;
  p.z = s.con.zpos; p.vz = 0.0; p.fz = 0.0;
  p.Nx = 0.0; p.Ny = 0.0; p.Nz = 1.0;
  if (s.periodic) enforce_periodic(s, p);
  else if (!s.con.unlimited) {
    if (p.x < xlo) { p.x = xlo; p.vx = -p.vx; } else if (p.x > xhi) { p.x = xhi; p.vx = -p.vx; }
    if (p.y < ylo) { p.y = ylo; p.vy = -p.vy; } else if (p.y > yhi) { p.y = yhi; p.vy = -p.vy; }
  }
  p.nz = 0.0;
  double len_n = std::sqrt(p.nx * p.nx + p.ny * p.ny);
  double inv_len = 1.0;
  if (len_n != 0.0) inv_len = 1.0 / std::sqrt(p.nx * p.nx + p.ny * p.ny);
  p.nx *= inv_len; p.ny *= inv_len;
}

// Implicit surfaces projected by the generic SHAKE iteration.
// ConstraintTorus (src/constraints/constraint_torus.cpp:26-52):
//   g = (c - sqrt(x^2 + y^2))^2 + z^2 - a^2
// ConstraintEllipsoid (src/constraints/constraint_ellipsoid.cpp:26-48):
//   g = x^2/a^2 + y^2/b^2 + z^2/c^2 - 1
double surface_value(const Sim& s, double x, double y, double z) {
  const Constraint& c = s.con;
  if (c.kind == 4) {
    double ft = c.tc - std::sqrt(x * x + y * y);
    return ft * ft + z * z - c.ta * c.ta;
  }
  return x * x / c.ea / c.ea + y * y / c.eb / c.eb + z * z / c.ec / c.ec - 1.0;
}

void surface_gradient(const Sim& s, double x, double y, double z, double g[3]) {
  const Constraint& c = s.con;
  if (c.kind == 4) {
    double rho = std::sqrt(x * x + y * y);
    double f = -2.0 * (c.tc - rho) / rho;
    g[0] = x * f; g[1] = y * f; g[2] = 2.0 * z;
  } else {
    g[0] = 2.0 * x / c.ea / c.ea; g[1] = 2.0 * y / c.eb / c.eb; g[2] = 2.0 * z / c.ec / c.ec;
  }
}

// compute_normal of the sphere, cylinder (x/r, y/r, 0), torus and ellipsoid
// (normalised gradient); records it in the particle's N like the reference.
void surface_normal(const Sim& s, Particle& p, double N[3]) {
  const Constraint& c = s.con;
  if (c.kind == 2) { N[0] = p.x / c.r; N[1] = p.y / c.r; N[2] = p.z / c.r; }
  else if (c.kind == 3) { N[0] = p.x / c.r; N[1] = p.y / c.r; N[2] = 0.0; }
  else {
    surface_gradient(s, p.x, p.y, p.z, N);
    double len = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    N[0] /= len; N[1] /= len; N[2] /= len;
  }
  p.Nx = N[0]; p.Ny = N[1]; p.Nz = N[2];
}

// Constraint::enforce (src/constraints/constraint.cpp:40-87): move along the
// gradient taken at the entry point by lambda = g / (grad g . grad g_0) until
// |g| <= tol or max_iter moves; then tangent projections of v and n, n
// renormalised, periodic wrap.
void shake_enforce(Sim& s, Particle& p) {
  double g0[3];
  surface_gradient(s, p.x, p.y, p.z, g0);
  int iter = 0;
  while (iter < s.con.max_iter) {
    ++iter;
    double gr[3];
    surface_gradient(s, p.x, p.y, p.z, gr);
    double dot = gr[0] * g0[0] + gr[1] * g0[1] + gr[2] * g0[2];
    double g = surface_value(s, p.x, p.y, p.z);
    if (!(std::fabs(g) > s.con.tol)) break;
    double lambda = g / dot;
    p.x -= lambda * g0[0]; p.y -= lambda * g0[1]; p.z -= lambda * g0[2];
  }
  double N[3];
  surface_normal(s, p, N);
  double vN = p.vx * N[0] + p.vy * N[1] + p.vz * N[2];
  double nN = p.nx * N[0] + p.ny * N[1] + p.nz * N[2];
  p.vx -= vN * N[0]; p.vy -= vN * N[1]; p.vz -= vN * N[2];
  p.nx -= nN * N[0]; p.ny -= nN * N[1]; p.nz -= nN * N[2];
  double inv_len = 1.0 / std::sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
  p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
  enforce_periodic(s, p);
}

// ConstraintCylinder::enforce (src/constraints/constraint_cylinder.cpp:30-66):
// radial scaling in xy, xy projections of v and n, n renormalised over all
// three components, z wrapped (strict bounds, image flag untouched).
void cylinder_enforce(Sim& s, Particle& p) {
  double R = std::sqrt(p.x * p.x + p.y * p.y);
  double sc = s.con.r / R;
  p.x *= sc; p.y *= sc;
  double Nx = p.x / s.con.r, Ny = p.y / s.con.r;
  double vN = p.vx * Nx + p.vy * Ny;
  double nN = p.nx * Nx + p.ny * Ny;
  p.vx -= vN * Nx; p.vy -= vN * Ny;
  p.nx -= nN * Nx; p.ny -= nN * Ny;
  double inv_len = 1.0 / std::sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
  p.nx *= inv_len; p.ny *= inv_len; p.nz *= inv_len;
  if (s.periodic) {
    if (p.z > s.box.zhi) p.z -= s.box.Lz;
    else if (p.z < s.box.zlo) p.z += s.box.Lz;
  }
}

// Constrainer::enforce (src/constraints/constrainer.hpp:73-78) with one constraint;
// ConstraintNone::enforce (src/constraints/constraint_none.hpp:58-62)
void enforce(Sim& s, Particle& p) {
  switch (s.con.kind) {
    case 0: enforce_periodic(s, p); break;
    case 1: plane_enforce(s, p); break;
    case 2: sphere_enforce(s, p); break;
    case 3: cylinder_enforce(s, p); break;
    case 4: case 5: shake_enforce(s, p); break;
    default: break;
  }
}

// project_torque: sphere via Constraint::project_torque (src/constraints/constraint.cpp:158-173),
// plane (src/constraints/constraint_plane.cpp:146-160), none returns 0.
double project_torque(Sim& s, Particle& p) {
  switch (s.con.kind) {
    case 1: return p.tau_z;
    case 2: case 3: case 4: case 5: {
      double N[3];
      surface_normal(s, p, N);
      return p.tau_x * N[0] + p.tau_y * N[1] + p.tau_z * N[2];
    }
    default: return 0.0;
  }
}

// rotate_director: sphere via Constraint::rotate_director (src/constraints/constraint.cpp:96-119),
// plane (src/constraints/constraint_plane.cpp:95-115), none is a no-op.
void rotate_director(Sim& s, Particle& p, double phi) {
  if (s.con.kind == 1) {
    double c = std::cos(phi), sn = std::sin(phi);
    double Rxx = c, Rxy = -sn, Ryx = sn, Ryy = c;
    double nx = Rxx * p.nx + Rxy * p.ny;
    double ny = Ryx * p.nx + Ryy * p.ny;
    double len = std::sqrt(nx * nx + ny * ny);
    p.nx = nx / len; p.ny = ny / len;
  } else if (s.con.kind >= 2) {
    double N[3];
    surface_normal(s, p, N);
    double U = N[0], V = N[1], W = N[2];
    double c = std::cos(phi), sn = std::sin(phi);
    double nx = U * (U * p.nx + V * p.ny + W * p.nz) * (1.0 - c) + p.nx * c + (-W * p.ny + V * p.nz) * sn;
    double ny = V * (U * p.nx + V * p.ny + W * p.nz) * (1.0 - c) + p.ny * c + (W * p.nx - U * p.nz) * sn;
    double nz = W * (U * p.nx + V * p.ny + W * p.nz) * (1.0 - c) + p.nz * c + (-V * p.nx + U * p.ny) * sn;
    double len = std::sqrt(nx * nx + ny * ny + nz * nz);
    p.nx = nx / len; p.ny = ny / len; p.nz = nz / len;
  }
}

// rotate_velocity: sphere (src/constraints/constraint.cpp:128-150), plane (:118-140 of constraint_plane.cpp)
void rotate_velocity(Sim& s, Particle& p, double phi) {
  if (s.con.kind == 1) {
    double c = std::cos(phi), sn = std::sin(phi);
    double vx = c * p.vx + (-sn) * p.vy;
    double vy = sn * p.vx + c * p.vy;
    p.vx = vx; p.vy = vy;
  } else if (s.con.kind >= 2) {
    double N[3];
    surface_normal(s, p, N);
    double U = N[0], V = N[1], W = N[2];
    double c = std::cos(phi), sn = std::sin(phi);
    double vx = U * (U * p.vx + V * p.vy + W * p.vz) * (1.0 - c) + p.vx * c + (-W * p.vy + V * p.vz) * sn;
    double vy = V * (U * p.vx + V * p.vy + W * p.vz) * (1.0 - c) + p.vy * c + (W * p.vx - U * p.vz) * sn;
    double vz = W * (U * p.vx + V * p.vy + W * p.vz) * (1.0 - c) + p.vz * c + (-V * p.vx + U * p.vy) * sn;
    p.vx = vx; p.vy = vy; p.vz = vz;
  }
}

// ---------------------------------------------------------------------------
// Integrators
// ---------------------------------------------------------------------------
// Noise draws of one integrator.  Philox: keyed by (seed, tag, particle id,
// call), so the order of evaluation does not matter.  GslStream: drawn in call
// order, which the integrators below make the reference's per-particle order
// (flip uniform, then [three translational Gaussians], then the rotational one).
inline void noise3(Integ& I, uint32_t tag, uint32_t pid, uint64_t call, double g[3]) {
  if (I.stream) {
    g[0] = I.gsl.gauss(1.0); g[1] = I.gsl.gauss(1.0); g[2] = I.gsl.gauss(1.0);
  } else {
    gauss3(I.seed, tag, pid, call, g);
  }
}
inline double noise_rot(Integ& I, uint32_t tag, uint32_t pid, uint64_t call) {
  if (I.stream) return I.gsl.gauss(1.0);
  double g, u; gauss1_u1(I.seed, tag, pid, call, g, u);
  return g;
}
inline double noise_flip(Integ& I, uint32_t tag, uint32_t pid, uint64_t call) {
  if (I.stream) return I.gsl.uniform();
  double g, u; gauss1_u1(I.seed, tag, pid, call, g, u);
  return u;
}

// IntegratorBrownian::integrate (src/integrators/integrator_brownian.cpp:42-124)
void brownian(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double T = I.Tval(s.run_step);
  double B = std::sqrt(2.0 * I.mu * T);
  double sqrt_dt = std::sqrt(I.dt);
  uint64_t call = I.calls++;
  if (I.nematic) {
    for (int i : G) {
      Particle& p = s.p[i];
      if (noise_flip(I, TAG_BROWNIAN, (uint32_t)p.id, call) < I.tau) {
        p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz;
        if (I.velocity) { p.vx = -p.vx; p.vy = -p.vy; p.vz = -p.vz; }
      }
    }
  }
  for (int i : G) {
    Particle& p = s.p[i];
    double fdx = I.v0 * p.nx + I.mu * p.fx;
    double fdy = I.v0 * p.ny + I.mu * p.fy;
    double fdz = I.v0 * p.nz + I.mu * p.fz;
    p.vx = fdx; p.vy = fdy; p.vz = fdz;
    p.x += I.dt * fdx; p.y += I.dt * fdy; p.z += I.dt * fdz;
    if (T > 0.0) {
      double g[3]; noise3(I, TAG_BROWNIAN, (uint32_t)p.id, call, g);
      double frx = B * g[0], fry = B * g[1], frz = B * g[2];
      p.vx += frx; p.vy += fry; p.vz += frz;
      p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
    }
    enforce(s, p);
    p.omega = I.mur * project_torque(s, p);
    double g = noise_rot(I, TAG_BROWNIAN, (uint32_t)p.id, call);
    double dtheta = I.dt * p.omega + I.stoch * g;
    rotate_director(s, p, dtheta);
    if (I.velocity) rotate_velocity(s, p, dtheta);
    p.age += I.dt;
  }
}

// IntegratorBrownianPos::integrate (src/integrators/integrator_brownian_pos.cpp:37-84)
void brownian_pos(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double T = I.Tval(s.run_step);
  double B = std::sqrt(2.0 * I.mu * T);
  double sqrt_dt = std::sqrt(I.dt);
  uint64_t call = I.calls++;
  for (int gi : G) {
    Particle& p = s.p[gi];
    double fdx = I.mu * p.fx, fdy = I.mu * p.fy, fdz = I.mu * p.fz;
    p.vx = fdx; p.vy = fdy; p.vz = fdz;
    p.x += I.dt * p.vx; p.y += I.dt * p.vy; p.z += I.dt * p.vz;
    if (T > 0.0) {
      double g[3]; noise3(I, TAG_BROWNIAN_POS, (uint32_t)p.id, call, g);
      double frx = B * g[0], fry = B * g[1], frz = B * g[2];
      p.vx += frx; p.vy += fry; p.vz += frz;
      p.x += sqrt_dt * frx; p.y += sqrt_dt * fry; p.z += sqrt_dt * frz;
    }
    enforce(s, p);
    p.age += I.dt;
  }
}

// IntegratorBrownianAlign::integrate (src/integrators/integrator_brownian_align.cpp:41-76)
void brownian_align(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  uint64_t call = I.calls++;
  if (I.nematic) {
    for (int gi : G) {
    Particle& p = s.p[gi];
      if (noise_flip(I, TAG_BROWNIAN_ALIGN, (uint32_t)p.id, call) < I.tau) { p.nx = -p.nx; p.ny = -p.ny; p.nz = -p.nz; }
    }
  }
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.omega = I.mur * project_torque(s, p);
    double g = noise_rot(I, TAG_BROWNIAN_ALIGN, (uint32_t)p.id, call);
    double dtheta = I.dt * p.omega + I.stoch * g;
    rotate_director(s, p, dtheta);
  }
}

// IntegratorLangevin::integrate_baoab (src/integrators/integrator_langevin.cpp:54-118)
void langevin_baoab(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double T = I.Tval(s.run_step);
  double B = std::sqrt(T * (1.0 - std::exp(-2.0 * I.gamma * I.dt)));
  double exp_dt = std::exp(-I.gamma * I.dt);
  double dt2 = 0.5 * I.dt;
  uint64_t call = I.calls++;
  for (int gi : G) {
    Particle& p = s.p[gi];
    double fact = dt2 / p.mass;
    p.vx += fact * p.fx; p.vy += fact * p.fy; p.vz += fact * p.fz;
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    enforce(s, p);
    p.vx *= exp_dt; p.vy *= exp_dt; p.vz *= exp_dt;
    if (B != 0.0) {
      double sf = B / std::sqrt(p.mass);
      double g[3]; noise3(I, TAG_LANGEVIN, (uint32_t)p.id, call, g);
      p.vx += sf * g[0]; p.vy += sf * g[1]; p.vz += sf * g[2];
    }
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    enforce(s, p);
  }
  reset_forces(s);
  pot_compute(s);
  for (int gi : G) {
    Particle& p = s.p[gi];
    double fact = dt2 / p.mass;
    p.vx += fact * p.fx; p.vy += fact * p.fy; p.vz += fact * p.fz;
    p.age += I.dt;
  }
}

// IntegratorLangevin::integrate_spv (src/integrators/integrator_langevin.cpp:128-183)
void langevin_spv(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double T = I.Tval(s.run_step);
  double zeta = std::sqrt(T * (1.0 - std::exp(-2.0 * I.gamma * I.dt)));
  double eta = (I.gamma == 0.0) ? 0.0 : (1.0 - std::exp(-I.dt * I.gamma)) / I.gamma;
  double exp_dt = std::exp(-I.dt * I.gamma);
  double dt2 = 0.5 * I.dt;
  uint64_t call = I.calls++;
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    enforce(s, p);
  }
  reset_forces(s);
  pot_compute(s);
  for (int gi : G) {
    Particle& p = s.p[gi];
    double fact = eta / p.mass;
    p.vx = exp_dt * p.vx + fact * p.fx;
    p.vy = exp_dt * p.vy + fact * p.fy;
    p.vz = exp_dt * p.vz + fact * p.fz;
    if (zeta != 0.0) {
      double sf = zeta / std::sqrt(p.mass);
      double g[3]; noise3(I, TAG_LANGEVIN, (uint32_t)p.id, call, g);
      p.vx += sf * g[0]; p.vy += sf * g[1]; p.vz += sf * g[2];
    }
    p.x += dt2 * p.vx; p.y += dt2 * p.vy; p.z += dt2 * p.vz;
    enforce(s, p);
    p.age += I.dt;
  }
}

// IntegratorLangevin::integrate_bbk (src/integrators/integrator_langevin.cpp:193-253).
// The stored random vectors R are indexed by particle id; the initial draw
// (integrator_langevin.hpp:106-116) uses call counter 2^63.
void langevin_bbk(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double T = I.Tval(s.run_step);
  double B = std::sqrt(2.0 * T * I.dt * I.gamma);
  double dt2 = 0.5 * I.dt;
  double one_m_dt2 = 1.0 - I.gamma * dt2;
  double one_div = 1.0 / (1.0 + I.gamma * dt2);
  uint64_t call = I.calls++;
  for (int gi : G) {
    Particle& p = s.p[gi];
    double dm = dt2 / p.mass;
    p.vx = one_m_dt2 * p.vx + dm * p.fx;
    p.vy = one_m_dt2 * p.vy + dm * p.fy;
    p.vz = one_m_dt2 * p.vz + dm * p.fz;
    if (B != 0.0) {
      double sf = 0.5 * B / std::sqrt(p.mass);
      p.vx += sf * I.Rx[p.id]; p.vy += sf * I.Ry[p.id]; p.vz += sf * I.Rz[p.id];
    }
    p.x += I.dt * p.vx; p.y += I.dt * p.vy; p.z += I.dt * p.vz;
    enforce(s, p);
  }
  reset_forces(s);
  pot_compute(s);
  for (int gi : G) {
    Particle& p = s.p[gi];
    double dm = dt2 / p.mass;
    p.vx = one_div * (p.vx + dm * p.fx);
    p.vy = one_div * (p.vy + dm * p.fy);
    p.vz = one_div * (p.vz + dm * p.fz);
    if (B != 0.0) {
      double sf = 0.5 * B * one_div / std::sqrt(p.mass);
      double g[3]; noise3(I, TAG_LANGEVIN, (uint32_t)p.id, call, g);
      I.Rx[p.id] = g[0]; I.Ry[p.id] = g[1]; I.Rz[p.id] = g[2];
      p.vx += sf * g[0]; p.vy += sf * g[1]; p.vz += sf * g[2];
      p.age += I.dt;
    }
  }
}

// IntegratorFIRE::integrate (src/integrators/integrator_fire.cpp:34-148).  The
// per-iteration stdout line is not printed; its three numbers are kept instead.
void fire(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  int N = (int)G.size();
  double sqrt_ndof = std::sqrt(3 * N);
  double dt_2 = 0.5 * I.dt;
  double P = 0.0, Fn = 0.0, Vn = 0.0;
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.vx += dt_2 * p.fx; p.vy += dt_2 * p.fy; p.vz += dt_2 * p.fz;
    p.x += I.dt * p.vx; p.y += I.dt * p.vy; p.z += I.dt * p.vz;
    enforce(s, p);
    p.omega += dt_2 * project_torque(s, p);
    double dtheta = I.dt * p.omega;
    rotate_director(s, p, dtheta);
  }
  reset_forces(s);
  reset_torques(s);
  pot_compute(s);
  align_compute(s);
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.vx += dt_2 * p.fx; p.vy += dt_2 * p.fy; p.vz += dt_2 * p.fz;
    P += p.vx * p.fx + p.vy * p.fy + p.vz * p.fz;
    Fn += p.fx * p.fx + p.fy * p.fy + p.fz * p.fz;
    Vn += p.vx * p.vx + p.vy * p.vy + p.vz * p.vz;
    enforce(s, p);
    p.omega += dt_2 * project_torque(s, p);
  }
  Fn = std::sqrt(Fn);
  Vn = std::sqrt(Vn);
  double E = pot_energy(s);
  I.last_P = P; I.last_Fnorm = Fn; I.last_Vnorm = Vn; I.last_E = E;
  if (Fn / sqrt_ndof < I.F_tol && std::fabs(E - I.old_energy) < I.E_tol) {
    I.converged = true;
    return;
  }
  double inv_F = 1.0 / Fn;
  double f1 = 1.0 - I.alpha;
  double f2 = I.alpha * inv_F * Vn;
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.vx = f1 * p.vx + f2 * p.fx;
    p.vy = f1 * p.vy + f2 * p.fy;
    p.vz = f1 * p.vz + f2 * p.fz;
  }
  if (P > 0.0) {
    I.last_neg++;
    if (I.last_neg > I.N_min) { I.dt = std::min(I.dt * I.f_inc, I.dt_max); I.alpha *= I.f_alpha; }
  } else {
    I.dt *= I.f_dec;
    I.alpha = I.alpha_init;
    I.last_neg = 0;
    for (int gi : G) {
    Particle& p = s.p[gi]; p.vx = 0.0; p.vy = 0.0; p.vz = 0.0; }
  }
  I.converged = false;
  I.old_energy = E;
}

// IntegratorNVE::integrate (src/integrators/integrator_nve.cpp:37-141): velocity
// Verlet with optional displacement / rotation limits.  The reference clamps
// omega with its angle limit whenever a displacement limit is set, reading an
// unset member when no angle_limit was given; here that clamp needs both.
inline double sign_of(double x) { return x >= 0.0 ? 1.0 : -1.0; }   // integrator_nve.hpp:38

void nve(Sim& s, Integ& I) {
  const std::vector<int>& G = members(s, I);
  double dt_2 = 0.5 * I.dt;
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.vx += dt_2 * p.fx; p.vy += dt_2 * p.fy; p.vz += dt_2 * p.fz;
    enforce(s, p);
    p.omega += dt_2 * project_torque(s, p);
  }
  for (int gi : G) {
    Particle& p = s.p[gi];
    double dx = I.dt * p.vx, dy = I.dt * p.vy, dz = I.dt * p.vz;
    if (I.has_limit) {
      double dr = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (dr > I.limit) {
        dx = I.limit * dx / dr;
        dy = I.limit * dy / dr;
        dz = I.limit * dz / dr;
      }
    }
    p.x += dx; p.y += dy; p.z += dz;
    enforce(s, p);
  }
  for (int gi : G) {
    Particle& p = s.p[gi];
    double dtheta = I.dt * p.omega;
    if (I.has_theta_limit && std::fabs(dtheta) > I.theta_limit) dtheta = sign_of(dtheta) * I.theta_limit;
    rotate_director(s, p, dtheta);
  }
  reset_forces(s);
  reset_torques(s);
  pot_compute(s);
  align_compute(s);
  for (int gi : G) {
    Particle& p = s.p[gi];
    p.vx += dt_2 * p.fx; p.vy += dt_2 * p.fy; p.vz += dt_2 * p.fz;
    if (I.has_limit) {
      double v = std::sqrt(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
      if (v * I.dt > I.limit) {
        p.vx = p.vx / v * I.limit / I.dt;
        p.vy = p.vy / v * I.limit / I.dt;
        p.vz = p.vz / v * I.limit / I.dt;
      }
    }
    enforce(s, p);
    p.omega += dt_2 * project_torque(s, p);
    if (I.has_limit && I.has_theta_limit && std::fabs(p.omega) * I.dt > I.theta_limit)
      p.omega = sign_of(p.omega) * I.theta_limit / I.dt;
    p.age += I.dt;
  }
}

void integrate(Sim& s, Integ& I) {
  if (I.name == "brownian") brownian(s, I);
  else if (I.name == "brownian_pos") brownian_pos(s, I);
  else if (I.name == "brownian_align") brownian_align(s, I);
  else if (I.name == "langevin") {
    if (I.method == 0) langevin_baoab(s, I); else if (I.method == 1) langevin_spv(s, I); else langevin_bbk(s, I);
  } else if (I.name == "fire") fire(s, I);
  else if (I.name == "nve") nve(s, I);
}

bool needs_nlist(const Sim& s) { return s.has_nlist && (!s.pots.empty() || !s.als.empty()); }

// ConstraintSphere::rescale (src/constraints/constraint_sphere.cpp:87-104): at
// time steps that are multiples of rescale_freq and below rescale_steps the
// radius is multiplied by the per-rescale factor and every particle enforced.
bool sphere_rescale(Sim& s) {
  Constraint& c = s.con;
  if (c.kind != 2 || c.rescale == 1.0) return false;
  const int step = s.time_step;
  if (step % c.rescale_freq == 0 && step < c.rescale_steps) {
    c.r *= c.scale;
    for (auto& p : s.p) enforce(s, p);
    return true;
  }
  return false;
}

// Run loop (src/samos.cpp:872-970): N+1 iterations.
void run(Sim& s, int steps) {
  pot_compute(s);
  align_compute(s);
  for (int t = 0; t <= steps; t++) {
    s.run_step = t;
    if (sphere_rescale(s) && s.has_nlist) nlist_build(s);
    for (auto& p : s.p) enforce(s, p);
    reset_forces(s);
    reset_torques(s);
    pot_compute(s);
    align_compute(s);
    for (auto& kv : s.integs) integrate(s, kv.second);
    if (needs_nlist(s)) {
      bool rebuild = false;
      for (auto& p : s.p) if (need_update(s, p)) { rebuild = true; break; }
      if (rebuild) nlist_build(s);
    }
    s.time_step++;
  }
}

thread_local std::string g_err;

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI (ctypes).  Field codes match include/samos_hip.h's samos_field.
// ---------------------------------------------------------------------------
extern "C" {

enum { F_X = 0, F_Y, F_Z, F_VX, F_VY, F_VZ, F_NX, F_NY, F_NZ, F_FX, F_FY, F_FZ,
       F_TX, F_TY, F_TZ, F_OMEGA, F_AGE, F_RADIUS, F_MASS,
       F_TYPE = 32, F_IX, F_IY, F_IZ, F_COORD };

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(int n, double lx, double ly, double lz, int periodic) {
  Sim* s = new Sim();
  s->p.resize(n);
  for (int i = 0; i < n; i++) {
    Particle& q = s->p[i];
    std::memset(&q, 0, sizeof(Particle));
    q.id = i; q.type = 1; q.radius = 1.0; q.mass = 1.0;
  }
  s->box.set(lx, ly, lz);
  s->periodic = periodic != 0;
  std::vector<int>& all = s->groups["all"];
  for (int i = 0; i < n; i++) all.push_back(i);
  return s;
}

// `group name { ... }` (src/samos.cpp:1108-1140): member[i] != 0 puts particle i
// into the group; a group defined again gains the new members (the reference
// appends them to its list; a particle listed twice is kept once here).
int orc_group_define(void* h, const char* name, const int* member) {
  Sim* s = static_cast<Sim*>(h);
  const std::string nm = name;
  if (nm == "all") { g_err = "group all always holds every particle"; return -1; }
  std::vector<int>& g = s->groups[nm];
  std::vector<char> in(s->p.size(), 0);
  for (int i : g) in[i] = 1;
  for (int i = 0; i < (int)s->p.size(); i++) if (member[i] && !in[i]) g.push_back(i);
  return 0;
}

// The `group` parameter of the next integrator added (src/samos.cpp:654-690).
int orc_set_integrator_group(void* h, const char* group) {
  Sim* s = static_cast<Sim*>(h);
  if (!s->groups.count(group)) { g_err = std::string("Group ") + group + " does not exist."; return -1; }
  s->next_group = group;
  return 0;
}

// `disable type { group = g }`: removes that integrator.
int orc_integrator_disable(void* h, const char* type, const char* group) {
  Sim* s = static_cast<Sim*>(h);
  auto it = s->integs.find(std::string(type) + "_" + group);
  if (it == s->integs.end()) { g_err = "Integrator does not exist."; return -1; }
  s->integs.erase(it);
  return 0;
}

void orc_destroy(void* h) { delete static_cast<Sim*>(h); }

static double* dfield(Particle& q, int f) {
  switch (f) {
    case F_X: return &q.x; case F_Y: return &q.y; case F_Z: return &q.z;
    case F_VX: return &q.vx; case F_VY: return &q.vy; case F_VZ: return &q.vz;
    case F_NX: return &q.nx; case F_NY: return &q.ny; case F_NZ: return &q.nz;
    case F_FX: return &q.fx; case F_FY: return &q.fy; case F_FZ: return &q.fz;
    case F_TX: return &q.tau_x; case F_TY: return &q.tau_y; case F_TZ: return &q.tau_z;
    case F_OMEGA: return &q.omega; case F_AGE: return &q.age;
    case F_RADIUS: return &q.radius; case F_MASS: return &q.mass;
  }
  return nullptr;
}
static int* ifield(Particle& q, int f) {
  switch (f) {
    case F_TYPE: return &q.type; case F_IX: return &q.ix; case F_IY: return &q.iy; case F_IZ: return &q.iz;
    case F_COORD: return &q.coordination;
  }
  return nullptr;
}

int orc_set_field(void* h, int f, const void* data) {
  Sim* s = static_cast<Sim*>(h);
  int n = (int)s->p.size();
  if (f < 32) {
    const double* d = static_cast<const double*>(data);
    for (int i = 0; i < n; i++) { double* t = dfield(s->p[i], f); if (!t) return -1; *t = d[i]; }
  } else {
    const int* d = static_cast<const int*>(data);
    for (int i = 0; i < n; i++) { int* t = ifield(s->p[i], f); if (!t) return -1; *t = d[i]; }
    if (f == F_TYPE) { int mx = 1; for (int i = 0; i < n; i++) mx = std::max(mx, d[i]); s->ntypes = mx; }
  }
  return 0;
}

int orc_get_field(void* h, int f, void* out) {
  Sim* s = static_cast<Sim*>(h);
  int n = (int)s->p.size();
  if (f < 32) {
    double* d = static_cast<double*>(out);
    for (int i = 0; i < n; i++) { double* t = dfield(s->p[i], f); if (!t) return -1; d[i] = *t; }
  } else {
    int* d = static_cast<int*>(out);
    for (int i = 0; i < n; i++) { int* t = ifield(s->p[i], f); if (!t) return -1; d[i] = *t; }
  }
  return 0;
}

int orc_set_timestep(void* h, double dt) { static_cast<Sim*>(h)->dt = dt; return 0; }

int orc_nlist_create(void* h, double rcut, double pad) {
  nlist_create(*static_cast<Sim*>(h), rcut, pad);
  return 0;
}
int orc_nlist_build(void* h) { nlist_build(*static_cast<Sim*>(h)); return 0; }
int orc_nlist_uses_cells(void* h) { return static_cast<Sim*>(h)->use_cell ? 1 : 0; }
long orc_nlist_builds(void* h) { return static_cast<Sim*>(h)->builds; }
int orc_cell_dims(void* h, int* out) {
  Sim* s = static_cast<Sim*>(h);
  out[0] = s->cnx; out[1] = s->cny; out[2] = s->cnz;
  return 0;
}

// Half pair list as (i, j) rows with i < j, sorted; returns the row count
// (duplicates, which the reference produces when a grid side has two cells,
// are kept).  With out == nullptr only the count is returned.
long orc_nlist_pairs(void* h, int* out, long cap) {
  Sim* s = static_cast<Sim*>(h);
  std::vector<std::pair<int, int>> pr;
  for (int i = 0; i < (int)s->list.size(); i++)
    for (int j : s->list[i]) pr.push_back({std::min(i, j), std::max(i, j)});
  std::sort(pr.begin(), pr.end());
  if (out) {
    if ((long)pr.size() > cap) return -1;
    for (size_t k = 0; k < pr.size(); k++) { out[2 * k] = pr[k].first; out[2 * k + 1] = pr[k].second; }
  }
  return (long)pr.size();
}

// Brute-force O(N^2) reference pair set with the same d^2 expression.
long orc_brute_pairs(void* h, double cut, int* out, long cap) {
  Sim* s = static_cast<Sim*>(h);
  double cut2 = cut * cut;
  long cnt = 0;
  int n = (int)s->p.size();
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double dx = s->p[i].x - s->p[j].x, dy = s->p[i].y - s->p[j].y, dz = s->p[i].z - s->p[j].z;
      apply_periodic(*s, dx, dy, dz);
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < cut2) {
        if (out) { if (cnt >= cap) return -1; out[2 * cnt] = i; out[2 * cnt + 1] = j; }
        cnt++;
      }
    }
  return cnt;
}

int orc_need_update_any(void* h) {
  Sim* s = static_cast<Sim*>(h);
  for (auto& p : s->p) if (need_update(*s, p)) return 1;
  return 0;
}

int orc_add_pair_soft(void* h, double k, double a, int use_radii) {
  Sim* s = static_cast<Sim*>(h);
  PairPot P; P.kind = 0; P.k = k; P.a = a; P.use_radii = use_radii != 0;
  P.tab.init(s->ntypes, k, a, 0.0);
  s->pots["soft"] = P;
  return 0;
}

int orc_add_pair_lj(void* h, double eps, double sigma, double rcut, int shifted, int use_radii) {
  Sim* s = static_cast<Sim*>(h);
  PairPot P; P.kind = 1; P.eps = eps; P.sigma = sigma; P.rcut = rcut; P.shifted = shifted != 0;
  P.use_radii = use_radii != 0;
  P.tab.init(s->ntypes, eps, sigma, rcut);
  s->pots["lj"] = P;
  return 0;
}

// kind: 2 morse, 3 gaussian, 4 yukawa, 5 soft_attractive; q as in extra_compute.
int orc_add_pair_extra(void* h, int kind, const double* q, int use_radii, int shifted) {
  static const char* names[] = {"", "", "morse", "gaussian", "yukawa", "soft_attractive"};
  if (kind < 2 || kind > 5) { g_err = "unknown pair potential kind"; return -1; }
  Sim* s = static_cast<Sim*>(h);
  PairPot P; P.kind = kind;
  for (int i = 0; i < 8; i++) P.q[i] = q[i];
  P.use_radii = use_radii != 0; P.shifted = shifted != 0;
  s->pots[names[kind]] = P;
  return 0;
}

// Per-type pair parameters (set_pair_parameters, symmetric fill).
// soft: (k, a); lj: (eps, sigma, rcut, wca flag -> rcut = 2^(1/6) sigma, shifted).
int orc_pair_type_param(void* h, const char* name, int t1, int t2, double p0, double p1, double p2, int wca) {
  Sim* s = static_cast<Sim*>(h);
  auto it = s->pots.find(name);
  if (it == s->pots.end()) { g_err = "no such pair potential"; return -1; }
  PairPot& P = it->second;
  int nt = P.tab.ntypes;
  if (t1 < 1 || t2 < 1 || t1 > nt || t2 > nt) { g_err = "type out of range"; return -1; }
  if (P.kind == 1 && wca) { p2 = 1.122462048309373 * p1; P.shifted = true; }
  int a = (t1 - 1) * nt + (t2 - 1), b = (t2 - 1) * nt + (t1 - 1);
  P.tab.p0[a] = P.tab.p0[b] = p0;
  P.tab.p1[a] = P.tab.p1[b] = p1;
  P.tab.p2[a] = P.tab.p2[b] = p2;
  P.has_pair_params = true;
  return 0;
}

int orc_add_align(void* h, int kind, double J, double rcut) {
  Sim* s = static_cast<Sim*>(h);
  PairAl A; A.kind = kind; A.J = J; A.rcut = rcut;
  A.tab.init(s->ntypes, J, rcut, 0.0);
  s->als[kind == 0 ? "polar" : "nematic"] = A;
  return 0;
}

int orc_align_type_param(void* h, int kind, int t1, int t2, double J, double rcut) {
  Sim* s = static_cast<Sim*>(h);
  auto it = s->als.find(kind == 0 ? "polar" : "nematic");
  if (it == s->als.end()) { g_err = "no such aligner"; return -1; }
  PairAl& A = it->second;
  int nt = A.tab.ntypes;
  int a = (t1 - 1) * nt + (t2 - 1), b = (t2 - 1) * nt + (t1 - 1);
  A.tab.p0[a] = A.tab.p0[b] = J;
  A.tab.p1[a] = A.tab.p1[b] = rcut;
  A.has_pair_params = true;
  return 0;
}

int orc_set_self_propulsion(void* h, double alpha) {
  Sim* s = static_cast<Sim*>(h);
  s->has_self_prop = true; s->sp_alpha = alpha;
  return 0;
}

// kind: 0 none, 1 plane (param = zpos, flag = unlimited), 2 sphere (param = r).
// Adding a constraint enforces it on every particle (src/samos.cpp:510-522).
int orc_set_constraint(void* h, int kind, double param, int flag) {
  Sim* s = static_cast<Sim*>(h);
  s->con.kind = kind;
  if (kind == 1) { s->con.zpos = param; s->con.unlimited = flag != 0; }
  if (kind == 2) s->con.r = param;
  for (auto& p : s->p) enforce(*s, p);
  return 0;
}

// Sphere radius schedule {rescale; rescale_steps; rescale_freq}
// (constraint.hpp:85-108): the factor per rescale is rescale^(freq / steps).
int orc_set_constraint_rescale(void* h, double rescale, int steps, int freq) {
  Sim* s = static_cast<Sim*>(h);
  if (s->con.kind != 2) { g_err = "only the sphere constraint rescales"; return -1; }
  s->con.rescale = rescale; s->con.rescale_steps = steps; s->con.rescale_freq = freq;
  s->con.scale = std::pow(rescale, static_cast<double>(freq) / static_cast<double>(steps));
  return 0;
}

double orc_constraint_radius(void* h) { return static_cast<Sim*>(h)->con.r; }

// Cylinder {r}, torus {c, a}, ellipsoid {a, b, c}; enforced at once.
int orc_set_constraint_surface(void* h, int kind, const double* q, int max_iter, double tol) {
  Sim* s = static_cast<Sim*>(h);
  if (kind < 3 || kind > 5) { g_err = "unknown constraint kind"; return -1; }
  if (kind == 4 && q[1] >= q[0]) { g_err = "Toroidal constraint. Tube radius has to be smaller than torus radius."; return -1; }
  s->con.kind = kind;
  s->con.max_iter = max_iter; s->con.tol = tol;
  if (kind == 3) s->con.r = q[0];
  if (kind == 4) { s->con.tc = q[0]; s->con.ta = q[1]; }
  if (kind == 5) { s->con.ea = q[0]; s->con.eb = q[1]; s->con.ec = q[2]; }
  for (auto& p : s->p) enforce(*s, p);
  return 0;
}

int orc_enforce_all(void* h) {
  Sim* s = static_cast<Sim*>(h);
  for (auto& p : s->p) enforce(*s, p);
  return 0;
}

int orc_rotate_directors(void* h, const double* phi) {
  Sim* s = static_cast<Sim*>(h);
  for (size_t i = 0; i < s->p.size(); i++) rotate_director(*s, s->p[i], phi[i]);
  return 0;
}

int orc_project_torques(void* h, double* out) {
  Sim* s = static_cast<Sim*>(h);
  for (size_t i = 0; i < s->p.size(); i++) out[i] = project_torque(*s, s->p[i]);
  return 0;
}

static Integ make_integ(Sim* s, const char* name, double dt) {
  Integ I; I.name = name; I.dt = dt > 0.0 ? dt : s->dt;
  I.stream = s->rng_stream;
  I.group = s->next_group;
  s->next_group = "all";
  return I;
}

// Registers under type_group (src/samos.cpp:666-690); an existing key is an error.
static int put_integ(Sim* s, const Integ& I) {
  const std::string key = I.name + "_" + I.group;
  if (s->integs.count(key)) {
    g_err = "Integrator of type " + I.name + " (for group " + I.group + ") already exists.";
    return -1;
  }
  s->integs[key] = I;
  return 0;
}

// Noise stream for integrators added after this call: 0 counter-based Philox
// (the device stream, the default), 1 the reference's sequential MT19937 order.
int orc_set_rng_stream(void* h, int stream) {
  if (stream != 0 && stream != 1) { g_err = "unknown noise stream"; return -1; }
  static_cast<Sim*>(h)->rng_stream = stream;
  return 0;
}

// First draws of the reference-order stream: nu uniforms, then ng Gaussians.
int orc_gsl_stream_sample(uint32_t seed, int nu, double* uni, int ng, double* gau) {
  GslStream a; a.set(seed);
  for (int i = 0; i < nu; i++) uni[i] = a.uniform();
  for (int i = 0; i < ng; i++) gau[i] = a.gauss(1.0);
  return 0;
}

// Brownian family.  which: 0 brownian, 1 brownian_pos, 2 brownian_align.
// tau here is the flip time (the integrator stores dt/tau).
int orc_add_brownian(void* h, int which, double v0, double nu, double mu, double mur, int seed,
                     int nematic, double tau, int velocity_align, int tkind, double tmin, double tmax, int tsteps) {
  Sim* s = static_cast<Sim*>(h);
  const char* nm = which == 0 ? "brownian" : which == 1 ? "brownian_pos" : "brownian_align";
  Integ I = make_integ(s, nm, 0.0);
  if (I.dt <= 0.0) { g_err = "Integrator step has not been set."; return -1; }
  I.v0 = v0; I.nu = nu; I.mu = mu; I.mur = mur; I.seed = (uint32_t)seed;
  I.gsl.set(I.seed);
  I.nematic = nematic != 0; I.tau = I.dt / tau; I.velocity = velocity_align != 0;
  I.tkind = tkind; I.tmin = tmin; I.tmax = tmax; I.tsteps = tsteps;
  I.stoch = std::sqrt(I.nu * I.dt);
  return put_integ(s, I);
}

int orc_add_langevin(void* h, double gamma, int seed, int method, int tkind, double tmin, double tmax, int tsteps) {
  Sim* s = static_cast<Sim*>(h);
  Integ I = make_integ(s, "langevin", 0.0);
  if (I.dt <= 0.0) { g_err = "Integrator step has not been set."; return -1; }
  I.gamma = gamma; I.seed = (uint32_t)seed; I.method = method;
  I.gsl.set(I.seed);
  I.tkind = tkind; I.tmin = tmin; I.tmax = tmax; I.tsteps = tsteps;
  if (method == 2) {
    int n = (int)s->p.size();
    I.Rx.resize(n); I.Ry.resize(n); I.Rz.resize(n);
    for (int i = 0; i < n; i++) {
      double g[3]; noise3(I, TAG_BBK_INIT, (uint32_t)i, 0, g);
      I.Rx[i] = g[0]; I.Ry[i] = g[1]; I.Rz[i] = g[2];
    }
  }
  return put_integ(s, I);
}

int orc_add_fire(void* h, double alpha, double alpha_scale, double f_tol, double e_tol,
                 double dt_inc, double dt_dec, double dt_max, int min_neg) {
  Sim* s = static_cast<Sim*>(h);
  Integ I = make_integ(s, "fire", 0.0);
  if (I.dt <= 0.0) { g_err = "Integrator step has not been set."; return -1; }
  I.alpha = alpha; I.alpha_init = alpha; I.f_alpha = alpha_scale; I.F_tol = f_tol; I.E_tol = e_tol;
  I.f_inc = dt_inc; I.f_dec = dt_dec; I.dt_max = dt_max; I.N_min = min_neg;
  return put_integ(s, I);
}

// NVE; a negative limit or angle_limit means the parameter is not given.
int orc_add_nve(void* h, double limit, double angle_limit) {
  Sim* s = static_cast<Sim*>(h);
  Integ I = make_integ(s, "nve", 0.0);
  if (I.dt <= 0.0) { g_err = "Integrator step has not been set."; return -1; }
  I.has_limit = limit >= 0.0; I.limit = limit;
  I.has_theta_limit = angle_limit >= 0.0; I.theta_limit = angle_limit;
  return put_integ(s, I);
}

// FIRE state: dt, alpha, P, |F|, |v|, E, converged, last_neg
int orc_fire_state(void* h, double* out) {
  Sim* s = static_cast<Sim*>(h);
  auto it = s->integs.begin();
  while (it != s->integs.end() && it->second.name != "fire") ++it;
  if (it == s->integs.end()) return -1;
  Integ& I = it->second;
  out[0] = I.dt; out[1] = I.alpha; out[2] = I.last_P; out[3] = I.last_Fnorm; out[4] = I.last_Vnorm;
  out[5] = I.last_E; out[6] = I.converged ? 1.0 : 0.0; out[7] = I.last_neg;
  return 0;
}

// Energy of the further pair potentials (kinds 2-5) from the last evaluation.
double orc_extra_energy(void* h) {
  Sim* s = static_cast<Sim*>(h);
  double e = 0.0;
  for (auto& kv : s->pots) if (kv.second.kind >= 2) e += kv.second.energy;
  return e;
}

// Loop-head force/torque evaluation: reset, Potential::compute, Aligner::compute.
// energies: [sum of pair energies, sum of aligner energies]
int orc_compute_forces(void* h, double* energies) {
  Sim* s = static_cast<Sim*>(h);
  reset_forces(*s);
  reset_torques(*s);
  pot_compute(*s);
  align_compute(*s);
  if (energies) {
    energies[0] = pot_energy(*s);
    double ea = 0.0; for (auto& kv : s->als) ea += kv.second.energy;
    energies[1] = ea;
  }
  return 0;
}

int orc_integrate_once(void* h) {
  Sim* s = static_cast<Sim*>(h);
  for (auto& kv : s->integs) integrate(*s, kv.second);
  return 0;
}

int orc_run(void* h, int steps) {
  try { run(*static_cast<Sim*>(h), steps); }
  catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}

// Known-answer hook for the counter-based RNG.
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  U4 c; for (int i = 0; i < 4; i++) c.v[i] = ctr[i];
  U4 r = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; i++) out[i] = r.v[i];
}

void orc_gauss3(uint32_t seed, uint32_t tag, uint32_t pid, uint64_t call, double* g) {
  gauss3(seed, tag, pid, call, g);
}

// The Box-Muller transcendentals on n arguments in (0, 1): out[3i] = log u,
// out[3i+1] = sin 2 pi u, out[3i+2] = cos 2 pi u.
void orc_bm_eval(const double* u, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    out[3 * i] = bm_log(u[i]);
    bm_sincos_2pi(u[i], out[3 * i + 1], out[3 * i + 2]);
  }
}

}  // extern "C"
