// Device-side data model of replica ensembles and the host-side context,
// shared by the ensemble engine and the warm-start network.
#pragma once
#include "samos_device.hpp"

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

namespace samos {

struct alignas(32) D4 {
  double x, y, z, w;
};

struct RepD {
  BoxD box;
  ConstraintD con;
  int n0, n;                  // slot range
  int b0, nb;                 // bucket base and bucket count (one end entry follows)
  int hbits, ncx, ncy, ncz;
  double wx, wy, wz, cut2, pad2q;
  // soft repulsion and polar alignment
  double k, a, J, rc2_al;
  int radii, has_al;
  // Brownian
  double v0, mu, mur, stoch, dt;
  uint32_t seed;
  int pad0_;
  // FIRE
  double fdt, alpha, alpha0, f_alpha, f_inc, f_dec, dt_max, f_tol, e_tol, E_old, f1, f2, stop_fmax;
  int n_min, last_neg, frozen, zero_v;
  long long iters;
  double P, F2, V2, E, Fmax;
};

// Every column of an ensemble (slot order).
struct EState {
  D4* p[2];                   // x, y, z, radius
  D4* d[2];                   // nx, ny, nz, 0
  double *vx, *vy, *vz, *fx, *fy, *fz, *tx, *ty, *tz, *omega, *age;
  int *rep, *lid, *cellc;
};

// Replicas selected for a pass: ids and exclusive prefix sums of their
// particle and bucket-entry counts; cnt = {selected, particles, bucket entries}.
struct Work {
  int* rep;
  int* poff;
  int* boff;
  int* cnt;
};

}  // namespace samos

struct samos_ens {
  hipStream_t st = nullptr;
  std::string err;
  int nrep = 0, n = 0, np = 0;
  std::vector<samos::RepD> h_rep;
  std::vector<int> h_gid;        // slot -> replica-major input index
  samos::RepD* d_rep = nullptr;
  samos::EState e{};
  int cur = 0;                   // buffer holding the current positions and directors
  samos::NListD nl{};
  bool has_nlist = false;
  long long ntiles = 1;          // scan tiles over every bucket entry
  samos::Work wr{}, wa{};        // rebuild list, FIRE running list
  int* d_flags = nullptr;        // [0] rebuild needed, [2] overflow, [3] largest bucket, [4] builds,
                                 // [5] largest neighbour count, [8] replicas still relaxing
  int* rep_need = nullptr;
  unsigned long long* d_call = nullptr;
  double* cols = nullptr;        // 5 reduction columns of np
  int batch = 4;
  int kind = -1;                 // 0 Brownian, 1 FIRE
  double rcut = 2.4, pad = 0.3;
  // graphs: Brownian chunk (K steps) and FIRE chunk (check iterations)
  hipGraphExec_t g_brown = nullptr;
  int g_brown_cur = -1, g_brown_batch = 0;
  hipGraphExec_t g_fire = nullptr;
  int g_fire_k = 0, g_fire_cur = -1, g_fire_batch = 0;
  std::vector<void*> owned;      // every device allocation
  // timing: HIP events around each step kernel and each rebuild sequence
  // (launches go step by step while enabled)
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_step, ev_rebuild;
};

namespace samos {
// Rebuilds every replica's neighbour list from the current positions, growing
// the capacity until nothing overflows (host-synchronous).
int ens_build_all(samos_ens* E);
// Enforces the constraint of every particle (current buffer), on the stream.
void ens_enforce_async(samos_ens* E);
}  // namespace samos
