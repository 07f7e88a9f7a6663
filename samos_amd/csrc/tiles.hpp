// Cell-tiled engine for planar single-system runs: the state lives in
// cell order (row-major over a 2-D cell grid), each workgroup owns a run of
// cells in one row and stages the positions of the 3 x (run + 2) surrounding
// cells in LDS, and the Verlet list of a particle holds 16-bit indices into
// that stage instead of global indices.  The stage's layout depends only on
// the cell starts of the current set, which change only at builds, so an
// index written by a build stays valid for every step until the next one.
#pragma once
#include "kernels.hpp"

namespace samos {

constexpr int kTileThreads = 256;
constexpr int kTileMaxCells = 168;    // cells per tile along x (3 * (168 + 2) halo cells <= 2 * 256)
constexpr int kTileHaloMax = 1024;    // halo positions staged in LDS; larger halos read L2
constexpr int kTileStageMax = 65535;  // staged indices are 16-bit (the largest is the sentinel)
constexpr int kTileMaxOwn = 2 * kTileThreads;  // particles of one tile (two per step-kernel thread)
static_assert(kTileHaloMax % kTileThreads == 0, "the stage is loaded in whole rounds of the workgroup");
static_assert(3 * (kTileMaxCells + 2) <= 2 * kTileThreads, "two halo cells per thread in stage_index");
// Stage descriptor of a tile, written by each build for the new set:
//   [0] halo size (stage entries)  [1], [2] stage range of the tile's own
//   particles  [3] set index of the first of them  [4] runs
//   [5 + 2r], [6 + 2r] first stage entry and first set index of run r: the
//   stage is the concatenation of at most 9 runs of consecutive set indices
//   (each halo row is one run, or up to three where it wraps along x).
constexpr int kTileRuns = 9;
constexpr int kTileDesc = 5 + 2 * kTileRuns + 1;

// Control block in device memory.  `word` holds the current set and the
// build handshake, changed only by atomics:
//   bit 0  set of the last committed build
//   bit 1  a build finished and is not yet committed (the current set is
//          bit 0 ^ bit 1 for a step kernel; build kernels run only after a
//          step kernel committed, when bit 1 is clear)
//   bit 2 + p  a particle moved pad/2 in a step reading position buffer p:
//          build before the next step
// The first workgroup of every step kernel commits a finished build (flips
// bit 0, clears bit 1) and clears the other parity's request; both keep
// bit 0 ^ bit 1 unchanged for the workgroups reading it concurrently.
struct TileCtl {
  int word;
  int init;      // the pending build is a run's first (snapshot from the slot list)
  int rebuilds;  // builds during the run (the first one not counted)
  int overflow;  // some particle had more neighbours than the list holds, a
                 // halo more than kTileStageMax particles or a tile more than
                 // kTileMaxOwn
  int far;       // some particle changed cell by more than one between builds
  int maxcand;   // largest neighbour count of the builds
  int maxhalo;   // largest staged halo of the builds
  int pad_;
};

// One state set: the arrays of the cell-ordered state.
struct TileSet {
  double2* xy[2];        // positions, double-buffered by step
  double2* v;            // velocities
  double2* nd[2];        // directors: Langevin runs keep them in nd[0]; Brownian
                         // runs double-buffer them by step like the positions
  double* om;            // angular velocity omega (Brownian runs)
  double2* o;            // positions at the last build
  int* id;               // particle ids (noise keys)
  int* slot;             // slot of each particle in the slot-ordered columns
  int* cstart;           // ncell + 1 cell starts of this set's order
};

struct TileD {
  int n;                 // particles
  int ncx, ncy, ncell;   // 2-D cell grid (the reference's n_x, n_y; one z layer)
  int T;                 // cells per tile along x
  int ntx, ntiles;       // tiles per row, tiles
  int K;                 // list capacity per particle (a multiple of 4)
  int dirs2;             // directors double-buffered by step (Brownian runs)
  double wx, wy, xlo, ylo;
  double cut2;           // (rcut + pad)^2
  double pad2q;          // pad^2 / 4
  const TileSet* sets;   // [2], in device memory
  // Verlet lists of the current set: row r of particle q holds its entries
  // 4r .. 4r + 3 (staged indices, ascending staging order; the last row is
  // padded with the stage's sentinel index, the halo size), lst[r * n + q].
  ushort4* lst;          // [K / 4][n]
  int* len;              // [n] entries of each list
  int* desc;             // [ntiles][kTileDesc] stage descriptors of the current set
  int* count;            // [ncell] particles per cell of the build in progress
  int* tcount;           // [ntiles] particles per tile of the build in progress
  int* tstart;           // [ntiles + 1] exclusive scan of tcount
  int* cell_of;          // [n] new cell of each particle of the current set
  int* members;          // [n] current-set index of each new position
  TileCtl* ctl;
};

// Slot-ordered columns a step writes: closing forces (and torques), image-flag
// changes.
struct SlotOut {
  double *fx, *fy, *fz;
  int *ix, *iy, *iz;
  double *tx, *ty, *tz;
};

// Pair and self-propulsion constants of a one-type planar system.
struct PairConst {
  double soft_k, soft_a;          // pair_soft_potential: k, a
  double lj_eps, lj_sig2, lj_rc2; // pair_lj_potential: epsilon, sigma^2, rcut^2
  double sp_alpha;                // external self_propulsion alpha
  int has_sp;
};

// Copies the slot-ordered columns into set 1 (identity order) and arms the
// initial build: its snapshot is the slot list's (nl.ox, nl.oy), so the tiled
// run continues the reference's current neighbour list exactly.
void launch_tile_load(hipStream_t st, TileD td, Fields f, NListD nl);
// One build: bin by the snapshot, cell starts, member lists, then the new
// cell-ordered set with stencil masks.  Each kernel returns at once unless the
// request bit of parity par is set (a step that read buffer par asked for it);
// the current positions are in buffer par ^ 1, and so are the new set's.
// init: the first build of a run (scatter + sort).
void launch_tile_build(hipStream_t st, TileD td, BoxD box, int init, int par);
// The constant arguments of a run's step kernels, kept in device memory
// (the FusedArgs fields t_off and head_enforce are taken from the launch).
struct TileStepArgs {
  TileD td;
  SlotOut so;
  BoxD b;
  ConstraintD c;
  PairConst pc;
  FusedArgs fa;
};
// One fused BAOAB step (phases as launch_baoab_fused) on the current set,
// reading positions from buffer par and writing buffer par ^ 1; args: the
// run's TileStepArgs in device memory (its td equal to the td given).
void launch_tile_step(hipStream_t st, int pot_mask, int phase, int fp32, const TileD& td, const TileStepArgs* args,
                      int par, int t_off, int head_enforce);
// The constant arguments of a tiled Brownian run (one IntegratorBrownian on
// every particle of a planar system), kept in device memory.
struct TileBrownArgs {
  TileD td;
  SlotOut so;
  BoxD b;
  ConstraintD c;
  PairConst pc;
  double dt, v0, mu, mur, stoch, flip_p;  // IntegratorBrownian: dt, v0, mu, mur, sqrt(nu dt), dt / tau
  int nematic, velocity;                  //   nematic flips, velocity_align
  TempD temp;
  uint32_t seed;
  const ClockD* clock;
  int clock_slot;
  double al_J, al_rc2;                    // aligner: effective J and rcut^2 (al_mask 0: unused)
};
// One fused Brownian step (k_brownian_fused's physics, mode 0) on the current
// set: positions and directors from buffer par, written to par ^ 1.  td.dirs2
// must be set for the run.
// store: also write the loop-head forces and torques of this step to the
// slot-ordered columns (the last step of a range: the state the reference
// leaves behind).
void launch_tile_brownian(hipStream_t st, int pot_mask, int al_mask, int fp32, const TileD& td,
                          const TileBrownArgs* args, int par, int t_off, int head_enforce, int store);
// Writes the current set (positions, and with dirs2 directors, in buffer par)
// back into the slot-ordered columns (x, y, xy, v, n, omega) and the last
// build's snapshot into nl.ox / nl.oy.
void launch_tile_store(hipStream_t st, TileD td, Fields f, NListD nl, int par);

}  // namespace samos
