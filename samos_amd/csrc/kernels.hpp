// Host-side launchers for the device kernels.  Each launcher enqueues work on
// the given stream and returns immediately; none of them allocates memory.
#pragma once
#include "samos_device.hpp"

namespace samos {

// ---- neighbour list (nlist.hip) -------------------------------------------
// Rebuild stages.  `req` (may be null) points at a device flag: when it is
// zero the stage returns at once, so a rebuild can sit in a captured step
// sequence and cost only a dispatch when nothing moved far enough.
void launch_bin(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req);
void launch_scan(hipStream_t st, NListD nl, const int* req);
void launch_scatter(hipStream_t st, NListD nl, int n, const int* req);
void launch_cell_sort(hipStream_t st, NListD nl, const int* req);
void launch_build(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar);
void launch_build_nsq(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req);
// Whole rebuild (bin, scan, scatter, sort, build) or the N^2 build.  planar:
// read candidate positions from the interleaved (x, y) column.
void launch_rebuild(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar = 0);
// Displacement test of every particle against the last build; sets *req_out to 1
// when any particle moved by pad/2 or more.
void launch_need_update(hipStream_t st, Fields f, NListD nl, BoxD box, int n, int* req_out);
// Curve keys of each slot's cell, for the locality re-sort.
// Storage-order keys of every slot: plane = 1 orders each z layer of cells
// along a 2-D Hilbert curve (layer-major), plane = 0 the grid along a 3-D
// Morton curve.
void launch_morton_keys(hipStream_t st, Fields f, NListD nl, BoxD box, int n, unsigned long long* keys, int* slots,
                        int plane);
void launch_gather_d(hipStream_t st, double* dst, const double* src, const int* perm, int n);
void launch_gather_i(hipStream_t st, int* dst, const int* src, const int* perm, int n);
// Half pair export: for slot s, neighbours t with id[t] > id[s]; writes (id_lo, id_hi) rows.
void launch_export_pairs(hipStream_t st, Fields f, NListD nl, int n, int* out, unsigned long long* counter, long cap);
size_t radix_sort_temp_bytes(int n);
void radix_sort_pairs(hipStream_t st, void* temp, size_t temp_bytes, const unsigned long long* kin,
                      unsigned long long* kout, const int* vin, int* vout, int n);

// ---- pair interactions (interact.hip) -------------------------------------
// Forces and torques of every registered pair potential and aligner in one
// pass over the neighbour list.  mode: 0 overwrite f and tau, 1 add to them,
// 2 leave them untouched (energies only).
// Per-slot energies are written when e_out != nullptr: e_out[k * np + s] for
// k = soft, lj, polar, nematic.
void launch_interact(hipStream_t st, int pot_mask, int al_mask, int mode, Fields f, NListD nl, BoxD box,
                     InteractD ip, int n, double* e_out);
// Same with FP32 pair arithmetic (distances still formed in FP64, sums kept in FP64).
void launch_interact_f32(hipStream_t st, int pot_mask, int al_mask, int mode, Fields f, NListD nl, BoxD box,
                         InteractD ip, int n, double* e_out);

// Further pair potentials (morse, gaussian, yukawa, soft_attractive) in one pass,
// added to the forces when add_forces; per-slot energies (the reference's
// per-pair energy summed over listed neighbours) into e_col when not null.
void launch_interact_extra(hipStream_t st, Fields f, NListD nl, BoxD box, const ExtraD& ex, int n, double* e_col,
                           int add_forces);

// Deterministic two-pass sums of `ncols` columns of length n (column stride ld)
// into out[ncols] (device).  partial must hold ncols * reduce_blocks(n) doubles.
int reduce_blocks(int n);
void launch_reduce_sum(hipStream_t st, const double* cols, int ncols, long ld, int n, double* partial, double* out);

// ---- integrators and constraints (integrate.hip) --------------------------
void launch_enforce(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n);
// The integrator kernels act on the slots of the group with bit mask gbit only
// (Group::get_particles of the integrator's group; 1u: every slot).
void launch_brownian(hipStream_t st, int which, Fields f, BoxD box, ConstraintD con, int n, double dt,
                     double v0, double mu, double mur, double stoch, double flip_p, int nematic, int velocity,
                     TempD temp, uint32_t seed, unsigned long long call, const ClockD* clock, int clock_slot, int t_off,
                     unsigned gbit = 1u);
void launch_langevin_a(hipStream_t st, int method, Fields f, BoxD box, ConstraintD con, int n, double dt,
                       double gamma, TempD temp, uint32_t seed, unsigned long long call, const ClockD* clock,
                       int clock_slot, int t_off, double* R, unsigned gbit = 1u);
void launch_langevin_b(hipStream_t st, int method, Fields f, int n, double dt, double gamma, TempD temp,
                       uint32_t seed, unsigned long long call, const ClockD* clock, int clock_slot, int t_off,
                       double* R, BoxD box, ConstraintD con, unsigned gbit = 1u);
void launch_bbk_init(hipStream_t st, Fields f, int n, uint32_t seed, double* R);
// FIRE iteration (src/integrators/integrator_fire.cpp:34-148) with its scalar
// state on the device (dt read from `fs`).  fire_b writes P, |F|^2, |v|^2 per
// slot to `cols` (zeros outside the group); after their sums and the energy
// sums are in red[0..2] and red[3..5], fire_update applies the convergence
// test and the dt / alpha adaptation, and fire_mix mixes the velocities.
void launch_fire_a(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, const FireD* fs, unsigned gbit = 1u);
void launch_fire_b(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, const FireD* fs, double* cols, long ld,
                   unsigned gbit = 1u);
void launch_fire_update(hipStream_t st, FireD* fs, const double* red, int has_extra);
void launch_fire_mix(hipStream_t st, Fields f, int n, const FireD* fs, unsigned gbit = 1u);
// IntegratorNVE (src/integrators/integrator_nve.cpp:37-141): the part before the
// force evaluation and the part after it; limit / theta < 0: not given.
void launch_nve_a(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, double dt, double limit, double theta,
                  unsigned gbit = 1u);
void launch_nve_b(hipStream_t st, Fields f, BoxD box, ConstraintD con, int n, double dt, double limit, double theta,
                  unsigned gbit = 1u);
void launch_clock_advance(hipStream_t st, ClockD* clock, int nsteps);

// Fused Langevin BAOAB step over the neighbour list (one launch per step):
//   phase 0: forces F(x_t), first half of step t (B, A, enforce, O, A, enforce)
//   phase 1: forces F(x_t), closing B of step t-1, then phase-0 work for step t
//   phase 2: forces F(x_t), closing B of step t-1 only
// x_t must already carry step t's loop-head enforce; with head_enforce set the
// kernel applies step t+1's loop-head enforce to the positions it writes.
// Positions are read from `cur` and written to `nxt` (only x, y, z differ), so
// neighbour reads never see a half-updated step.  Sets flags[parity] when any
// particle passed pad/2 and clears flags[1 - parity].
struct FusedArgs {
  double dt, gamma;
  TempD temp;
  uint32_t seed;
  const ClockD* clock;
  int clock_slot, t_off, parity, head_enforce;
  double mass_u;   // common mass, or -1 to read each particle's
  double exp_dt;   // exp(-gamma dt)
  double Bc;       // O-step noise amplitude at constant temperature, or -1
  double fact_u;   // (dt / 2) / mass_u when the mass is common
  double sf_u;     // Bc / sqrt(mass_u) when both are known, or -1
  int batch;       // neighbour gathers in flight per thread (4 or 8)
  const AgeCtl* age = nullptr;  // common-age state (null or not uniform: each slot adds dt)
  ChunkCtl* chunk = nullptr;  // captured-chunk mode (null: direct launch with the parity flags)
  int slot = 0;               // slot index inside the chunk
  unsigned gbit = 1u;         // the integrator's group: other slots only take the loop-head enforce
  unsigned long long call0 = 0;  // noise call of a launch without the device clock (clock null)
  double* out_x = nullptr;    // closing phase of a planar run: x, y columns to write (or null)
  double* out_y = nullptr;
};
void launch_baoab_fused(hipStream_t st, int pot_mask, int phase, Fields cur, Fields nxt, NListD nl, BoxD box,
                        ConstraintD con, InteractD ip, int n, const FusedArgs& fa, int fp32, int planar,
                        int multitype);
// The whole rebuild as one cooperative launch (grid barriers between the
// phases); returns the launch status so the caller can fall back to
// launch_rebuild where cooperative launches are unavailable.
hipError_t launch_rebuild_coop(hipStream_t st, Fields f, NListD nl, BoxD box, int n, const int* req, int planar);
// Fused Brownian step (IntegratorBrownian, src/integrators/integrator_brownian.cpp:42-117)
// with the loop-head forces and torques of the same step; positions and
// directors read from `cur` and written to `nxt`.
// mode 0: IntegratorBrownian ("brownian"); mode 1: the pair of
// IntegratorBrownianAlign then IntegratorBrownianPos over the same particles
// (the map order of "brownian_align" and "brownian_pos"), the align half with
// dt, mur, stoch, flip_p, nematic, seed and clock_slot, the position half with
// the pos_* fields.
struct BrownArgs {
  double dt, v0, mu, mur, stoch, flip_p;
  int nematic, velocity;
  TempD temp;
  uint32_t seed;
  const ClockD* clock;
  int clock_slot, t_off, parity, head_enforce;
  int mode = 0;
  double pos_dt = 0.0, pos_mu = 0.0;
  TempD pos_temp{0, 0.0, 0.0, 1};
  uint32_t pos_seed = 0;
  int pos_slot = 0;
  ChunkCtl* chunk = nullptr;  // captured-chunk mode (null: direct launch with the parity flags)
  int slot = 0;
  int lanes = 1, lane_shift = 0;  // threads per slot (power of two up to 64) and its log2
  const AgeCtl* age = nullptr;    // common-age state (null or not uniform: each slot adds dt)
  // Several "brownian" integrators on disjoint groups (the type_group keys of
  // src/samos.cpp:654-690): with ngroups > 0 (multi-type kernels only) each
  // slot takes the parameters of the group that holds it, and a slot in none
  // of them is not integrated; the top-level fields are then unused.
  int ngroups = 0;
  struct Group {
    double dt, v0, mu, mur, stoch, flip_p;
    int nematic, velocity;
    TempD temp;
    uint32_t seed;
    int clock_slot;
    unsigned gbit;
  } grp[4];
};
bool brownian_fused_supported(int pot_mask, int al_mask);
void launch_brownian_fused(hipStream_t st, int pot_mask, int al_mask, Fields cur, Fields nxt, NListD nl, BoxD box,
                           ConstraintD con, InteractD ip, int n, const BrownArgs& ba, int fp32, int planar,
                           int multitype);

// Captured-chunk bookkeeping (see ChunkCtl): the settle step after the slots
// copies positions (and, with `directors`, directors) left in buffer 1 back
// to buffer 0 and forms the rebuild
// request; the finish step advances the clock and the chunk base by the slots
// that ran and clears the control words; init sets base, limit and the run's
// last iteration.
void launch_chunk_settle(hipStream_t st, ChunkCtl* ck, Fields f0, Fields f1, int n, int planar, int directors);
void launch_chunk_finish(hipStream_t st, ChunkCtl* ck, ClockD* clock, const int* nl_flags);
// The end of a captured step chunk in five dispatches: settle + binning, the
// scan, scatter and cell-sort stages, build + the chunk's bookkeeping (the
// stages of the rebuild do nothing unless a slot asked for one).
void launch_chunk_rebuild(hipStream_t st, ChunkCtl* ck, ClockD* clk, Fields f0, Fields f1, NListD nl, BoxD box,
                          int n, int planar, int directors);
// resort_at: builds count at which the slots stop for a storage re-sort.
// clk (optional): the device clock is set to noise call `calls` in slot
// `slot` and run step `base`.
void launch_chunk_init(hipStream_t st, ChunkCtl* ck, int base, int limit, int steps, int resort_at,
                       const int* pending = nullptr, ClockD* clk = nullptr, int slot = 0,
                       unsigned long long calls = 0);
// Common-age bookkeeping of fused ranges (AgeCtl): `check` records whether
// every age equals age[0]; `finish` advances the common value by `count`
// additions of dt and writes it to every slot (nothing when not uniform).
void launch_age_check(hipStream_t st, AgeCtl* ctl, const double* age, int n);
void launch_age_finish(hipStream_t st, AgeCtl* ctl, double* age, int n, double dt, int count);
// Interleaved (x, y) copy of the positions, read by the planar kernels.
void launch_pack_xy(hipStream_t st, Fields f, int n);
// Back from the interleaved column into the x, y columns.
void launch_unpack_xy(hipStream_t st, Fields f, int n);

}  // namespace samos
