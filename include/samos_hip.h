/*
 * samos_hip.h -- C-ABI of the MI355X (gfx950) SAMoS particle pipeline.
 *
 * One context owns one particle system resident in HBM on one GPU: the
 * structure-of-arrays state, the cell list and neighbour list, the registered
 * pair potentials / pair aligners / integrators / constraint, and one HIP
 * stream.  All functions return 0 on success and a negative status on error;
 * samos_last_error() then returns the message (the reference's
 * Messenger::ERROR text where one exists).  No function throws.  A context is
 * not re-entrant: one host thread drives it.
 *
 * Host data crosses this boundary as plain arrays in the reference's particle
 * id order (id = row of the input file).  Indices are 0-based; particle types
 * are 1-based as in the reference input format.
 *
 * Each entry point names the reference interface it replaces.  Paths are
 * relative to the SAMoS source tree.
 */
#ifndef SAMOS_HIP_H
#define SAMOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct samos_ctx samos_ctx;

/* Particle columns (src/system/particle.hpp:257-293).  Codes below 32 are
 * double columns, codes from 32 are int columns. */
enum samos_field {
  SAMOS_X = 0, SAMOS_Y, SAMOS_Z, SAMOS_VX, SAMOS_VY, SAMOS_VZ,
  SAMOS_NX, SAMOS_NY, SAMOS_NZ, SAMOS_FX, SAMOS_FY, SAMOS_FZ,
  SAMOS_TAUX, SAMOS_TAUY, SAMOS_TAUZ, SAMOS_OMEGA, SAMOS_AGE,
  SAMOS_RADIUS, SAMOS_MASS,
  SAMOS_TYPE = 32, SAMOS_IX, SAMOS_IY, SAMOS_IZ, SAMOS_COORD
};

enum samos_constraint_kind {
  SAMOS_CONSTRAINT_NONE = 0, SAMOS_CONSTRAINT_PLANE = 1, SAMOS_CONSTRAINT_SPHERE = 2,
  SAMOS_CONSTRAINT_CYLINDER = 3, SAMOS_CONSTRAINT_TORUS = 4, SAMOS_CONSTRAINT_ELLIPSOID = 5
};
enum samos_pair_kind {
  SAMOS_PAIR_SOFT = 0, SAMOS_PAIR_LJ = 1,
  SAMOS_PAIR_MORSE = 2, SAMOS_PAIR_GAUSSIAN = 3, SAMOS_PAIR_YUKAWA = 4, SAMOS_PAIR_SOFT_ATTRACTIVE = 5
};
enum samos_align_kind { SAMOS_ALIGN_POLAR = 0, SAMOS_ALIGN_NEMATIC = 1 };
enum samos_brownian_kind { SAMOS_BROWNIAN = 0, SAMOS_BROWNIAN_POS = 1, SAMOS_BROWNIAN_ALIGN = 2 };
enum samos_langevin_method { SAMOS_LANGEVIN_BAOAB = 0, SAMOS_LANGEVIN_SPV = 1, SAMOS_LANGEVIN_BBK = 2 };
enum samos_value_kind { SAMOS_VALUE_CONSTANT = 0, SAMOS_VALUE_LINEAR = 1 };

/* samos_run flags */
#define SAMOS_RUN_FAITHFUL   1  /* evaluate forces exactly where the reference does (no fused step) */
#define SAMOS_RUN_NO_GRAPH   2  /* launch step by step instead of replaying captured chunks */
#define SAMOS_RUN_FP32_FORCE 4  /* pair arithmetic in FP32, sums and state in FP64 */
#define SAMOS_RUN_TILES      8  /* planar fused runs on the cell-tiled engine instead of the neighbour-list kernels */

/* ---- context and state ---------------------------------------------------
 * System(input, msg, box) (src/system/system.cpp:64-452) and Box(lx, ly, lz)
 * (src/system/box.hpp:56-62): a system of n particles in a box centred at the
 * origin; periodic as set by `box periodic` / `box fixed` (src/samos.cpp:188-226).
 * New particles have type 1, radius 1, mass 1 and every other column 0. */
int samos_create(samos_ctx** out, int device, int n, double lx, double ly, double lz, int periodic);
int samos_destroy(samos_ctx* ctx);
const char* samos_last_error(const samos_ctx* ctx); /* ctx may be NULL: last create error */
int samos_num_particles(const samos_ctx* ctx);

/* Upload / download one column in particle-id order.  `host` holds n doubles
 * or n ints according to the field code.  Replaces direct Particle& access
 * (System::get_particle, src/system/system.hpp:101). */
int samos_set_field(samos_ctx* ctx, int field, const void* host);
int samos_get_field(samos_ctx* ctx, int field, void* host);

/* `timestep` command (src/samos.cpp:1332-1356, System::set_integrator_step). */
int samos_set_timestep(samos_ctx* ctx, double dt);

/* ---- neighbour list (src/system/neighbour_list.hpp:124-272) -------------- */
/* NeighbourList(sys, msg, rcut, pad, param): chooses the cell build when every
 * box side exceeds 2 (rcut + pad) (neighbour_list.hpp:138-150), else N^2,
 * re-sorts particles along a curve of the cell grid (Hilbert per cell layer
 * under the plane constraint, else Morton), and builds.  Setting the plane
 * constraint (or leaving it) with a list present re-sorts into that order and
 * rebuilds from the list's snapshot (the same pair set). */
int samos_nlist_create(samos_ctx* ctx, double rcut, double pad);
/* NeighbourList::build (src/system/neighbour_list.cpp:117-149). */
int samos_nlist_build(samos_ctx* ctx);
/* any(NeighbourList::need_update(p)) over all particles (src/samos.cpp:944-965). */
int samos_nlist_need_update(samos_ctx* ctx, int* out);
/* info[8]: uses_cells, ncx, ncy, ncz, stride, builds, max_cell_count, overflow */
int samos_nlist_info(samos_ctx* ctx, long long* info);
/* The half-pair set of NeighbourList::get_neighbours (neighbour_list.hpp:233):
 * rows (i, j), i < j in id order, sorted ascending.  With out == NULL only
 * *count is set.  Pairs are kept once per stencil visit, as in the reference. */
int samos_nlist_pairs(samos_ctx* ctx, int* out, long long cap, long long* count);
/* No counterpart in the reference: samos_nlist_pairs for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc, hipFree or other allocating
 * routine: the device staging of the rows lives in a fixed file-scope array of
 * the library, one per GPU (16 MiB plus 256 bytes, held until the process
 * ends), at most 2,097,152 half pairs; a list with more returns an error.
 * The rows are sorted in place in `out`.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (several host threads or contexts
 * at once) race; calls one after another are independent. */
int samos_nlist_pairs_noalloc(samos_ctx* ctx, int* out, long long cap, long long* count);
/* Re-sort particle storage along the curve of the cell grid (Hilbert per
 * layer under the plane constraint, else Morton) and rebuild. */
int samos_resort(samos_ctx* ctx);

/* ---- pair potentials (src/potentials/pair/pair_potential.hpp:52-158) ----- */
/* PairSoftPotential(..., {k; a; use_particle_radii}) (pair_soft_potential.hpp:63-117) */
int samos_pair_soft(samos_ctx* ctx, double k, double a, int use_radii);
/* PairLJPotential(..., {epsilon; sigma; rcut; shifted; use_particle_radii})
 * (pair_lj_potential.hpp:61-145); WCA is resolved by the caller into
 * rcut = 2^(1/6) sigma and shifted = 1. */
int samos_pair_lj(samos_ctx* ctx, double eps, double sigma, double rcut, int shifted, int use_radii);
/* set_pair_parameters (pair_soft_potential.hpp:125-176, pair_lj_potential.hpp:153-232):
 * soft p0 = k, p1 = a; lj p0 = epsilon, p1 = sigma, p2 = rcut, shifted (0/1, ORed in). */
int samos_pair_type_param(samos_ctx* ctx, int kind, int type1, int type2, double p0, double p1, double p2, int shifted);

/* ---- pair aligners (src/aligner/pair/pair_aligner.hpp:53-89) ------------- */
/* PairPolarAlign / PairNematicAlign(..., {J; rcut}); rcut <= 0 selects the
 * neighbour list cutoff (pair_polar_aligner.hpp:81-90). */
int samos_align_polar(samos_ctx* ctx, double J, double rcut);
int samos_align_nematic(samos_ctx* ctx, double J, double rcut);
/* set_pair_parameters (pair_polar_aligner.hpp:120-163, pair_nematic_aligner.hpp:110-155). */
int samos_align_type_param(samos_ctx* ctx, int kind, int type1, int type2, double J, double rcut);

/* Further pair potentials with global parameters (phase_in not supported),
 * params in this order:
 *   MORSE            {D, a, re, rcut}                 pair_morse_potential.hpp:62-160, .cpp:32-107
 *   GAUSSIAN         {A, B, alpha, beta, rA, rB, rcut} pair_gaussian_potential.hpp:60-200, .cpp:32-111
 *   YUKAWA           {g, kappa, rcut}                 pair_yukawa_potential.hpp:60-120, .cpp:32-80
 *   SOFT_ATTRACTIVE  {k, a, re_fact}                  pair_soft_attractive_potential.hpp:62-140, .cpp:32-118
 * use_radii: use_particle_radii (morse: re, gaussian: rB, soft_attractive: a
 * become a_i + a_j); shifted: subtract the value at rcut (morse, gaussian).
 * soft_attractive recounts each particle's coordination on every evaluation. */
int samos_pair_extra(samos_ctx* ctx, int kind, const double* params, int nparams, int use_radii, int shifted);
/* Sum of the further pair potentials' energies (PairPotential::get_potential_energy). */
int samos_pair_extra_energy(samos_ctx* ctx, double* energy);
/* ExternalSelfPropulsion {alpha} (src/potentials/external/external_self_propulsion.cpp:33-50). */
int samos_external_self_propulsion(samos_ctx* ctx, double alpha);

/* ---- constraint (src/constraints/constraint.hpp:47-161) ------------------ */
/* kind NONE; PLANE with param = zpos, flag = unlimited; SPHERE with param = r.
 * Like the `constraint` command (src/samos.cpp:510-522) it enforces at once.
 * Entering or leaving the plane with a cell list present re-sorts the storage
 * (see samos_nlist_create); ids, the pair set and the trajectory are
 * unchanged up to the order of the force sums. */
int samos_constraint(samos_ctx* ctx, int kind, double param, int flag);
/* Further surfaces, enforced at once like samos_constraint:
 *   CYLINDER  params {r}        (src/constraints/constraint_cylinder.hpp:42-67, .cpp:30-66)
 *   TORUS     params {c, a}     (src/constraints/constraint_torus.hpp:42-77, .cpp:26-52)
 *   ELLIPSOID params {a, b, c}  (src/constraints/constraint_ellipsoid.hpp:42-88, .cpp:26-48)
 * The torus and the ellipsoid are projected by the generic SHAKE iteration of
 * Constraint::enforce (src/constraints/constraint.cpp:40-87): at most max_iter
 * steps along the gradient taken at the entry point, until |g(x)| <= tol
 * (the reference's defaults: maxiter 100, tol 1e-6). Fails when the tube
 * radius a is not smaller than the ring radius c. */
int samos_constraint_surface(samos_ctx* ctx, int kind, const double* params, int nparams, int max_iter, double tol);
/* Sphere radius schedule {rescale; rescale_steps; rescale_freq} (Constraint
 * constructor, src/constraints/constraint.hpp:85-108): at every loop head whose
 * time step is a multiple of rescale_freq and below rescale_steps the radius is
 * multiplied by rescale^(rescale_freq / rescale_steps), every particle is
 * enforced and the neighbour list rebuilt (ConstraintSphere::rescale,
 * src/constraints/constraint_sphere.cpp:87-104; src/samos.cpp:882-886).
 * Sphere only. */
int samos_constraint_rescale(samos_ctx* ctx, double rescale, int rescale_steps, int rescale_freq);
/* The sphere's current radius. */
int samos_constraint_radius(samos_ctx* ctx, double* r);
/* Constrainer::enforce on every particle (src/samos.cpp:887-888). */
int samos_enforce_all(samos_ctx* ctx);

/* ---- integrators (src/integrators/integrator.hpp:54-138) ----------------- */
/* IntegratorBrownian / BrownianPos / BrownianAlign (integrator_brownian.hpp:55-160):
 * tau is the flip time (the reference stores dt/tau); temperature Value
 * (src/system/value.hpp) as tkind/tmin/tmax/tsteps. */
int samos_integrator_brownian(samos_ctx* ctx, int which, double v0, double nu, double mu, double mur, int seed,
                              int nematic, double tau, int velocity_align, int tkind, double tmin, double tmax,
                              int tsteps);
/* IntegratorLangevin {gamma; seed; method} (integrator_langevin.hpp:64-121). */
int samos_integrator_langevin(samos_ctx* ctx, double gamma, int seed, int method, int tkind, double tmin,
                              double tmax, int tsteps);
/* IntegratorFIRE (integrator_fire.hpp:58-154). */
int samos_integrator_fire(samos_ctx* ctx, double alpha, double alpha_scale, double f_tol, double e_tol,
                          double dt_inc, double dt_dec, double dt_max, int min_neg_steps);
/* IntegratorNVE {limit; angle_limit} (integrator_nve.hpp:59-86,
 * integrator_nve.cpp:37-141): velocity Verlet with an optional displacement
 * limit per step and director rotation limit; a negative value means the
 * parameter is not given. */
int samos_integrator_nve(samos_ctx* ctx, double limit, double angle_limit);
/* `disable <type> { group = all }` (src/samos.cpp, disable command): removes
 * the integrator registered under type. */
int samos_integrator_remove(samos_ctx* ctx, const char* type);
/* `disable type { group = g }` (src/samos.cpp:1141-1170): removes the
 * integrator registered under type_group. */
int samos_integrator_disable(samos_ctx* ctx, const char* type, const char* group);
/* `group name { type = ...; ... }` (src/samos.cpp:1108-1140, System::make_group
 * src/system/system.cpp:462-570): member[n] (id order) non-zero puts that
 * particle into the group; the caller evaluates the criteria (type, not_type,
 * r_eq, r_le, r_ge, x/y/z_le, x/y/z_ge, all of those given) on the current
 * state.  A group defined again gains the new members.  "all" (every particle)
 * always exists; "boundary" and "tissue" are reserved; at most 31 groups. */
int samos_group_define(samos_ctx* ctx, const char* name, const int* member);
int samos_group_size(samos_ctx* ctx, const char* name, int* size);
/* The `group` parameter of the next integrator created (src/samos.cpp:654-690):
 * it acts on that group's particles only (Integrator::m_group_name,
 * src/integrators/integrator.hpp:98-108) and is keyed type_group. */
int samos_set_integrator_group(samos_ctx* ctx, const char* group);
/* FIRE state out[8]: dt, alpha, P, |F|, |v|, E, converged, last_neg. */
int samos_fire_state(samos_ctx* ctx, double* out);

/* ---- one-call components ------------------------------------------------- */
/* Loop head: reset_forces + reset_torques (system.hpp:136-157), Potential::compute
 * (potential.cpp:36-52) and Aligner::compute (aligner.cpp:35-45).  energies[4]
 * (may be NULL): soft, lj, polar, nematic totals, as get_potential_energy(). */
int samos_compute_forces(samos_ctx* ctx, double* energies);
/* PairPotential::compute of one potential, added onto the current forces. */
int samos_pair_compute(samos_ctx* ctx, int kind, double* energy);
/* PairAlign::compute of one aligner, added onto the current torques. */
int samos_align_compute(samos_ctx* ctx, int kind, double* energy);
/* Integrator::integrate of every registered integrator once, in the
 * reference's map order (type_group key, src/samos.cpp:666-690). */
int samos_integrate(samos_ctx* ctx);
/* The `run N` loop (src/samos.cpp:872-970): N+1 iterations of enforce, reset,
 * forces, torques, integrators, neighbour-list check.  flags: SAMOS_RUN_*. */
int samos_run(samos_ctx* ctx, int steps, int flags);
/* Iterations t_begin .. t_end-1 of a `run steps` (t_end <= steps + 1), so a
 * driver can stop where the reference calls dumps and logs: on return the
 * state is the loop-head state of iteration t_end (after enforce, reset and
 * the force/torque evaluation, src/samos.cpp:887-898).  Ranges of one run
 * must be called in order; samos_run(steps) is samos_run_range(0, steps+1).
 * samos_run_range(0, 0, steps) performs only the loop head of iteration 0. */
int samos_run_range(samos_ctx* ctx, int t_begin, int t_end, int steps, int flags);
/* Energies of the current configuration without touching forces/torques
 * (Potential/Aligner energy getters used by the pair-energy logs,
 * src/log/log_pair_eng.hpp:56, src/log/log_pair_align.hpp:56). */
int samos_energies(samos_ctx* ctx, double* energies);
/* Scalar observables: what = 0 max|F|, 1 sum |F|^2, 2 kinetic energy as the
 * reference logs it (sum v^2 / (2 m), src/log/log_kinetic_eng.hpp:58-64), 3 |sum v| / N, 4 <|v|>, 5 |sum n| / N. */
int samos_reduce(samos_ctx* ctx, int what, double* out);

/* ---- timing -------------------------------------------------------------- */
/* When enabled, HIP events bracket every launch of the named kernel classes
 * on the context's stream ("interact", "fused", "rebuild", "integrate").
 * Reading returns the summed duration (ms) and launch count since enabling. */
int samos_profile(samos_ctx* ctx, int enable);
int samos_profile_read(samos_ctx* ctx, const char* kernel_class, double* ms_total, long long* launches);
int samos_sync(samos_ctx* ctx);
/* Tuning switches (no counterpart in the reference; results do not depend on
 * them beyond summation order): "rebuild_coop" 1 = one cooperative launch per
 * neighbour-list rebuild check, 0 = stage kernels replayed in graphs
 * (default); "gather_batch" 4 or 8 = neighbour gathers in flight per thread in
 * the fused step; "chunk_slots" (even, 2..256, default 32) = fused Langevin
 * steps per captured chunk, each chunk ending in one device-conditional
 * rebuild; "tile_cells" 0 = automatic, 1..168 = cells per workgroup tile of the
 * cell-tiled planar engine; "brownian_lanes" 0 = automatic, 1, 2, 4, 8 or 16 =
 * threads sharing one particle's neighbour list in the fused Brownian step. */
int samos_set_option(samos_ctx* ctx, const char* name, double value);
/* "neighbour_capacity" n (no counterpart in the reference, whose lists are
 * vectors): neighbour slots per particle, at least the current largest count;
 * the list keeps its pairs.  A device rebuild that finds more neighbours than
 * the capacity stops the run's captured steps before any of them reads the
 * list; the capacity then grows and the list is rebuilt from the same
 * positions, so the run continues exactly as with room to spare.
 * "resort_every" k (no counterpart in the reference): inside a run, re-sort
 * the particles' storage along the curve order of their cells after every k
 * rebuilds of the captured steps (default 200; 0 = never), keeping the
 * neighbour gathers local over long runs; the pair set and the trajectory
 * are unchanged up to the order of the force sums.
 * Run statistics, no counterpart in the reference: out[5] = step-chunk graphs
 * instantiated, chunk graphs reused, list overflows recovered, capacity,
 * storage re-sorts inside runs. */
int samos_run_stats(samos_ctx* ctx, long long* out);
/* Cell-tiled planar engine (no counterpart in the reference): planar fused
 * Langevin runs keep the state in cell order and the Verlet list as 16-bit
 * indices into each tile's LDS-staged halo (see SAMOS_RUN_TILES).  info[6]:
 * tiled runs, rebuilds inside them, largest neighbour count, largest staged
 * halo, cells per tile and list capacity of the last run. */
int samos_tile_info(samos_ctx* ctx, long long* info);

/* ---- replica ensembles ---------------------------------------------------
 * Many independent systems (replicas) of uneven size and parameters resident
 * together and advanced by one fused launch per step.  Each replica behaves
 * as the reference run alone on its own conf (System, NeighbourList,
 * PairSoftPotential, PairPolarAlign, IntegratorBrownian / IntegratorFIRE,
 * one ConstraintSphere / ConstraintPlane / none): the reference has one
 * system per process, so a study of R parameter sets is R invocations of
 * `samos run.conf`; an ensemble replaces those invocations.
 * Columns cross the boundary replica-major: replica 0's particles in its
 * input order, then replica 1's, ... */
typedef struct samos_ens samos_ens;
const char* samos_ens_last_error(const samos_ens* ens); /* NULL: error of the last create */
/* counts[nrep]; boxes[3 nrep] (Lx, Ly, Lz per replica, centred at the origin);
 * periodic[nrep]; constraint = SAMOS_CONSTRAINT_*, the same for every replica;
 * con_param[nrep]: sphere radius R (ConstraintSphere, constraint_sphere.hpp:58-62)
 * or plane height (may be NULL for the plane at z = 0). */
int samos_ens_create(samos_ens** out, int device, int nrep, const int* counts, const double* boxes,
                     const int* periodic, int constraint, const double* con_param);
int samos_ens_destroy(samos_ens* ens);
int samos_ens_num_particles(const samos_ens* ens);
/* Columns SAMOS_X .. SAMOS_RADIUS (not MASS or the int columns). */
int samos_ens_set_field(samos_ens* ens, int field, const double* host);
int samos_ens_get_field(samos_ens* ens, int field, double* host);
/* Per-replica parameters, 12 doubles each:
 *   0 k, 1 a, 2 use_particle_radii   pair_potential soft (pair_soft_potential.hpp:63-110)
 *   3 J, 4 rcut (<= 0: no aligner)    pair_align polar (pair_polar_aligner.hpp:65-110)
 *   5 v0, 6 mu, 7 mur, 8 nu, 9 seed   integrator brownian at T = 0 (integrator_brownian.hpp:60-120)
 *   10 dt                             timestep
 *   11 unused */
int samos_ens_set_params(samos_ens* ens, const double* params);
/* NeighbourList(rcut, pad) of every replica (neighbour_list.hpp:56-95): hashed
 * per-replica cell grids, each box side at least three cells of rcut + pad. */
int samos_ens_nlist(samos_ens* ens, double rcut, double pad);
/* out[7]: builds, capacity per particle, buckets, largest bucket, largest
 * neighbour count, particles, replicas. */
int samos_ens_nlist_info(samos_ens* ens, long long* out);
/* Sorted (i, j), i < j, pairs of one replica in its input order, as samos_nlist_pairs. */
int samos_ens_pairs(samos_ens* ens, int rep, int* out, long long cap, long long* count);
int samos_ens_enforce(samos_ens* ens);
/* Forces and torques of every replica; energies[nrep] (may be NULL): soft energy. */
int samos_ens_forces(samos_ens* ens, double* energies);
/* `run steps` with the Brownian integrator in every replica (steps + 1
 * iterations, src/samos.cpp:878-970).  flags: SAMOS_RUN_NO_GRAPH. */
int samos_ens_run_brownian(samos_ens* ens, int steps, int flags);
/* Autotuner (no counterpart in the reference): times trial_steps Brownian
 * steps for every candidate skin pads[i] (cell size rcut + pad) and neighbour
 * gather batch 4 / 8, each from a saved copy of the state, then restores the
 * state and keeps the fastest.  out[3] (may be NULL): pad, batch, ms per step.
 * Trajectories differ between candidates only by summation order. */
int samos_ens_autotune(samos_ens* ens, int trial_steps, const double* pads, int npads, double* out);
/* IntegratorFIRE parameters for every replica (integrator_fire.hpp:58-154) and
 * the stop rule max|F| < stop_fmax that freezes a replica (no counterpart:
 * the reference's FIRE never stops the run loop). */
int samos_ens_fire_setup(samos_ens* ens, double dt, double alpha, double alpha_scale, double f_tol, double e_tol,
                         double dt_inc, double dt_dec, double dt_max, int min_neg_steps, double stop_fmax);
/* Up to max_iter FIRE run-loop iterations, stopping once every replica has met
 * the stop rule (tested every `check` iterations). */
int samos_ens_fire_run(samos_ens* ens, int max_iter, int check, int* remaining, int* iters);
/* out[8 nrep]: iterations, stopped (1/0), energy, max|F|, |F|, dt, alpha, P per replica. */
int samos_ens_fire_state(samos_ens* ens, double* out);
/* "gather_batch" 4 or 8. */
int samos_ens_set_option(samos_ens* ens, const char* name, double value);
/* HIP events around every launch of the per-particle step kernel ("step":
 * the Brownian step or the FIRE force pass) and every rebuild sequence
 * ("rebuild"); launches go step by step while enabled. */
int samos_ens_profile(samos_ens* ens, int enable);
int samos_ens_profile_read(samos_ens* ens, const char* kernel_class, double* ms_total, long long* launches);

/* ---- warm-start network for FIRE relaxation --------------------------------
 * No counterpart in the reference (which relaxes with IntegratorFIRE only,
 * src/integrators/integrator_fire.cpp:34-148): a message-passing network on
 * an ensemble's neighbour list that predicts each particle's displacement to
 * its FIRE minimum, trained on (start, relaxed) pairs from batched FIRE.  The
 * ensemble's replicas are 2-D configurations (plane constraint); the network
 * sees radii and pair geometry only.  Bfloat16 MFMA products with FP32
 * accumulation; FP32 parameters (flat order in samos_amd/csrc/mpnn.hip). */
typedef struct samos_mpnn samos_mpnn;
const char* samos_mpnn_last_error(const samos_mpnn* net); /* NULL: error of the last create */
int samos_mpnn_num_params(void);
int samos_mpnn_create(samos_mpnn** out, int device, unsigned long long seed);
int samos_mpnn_destroy(samos_mpnn* net);
int samos_mpnn_get_params(samos_mpnn* net, float* out);
int samos_mpnn_set_params(samos_mpnn* net, const float* in);
/* Displacements of the ensemble's current configuration; dxdy[2 n] (x, y per
 * particle, replica-major input order) may be NULL. */
int samos_mpnn_predict(samos_mpnn* net, samos_ens* ens, double* dxdy);
/* Loss (1/2N) sum |delta - target|^2 and its gradient (flat order) for the
 * current configuration; no update. */
int samos_mpnn_gradients(samos_mpnn* net, samos_ens* ens, const double* target_dxdy, float* grad, double* loss);
/* x <- x + scale * delta for every particle, then enforce and rebuild lists. */
int samos_mpnn_warm_start(samos_mpnn* net, samos_ens* ens, double scale);
/* Training pairs held on the device: reserve, then for each batch call
 * dataset_begin before relaxing the ensemble and dataset_end after. */
int samos_mpnn_dataset_reserve(samos_mpnn* net, int configs, int per_config);
int samos_mpnn_dataset_begin(samos_mpnn* net, samos_ens* ens);
int samos_mpnn_dataset_end(samos_mpnn* net, samos_ens* ens);
int samos_mpnn_dataset_size(const samos_mpnn* net);
/* Adam steps on minibatches of `work`'s replica count drawn from dataset
 * configurations [first, first + count); losses[steps] may be NULL. */
int samos_mpnn_train(samos_mpnn* net, samos_ens* work, int steps, double lr, unsigned long long seed,
                     int first_config, int num_configs, double* losses);
int samos_mpnn_evaluate(samos_mpnn* net, samos_ens* work, int first_config, int num_configs, double* mean_loss);
/* HIP events around every MFMA product launch: total ms, flops and launches. */
int samos_mpnn_profile(samos_mpnn* net, int enable);
int samos_mpnn_profile_read(samos_mpnn* net, double* ms_total, double* flops, long long* launches);
int samos_ens_sync(samos_ens* ens);

#ifdef __cplusplus
}
#endif

#endif /* SAMOS_HIP_H */
