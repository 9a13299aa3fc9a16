/*
 * tfg.h -- C ABI of the MI355X engine_step module (libtfg.so).
 *
 * Plain C types only.  Every function returns an HRESULT-style code
 * (0 = success, negative = failure; reference include/tf_port.h:255-282)
 * unless stated otherwise, and leaves a message for tfg_last_error()
 * (the role of the reference's error stack, source/tfError.h:78).
 *
 * Each declaration names the reference interface whose role it takes
 * (paths relative to the reference repository).
 */
#ifndef TFG_H
#define TFG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tfg_engine tfg_engine;
typedef struct tfg_potential tfg_potential;
typedef struct tfg_ensemble tfg_ensemble;

#define TFG_OK 0
#define TFG_E_FAIL ((int)0x80004005)
#define TFG_E_INVALIDARG ((int)0x80070057)
#define TFG_E_OUTOFMEMORY ((int)0x8007000E)
#define TFG_E_RANGE ((int)0x8000000B)
#define TFG_E_STATE ((int)0x8000FFFF)
#define TFG_E_DEVICE ((int)0x80004006)
#define TFG_E_TOOFAST ((int)0x80004007)

/* Precision of the device path. */
#define TFG_PREC_MIXED 0 /* FP32 pair/bonded/integrator arithmetic, FP64 reductions */
#define TFG_PREC_FP64 1  /* all FP64 */

/* Dynamics (source/mdcore/include/tfParticle.h:63-64). */
#define TFG_NEWTONIAN 0
#define TFG_OVERDAMPED 1
/* Integrators (source/mdcore/include/tfEngine.h:89-92). */
#define TFG_FORWARD_EULER 0
#define TFG_RUNGE_KUTTA_4 1
/* Sum rules of tfg_engine_set_sum_rule. */
#define TFG_SUM_FIRST_IN_RANGE 0
#define TFG_SUM_ALL 1
/* Boundary condition kinds (BoundaryConditionKind,
 * source/mdcore/include/tfBoundaryConditions.h:38-46) and faces. */
#define TFG_BOUNDARY_VELOCITY 1u
#define TFG_BOUNDARY_PERIODIC 2u
#define TFG_BOUNDARY_FREESLIP 4u
#define TFG_BOUNDARY_POTENTIAL 8u
#define TFG_BOUNDARY_NO_SLIP 16u
#define TFG_FACE_LEFT 0
#define TFG_FACE_RIGHT 1
#define TFG_FACE_FRONT 2
#define TFG_FACE_BACK 3
#define TFG_FACE_BOTTOM 4
#define TFG_FACE_TOP 5
/* Flux kinds (source/mdcore/include/tfFlux.h:40-42). */
#define TFG_FLUX_FICK 0
#define TFG_FLUX_SECRETE 1
#define TFG_FLUX_UPTAKE 2
/* DPD random term: unit (the reference's sigma*omega/sqrt(dt),
 * source/mdcore/src/tf_dpd_eval.h:64) or a Philox4x32-10 draw per pair and step. */
#define TFG_DPD_NOISE_UNIT 0
#define TFG_DPD_NOISE_PHILOX 1

/* Last error message of the calling thread (source/tfError.h:78). */
const char *tfg_last_error(void);
const char *tfg_version(void);

/* ---- engine lifecycle: engine_init (source/mdcore/src/tfEngine.cpp:1702-1874)
 *      as called by universe_init (source/tfSimulator.cpp:629-666). ---- */
int tfg_engine_create(tfg_engine **out, const double origin[3], const double dim[3],
                      const int cells[3], double cutoff, int max_types, double dt,
                      int nr_fluxsteps);
int tfg_engine_destroy(tfg_engine *e);
/* space_init geometry (source/mdcore/src/tfSpace.cpp:737-770, 893-895). */
int tfg_engine_geometry(const tfg_engine *e, int cdim[3], double h[3], int span[3]);
/* Universe max_distance -> particle_max_dist_fraction (source/tfSimulator.cpp:661-666). */
int tfg_engine_set_max_distance_fraction(tfg_engine *e, double fraction);
/* Universe integrator (source/tfSimulator.cpp:218-229, EngineIntegrator of
 * source/mdcore/include/tfEngine.h:89-92): TFG_FORWARD_EULER or
 * TFG_RUNGE_KUTTA_4 (engine_advance_runge_kutta_4,
 * source/mdcore/src/tf_engine_advance.cpp:499-763); any other kind is
 * TFG_E_INVALIDARG, as the reference rejects it. */
int tfg_engine_set_integrator(tfg_engine *e, int kind);
int tfg_engine_integrator(const tfg_engine *e);
/* How a sum bound as a pair potential adds up: TFG_SUM_FIRST_IN_RANGE, the
 * CPU engine's potential_eval_super_ex (source/mdcore/src/tf_potential_eval.h:
 * 211-215: pca || pcb, so the first constituent whose range holds the pair
 * is the only one evaluated), or TFG_SUM_ALL, the reference's GPU path
 * (potential_eval_super_ex_cuda, source/mdcore/src/tfRunner_cuda.cu:1432-1474,
 * every constituent packed by engine_cuda_build_pots_pack, :436-550, each in
 * range adding its term).  Default TFG_SUM_FIRST_IN_RANGE. */
int tfg_engine_set_sum_rule(tfg_engine *e, int rule);
int tfg_engine_sum_rule(const tfg_engine *e);
/* One face's boundary condition: BoundaryCondition::init
 * (source/mdcore/src/tfBoundaryConditions.cpp:151-171; NO_SLIP becomes a
 * velocity face with zero velocity), restore and the wall radius
 * (tfBoundaryConditions.h:57-74).  Faces start periodic; when one face of an
 * axis is periodic both are (check_periodicy, tfBoundaryConditions.cpp:192-217).
 * Kinds are bit sets of TFG_BOUNDARY_*; "potential" faces of the reference's
 * string form are TFG_BOUNDARY_POTENTIAL | TFG_BOUNDARY_FREESLIP (:112).
 * Resetting faces are not supported (TFG_E_INVALIDARG). */
int tfg_engine_set_boundary(tfg_engine *e, int face, uint32_t kind, const double velocity[3],
                            double restore, double radius);
/* BoundaryCondition::set_potential / BoundaryConditions::set_potential
 * (source/mdcore/src/tfBoundaryConditions.cpp:504-525): the wall potential of
 * a particle type on one face, or on all six with face = -1; NULL clears it.
 * Evaluated by boundary_eval (source/mdcore/src/tf_boundary_eval.h:204-254). */
int tfg_engine_set_wall_potential(tfg_engine *e, int face, int type, const tfg_potential *p);
/* Neighbour lists with a skin: the role of engine_flag_verlet /
 * engine_verlet_update (source/mdcore/src/tfEngine.cpp:214-354,
 * tf_runner_verlet.cpp).  Lists hold every pair within cutoff + skin at their
 * build and are rebuilt, on the device, once a particle has moved more than
 * skin / 2 since (the reference's 2 maxdx > skin, tfEngine.cpp:287-288).  The
 * pair set of every step is the one engine_step uses, r^2 <= cutoff^2, and
 * the integrator is unchanged (the reference's Verlet mode also swaps the
 * integrator; that is not reproduced).  skin = 0 turns lists off; capacity
 * is the entries per particle (0: from the density); a step that would
 * exceed it fails with TFG_E_RANGE.  Lists are used when every engine of the
 * device copy asks for them, integrates with forward Euler, has tabulated
 * and DPD potentials (with species fluxes, but no radius transforms, sums or
 * periodic images) whose tables fit the LDS cache, and its cutoff + skin
 * spans at most two cells per axis and less than half the box along a
 * periodic axis; otherwise the cell pass runs. */
int tfg_engine_set_verlet(tfg_engine *e, double skin, int capacity);
/* Whether the device copy steps with lists, the rebuilds so far and the
 * capacity; zeros when the engine is not on the device. */
int tfg_engine_verlet_state(const tfg_engine *e, int *active, int64_t *rebuilds, int *capacity);
/* Mean and largest number of list entries per particle at the last build
 * (zeros without lists).  No counterpart in the reference: diagnostics of
 * the list sizes for choosing the skin and the capacity. */
int tfg_engine_verlet_entries(const tfg_engine *e, double *mean, int *maxn);
/* Species count per particle (the state vector of source/state/, flat here). */
int tfg_engine_set_nspecies(tfg_engine *e, int nspecies);
/* DPD random term: TFG_DPD_NOISE_UNIT or TFG_DPD_NOISE_PHILOX with its key. */
int tfg_engine_set_dpd_noise(tfg_engine *e, int mode, uint64_t seed);
/* Step counter (engine->time, source/mdcore/include/tfEngine.h:183). */
int64_t tfg_engine_time(const tfg_engine *e);
/* Species count per particle, as set (for an engine made by tfg_engine_load). */
int tfg_engine_nspecies(const tfg_engine *e);

/* ---- engine state checkpoint: io::FIO::toFile / fromFile / toString
 *      (source/io/tfFIO.h:247-276).  The document is JSON: geometry as
 *      requested, Universe settings (dt, flux sub-steps, integrator, distance
 *      fraction, step counter, DPD noise, list settings), boundary faces and
 *      wall potentials, types, every table with its coefficients, bindings,
 *      fluxes, particles in id order (type, flags, cell id, position relative
 *      to the cell origin, velocity, last force, inverse mass, radius),
 *      species and pending flux, bonds, angles and the last energies.  Numbers
 *      are written so that they read back to the same double: a save / load
 *      round trip is exact. ---- */
/* Writes the state to path.  An engine on the device is read back first and
 * stays on the device. */
int tfg_engine_save(tfg_engine *e, const char *path);
/* The same document in buf (zero-terminated).  *len receives its length
 * without the terminator; with buf NULL only *len is set.  A capacity below
 * *len + 1 fails with TFG_E_RANGE and sets *len. */
int tfg_engine_to_string(tfg_engine *e, char *buf, int64_t cap, int64_t *len);
/* A new host engine from a saved document (not on the device: call
 * tfg_engine_to_device to continue stepping).  A document that is not an
 * engine state, or whose contents are inconsistent, fails with
 * TFG_E_INVALIDARG and names what is wrong. */
int tfg_engine_load(tfg_engine **out, const char *path);
int tfg_engine_from_string(tfg_engine **out, const char *text, int64_t len);

/* ---- particle types: engine_addtype (source/mdcore/src/tfEngine.cpp:1040-1052).
 *      Returns the new type id (>= 0) or an error. ---- */
int tfg_engine_addtype(tfg_engine *e, const char *name, double mass, double radius, int dynamics);

/* ---- potentials: Potential::morse / lennard_jones_12_6 / harmonic /
 *      harmonic_angle / dpd (source/mdcore/src/tfPotential.cpp:2741-2751,
 *      2961-2975, 2991-3003, 3068-3084, 3128-3142).  Tables are built by
 *      potential_init (:1396-1634).  Negative tol selects the reference default. ---- */
int tfg_potential_morse(tfg_potential **out, double d, double a, double r0, double min,
                        double max, double tol, int shifted);
int tfg_potential_lennard_jones_12_6(tfg_potential **out, double min, double max, double A,
                                     double B, double tol);
int tfg_potential_harmonic(tfg_potential **out, double k, double r0, double min, double max,
                           double tol);
int tfg_potential_harmonic_angle(tfg_potential **out, double k, double theta0, double min,
                                 double max, double tol);
int tfg_potential_dpd(tfg_potential **out, double alpha, double gamma, double sigma,
                      double cutoff, int shifted);
/* Potential::overlapping_sphere (source/mdcore/src/tfPotential.cpp:1224-1307,
 * 3086-3102; reference defaults mu 1, kc 1, kh 0, r0 0, min 0.001, max 10,
 * tol 0.001): a table in the scaled distance r / (ri + rj) (POTENTIAL_SCALED).
 * Negative tol selects the default. */
int tfg_potential_overlapping_sphere(tfg_potential **out, double mu, double kc, double kh, double r0,
                                     double min, double max, double tol);
/* Potential::glj (source/mdcore/src/tfPotential.cpp:2085-2185, 3047-3066;
 * reference defaults m 3, n 2m, k 0, r0 1, min 0.05 r0, max 3 r0, tol 0.01,
 * not shifted): shifted evaluates at r - (ri + rj) + r0 + 1, otherwise at the
 * scaled distance r / (ri + rj).  Negative tol selects the default. */
int tfg_potential_glj(tfg_potential **out, double e, double m, double n, double k, double r0, double min,
                      double max, double tol, int shifted);
/* A table potential from its own record: the fields of struct Potential that
 * potential_eval_ex reads (source/mdcore/include/tfPotential.h:217-252) and
 * that the reference's GPU path packs per potential
 * (engine_cuda_build_pots_pack, source/mdcore/src/tfRunner_cuda.cu:391-550):
 * flags (PotentialFlags, tfPotential.h:56-108; SHIFTED and SCALED select the
 * distance transform), the range [a, b], r0_plusone, alpha[0..2], n
 * intervals and the 8 (n + 1) coefficients c.  The table is evaluated as
 * given, so a binding hands over the reference's own Potential bit for bit.
 * Fails with TFG_E_INVALIDARG unless 1 <= n <= 3048, ncoef = 8 (n + 1), a < b
 * and every number is finite (except the 8 coefficients of interval 0, the
 * continuation left of a, which the evaluation never reaches for r >= a). */
int tfg_potential_from_table(tfg_potential **out, uint32_t flags, double a, double b, double r0_plusone,
                             const double alpha[3], int n, const double *c, int64_t ncoef);
int tfg_potential_destroy(tfg_potential *p);
/* Table description: kind, flags, n, a, b, r0_plusone, alpha[0..2]. */
int tfg_potential_info(const tfg_potential *p, int *kind, uint32_t *flags, int *n,
                       double *a, double *b, double *r0_plusone, double alpha[3]);
/* Copies up to cap coefficients (8*(n+1) in total); returns the count. */
int tfg_potential_coeffs(const tfg_potential *p, double *out, int cap);
/* Potential::operator() / force at distance r (potential_eval_ex,
 * source/mdcore/src/tf_potential_eval.h:104-146): energy and force/r. */
int tfg_potential_eval(const tfg_potential *p, double r, double ri, double rj, double *e,
                       double *f_over_r);
/* Potential::operator+ (source/mdcore/src/tfPotential.cpp:2706-2727): the sum
 * of a and b, flattened into its constituents as Potential::constituents
 * (:2684-2704) lists them (a's, then b's).  Fails with TFG_E_INVALIDARG when
 * a is an angle potential and b is not ("Incompatible potentials") or when
 * the sum would hold more than 7 constituents (a base and its images on
 * both sides of three periodic axes, as create_periodic_potential makes
 * them at order 1, tfPotential.cpp:158-179).  Constituents are tables or
 * DPD; a and b are copied, so either may be destroyed afterwards.  How the
 * pair pass adds the constituents up is the engine's sum rule
 * (tfg_engine_set_sum_rule). */
int tfg_potential_sum(tfg_potential **out, const tfg_potential *a, const tfg_potential *b);
/* The periodic image offset of a potential (Potential::offset with
 * POTENTIAL_PERIODIC, the record _do_create_periodic_potential writes,
 * source/mdcore/src/tfPotential.cpp:136-155): the pair is evaluated at
 * dx - offset (tf_potential_eval.h:220-225; tfRunner_cuda.cu:1381-1391),
 * with dx taken from the particle of lower id to the other.  (The reference
 * takes dx in its traversal order; the two agree whenever the result does
 * not depend on that order, as for the mirrored images the reference
 * creates, evaluated under TFG_SUM_ALL.)  Sets the flag; a sum takes offsets
 * per constituent (TFG_E_INVALIDARG). */
int tfg_potential_set_offset(tfg_potential *p, const double offset[3]);
int tfg_potential_offset(const tfg_potential *p, double offset[3]);
/* Constituent count of a sum (0 for a single potential) and a copy of
 * constituent k (destroy it with tfg_potential_destroy). */
int tfg_potential_terms(const tfg_potential *p);
int tfg_potential_term(const tfg_potential *p, int k, tfg_potential **out);

/* ---- bindings: engine_addpot (source/mdcore/src/tfEngine.cpp:1054-1077). ---- */
int tfg_engine_addpot(tfg_engine *e, const tfg_potential *p, int type_i, int type_j);
/* Fluxes::create (source/mdcore/src/tfFlux.cpp:56-112); cutoff <= 0 means global. */
int tfg_engine_addflux(tfg_engine *e, int kind, int type_a, int type_b, int species_a,
                       int species_b, double k, double decay, double target, double cutoff);

/* ---- particles: space_addpart / space_setpartp (source/mdcore/src/tfSpace.cpp:432-475).
 *      x, v: n*3 global coordinates; species: n*nspecies or NULL.
 *      ids_out receives the new particle ids. ---- */
int tfg_engine_addparts(tfg_engine *e, int n, const int *types, const double *x,
                        const double *v, const double *species, int32_t *ids_out);
int tfg_engine_set_particle_flags(tfg_engine *e, int n, const int32_t *ids, const uint32_t *flags);
int tfg_engine_nparts(const tfg_engine *e);

/* ---- bonded topology: Bond::create / engine_bond_alloc
 *      (source/mdcore/src/tfBond.cpp:636-654, tf_engine_bonded.cpp:1140) and
 *      Angle::create / engine_angle_alloc (tf_engine_bonded.cpp:975). ---- */
int tfg_engine_addbonds(tfg_engine *e, const tfg_potential *p, int n, const int32_t *ij);
int tfg_engine_addangles(tfg_engine *e, const tfg_potential *p, int n, const int32_t *ijk);

/* ---- device toggle: cuda::engine_toCUDA / engine_fromCUDA
 *      (source/mdcore/src/tfEngine.cpp:1136-1197) as wrapped by
 *      cuda::EngineConfig::toDevice / fromDevice / setDevice
 *      (source/cuda/tfEngineConfig.cpp:67-172). ---- */
int tfg_device_count(void);
int tfg_device_set(int ordinal);
/* cuda::EngineConfig::getDevice (source/cuda/tfEngineConfig.cpp:42-45): the device
 * that uploads from the calling thread use; -1 without a device. */
int tfg_device_get(void);
/* cuda::EngineConfig::clearDevice (source/cuda/tfEngineConfig.cpp:57-65): back to device 0. */
int tfg_device_clear(void);
/* Uploads the engine; a second call while on the device is a no-op.  An
 * engine with no particles fails with TFG_E_INVALIDARG. */
int tfg_engine_to_device(tfg_engine *e, int precision);
/* Downloads x, v, f, cell ids and species, frees the device copy; a no-op
 * when the engine is not on the device (source/mdcore/src/tfEngine.cpp:1178-1179). */
int tfg_engine_from_device(tfg_engine *e);
/* cuda::EngineConfig::refresh / refreshPotentials / refreshFluxes
 * (source/cuda/tfEngineConfig.cpp:126-172): the device copy is rebuilt from
 * the current state.  Every setter above (types, potentials, fluxes,
 * particles, flags, bonds, angles, noise) does this by itself when the engine
 * is on the device, as engine_addpot / engine_addfluxes / space_addpart do
 * (source/mdcore/src/tfEngine.cpp:1070-1073, 1093-1096, tfSpace.cpp:493-496). */
int tfg_engine_refresh(tfg_engine *e);
int tfg_engine_on_device(const tfg_engine *e);
/* cuda::EngineConfig::setBlocks / setThreads (source/cuda/tfEngineConfig.cpp:98-124):
 * workgroups of the pair kernel, threads per workgroup of the streaming
 * kernels (a multiple of 64 up to 1024), applied at the next
 * tfg_engine_to_device (0 keeps the default).  As in the reference they fail
 * while the engine is on the device (TFG_E_STATE, the reference's
 * TFCUDAERR_ondevice). */
int tfg_engine_set_blocks(tfg_engine *e, int blocks);
int tfg_engine_set_threads(tfg_engine *e, int threads);
int tfg_engine_launch_config(const tfg_engine *e, int *blocks, int *threads);
/* Launch shape of the device copy itself, while on the device (no reference
 * counterpart: the reference sets the shape only before the upload).  The
 * autotuner's results can be replaced with it; no shape changes a result bit. */
int tfg_set_launch(tfg_engine *e, int blocks, int threads);
/* Launch-shape autotuner (no reference counterpart: the reference takes
 * setBlocks / setThreads by hand).  Times reps force-only passes per
 * candidate grid of the pair kernel, then of the streaming kernels, and keeps
 * the fastest; the chosen workgroup counts are returned.  Only shapes that
 * cannot change a result bit are tuned.  The state is left as it was. */
int tfg_autotune(tfg_engine *e, int reps, int *blocks, int *stream_blocks);

/* ---- step: engine_step (source/mdcore/src/tfEngine.cpp:1378-1429), whole step on
 *      the device (binning, nonbonded, bonded, flux, forward Euler).  A step
 *      that fails (a particle moving two cells or more: TFG_E_TOOFAST; a
 *      neighbour list over its capacity: TFG_E_RANGE) ends the call there, as
 *      Universe_Step stops at the first failing engine_step
 *      (source/tfUniverse.cpp:368-370): later steps do nothing, the step
 *      counter includes the failing step (engine_step increments time first,
 *      tfEngine.cpp:1388), and the state is the one the failing step left
 *      (after its update when a particle moved too fast, that of the step
 *      before for a list overflow). ---- */
int tfg_engine_step(tfg_engine *e, int nsteps);
/* engine_force only (source/mdcore/src/tfEngine.cpp:1504-1557): binning, nonbonded
 * and bonded forces, no update.  The role of cuda::engine_nonbond_cuda
 * (source/mdcore/src/tfRunner_cuda.cu:2940-3354) plus the bonded pass. */
int tfg_nonbond_eval(tfg_engine *e);
/* The pass of tfg_nonbond_eval for an engine on the host, in one call that
 * allocates no device memory: upload, binning, nonbonded and bonded forces
 * and the energies, with the forces and number density of each particle by
 * id into f (n*4) and the six values of tfg_read_energies into energies (may
 * be NULL).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Every device buffer of the pass is carved from one fixed
 * 64 MiB array that each GPU the library runs on holds until the process
 * ends: about 300,000 particles in mixed precision, 200,000 in FP64 (fewer
 * with species or bonds); a larger system fails with TFG_E_RANGE before any
 * GPU work.  The pass is the cell kernel (neighbour-list settings do not
 * apply) on the null stream; the engine's state is unchanged and the call
 * returns after the GPU work has drained.  Calls on one GPU share the
 * array, so a per-device lock makes overlapping calls from several threads
 * run one after the other.  TFG_E_STATE while the engine is on the device,
 * TFG_E_INVALIDARG for an engine with no particles. */
int tfg_nonbond_eval_noalloc(tfg_engine *e, int precision, double *f, double energies[6]);

/* ---- readback (synchronous). ---- */
/* out[0..5] = epot, epot_nonbond, epot_bond, epot_angle, ekin, virial of the
 * last evaluated step (space epot fields, source/mdcore/include/tfSpace.h:207;
 * engine_kinetic_energy, source/mdcore/src/tfEngine.cpp:1893-1927). */
int tfg_read_energies(tfg_engine *e, double out[6]);
/* Per particle, indexed by id (any pointer may be NULL): cell id, position
 * relative to the cell origin (n*3), velocity (n*3), force and number
 * density (n*4), species (n*nspecies), pending flux (n*nspecies).  Downloads
 * from the device without leaving it. */
int tfg_read_particles(tfg_engine *e, int32_t *cid, double *x, double *v, double *f,
                       double *species, double *species_q);
/* Global positions origin + loc*h + x (n*3). */
int tfg_read_positions(tfg_engine *e, double *xg);
/* FP64 per-species totals (flux conservation); returns the count. */
int tfg_read_species_totals(tfg_engine *e, double *out, int cap);
/* Device time of the last tfg_engine_step call in ms. */
double tfg_last_step_ms(const tfg_engine *e);
/* Per-phase device time: [bin, list build, pair pass, advance] of the last
 * step, then the averages over all steps of the last call when profiling is
 * on (the list build is 0 without neighbour lists); returns the count. */
int tfg_kernel_times(const tfg_engine *e, double *ms, int cap);
/* Records HIP events around every phase of every step (the role of the
 * reference's engine timers, source/mdcore/include/tfEngine.h:95-145). */
int tfg_set_profiling(tfg_engine *e, int on);

/* ---- batched ensemble: independent systems advanced by the same launches
 *      (no reference counterpart: the reference engine is a process-global
 *      singleton, source/mdcore/include/tfEngine.h:1390).  Every member
 *      needs at least one particle (TFG_E_INVALIDARG otherwise). ---- */
int tfg_ensemble_create(tfg_ensemble **out, tfg_engine *const *engines, int n, int precision);
int tfg_ensemble_step(tfg_ensemble *ens, int nsteps, int use_graph);
int tfg_ensemble_destroy(tfg_ensemble *ens, int download);
double tfg_ensemble_last_ms(const tfg_ensemble *ens);
int tfg_ensemble_set_launch(tfg_ensemble *ens, int blocks, int threads);
int tfg_ensemble_autotune(tfg_ensemble *ens, int reps, int *blocks, int *stream_blocks);

#ifdef __cplusplus
}
#endif

#endif /* TFG_H */
