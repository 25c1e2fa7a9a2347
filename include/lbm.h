/*
 * liblbm -- MI355X (gfx950) D3Q19 MRT channel-flow core, C ABI.
 *
 * Drop-in boundary for the Channel-Flow hot path of UDel-CFD/D3Q19-Single-Phase.
 * The reference's entry points are argument-free Fortran subroutines acting
 * on module globals (Channel-Flow/var_inc.f90).  Here the lattice lives on the
 * GPU inside an opaque handle; every reference subroutine on the path maps to
 * one function below that takes the handle plus plain pointers and sizes.
 * A Fortran ISO_C_BINDING shim keeps the original subroutine names.
 *
 * Axes follow the reference: x = wall-normal (lx nodes, halfway bounce-back
 * walls beyond both ends), y = streamwise (ly, periodic), z = spanwise (lz,
 * periodic).  Host arrays use the reference layout: f(0:18,lx,ly,lz) and
 * macroscopic fields (lx,ly,lz), Fortran column-major.
 *
 * Every function returns 0 (LBM_OK) on success or a positive LBM_ERR_* code;
 * lbm_last_error() then describes the failure.  Nothing aborts inside the
 * library.  One host thread per handle; calls on a handle are enqueued on the
 * handle's own HIP stream, and functions that return host data synchronise.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_ARG 1     /* invalid argument or configuration */
#define LBM_ERR_HIP 2     /* HIP runtime error (message names the call) */
#define LBM_ERR_NOMEM 3   /* device or host allocation failed */
#define LBM_ERR_STATE 4   /* call not valid in the handle's current state */

/* Storage precision of the populations.  The reference's f is already the
 * shifted population f - w_i*rho0 (Channel-Flow/initial.f90:29-44: rho holds
 * the density fluctuation only); FP32/FP16/FP16C keep that same quantity at
 * lower precision.  LBM_FP64 computes in FP64; the others compute in FP32.
 * LBM_FP16 stores IEEE half values with per-direction power-of-two scaling;
 * LBM_FP16C stores the same scaled values in a 16-bit format with 1 sign, 4
 * exponent and 11 significand bits (normal range [2^-13, 4 - 2^-10],
 * denormals to 2^-24, saturating above; Lehmann et al., Phys. Rev. E 106,
 * 015308, 2022): one significand bit more than FP16, for the same traffic. */
#define LBM_FP64 0
#define LBM_FP32 1
#define LBM_FP16 2
#define LBM_FP16C 3

/* Run-time constants.  Field names and meanings follow
 * Channel-Flow/var_inc.f90:85-93 as filled by Channel-Flow/para.f90:106-174;
 * the driver computes them exactly as para.f90 does and passes them here, so
 * nothing is recomputed differently.  The structural constants (coef*, val*,
 * ww*) must equal para.f90's values: lbm_create rejects others. */
typedef struct lbm_config {
    int lx, ly, lz;     /* node counts: var_inc.f90:51-52 (lx), para.f90:234-245 (ly, lz) */
    int precision;      /* LBM_FP64 / LBM_FP32 / LBM_FP16 / LBM_FP16C */
    int device;         /* HIP device ordinal */
    int reserved;       /* must be 0 */
    double s1, s2, s4, s9, s10, s13, s16;             /* para.f90:106-141 */
    double omegepsl, omegepslj, omegxx;               /* para.f90:118-140 */
    double coef1, coef2, coef3, coef4, coef5, coef3i, coef4i; /* para.f90:143-150 */
    double val8;                                      /* para.f90:159 */
    double val1i, val2i, val3i, val4i, val5i, val6i, val7i, val8i, val9i; /* para.f90:162-170 */
    double ww0, ww1, ww2;                             /* para.f90:172-174 */
} lbm_config;

typedef struct lbm_channel lbm_channel;

/* Message for the last failing call on this host thread ("" if none). */
const char *lbm_last_error(void);

/* Allocates the device lattices for one rank's (lx,ly,lz) block -- replaces
 * Channel-Flow/para.f90:418-447 (allocarray) for f, rho, ux, uy, uz, force.
 * The populations start at zero; the macroscopic fields read as macrovar of
 * the lattice until a call sets them.  The step runs one node per thread;
 * the environment variable LBM_STEP_KERNEL=runs, read here, makes FP64 and
 * FP32 handles step two nodes per thread with 2-wide vector accesses when lx
 * is even (FP64 results identical bit for bit, FP32 to round-off; measured
 * slower on the 192x768x384 channel). */
int lbm_create(const lbm_config *cfg, lbm_channel **out);
int lbm_destroy(lbm_channel *h);

/* Host <-> device copies of f(0:18,lx,ly,lz) in FP64, converting to and from
 * the storage precision.  Upload replaces Channel-Flow/saveload.f90:152-156
 * (loadinitflow) and :327-330 (loadcntdflow) and initial.f90:34-44 (initpop
 * from host data); download feeds saveinitflow/savecntdflow
 * (saveload.f90:119-120, :226-227). */
int lbm_upload_f(lbm_channel *h, const double *f);
int lbm_download_f(lbm_channel *h, double *f);

/* FORCING, Channel-Flow/collision.f90:515-527: a uniform body force
 * (force_realx, force_realy, force_realz).  Fields formed by an earlier
 * macrovar keep the force they were formed with, as the reference's arrays
 * do, until the next macrovar. */
int lbm_set_force(lbm_channel *h, double fx, double fy, double fz);

/* FORCINGP, Channel-Flow/collision.f90:529-602 (one rank): the force arrays
 * become the mean force (0, force_in_y, 0) plus the perturbation of time step
 * istep, a divergence-free field of 3 streamwise and 2 spanwise waves in two
 * bands of lx/4 wall-normal nodes, amplitude 40*3/ly*sin(2 pi istep/2000).
 * The driver calls it before collision_MRT while istep < npforcing and
 * FORCING at istep == npforcing (main.f90:148-153).  The force is evaluated
 * per node inside the step from tables of one coordinate each (no 3 x N
 * force arrays).  Not combined with particles (main.f90:148). */
int lbm_forcingp(lbm_channel *h, int istep, double force_in_y);

/* nsteps iterations of the main loop with the perturbation forcing switched
 * on (main.f90:142-161 with the block at :148-153), for istep = istep0 + 1 ..
 * istep0 + nsteps: lbm_forcingp(istep, force_in_y) while istep < npforcing,
 * lbm_set_force(0, fy_forcing, 0) at istep == npforcing (FORCING, with
 * fy_forcing = force_in_y*force_mag), then collision_MRT and macrovar.  Pairs
 * of steps under one uniform force run as one blocked sweep (see lbm_run);
 * FORCINGP steps run one by one, or also in pairs with the force evaluated
 * per node when LBM_BLOCK_FORCED=1 at lbm_create (measured slower).  Results
 * equal those calls one by one, bit for bit.  No reference counterpart as
 * one call: it batches the driver's loop. */
int lbm_run_forcingp(lbm_channel *h, int istep0, int nsteps, int npforcing, double force_in_y,
                     double fy_forcing);

/* Sets the macroscopic fields the next collision reads, as the driver does by
 * assigning the module arrays (initial.f90:94-115 initvel, :32 rho = 0).
 * Any pointer may be NULL to keep that field. */
int lbm_set_macros(lbm_channel *h, const double *rho, const double *ux, const double *uy,
                   const double *uz);

/* Sets the fields the next collision reads to rho = 0 (initial.f90:32) and a
 * velocity that is a sum of separable terms,
 *   u_c(ix, iy, iz) = sum over k with comp[k] == c of planes[k][iy + ly*iz] * profiles[k][ix],
 * each component summed in increasing k from 0 (comp[k] = 0, 1, 2 for ux,
 * uy, uz; planes nterms x ly*lz, profiles nterms x lx, nterms <= 64).  For
 * initial fields of large lattices (initvel, initial.f90:75-147: a mean
 * profile in x plus perturbations) without 3 x N host arrays, which take
 * 25.8 GB for the 2048x1024x512 channel.  No reference counterpart as one
 * call. */
int lbm_set_velocity_separable(lbm_channel *h, int nterms, const int *comp, const double *planes,
                               const double *profiles);

/* initpop, Channel-Flow/initial.f90:19-46: rho = 0 and f = equilibrium of
 * (rho = 0, u) from the current velocity fields, on the device. */
int lbm_initpop(lbm_channel *h);

/* collision_MRT, Channel-Flow/collision.f90:24-273, including
 * collisionExchnge (:281-372), which on one GPU is the periodic y/z wrap.
 * Reads the macroscopic fields as the reference does: the values left by the
 * last macrovar / rhoupdat / avedensity / lbm_set_macros.
 * A driver calling collision_MRT and macrovar once per step (main.f90:157-161)
 * gets lbm_run's blocked pairs: when the step could be paired (fields from
 * macrovar, uniform force, no particles or closure) its launch is deferred,
 * and collision, macrovar, collision, macrovar runs as one pair at the second
 * macrovar.  Every other entry point first launches what is deferred, so
 * results, fields and downloads are those of the calls one by one, bit for
 * bit; a launch error of a deferred step is reported by the call that
 * launches it.  LBM_DEFER=0 at lbm_create turns the deferral off. */
int lbm_collision_mrt(lbm_channel *h);

/* macrovar, Channel-Flow/collision.f90:378-463: rho, u of the current lattice
 * (u includes F/2).  Pointers may be NULL: the fields then stay on the device
 * (computed lazily, fused into the next collision). */
int lbm_macrovar(lbm_channel *h, double *rho, double *ux, double *uy, double *uz);

/* rhoupdat, Channel-Flow/collision.f90:469-480, plus the convergence measure
 * of the pre-relaxation loop, Channel-Flow/main.f90:71-81:
 * *rhoerrmax = max |rho_new - rho_old| (may be NULL). */
int lbm_rhoupdat(lbm_channel *h, double *rhoerrmax);

/* avedensity, Channel-Flow/collision.f90:487-513: subtracts the mean density
 * fluctuation (FP64 reduction) from rho; *rhomean receives it (may be NULL). */
int lbm_avedensity(lbm_channel *h, double *rhomean);

/* nsteps iterations of the main loop body, Channel-Flow/main.f90:157-161
 * (collision_MRT then macrovar), without host I/O.  Pairs of steps run as one
 * blocked sweep when the fields come from macrovar, the force is uniform and
 * there are no particles or closure: step t of a few z or y slabs goes to a
 * ring that the last-level cache holds, step t+1 follows one slab behind, so
 * HBM carries one lattice read and write per two steps.  The results equal
 * the single-step path's bit for bit.  The environment variables
 * LBM_BLOCK_AXIS=z|y and LBM_BLOCK_PLANES (read by lbm_create) set the axis
 * and the slabs per chunk (at most 32), 0 turns the sweep off; lbm_block_info
 * and lbm_block_rows report the choice. */
int lbm_run(lbm_channel *h, int nsteps);

/* Sum of all populations in FP64 (mass monitor; the sum is invariant under
 * streaming).  Synchronises. */
int lbm_mass_sum(lbm_channel *h, double *mass);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same FP64 sum as lbm_mass_sum, bit for bit, written
 * to mass_dev (one double of device memory on the handle's GPU) on the
 * handle's stream.  Returns after enqueue, without synchronising, and
 * allocates nothing: its per-block partial sums (at most 2048 doubles) live in
 * a 16 KiB static device array per GPU, shared by every handle on that GPU.
 * Calls that overlap on one GPU (concurrent host threads, or handles whose
 * streams run at the same time) race on it.  A host pointer for mass_dev is
 * rejected with LBM_ERR_ARG. */
int lbm_mass_sum_noalloc(lbm_channel *h, double *mass_dev);

/* Plane statistics of Channel-Flow/saveload.f90:1382-1470 (statistc2) for the
 * current macroscopic fields: out is (15, lx) column-major, the 15 columns
 * written at saveload.f90:1492.  ustar and ystar are the driver's values
 * (para.f90:64-66). */
int lbm_statistc2(lbm_channel *h, double ustar, double ystar, double *out);

/* Run monitor of Channel-Flow/saveload.f90:1507-1676 (diag) from deterministic
 * device reductions over the fluid nodes of the current fields: out[13] =
 * vmax, its node (imout, jmout, kmout, 1-based, the first in x, y, z order),
 * umean, vmean, wmean, urms, vrms, wrms (each / ustar), the solid volume
 * fraction, rhomax, rhomin -- the values diag.dat carries after the step
 * number.  Synchronises; no field download. */
int lbm_diag(lbm_channel *h, double ustar, double *out);

/* vortcalc, Channel-Flow/saveload.f90:3929-4023: the vorticity of the current
 * fields (central differences, periodic in y and z, the reference's one-sided
 * forms at the walls; twice the angular velocity inside particles) into the
 * host arrays ox, oy, oz (lx,ly,lz each).  Synchronises. */
int lbm_vortcalc(lbm_channel *h, double *ox, double *oy, double *oz);

/* dissipation_calc, Channel-Flow/saveload.f90:4644-4882: per wall-normal plane
 * the fluid-node means of S_ij S_ij and S_xy from the non-equilibrium moments
 * and from the non-equilibrium populations of the current f and fields.
 * out is (6, lx) column-major: x, y+, <SS>_moments, <SS>_populations,
 * <S_xy>_moments, <S_xy>_populations (the columns of dissipation.dat).  tau
 * and ystar are the driver's (para.f90:66,106).  Synchronises. */
int lbm_dissipation(lbm_channel *h, double tau, double ystar, double *out);

/* Time-averaged statistics: lbm_stats_accumulate adds the current plane sums
 * of u_i, u_i u_j and rho, rho^2 (FP64) to running device sums;
 * lbm_stats_read returns the statistc2 columns of the running average
 * (fluctuations about the time-and-plane mean) and the sample count;
 * lbm_stats_reset clears them. */
int lbm_stats_accumulate(lbm_channel *h);
int lbm_stats_read(lbm_channel *h, double ustar, double ystar, double *out, long long *nsamples);
int lbm_stats_reset(lbm_channel *h);

/* The same running statistics for the 2x2x2 box-filtered fields (no
 * reference counterpart: the filtered fine statistics a 2x-coarsened run is
 * compared with).  Coarse node (i, j, k) carries the means of rho and u over
 * fine nodes (2i..2i+1, 2j..2j+1, 2k..2k+1) of the current populations; out
 * has lx/2 rows of 15 columns as statistc2 writes them for the coarse grid,
 * with ustar, ystar in that grid's units (ystar of the fine grid / 2).  Even
 * lx, ly, lz, no particles.  lbm_stats_reset clears these sums too. */
int lbm_stats_accumulate_filtered(lbm_channel *h);
int lbm_stats_read_filtered(lbm_channel *h, double ustar, double ystar, double *out, long long *nsamples);

/* Blocks until all work on the handle's stream has finished. */
int lbm_sync(lbm_channel *h);

/* Measurement: runs nsteps of lbm_run, timing each step (or blocked pair of
 * steps) with HIP events on the handle's stream; *avg_ms is the device time
 * per step and *total_ms the event time from first launch to last
 * completion. */
int lbm_time_steps(lbm_channel *h, int nsteps, double *avg_ms, double *total_ms);

/* Device bytes held by the handle (lattices, fields, scratch). */
int lbm_device_bytes(lbm_channel *h, size_t *bytes);

/* The two-step blocked sweep lbm_run uses for pairs of steps, as lbm_create
 * chose it: *axis = 'z' (slabs of lx*ly nodes) or 'y' (lx*lz), *slabs = slabs
 * per chunk (0: single steps only).  Either pointer may be null.  No
 * reference counterpart (measurement and tests). */
int lbm_block_info(lbm_channel *h, int *axis, int *slabs);

/* Rows per window of that sweep's slabs: 0 when it sweeps whole slabs.  A
 * lattice whose slabs are too large for a chunk of four in the ring (the
 * 2048x1024x512 channel) is swept in row windows, each window on its own,
 * the two rows at its edges computed twice; the environment variable
 * LBM_BLOCK_ROWS=W sets the window.  No reference counterpart (measurement
 * and tests). */
int lbm_block_rows(lbm_channel *h, int *rows);

/* The stored post-collision lattice as FP64 structure-of-arrays (19*N values,
 * population-major, node index ix + lx*(iy + ly*iz)), for checkpoint-free
 * inspection and tests.  Synchronises. */
int lbm_download_post(lbm_channel *h, double *post);

/* The nodes [first, first + count) of the same stored lattice, FP64,
 * population-major: post[i*count + k] is population i of node first + k.
 * For a window of a lattice too large to download whole (the 2048x1024x512
 * channel's FP32 lattice is 81.6 GB).  Synchronises. */
int lbm_download_post_range(lbm_channel *h, long long first, long long count, double *post);

/* ---------------------------------------------------------------------------
 * Rigid spherical particles.  The reference keeps only the data structures of
 * its particle subsystem (Channel-Flow/var_inc.f90:30-38,120-150,
 * para.f90:332-379,449-501) and the routine names beads_links,
 * beads_collision, beads_filling, beads_move (var_inc.f90:166-168); the
 * entry points below carry those names.  Positions use the reference's node
 * coordinates: node ix (1-based) sits at x = ix - 0.5 (saveload.f90:1322),
 * walls at x = 0 and x = lx, periodic y in [0, ly) and z in [0, lz).
 * Particle state and all particle arithmetic are FP64.
 */
typedef struct lbm_particle_config {
    int npart;             /* number of spheres */
    int maxlink;           /* per-particle cut-link capacity (0 = from the radius) */
    double rad;            /* radius (var_inc.f90:67 rad) */
    double rhopart;        /* particle / fluid density ratio (var_inc.f90:65) */
    double mingap;         /* particle-particle contact range zeta (var_inc.f90:67) */
    double mingap_w;       /* particle-wall contact range (var_inc.f90:67) */
    double stf0, stf1;     /* particle-particle contact force scales of Feng & Michaelides,
                              J. Comput. Phys. 202 (2005) 20, eq. 28: approach and overlap
                              terms (para.f90:357-358: 0.025, 0.002) */
    double stf0_w, stf1_w; /* the same against the walls, mirror-image particles (para.f90:359-360) */
    double visc;           /* kinematic viscosity for the lubrication correction (flubp,
                              var_inc.f90:150; Nguyen & Ladd, Phys. Rev. E 66 (2002) 046708);
                              0 = none */
    double lub_cut;        /* lubrication cut-off gap hc (e.g. 2/3) */
    double lub_min;        /* smallest gap the lubrication force uses (e.g. 0.01) */
    int pressure_force;    /* 1: each particle also feels volume x body force */
    int nave;              /* lbm_run_particles: avedensity after every step that is a multiple
                              of nave (main.f90:163-167: 100); 0 = never */
} lbm_particle_config;

/* Creates the particle set (positions, velocities, angular velocities, each
 * 3*npart, particle-major) and classifies the lattice (beads_links). */
int lbm_particles_create(lbm_channel *h, const lbm_particle_config *pc, const double *pos, const double *vel,
                         const double *omg);

/* beads_links: marks nodes inside particles (bit-exact FP64 test
 * |x - X|^2 < R^2), lists every cut link with its wall fraction q, and
 * records the nodes the last move uncovered. */
int lbm_beads_links(lbm_channel *h);

/* beads_filling: gives nodes uncovered by the last move the equilibrium of
 * the local wall velocity and the mean density of their fluid neighbours. */
int lbm_beads_filling(lbm_channel *h);

/* beads_collision: interpolated bounce-back on every cut link (written where
 * the next collision_MRT pulls it) and the momentum-exchange force and torque
 * on every particle (FP64, summed link by link in a fixed order). */
int lbm_beads_collision(lbm_channel *h);

/* beads_move: contact and lubrication forces and FP64 integration of the
 * rigid-body motion (Adams-Bashforth 2 velocities on the hydrodynamic force
 * and torque averaged over the last two steps, trapezoidal positions). */
int lbm_beads_move(lbm_channel *h);

/* Steps istep0+1 .. istep0+nsteps of the particle-laden main loop:
 * beads_collision, collision_MRT, macrovar, beads_move, beads_links,
 * beads_filling.  After a step that is a multiple of the configuration's
 * nave, avedensity (main.f90:163-167) is taken once the next step's
 * bounce-back has completed the fluid populations, and that step's
 * collision subtracts the mean; a mean still due at the end of a call is
 * taken in the next call. */
int lbm_run_particles(lbm_channel *h, int istep0, int nsteps);

/* Particle state (each 3*npart, any pointer may be NULL).  force/torque are
 * the last momentum-exchange sums, contact the last contact forces, lub the
 * last lubrication forces (the reference's fHIp and flubp,
 * var_inc.f90:149-150). */
int lbm_particles_get(lbm_channel *h, double *pos, double *vel, double *omg, double *force, double *torque,
                      double *contact, double *lub);

/* Cut links: counts[npart]; node/dir/q hold npart*maxlink entries
 * (particle-major, counts[p] valid each).  *maxlink receives the capacity. */
int lbm_particles_links(lbm_channel *h, int *maxlink, int *counts, int *node, int *dir, double *q);

/* Node flags (N ints): -1 fluid, p >= 0 inside particle p. */
int lbm_particles_solid(lbm_channel *h, int *solid);

/* ---------------------------------------------------------------------------
 * Learned relaxation-rate closure (no counterpart in the reference: the
 * target of a 2x-coarsened channel that reproduces filtered fine statistics).
 * A 10-32-32-1 tanh network maps the nine non-equilibrium moments the MRT
 * relaxes with s1, s2, s9, s10, s13 (scaled) and y+ to an eddy viscosity
 * nu_t = nu * (softplus(o) + m); the node's shear rates become
 * s9 = s13 = 1 / (3 (nu + nu_t) + 0.5).  Parameters, FP32, LBM_MLP_TOTAL
 * values: W1[32][10], b1[32], W2[32][32], b2[32], W3[32], b3 (the
 * LBM_MLP_NPARAM trained values), then 9 input scales, the y+ divisor, nu,
 * and the floor m of nu_t/nu, in (-1, 0] (0: a non-negative eddy viscosity;
 * negative: the closure may also lower the viscosity, tau staying above 1/2).
 */
#define LBM_MLP_NPARAM 1441
#define LBM_MLP_TOTAL 1453

/* Evaluates the closure inside every following collision_MRT (FP32 and FP16
 * storage only; FP64 stays the reference's plain MRT).  ystar = nu/u* of the
 * run (para.f90:66).  Hidden-layer biases must lie within +-32 and the floor
 * within (-1, 0] (LBM_ERR_ARG otherwise). */
int lbm_closure_set(lbm_channel *h, const float *params, double ystar);
/* Back to plain MRT. */
int lbm_closure_clear(lbm_channel *h);

/* What the closure sees and applies at every node of the current lattice, for
 * the next collision_MRT (the fields must come from macrovar): the ten scaled
 * network inputs (feat[10*n + k], n = ix + lx*(iy + ly*iz)), or the eddy
 * viscosity ratio nu_t/nu = softplus(o) + m (nut[n]) from the same matrix-core
 * evaluation the fused collision uses. */
int lbm_closure_inputs(lbm_channel *h, float *feat);
int lbm_closure_nut(lbm_channel *h, float *nut);

/* Trainer: a sample store and the network on one device.  Samples are
 * (10 features, target nu_t/nu).  lbm_trainer_add_filtered box-filters the
 * fine channel's current lattice 2x2x2 and appends one sample per coarse node
 * (target: the eddy viscosity nu_t = -tau:S/(2 S:S) of the subgrid stress,
 * features: the first-order non-equilibrium moments of the coarse MRT with
 * shear rate s9_coarse, scaled by the parameters' input scales).  Gradients
 * are of mean((softplus(o) + m - target)^2) over a sample range; training
 * uses Adam (beta 0.9, 0.999) over consecutive batches.
 *
 * lbm_trainer_set_plane_targets overwrites the targets of samples
 * [first, first + count): sample first + i gets values[i mod nx].  With first
 * at the start of an add_filtered snapshot (its samples run x fastest) and nx
 * the coarse lx, every sample gets the value of its wall-normal plane: a
 * target fixed by the mean momentum balance of the coarse run, for example.
 *
 * lbm_trainer_set_target chooses the target of later add_filtered calls:
 * LBM_TARGET_LOCAL (default) the node's own -tau:S/(2 S:S)/nu, clipped to
 * [-10, 10]; LBM_TARGET_PLANE the least-squares eddy viscosity of the node's
 * wall-parallel plane, -sum(tau:S)/(2 sum(S:S))/nu (the planar averaging of
 * dynamic subgrid models), the same for every node of the plane. */
#define LBM_TARGET_LOCAL 0
#define LBM_TARGET_PLANE 1
typedef struct lbm_trainer lbm_trainer;
int lbm_trainer_create(int device, long long capacity, const float *params, lbm_trainer **out);
int lbm_trainer_destroy(lbm_trainer *t);
int lbm_trainer_set_target(lbm_trainer *t, int mode);
int lbm_trainer_add_filtered(lbm_trainer *t, lbm_channel *fine, double nu_coarse, double s9_coarse,
                             double ystar_coarse, long long *added);
int lbm_trainer_set_plane_targets(lbm_trainer *t, long long first, long long count, int nx, const float *values);
int lbm_trainer_add_samples(lbm_trainer *t, const float *features, const float *targets, long long n);
int lbm_trainer_samples(lbm_trainer *t, long long *n, float *features, float *targets);
int lbm_trainer_grad(lbm_trainer *t, long long first, long long count, float *grad, float *loss);
int lbm_trainer_train(lbm_trainer *t, long long first, long long count, int iterations, long long batch, float lr,
                      float *loss);
int lbm_trainer_predict(lbm_trainer *t, long long first, long long count, float *out);
int lbm_trainer_params(lbm_trainer *t, float *params);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
