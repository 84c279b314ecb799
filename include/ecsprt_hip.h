/*
 * ecsprt_hip.h -- C ABI of the MI355X (gfx950) implementation of ECSPRT's
 * per-timestep hybrid-PIC hot path.
 *
 * Every entry point returns an int error code (0 = success), the analogue of
 * PetscErrorCode; the library never aborts the process.  Paths in citations are
 * relative to the ECSPRT reference repository root.
 *
 * Ownership: the library owns the particle state (tile-binned structure of arrays)
 * and the grid vectors on the device between calls.  Host arrays cross the boundary
 * only through the explicit get/set calls, in the reference's layouts:
 *   particles  AoS double[np][3] (DMSwarm coordinate and "velocity" fields, src/setup.c:494)
 *   moments    double[Nz][Ny][Nx][4] = {n, Gx, Gy, Gz} (FluidNode, src/ecsprt.h:194-199)
 *   potential  double[Nz][Ny][Nx]
 * Device-vector entry points (matmult, pcapply, remove_mean, rhs) take device
 * pointers to [Nz][Ny][Nx] vectors.  All work is ordered on the context's stream;
 * a context is used by one host thread at a time.
 */
#ifndef ECSPRT_HIP_H
#define ECSPRT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define ECSPRT_OK                 0
#define ECSPRT_ERR_ARG            1 /* SETERRQ(PETSC_ERR_ARG_WRONG) analogue */
#define ECSPRT_ERR_COLLISION      2 /* >=10 refused collisions: src/collisions.c:141-148 Abort */
#define ECSPRT_ERR_COUNT_CHANGED  3 /* particle lost under periodic BC: src/boundaries.c:101-109 Abort */
#define ECSPRT_ERR_HIP            4 /* HIP runtime error */
#define ECSPRT_ERR_NOT_CONVERGED  5 /* KSP reached max_it without meeting rtol */
#define ECSPRT_ERR_ATTEMPTS       6 /* pre-drawn collision records exhausted before Nc successes */
#define ECSPRT_ERR_NOMEM          7 /* device allocation failed */
#define ECSPRT_ERR_STATE          8 /* call made before the state it needs was set */

/* Initial position distributions, numbered as PDistType (src/positions.c:8-10). */
#define ECSPRT_PDIST_SOBOL      0
#define ECSPRT_PDIST_REVERSE    1 /* not implemented by the reference (SETERRQ) */
#define ECSPRT_PDIST_NORMAL     2
#define ECSPRT_PDIST_UNIFORM    3 /* PETSc's DMSwarmSetPointsUniformCoordinates lattice: not provided */
#define ECSPRT_PDIST_SINUSOIDAL 4
#define ECSPRT_PDIST_GAUSSIAN   5 /* not implemented by the reference (SETERRQ) */

/* Command-line parameters of the hot path (subset of CLI, src/ecsprt.h:137-191). */
typedef struct {
  int    Nx, Ny, Nz;
  double dx, dy, dz;
  double x0, y0, z0, x1, y1, z1;
  double B0, E0, n0;
  double qi, mi, nui;
  double vi0x, vi0y, vi0z;
  double viTx, viTy, viTz;
  double Ti;
  double nue, Te;
  double veTx, veTy, veTz;
  double mn, vn0x, vn0y, vn0z, Tn;
  int    ndim;  /* 3 (0 also means 3) or 2: -ndim, src/setup.c:31-112 (in 2-D, Nz = 1) */
} ecsprt_cli;

/* Derived context scalars (Context, src/ecsprt.h:118-134, as SetUpContext fills them). */
typedef struct {
  int    Nx, Ny, Nz;
  double dx, dy, dz, Lx, Ly, Lz, x0, y0, z0, x1, y1, z1;
  double B0, E0, n0;
  double qi, mi, nui, vi0x, vi0y, vi0z, vi0, viTx, viTy, viTz, viT, Ti, Omega_i, kappa_i;
  double qe, me, nue, Te, veT, Omega_e, kappa_e;
  double mn, vn0x, vn0y, vn0z, vn0, vnT, Tn;
  double scale;
  double gammaT;
  int    ndim;  /* 2 or 3 */
} ecsprt_params;

/* One collision attempt's random numbers in the reference's draw order
 * (src/collisions.c:72,80,88,100,102).  u_acc1*u_acc2 is the acceptance draw:
 * PetscSqr(ran3(&seed)) expands to two ran3 calls. */
typedef struct {
  float u_ip;
  float g[3];
  float u_acc1, u_acc2;
  float u_z, u_phi;
} ecsprt_attempt;

/* Collision modes. */
#define ECSPRT_COLLIDE_OFF      0
#define ECSPRT_COLLIDE_PREDRAWN 1 /* records from ecsprt_hip_set_attempts; per-particle parity */
#define ECSPRT_COLLIDE_PHILOX   2 /* counter-based Philox4x32-10 stream; statistical parity */

/* Operator modes for the potential solve. */
#define ECSPRT_OPERATOR_FROZEN  0 /* built at the first solve only (KSPSetComputeOperators, src/pic.c:192) */
#define ECSPRT_OPERATOR_REBUILD 1 /* rebuilt from the current density at every solve */

/* Preconditioners. */
#define ECSPRT_PC_NONE   0
#define ECSPRT_PC_JACOBI 1
#define ECSPRT_PC_GMG    2

/* Precision of the potential solve. */
#define ECSPRT_SOLVE_FP64  0
#define ECSPRT_SOLVE_MIXED 1 /* FP32 GMRES+GMG inside FP64 iterative refinement */

/* Particle storage and arithmetic. */
#define ECSPRT_PARTICLES_FP64  0 /* FP64 positions and velocities, bit-compatible with the reference */
#define ECSPRT_PARTICLES_MIXED 1 /* int32 packed cell + FP32 in-cell offset, FP32 velocity;
                                    FP32 push, drift and deposit, FP64 moment fold.  Grids up
                                    to 2048 x 2048 x 1024 cells with the origin at 0. */

typedef struct {
  int      tile;           /* cells per tile edge along x and y: 4 or 8 (0 = by grid size) */
  int      tile_z;         /* cells per tile along z: 4 or 8 (0 = tile) */
  int      halo;           /* deposit halo in cells for the fused step (0 = 2) */
  int      block;          /* threads per tile workgroup (0 = 1024) */
  int      split;          /* workgroups per tile (0 = auto) */
  double   slack;          /* tile capacity slack fraction (0 = 0.25) */
  int      collide_mode;   /* ECSPRT_COLLIDE_* */
  uint64_t seed;           /* Philox key */
  int      operator_mode;  /* ECSPRT_OPERATOR_* */
  int      initial_guess;  /* 0 = zero every solve (ComputeInitialPhi, src/potential.c:20-29), 1 = previous phi */
  int      pc_type;        /* ECSPRT_PC_* */
  int      solve_precision;/* ECSPRT_SOLVE_* */
  int      ksp_restart;    /* GMRES restart (0 = 30) */
  int      ksp_max_it;     /* (0 = 10000, PETSc's default) */
  double   ksp_rtol;       /* (0 = 1e-5, PETSc's default) */
  double   ksp_atol;       /* (0 = 1e-50) */
  int      mg_pre, mg_post;/* smoothing sweeps (0 = 2 each) */
  int      mg_coarse_sweeps;/* sweeps on the coarsest level (0 = 50) */
  double   inner_rtol;     /* mixed precision: inner FP32 solve tolerance (0 = ksp_rtol^(1/k), the fewest
                              k refinement cycles of at most 1e-5 each) */
  int      timing;         /* 1 = record per-stage HIP events */
  int      particle_precision; /* ECSPRT_PARTICLES_* */
  int      mg_sweep;       /* smoother kernels: 0 = by level size, 1 = box sweeps, 2 = marching sweeps
                              (identical arithmetic; a choice of speed only) */
  int      gmres_refine;   /* 0 = classical Gram-Schmidt once per Arnoldi step (PETSc's default,
                              KSP_GMRES_CGS_REFINE_NEVER), 1 = with one reorthogonalisation */
} ecsprt_options;

typedef struct ecsprt_hip_ctx ecsprt_hip_ctx;

/* ---- set-up --------------------------------------------------------------- */

/* SetUpContext + NormalizeGrid (src/setup.c:116-271, 296-346). */
int ecsprt_hip_setup_context(const ecsprt_cli *cli, ecsprt_params *out);
/* Fill opt with defaults. */
int ecsprt_hip_default_options(ecsprt_options *opt);
/* Create a context on the current HIP device: grid vectors, solver, no particles.
 * Replaces CreateSwarmDM/CreateFluidDM/CreatePotentialDM/KSPCreate (src/pic.c:179-192). */
int ecsprt_hip_create(const ecsprt_params *p, const ecsprt_options *opt, ecsprt_hip_ctx **out);
int ecsprt_hip_destroy(ecsprt_hip_ctx *ctx);

/* ---- particle state -------------------------------------------------------- */

/* Upload np particles from host AoS arrays (DMSwarm layout); particle ids are the
 * array indices.  Returns ECSPRT_ERR_COUNT_CHANGED if any coordinate is outside
 * the domain (what DMSwarmMigrate would drop). */
int ecsprt_hip_set_particles(ecsprt_hip_ctx *ctx, long np, const double *pos, const double *vel);
/* Download particles to host AoS arrays in id order. */
int ecsprt_hip_get_particles(ecsprt_hip_ctx *ctx, double *pos, double *vel);
/* Generate np particles on the device: positions uniform in the box with an optional
 * x-density perturbation 1 + amp*sin(2*pi*k*x/Lx) (inverse CDF), velocities
 * Maxwellian with per-component thermal speeds viT{x,y,z} and drift vi0{x,y,z}. */
int ecsprt_hip_init_particles(ecsprt_hip_ctx *ctx, long np, uint64_t seed, double amp, int k);

/* The reference's own initial conditions on one rank: positions ran3 * L per coordinate
 * (NormalDistribution, src/positions.c:42-137 -- uniform despite the name; no domain
 * origin offset, as there), then velocities vT * gasdev + v0 per component
 * (NormalVelocities, src/velocities.c:14-47), each stream seeded with getseed() = -12345
 * (src/random.c:292-295).  The Numerical Recipes streams are serial and run on the host;
 * the arrays are uploaded and binned. */
int ecsprt_hip_init_reference(ecsprt_hip_ctx *ctx, long np);
/* InitializePositions (src/positions.c:390-433) for pdist, then the velocities of
 * ecsprt_hip_init_reference: ECSPRT_PDIST_NORMAL is ecsprt_hip_init_reference;
 * ECSPRT_PDIST_SOBOL the reference's Sobseq (src/random.c:233-289), one 3-D point per
 * ion times L (SobolDistribution, src/positions.c:139-234; np < 2^30);
 * ECSPRT_PDIST_SINUSOIDAL rejection against 1 + 0.25 sin(2 pi x/Lx) over PETSc's
 * default PetscRandom, the rand48 generator seeded with 7 (Rejection,
 * src/positions.c:236-388).  Other values return ECSPRT_ERR_ARG. */
int ecsprt_hip_init_distribution(ecsprt_hip_ctx *ctx, long np, int pdist);
long ecsprt_hip_num_particles(const ecsprt_hip_ctx *ctx);
/* Per-cell counts of (int)(pos/d) per axis, host int[Nz][Ny][Nx]. */
int ecsprt_hip_cell_counts(ecsprt_hip_ctx *ctx, int *counts);

/* ---- plugin slots (src/global.h:27-32, bound in src/setup.c:99-114) ----------- */

/* PushFunc: BorisMoverBz3D (src/velocities.c:281-371). */
int ecsprt_hip_push(double dt, void *ctx);
/* CollideFunc: ScatterElastic (src/collisions.c:9-167), in the context's collide mode:
 * pre-drawn records replayed per particle, or the Philox mode's Poisson attempts on every
 * ion (the same draws as the fused step's, so push + collide equals the fused step's
 * velocities); collide mode off does nothing.  Advances the Philox step counter. */
int ecsprt_hip_collide(int ndim, double dt, void *ctx);
/* BCFunc: ApplyBC (src/boundaries.c:6-34) on AoS positions pos[np][ndim] (vel unused).
 * pos may be device memory (wrapped in place by a kernel) or host memory, as the
 * reference's caller passes the swarm's array (src/boundaries.c:67-80): host arrays are
 * copied to the device, wrapped and copied back; no kernel reads host memory. */
int ecsprt_hip_apply_bc(int np, double *pos, double *vel, void *ctx);
/* CollectFunc: Collect3DFluidMoments (src/moments.c:108-226). */
int ecsprt_hip_collect(void *ctx);

/* ---- direct entry points ---------------------------------------------------- */

/* UpdateVelocities (src/velocities.c:375-390): push, then ecsprt_hip_collide if mn > 0
 * (src/velocities.c:384-386) and the collide mode is not off. */
int ecsprt_hip_update_velocities(int ndim, double dt, void *ctx);
/* UpdatePositions (src/positions.c:436-472): drift, ApplyBC, re-binning and the
 * count-conservation check of ApplyBCAndMigrate (src/boundaries.c:57-113). */
int ecsprt_hip_update_positions(int ndim, double dt, void *ctx);
/* ComputePotential (src/potential.c:6-16): operator (per operator mode), RHS, KSP solve. */
int ecsprt_hip_compute_potential(void *ctx);
/* One fused timestep in the order of src/pic.c:246-258: push, collide (Philox
 * mode only), drift, wrap, deposit in one pass over the particles, then the solve. */
int ecsprt_hip_step(void *ctx, double dt);

/* ---- linear system (src/lhs.c, src/rhs.c, src/potential.c) ------------------ */

/* LHSFunc analogue: ComputeFullLHS3D (src/lhs.c:337-505) captured from the current
 * moments as the matrix-free operator used by matmult and the solver. */
int ecsprt_hip_lhs(void *ctx);
/* RHSFunc analogue: ComputeFullRHS3D (src/rhs.c:259-373) into device vector b, mean removed. */
int ecsprt_hip_rhs(void *ctx, double *b);
/* MATSHELL MatMult: y = A x with the captured operator (device vectors). */
int ecsprt_hip_matmult(void *shell, const double *x, double *y);
/* PCShell apply: z = M^-1 r, one geometric-multigrid V-cycle (device vectors). */
int ecsprt_hip_pcapply(void *shell, const double *r, double *z);
/* MatNullSpaceRemove with the constant null space (device vector). */
int ecsprt_hip_remove_mean(void *ctx, double *v);
/* KSPSolve analogue on device vectors: A phi = b, phi overwritten (initial guess per options). */
int ecsprt_hip_ksp_solve(void *ctx, const double *b, double *phi);

/* ---- host access to grid vectors --------------------------------------------- */

int ecsprt_hip_get_moments(ecsprt_hip_ctx *ctx, double *moments);
int ecsprt_hip_set_moments(ecsprt_hip_ctx *ctx, const double *moments);
int ecsprt_hip_get_potential(ecsprt_hip_ctx *ctx, double *phi);
int ecsprt_hip_set_potential(ecsprt_hip_ctx *ctx, const double *phi);
int ecsprt_hip_get_forcing(ecsprt_hip_ctx *ctx, double *b);

/* ---- collisions ---------------------------------------------------------------- */

/* Set pre-drawn attempt records (host array, attempt order) for the next collide. */
int ecsprt_hip_set_attempts(ecsprt_hip_ctx *ctx, const ecsprt_attempt *att, long n);
/* Nc from the reference formula (src/collisions.c:58-63, x86 int-conversion semantics). */
int ecsprt_hip_collision_count(const ecsprt_params *p, long np, double dt);
/* Counters of the last collide: Nc, successes, refusals, attempts consumed. */
int ecsprt_hip_collision_stats(ecsprt_hip_ctx *ctx, int *Nc, long *Ns, long *Nf, long *attempts);

/* ---- diagnostics ----------------------------------------------------------------- */

/* Last solve: iterations, final true-residual norm, ||b||. */
int ecsprt_hip_solver_stats(ecsprt_hip_ctx *ctx, int *iters, double *rnorm, double *bnorm);
/* Algorithmic HBM bytes of the last ComputePotential (forcing, mean removal and KSP
 * solve): every kernel's compulsory reads and writes, each array once per point (the
 * numerator of the solve's roofline fraction). */
int ecsprt_hip_solver_traffic(ecsprt_hip_ctx *ctx, double *bytes);
/* Per-stage times of the last call that recorded them (ms, HIP events on the context's
 * stream, options.timing = 1).  ecsprt_hip_step: [0] fused particle kernel, [1] set-up
 * before it, [2] re-binning and fold, [4] operator + RHS, [5] KSP solve.  Stage calls:
 * [0] push, [3] deposit + fold, [4] operator + RHS, [5] KSP solve. */
int ecsprt_hip_stage_times(ecsprt_hip_ctx *ctx, double *ms);
/* Block until the context's stream is idle. */
int ecsprt_hip_synchronize(ecsprt_hip_ctx *ctx);
/* Text of the last error on this context. */
const char *ecsprt_hip_last_error(const ecsprt_hip_ctx *ctx);

/* *is_device = 1 when ptr is device (or managed) memory a kernel may dereference, 0 for
 * host memory, pinned or pageable: the test ecsprt_hip_apply_bc applies before launching
 * a kernel on a caller's array.  No counterpart in the reference. */
int ecsprt_hip_pointer_is_device(const void *ptr, int *is_device);

/* Device memory helpers for callers without their own allocator. */
int ecsprt_hip_dmalloc(void **ptr, long bytes);
int ecsprt_hip_dfree(void *ptr);
int ecsprt_hip_h2d(void *dst, const void *src, long bytes);
int ecsprt_hip_d2h(void *dst, const void *src, long bytes);
/* Device-to-device copy, timed with HIP events: *ms receives its duration (the bandwidth
 * probe used to calibrate HBM byte counters). */
int ecsprt_hip_d2d(void *dst, const void *src, long bytes, double *ms);

/* Philox4x32-10 and Philox2x32-10 (Salmon et al., SC'11) on the host, for known-answer
 * tests of the generators the statistical collision mode draws from. */
void ecsprt_hip_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ecsprt_hip_philox2x32_10(const uint32_t ctr[2], uint32_t key, uint32_t out[2]);

/* The statistical collision mode's number of attempts per ion and step: k[i] = the
 * smallest k with P(K <= k) >= u[i] for K ~ Poisson(lam) (inversion searched from the
 * mode), the function every Philox-mode kernel applies to its 24-bit uniform draw.  Host
 * code, for known-answer tests against a reference Poisson quantile.  lam in [0, 1e6]. */
int ecsprt_hip_poisson_attempts(double lam, long n, const double *u, int *k);

/* The Numerical Recipes streams the library draws the reference's initial conditions
 * from (src/random.c:24-204), for comparison with the reference's generators: n values of
 * ran1 (generator 0), ran3 (1) or gasdev (2) from fresh static state and seed `seed`
 * (getseed() gives -12345, src/random.c:292-295).  Host code; no device is used. */
int ecsprt_hip_nr_stream(int generator, long seed, long n, float *out);

/* ---- Batched standalone potential solves -------------------------------------------
 * No counterpart entry point in the reference, which solves one field per KSPSolve
 * (src/potential.c:11); each problem of a batch is that solve: the operator of
 * ComputeFullLHS3D (src/lhs.c:337-505) from the problem's density, the RHS of
 * ComputeFullRHS3D (src/rhs.c:259-373) with the mean removed, and the configured
 * Krylov/multigrid solve (ksp_rtol etc. from the options).  Problems share the physical
 * parameters of `p` and differ in grid shape (dims = Nx, Ny, Nz per problem). */
typedef struct ecsprt_hip_batch ecsprt_hip_batch;
/* nslots > 0: that many concurrent solver contexts, each with its own stream and host
 * thread, take the problems in turn.  nslots == 0: all problems in lockstep from one
 * host thread -- every kernel of the V-cycle, the operator and the Krylov updates is
 * launched once for all problems still iterating (multigrid preconditioning only; each
 * problem's coarsest level must have at most 1024 nodes, else ECSPRT_ERR_ARG). */
int ecsprt_hip_batch_create(const ecsprt_params *p, const ecsprt_options *o, int nprob, const int *dims,
                            int nslots, ecsprt_hip_batch **out);
int ecsprt_hip_batch_destroy(ecsprt_hip_batch *b);
/* moments[q]: device FluidNode [Nz][Ny][Nx][4] of problem q; phi[q]: device [Nz][Ny][Nx] that
 * receives its potential.  Problems run concurrently on the batch's slots (streams). */
int ecsprt_hip_batch_solve(ecsprt_hip_batch *b, double *const *moments, double *const *phi);
/* Per-problem Krylov iterations and final residual norms of the last batch_solve. */
int ecsprt_hip_batch_stats(const ecsprt_hip_batch *b, int *iters, double *rnorm);
const char *ecsprt_hip_batch_last_error(const ecsprt_hip_batch *b);
/* Algorithmic HBM bytes of the last batch_solve in lockstep mode (each kernel's
 * compulsory reads and writes of the problems it worked on); 0 with slots. */
int ecsprt_hip_batch_traffic(const ecsprt_hip_batch *b, double *bytes);

/* ---- Ensembles of independent runs --------------------------------------------------
 * One ecsprt_hip_step (the timestep of src/pic.c:246-258) on each of n contexts, run
 * concurrently (one host thread per context, each context on its own stream).  No
 * counterpart in the reference, which runs one simulation per process.  Returns the first
 * non-zero code; each context's message is in its ecsprt_hip_last_error. */
int ecsprt_hip_ensemble_step(ecsprt_hip_ctx *const *ctxs, int n, double dt);

/* Lockstep ensemble: the same timestep on each of n contexts (three-dimensional, Philox
 * collisions or none, equal solver options, zero initial guess, multigrid
 * preconditioning), driven from one host thread.  The fused particle passes of runs that
 * share a precision, tile shape and block size run as one kernel; each run's rate
 * prediction passes and re-binning go on that context's own stream, concurrently, and
 * the rate predictions and counter checks of all runs share each host synchronisation; the potentials are solved by one lockstep
 * GMRES/multigrid over all runs (the lockstep batch solver).  From the same state, each
 * run's ions and moments equal those of ecsprt_hip_step bit for bit; its potential
 * solves the same system to the same tolerance (the same method with its partial sums
 * in another order).  The contexts keep their own state and may still be stepped alone
 * between ensemble steps.  No counterpart in the reference.  advance returns the first
 * non-zero code of a run (its message in that context's ecsprt_hip_last_error) or of
 * the ensemble (ensemble_last_error); runs that failed in the particle stage are left
 * out of that step's solve. */
typedef struct ecsprt_hip_ensemble ecsprt_hip_ensemble;
int ecsprt_hip_ensemble_create(ecsprt_hip_ctx *const *ctxs, int n, ecsprt_hip_ensemble **out);
int ecsprt_hip_ensemble_advance(ecsprt_hip_ensemble *e, double dt);
int ecsprt_hip_ensemble_destroy(ecsprt_hip_ensemble *e);
const char *ecsprt_hip_ensemble_last_error(const ecsprt_hip_ensemble *e);
/* Algorithmic HBM bytes of the last advance's lockstep solve (every run). */
int ecsprt_hip_ensemble_traffic(const ecsprt_hip_ensemble *e, double *bytes);
/* HIP-event times (ms) of the last advance: [0] the fused tile passes (all launches),
 * [1] the stages before them (counters, rate predictions), [2] re-binning, late deposits
 * and folds with the counter checks, [3] the forcing and the lockstep solve. */
int ecsprt_hip_ensemble_stage_times(const ecsprt_hip_ensemble *e, double *ms);

/* Synthetic density field on the device: n = exp(sigma g - sigma^2/2), g ~ N(0,1) per node
 * (Philox4x32-10 keyed by seed), smoothed by a periodic Gaussian of `smooth` cells (radius
 * ceil(4 smooth)), normalised to mean 1; written as FluidNode moments (n, s n, s n, s n) with
 * s = flux_scale, as the reference's density input sets the flux (src/file-io.c:51-59). */
int ecsprt_hip_lognormal_moments(int Nx, int Ny, int Nz, uint64_t seed, double sigma, double smooth,
                                 double flux_scale, double *moments);

/* The allocation-free variant of ecsprt_hip_lognormal_moments.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and the same result bit
 * for bit, plus `stream` (a hipStream_t; NULL = the null stream): the kernels are
 * enqueued there and the call returns without synchronising, allocating nothing -- the
 * field lives in the flux components of `moments` until they are overwritten, and the
 * partial sums and smoothing weights in an 8 KB device array of the library that each
 * device holds until the process ends.  Limit: ceil(4 smooth) <= 383 cells (smooth <=
 * 95.75); larger returns ECSPRT_ERR_ARG.  Calls that overlap on one device race for that
 * array. */
int ecsprt_hip_lognormal_moments_noalloc(int Nx, int Ny, int Nz, uint64_t seed, double sigma, double smooth,
                                         double flux_scale, double *moments, void *stream);

/* ---- On-disk formats (src/file-io.c) ----------------------------------------------
 * HDF5 is loaded at run time (ECSPRT_HDF5_LIBRARY alone when set, else libhdf5.so on the
 * loader's path, else /opt/conda/lib/libhdf5.so.103; release 1.10 or later); without it the HDF5 entry
 * points return ECSPRT_ERR_STATE.  Files follow PETSc 3.20's viewers: HDF5 datasets in the
 * root group shaped [Nz][Ny][Nx] ([Ny][Nx] in 2-D), native doubles, attribute
 * "timestepping" = 0; binary vectors as big-endian int32 classid 1211214, int32 length and
 * big-endian doubles, with "-vecload_block_size <bs>" in "<file>.info". */

/* OutputFluidHDF5 (src/file-io.c:164-222): writes "fluid<insert>.hdf" in the working
 * directory with the datasets "density", "x flux", "y flux", "z flux" (the FluidNode
 * components), "rhs" (the last forcing) and "potential". */
int ecsprt_hip_output_fluid_hdf5(ecsprt_hip_ctx *ctx, const char *insert);

/* OutputSwarmBinary (src/file-io.c:103-161): writes "position<insert>.bin" and
 * "velocity<insert>.bin" (ndim values per ion, block size ndim; ions in the library's
 * order, as the reference's are in its swarm's order). */
int ecsprt_hip_output_swarm_binary(ecsprt_hip_ctx *ctx, const char *insert);

/* LoadFluidQuantities (src/file-io.c:8-75): reads the density from dataset "density-kji" of
 * the HDF5 file inpath (shape of the grid) and sets the moments to (n, s0 n, s1 n, s2 n)
 * with s = flux_scale; the next potential solve builds its operator from them.  An empty
 * or NULL inpath switches the forcing to ComputeSinusoidalRHS3D (src/rhs.c:86-149)
 * instead, as the reference does. */
int ecsprt_hip_load_fluid_quantities(ecsprt_hip_ctx *ctx, const double flux_scale[3], const char *inpath);

/* Host-buffer forms of the same formats, for callers that hold their own arrays.
 * ecsprt_io_hdf5_available returns 1 when HDF5 loads, else 0 with the reason in msg;
 * ecsprt_io_last_error is the message of this thread's last failed ecsprt_io call. */
int ecsprt_io_hdf5_available(char *msg, int cap);
const char *ecsprt_io_last_error(void);
/* moments: FluidNode array [N][4]; rhs, phi: [N] (N = Nx Ny Nz, or Nx Ny in 2-D). */
int ecsprt_io_write_fluid_hdf5(const char *path, int ndim, int Nx, int Ny, int Nz, const double *moments,
                               const double *rhs, const double *phi);
/* One dataset of ndims (<= 4) dimensions in a new file, e.g. a "density-kji" input. */
int ecsprt_io_write_hdf5_dataset(const char *path, const char *name, int ndims, const long *dims, const double *data);
/* Shape of dataset `name` into ndims and dims[4]; with out, its values (at most cap). */
int ecsprt_io_read_hdf5_dataset(const char *path, const char *name, int *ndims, long *dims, double *out, long cap);
/* PETSc binary vector of n values with block size bs (VecView / VecLoad layout). */
int ecsprt_io_write_vec_binary(const char *path, long n, int bs, const double *v);
/* Its length into n and, with v, its values (at most cap); bs from the .info file (1 when
 * absent; may be NULL). */
int ecsprt_io_read_vec_binary(const char *path, double *v, long cap, long *n, int *bs);

#ifdef __cplusplus
}
#endif
#endif
