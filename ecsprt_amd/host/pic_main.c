/*
 * pic_main.c -- C host driver of the GPU hot path, the counterpart of the reference's
 * PIC application loop (src/pic.c:148-288) with the reference's option names.
 *
 * The plugin slots of the reference (Species.push/collect/collide, src/ecsprt.h:89-92)
 * are function pointers here too and are bound to the libecsprt_hip entry points, so
 * the time loop reads like src/pic.c:238-269.  Options use the reference spellings:
 *   -Nx -Ny -Nz -dx -dy -dz -Np (negative = per cell) -Nt -dt -B0 -E0 -qi -mi -nui
 *   -nue -Te -viT -mn -Tn -vn0 ; plus -rtol, -fused (one-pass timestep), -seed,
 *   -no-collide (collision mode off; by default the Philox statistical mode),
 *   -mixed (mixed-precision particles and solve), -ndim 2 (a two-dimensional run:
 *   BorisMoverBz2D, Collect2DFluidMoments, ComputeFullLHS2D/RHS2D; Nz = 1).
 * Initial conditions come from the device generator (uniform positions, Maxwellian
 * velocities), the analogue of InitializePositions/InitializeVelocities; with
 * --position-dist normal|sobol|sinusoidal they are the reference's own
 * (NormalDistribution, SobolDistribution or the sinusoidal Rejection, then
 * NormalVelocities: gasdev velocities, seed -12345).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ecsprt_hip.h"

typedef int (*PushFunc)(double dt, void *opts);
typedef int (*CollideFunc)(int ndim, double dt, void *opts);
typedef int (*CollectFunc)(void *opts);

typedef struct {
  PushFunc push;
  CollideFunc collide;
  CollectFunc collect;
} Species;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int get_real(int argc, char **argv, const char *name, double *v) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!strcmp(argv[i], name)) {
      *v = atof(argv[i + 1]);
      return 1;
    }
  return 0;
}

static const char *get_str(int argc, char **argv, const char *name) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!strcmp(argv[i], name)) return argv[i + 1];
  return NULL;
}

static int has_flag(int argc, char **argv, const char *name) {
  for (int i = 1; i < argc; ++i)
    if (!strcmp(argv[i], name)) return 1;
  return 0;
}

#define CHECK(call)                                                                  \
  do {                                                                               \
    int rc_ = (call);                                                                \
    if (rc_ != ECSPRT_OK) {                                                          \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, ecsprt_hip_last_error(ctx)); \
      return rc_;                                                                    \
    }                                                                                \
  } while (0)

int main(int argc, char **argv) {
  const double Q = 1.6022e-19, KB = 1.3807e-23, ME = 9.1094e-31;
  ecsprt_cli cli;
  memset(&cli, 0, sizeof(cli));
  /* Reference defaults (src/parameters.c:25-346), in grid units for the ions. */
  double v;
  cli.Nx = cli.Ny = cli.Nz = 32;
  cli.dx = cli.dy = cli.dz = 1.0;
  cli.x1 = cli.y1 = cli.z1 = 1.0;
  cli.n0 = 1.0;
  cli.qi = 1.0;
  cli.mi = 1.0;
  cli.nui = 0.01;
  cli.B0 = 0.05;
  cli.Te = Q / KB;
  cli.Tn = -1.0;
  double viT = 1.0, Np = -64, Nt = 10, dt = 1.0, rtol = 1e-8, seed = 2024, kappa_e = 4.0;
  if (get_real(argc, argv, "-Nx", &v)) cli.Nx = (int)v;
  if (get_real(argc, argv, "-Ny", &v)) cli.Ny = (int)v;
  if (get_real(argc, argv, "-Nz", &v)) cli.Nz = (int)v;
  get_real(argc, argv, "-dx", &cli.dx);
  get_real(argc, argv, "-dy", &cli.dy);
  get_real(argc, argv, "-dz", &cli.dz);
  get_real(argc, argv, "-B0", &cli.B0);
  get_real(argc, argv, "-E0", &cli.E0);
  get_real(argc, argv, "-qi", &cli.qi);
  get_real(argc, argv, "-mi", &cli.mi);
  get_real(argc, argv, "-nui", &cli.nui);
  get_real(argc, argv, "-Te", &cli.Te);
  get_real(argc, argv, "-Tn", &cli.Tn);
  get_real(argc, argv, "-viT", &viT);
  get_real(argc, argv, "-Np", &Np);
  get_real(argc, argv, "-Nt", &Nt);
  get_real(argc, argv, "-dt", &dt);
  get_real(argc, argv, "-rtol", &rtol);
  get_real(argc, argv, "-seed", &seed);
  get_real(argc, argv, "-kappa_e", &kappa_e);
  cli.viTx = cli.viTy = cli.viTz = viT;
  cli.mn = cli.mi;
  get_real(argc, argv, "-mn", &cli.mn);
  if (get_real(argc, argv, "-vn0", &v)) cli.vn0x = cli.vn0y = cli.vn0z = v;
  /* nu_e from the requested electron magnetisation kappa_e = Omega_e / nu_e. */
  cli.nue = fabs(Q * cli.B0 / ME) / kappa_e;
  if (!(cli.nue > 0)) cli.nue = 1.0;
  get_real(argc, argv, "-nue", &cli.nue);
  double ndim_arg = 3;
  get_real(argc, argv, "-ndim", &ndim_arg);
  const int ndim = (int)ndim_arg;
  cli.ndim = ndim;
  const int fused = has_flag(argc, argv, "-fused");
  const int mixed = has_flag(argc, argv, "-mixed");
  const char *pdist = get_str(argc, argv, "--position-dist");
  int pdist_id = -1; /* PDistType order, src/positions.c:8-10 */
  if (pdist) {
    static const char *names[] = {"sobol", "reverse", "normal", "uniform", "sinusoidal", "gaussian"};
    for (int i = 0; i < 6; ++i)
      if (!strcmp(pdist, names[i])) pdist_id = i;
    if (pdist_id < 0) {
      fprintf(stderr, "unknown --position-dist %s\n", pdist);
      return 1;
    }
  }
  const int reference_ics = pdist_id >= 0;

  ecsprt_params p;
  ecsprt_hip_ctx *ctx = NULL;
  if (ecsprt_hip_setup_context(&cli, &p) != ECSPRT_OK) return 1;
  ecsprt_options o;
  ecsprt_hip_default_options(&o);
  o.ksp_rtol = rtol;
  o.seed = (uint64_t)seed;
  o.timing = 1;
  o.collide_mode = has_flag(argc, argv, "-no-collide") ? ECSPRT_COLLIDE_OFF : ECSPRT_COLLIDE_PHILOX;
  if (mixed) {
    o.particle_precision = ECSPRT_PARTICLES_MIXED;
    o.solve_precision = ECSPRT_SOLVE_MIXED;
  }
  CHECK(ecsprt_hip_create(&p, &o, &ctx));
  const long ncell = (long)p.Nx * p.Ny * p.Nz;
  const long np = Np > 0 ? (long)Np : (long)(-Np) * ncell; /* src/pic.c:100-106 */

  /* Plugin slots bound as in SetUpContext (src/setup.c:106-114). */
  Species ions = {ecsprt_hip_push, ecsprt_hip_collide, ecsprt_hip_collect};

  printf("grid %dx%dx%d, %ld ions, %d steps, dt %g, kappa_e %g, %s timestep, %s precision, %s initial conditions%s%s\n",
         p.Nx, p.Ny, p.Nz, np, (int)Nt, dt, p.kappa_e, fused ? "fused" : "staged", mixed ? "mixed" : "FP64",
         reference_ics ? "reference" : "device", reference_ics ? ", positions " : "", reference_ics ? pdist : "");
  if (reference_ics)
    CHECK(ecsprt_hip_init_distribution(ctx, np, pdist_id)); /* InitializePositions/Velocities, src/pic.c:203-207 */
  else
    CHECK(ecsprt_hip_init_particles(ctx, np, (uint64_t)seed, 0.0, 0));
  CHECK(ions.collect(ctx));                 /* src/pic.c:211 */
  CHECK(ecsprt_hip_compute_potential(ctx)); /* src/pic.c:215 */

  double t0 = now_s(), ms[6], acc[6] = {0, 0, 0, 0, 0, 0};
  for (int it = 0; it < (int)Nt; ++it) {
    if (fused) {
      CHECK(ecsprt_hip_step(ctx, dt));
      ecsprt_hip_stage_times(ctx, ms);
      for (int k = 0; k < 6; ++k) acc[k] += ms[k];
    } else {
      /* UpdateVelocities (src/pic.c:246, src/velocities.c:375-390): push, then the
       * collision slot when the neutrals have mass. */
      CHECK(ions.push(dt, ctx));
      if (p.mn > 0.0) CHECK(ions.collide(ndim, dt, ctx));
      CHECK(ecsprt_hip_update_positions(ndim, dt, ctx)); /* src/pic.c:250 */
      CHECK(ions.collect(ctx));                       /* src/pic.c:254 */
      CHECK(ecsprt_hip_compute_potential(ctx));       /* src/pic.c:258 */
    }
    int iters;
    double rn, bn;
    ecsprt_hip_solver_stats(ctx, &iters, &rn, &bn);
    int Nc;
    long Ns, Nf, natt;
    ecsprt_hip_collision_stats(ctx, &Nc, &Ns, &Nf, &natt);
    printf("< Time step %d > KSP iterations %d, ||r||/||b|| = %.2e, collisions %ld of Nc %d (%ld attempts)\n", it,
           iters, bn > 0 ? rn / bn : 0.0, Ns, Nc, natt);
  }
  ecsprt_hip_synchronize(ctx);
  double el = now_s() - t0;
  printf("%.3f s for %d steps: %.3e particle-steps/s\n", el, (int)Nt, np * Nt / el);
  if (fused)
    printf("mean ms: particle kernel %.3f, re-bin+fold %.3f, rhs %.3f, solve %.3f\n", acc[0] / Nt, acc[2] / Nt,
           acc[4] / Nt, acc[5] / Nt);
  ecsprt_hip_destroy(ctx);
  return 0;
}
