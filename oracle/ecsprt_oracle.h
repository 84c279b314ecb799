/*
 * ecsprt_oracle.h -- CPU restatement of ECSPRT's per-timestep hybrid-PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP path
 * (libecsprt_hip.so).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the CPU baseline.
 * Nothing in the product links it.
 *
 * Every function restates one reference function, cited as file:line relative to
 * the reference repository root.  Arithmetic is FP64 with the same operation order
 * as the reference; the library is compiled with -ffp-contract=off so that no FMA
 * contraction changes a rounding.  Layouts follow the reference:
 *   particles  : AoS double[np][3]  (DMSwarm "DMSwarmPIC_coor" and "velocity",
 *                src/setup.c:494)
 *   moments    : double[Nz][Ny][Nx][4] = {n, Gx, Gy, Gz} (FluidNode, src/ecsprt.h:194-199)
 *   potential  : double[Nz][Ny][Nx]
 * Ghosted DMDA local arrays are emulated by periodic index wrap (all-periodic
 * boundaries, one MPI rank).
 *
 * Pinning status of each function (what checks it, besides the HIP path) is listed
 * in the header comment of ecsprt_oracle.c.
 */
#ifndef ECSPRT_ORACLE_H
#define ECSPRT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants, src/constants.h:4-7. */
#define OR_Q  1.6022e-19
#define OR_MP 1.6726e-27
#define OR_ME 9.1094e-31
#define OR_KB 1.3807e-23

/* Subset of CLI (src/ecsprt.h:137-191) that the hot path reads. */
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
} oracle_cli;

/* Derived scalars of Context (src/ecsprt.h:118-134) that the hot path reads. */
typedef struct {
  int    Nx, Ny, Nz;
  double dx, dy, dz, Lx, Ly, Lz, x0, y0, z0, x1, y1, z1;
  double B0, E0, n0;
  /* ions */
  double qi, mi, nui, vi0x, vi0y, vi0z, vi0, viTx, viTy, viTz, viT, Ti, Omega_i, kappa_i;
  /* electrons */
  double qe, me, nue, Te, veT, Omega_e, kappa_e;
  /* neutrals */
  double mn, vn0x, vn0y, vn0z, vn0, vnT, Tn;
  double scale;   /* Linear.scale = 1/(1+kappa_e^2) */
  double gammaT;  /* electron thermal coefficient (1) */
} oracle_params;

/* One collision attempt's random numbers, in the order ScatterElastic draws them
 * (src/collisions.c:72,80,88,100,102).  u_acc1*u_acc2 is the acceptance draw:
 * PetscSqr(ran3(&seed)) is the macro ((a)*(a)) and so calls ran3 twice. */
typedef struct {
  float u_ip;
  float g[3];
  float u_acc1, u_acc2;
  float u_z, u_phi;
} oracle_attempt;

/* Static state of ran1/ran3/gasdev (src/random.c:28-29,96-98,174-175). */
typedef struct {
  long  r1_iy;
  long  r1_iv[32];
  int   r3_inext, r3_inextp, r3_iff;
  long  r3_ma[56];
  int   g_iset;
  float g_gset;
} oracle_nr_state;

/* setup.c */
void oracle_setup_context(const oracle_cli *cli, oracle_params *p);

/* velocities.c / vector-math.c */
void oracle_difference_vector3d(const double *phi, int Nx, int Ny, int Nz,
                                double x, double y, double z, double f[3]);
void oracle_boris_bz3d(const oracle_params *p, double dt, long np,
                       const double *pos, double *vel, const double *phi);

/* positions.c / boundaries.c */
void oracle_update_positions(double dt, long np, double *pos, const double *vel);
void oracle_apply_bc(const oracle_params *p, long np, double *pos);
long oracle_count_outside(const oracle_params *p, long np, const double *pos);

/* moments.c */
void oracle_collect_moments(const oracle_params *p, long np, const double *pos,
                            const double *vel, double *moments);
void oracle_cell_counts(const oracle_params *p, long np, const double *pos, int *counts);

/* random.c */
void  oracle_nr_init(oracle_nr_state *s);
float oracle_ran1(oracle_nr_state *s, long *idum);
float oracle_ran3(oracle_nr_state *s, long *idum);
float oracle_gasdev(oracle_nr_state *s, long *idum);
int   oracle_nr_stream(int which, long seed, long n, float *out);
long  oracle_getseed(int rank);

/* collisions.c */
int  oracle_collision_count(const oracle_params *p, long np, double dt);
int  oracle_scatter_elastic_nr(const oracle_params *p, double dt, long np, double *vel,
                               oracle_nr_state *st, long seed, oracle_attempt *log,
                               long log_cap, long *n_log, int *Ns, int *Nf);
int  oracle_scatter_elastic_predrawn(const oracle_params *p, double dt, long np, double *vel,
                                     const oracle_attempt *att, long n_att, int Nc,
                                     long *n_used, int *Ns, int *Nf);
void oracle_collide_one(const oracle_params *p, const oracle_attempt *a, const double vi_in[3],
                        double vf_out[3], int *accepted, int *failed);

/* ndim-generic forms (2 or 3): vel is double[np][ndim]; vi/vf of the single attempt
 * hold ndim components (the rest ignored). */
void oracle_collide_one_nd(const oracle_params *p, int ndim, const oracle_attempt *a, const double vi_in[3],
                           double vf_out[3], int *accepted, int *failed);
int  oracle_scatter_elastic_nr_nd(const oracle_params *p, int ndim, double dt, long np, double *vel,
                                  oracle_nr_state *st, long seed, oracle_attempt *log,
                                  long log_cap, long *n_log, int *Ns, int *Nf);
int  oracle_scatter_elastic_predrawn_nd(const oracle_params *p, int ndim, double dt, long np, double *vel,
                                        const oracle_attempt *att, long n_att, int Nc,
                                        long *n_used, int *Ns, int *Nf);

/* Two-dimensional variants (ndim = 2): particles double[np][2], moments [Ny][Nx][4],
 * phi and right-hand side [Ny][Nx]. */
void oracle_difference_vector2d(const double *phi, int Nx, int Ny, double x, double y, double f[2]);
void oracle_boris_bz2d(const oracle_params *p, double dt, long np, const double *pos, double *vel,
                       const double *phi);
void oracle_update_positions2d(double dt, long np, double *pos, const double *vel);
void oracle_apply_bc2d(const oracle_params *p, long np, double *pos);
void oracle_collect_moments2d(const oracle_params *p, long np, const double *pos, const double *vel,
                              double *moments);
void oracle_full_lhs2d_coeffs(const oracle_params *p, const double *moments, double *coef);
void oracle_full_lhs2d_csr(const oracle_params *p, const double *moments, int *rowptr, int *cols, double *vals);
void oracle_full_lhs2d_apply(const oracle_params *p, const double *moments, const double *x, double *y);
void oracle_full_rhs2d(const oracle_params *p, const double *moments, double *b);

/* lhs.c / rhs.c */
void oracle_full_lhs3d_coeffs(const oracle_params *p, const double *moments, double *coef);
void oracle_full_lhs3d_csr(const oracle_params *p, const double *moments,
                           int *rowptr, int *cols, double *vals);
void oracle_full_lhs3d_apply(const oracle_params *p, const double *moments,
                             const double *x, double *y);
void oracle_laplacian_lhs3d_apply(const oracle_params *p, const double *x, double *y);
void oracle_full_rhs3d(const oracle_params *p, const double *moments, double *b);
void oracle_sinusoidal_rhs3d(const oracle_params *p, double *b);
void oracle_remove_mean(long n, double *v);

/* random.c:233-289 (Sobseq, Numerical Recipes' sobseq with MAXBIT 30, MAXDIM 6) with its
 * static state held explicitly. */
typedef struct {
  unsigned long in;
  unsigned long ix[7];
  unsigned long iv[181];
  double fac;
} oracle_sobol_state;
void oracle_sobseq_init(oracle_sobol_state *s);
int  oracle_sobseq_next(oracle_sobol_state *s, int n, double *x);

/* PETSc 3.20.2's default PetscRandom (the rand48 linear congruential generator). */
typedef struct {
  unsigned long long x;
} oracle_rand48;
void   oracle_rand48_seed(oracle_rand48 *r, long seed);
double oracle_rand48_next(oracle_rand48 *r);

/* positions.c:139-234 (SobolDistribution) and positions.c:236-388 (Rejection of
 * SinusoidalDistribution), one rank: positions of np ions, pos[3*np]. */
int  oracle_sobol_positions(const oracle_params *p, long np, double *pos);
long oracle_sinusoidal_positions(const oracle_params *p, long np, double *pos);
long oracle_sinusoidal_positions_nd(const oracle_params *p, int ndim, long np, double *pos);

/* CPU baseline: Nt steps of push + collide + drift/wrap + deposit in the
 * reference loop order (src/pic.c:246-254), per-stage seconds returned in t[4]. */
int oracle_particle_steps(const oracle_params *p, double dt, int nt, long np, double *pos,
                          double *vel, const double *phi, double *moments,
                          oracle_nr_state *st, int collide, double t[4]);

#ifdef __cplusplus
}
#endif
/* ecsprt_solver.c: FP64 GMRES(restart) + geometric multigrid on one core, A from the
 * density of m_op (FluidNode moments), x overwritten (zero initial guess), converged on
 * ||b - A x|| <= rtol ||b||; returns 0 on convergence. */
int oracle_solve_gmres_mg(const oracle_params *p, const double *m_op, const double *b, double *x, double rtol,
                          int restart, int maxit, int *iters, double *rnorm);

#endif
