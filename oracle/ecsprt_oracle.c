/*
 * ecsprt_oracle.c -- CPU restatement of ECSPRT's hybrid-PIC hot path (TEST INFRASTRUCTURE).
 *
 * This file is the parity oracle for the HIP library.  It is loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product never links it.
 * Each function cites the reference function it restates (paths relative to the
 * reference repository root).  Build: oracle/Makefile (gcc -O3 -ffp-contract=off).
 *
 * How each function is pinned (the reference itself cannot be built here: every
 * hot-path file includes petsc.h and PETSc is absent, so no oracle/_ref exists):
 *   oracle_difference_vector3d  closed forms: phi = x^2 on nodes exposes the x/z
 *                               weight swap; quadratic phi exposes the one-sided
 *                               last-cell branch; linear phi gives 2*slope*dx.
 *   oracle_boris_bz3d           constant phi: exact rotation by 2*atan(t), |v_perp|
 *                               preserved, vz unchanged; t=0: pure half-kick.
 *   oracle_update_positions /   exact arithmetic; edge cases L-ulp, -ulp (wraps to L),
 *   oracle_apply_bc             >= L, NaN.
 *   oracle_collect_moments      sum(n)=np, sum(G)=sum(v), node-centred particle hits one
 *                               node, brute-force numpy CIC.
 *   oracle_cell_counts          numpy histogram of floor(pos/dx).
 *   oracle_ran1                 Park-Miller minimal standard: 10000th state from seed 1
 *                               is 1043618065 (Park & Miller 1988); outputs are a
 *                               shuffle of the LCG stream.
 *   oracle_ran3                 lagged-subtraction identity x_n = x_{n-55} - x_{n-24}
 *                               (mod 1e9) on consecutive outputs.
 *   oracle_gasdev               polar-method identity and N(0,1) moments.
 *   oracle_collision_count      hand-computed cases incl. int overflow -> 0.
 *   oracle_collide_one          centre-of-mass invariants: |vf-vc| = vcr, pair momentum
 *                               and energy conserved by the implied neutral velocity.
 *   oracle_full_lhs3d_*         zero row and column sums; uniform n reduces to
 *                               scale*n0*(Lxx+Lyy+(1+k^2)Lzz) solved exactly by FFT;
 *                               scipy sparse solve of the assembled matrix.
 *   oracle_full_rhs3d           closed form for sinusoidal n and G; zero mean.
 *   oracle_sinusoidal_rhs3d     closed form cos products.
 *   oracle_setup_context        hand-computed derived coefficients.
 *   oracle_scatter_elastic_*    predrawn and NR-stream paths agree when fed the NR
 *                               log; invariants above per collision.
 *   oracle_sobseq_*             dimensions 1-2 (where the reference's recurrence and
 *                               Numerical Recipes' agree) against an independent
 *                               Bratley-Fox implementation of the published
 *                               algorithm; every dimension's first 2^m points put one
 *                               point in each interval of width 2^-m.
 *   oracle_rand48_*             glibc's srand48/drand48 stream, value for value.
 *   oracle_*_positions          Sobol: the sequence times L; sinusoidal: acceptance
 *                               ratio and x histogram against 1 + 0.25 sin(2 pi x).
 */
#include "ecsprt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <float.h>

static inline int wrapi(int i, int n)
{
  int r = i % n;
  return r < 0 ? r + n : r;
}

static inline long nidx(int i, int j, int k, int Nx, int Ny, int Nz)
{
  return ((long)wrapi(k, Nz) * Ny + wrapi(j, Ny)) * Nx + wrapi(i, Nx);
}

/* ------------------------------------------------------------------------- */
/* setup.c:116-271 (SetUpContext) and setup.c:296-346 (NormalizeGrid).        */
/* ------------------------------------------------------------------------- */
void oracle_setup_context(const oracle_cli *c, oracle_params *p)
{
  double tmp;
  memset(p, 0, sizeof(*p));
  p->qe = -OR_Q;
  p->me = OR_ME;
  p->n0 = c->n0;
  p->B0 = c->B0;
  p->E0 = c->E0;
  p->qi = c->qi;
  p->mi = c->mi;
  p->vi0x = c->vi0x; p->vi0y = c->vi0y; p->vi0z = c->vi0z;
  p->viTx = c->viTx; p->viTy = c->viTy; p->viTz = c->viTz;
  p->Ti = c->Ti;
  p->nui = c->nui;
  p->nue = c->nue;
  p->Te = c->Te;
  p->mn = c->mn;
  p->vn0x = c->vn0x; p->vn0y = c->vn0y; p->vn0z = c->vn0z;
  p->Tn = c->Tn;

  p->Nx = c->Nx > 0 ? c->Nx : 7;
  p->Ny = c->Ny > 0 ? c->Ny : 7;
  p->Nz = c->Nz > 0 ? c->Nz : 7;
  p->dx = c->dx; p->dy = c->dy; p->dz = c->dz;
  p->x0 = c->x0; p->y0 = c->y0; p->z0 = c->z0;
  p->x1 = c->x1; p->y1 = c->y1; p->z1 = c->z1;
  /* NormalizeGrid, setup.c:319-342 */
  if (p->dx <= 0.0) { p->Lx = p->x1 - p->x0; p->dx = p->Lx / (double)p->Nx; }
  else { p->Lx = p->dx * (double)p->Nx; p->x0 = 0.0; p->x1 = p->Lx; }
  if (p->dy <= 0.0) { p->Ly = p->y1 - p->y0; p->dy = p->Ly / (double)p->Ny; }
  else { p->Ly = p->dy * (double)p->Ny; p->y0 = 0.0; p->y1 = p->Ly; }
  if (p->dz <= 0.0) { p->Lz = p->z1 - p->z0; p->dz = p->Lz / (double)p->Nz; }
  else { p->Lz = p->dz * (double)p->Nz; p->z0 = 0.0; p->z1 = p->Lz; }

  /* setup.c:222-230 */
  p->Omega_e = fabs(p->qe * p->B0 / p->me);
  p->Omega_i = fabs(p->qi * p->B0 / p->mi);
  p->kappa_e = p->Omega_e / p->nue;
  p->kappa_i = p->Omega_i / p->nui;
  p->scale = 1.0 / (1 + p->kappa_e * p->kappa_e);

  /* setup.c:233-235 */
  p->vi0 = sqrt(p->vi0x * p->vi0x + p->vi0y * p->vi0y + p->vi0z * p->vi0z);
  p->vn0 = sqrt(p->vn0x * p->vn0x + p->vn0y * p->vn0y + p->vn0z * p->vn0z);

  /* setup.c:238-247 (electrons) */
  if (c->veTx != 0.0 || c->veTy != 0.0 || c->veTz != 0.0) {
    p->veT = sqrt(c->veTx * c->veTx + c->veTy * c->veTy + c->veTz * c->veTz);
    p->Te = (0.5 * p->me / OR_KB) * (p->veT * p->veT);
  } else {
    tmp = sqrt(2.0 * OR_KB * p->Te / p->me) / 3.0;
    p->veT = sqrt(tmp * tmp + tmp * tmp + tmp * tmp);
  }
  /* setup.c:250-259 (ions) */
  if (p->viTx != 0.0 || p->viTy != 0.0 || p->viTz != 0.0) {
    p->viT = sqrt(p->viTx * p->viTx + p->viTy * p->viTy + p->viTz * p->viTz);
    p->Ti = (0.5 * p->mi / OR_KB) * (p->viT * p->viT);
  } else {
    tmp = sqrt(2.0 * OR_KB * p->Ti / p->mi) / 3.0;
    p->viTx = tmp; p->viTy = tmp; p->viTz = tmp;
    p->viT = sqrt(p->viTx * p->viTx + p->viTy * p->viTy + p->viTz * p->viTz);
  }
  /* setup.c:262-268 (neutrals) */
  if (p->Tn == -1.0) p->Tn = p->Ti;
  p->vnT = sqrt(2.0 * OR_KB * p->Tn / p->mn);
  /* setup.c:271 */
  p->gammaT = 1.0;
}

/* ------------------------------------------------------------------------- */
/* vector-math.c:596-722 (DifferenceVector3D).  F is the unghosted potential;  */
/* the ghost width-2 local array of the reference is emulated by wrapping.    */
/* The interpolation keeps the reference's ordering: y with wyh, then z with  */
/* wxh, then x with wzh (vector-math.c:645-651).                              */
/* ------------------------------------------------------------------------- */
#define PHI(kk, jj, ii) F[nidx((ii), (jj), (kk), Nx, Ny, Nz)]

void oracle_difference_vector3d(const double *F, int Nx, int Ny, int Nz,
                                double x0, double y0, double z0, double f[3])
{
  int ixl, ixh, iyl, iyh, izl, izh;
  double wxh, wyh, wzh;
  double hhh, lhh, hlh, llh, hhl, lhl, hll, lll;
  double whh, whl, wlh, wll, Ewh, Ewl;

  ixl = (int)x0; ixh = ixl + 1; wxh = x0 - (double)ixl;
  iyl = (int)y0; iyh = iyl + 1; wyh = y0 - (double)iyl;
  izl = (int)z0; izh = izl + 1; wzh = z0 - (double)izl;

  /* x component, vector-math.c:625-657 */
  if (ixl >= 0) {
    hhl = PHI(izh, iyh, ixl + 1) - PHI(izh, iyh, ixl - 1);
    lhl = PHI(izl, iyh, ixl + 1) - PHI(izl, iyh, ixl - 1);
    hll = PHI(izh, iyl, ixl + 1) - PHI(izh, iyl, ixl - 1);
    lll = PHI(izl, iyl, ixl + 1) - PHI(izl, iyl, ixl - 1);
  } else {
    hhl = -1.0 * PHI(izh, iyh, ixl + 2) + 4.0 * PHI(izh, iyh, ixl + 1) - 3.0 * PHI(izh, iyh, ixl);
    lhl = -1.0 * PHI(izl, iyh, ixl + 2) + 4.0 * PHI(izl, iyh, ixl + 1) - 3.0 * PHI(izl, iyh, ixl);
    hll = -1.0 * PHI(izh, iyl, ixl + 2) + 4.0 * PHI(izh, iyl, ixl + 1) - 3.0 * PHI(izh, iyl, ixl);
    lll = -1.0 * PHI(izl, iyl, ixl + 2) + 4.0 * PHI(izl, iyl, ixl + 1) - 3.0 * PHI(izl, iyl, ixl);
  }
  if (ixh < Nx) {
    hhh = PHI(izh, iyh, ixh + 1) - PHI(izh, iyh, ixh - 1);
    lhh = PHI(izl, iyh, ixh + 1) - PHI(izl, iyh, ixh - 1);
    hlh = PHI(izh, iyl, ixh + 1) - PHI(izh, iyl, ixh - 1);
    llh = PHI(izl, iyl, ixh + 1) - PHI(izl, iyl, ixh - 1);
  } else {
    hhh = +3.0 * PHI(izh, iyh, ixh) - 4.0 * PHI(izh, iyh, ixh - 1) + 1.0 * PHI(izh, iyh, ixh - 2);
    lhh = +3.0 * PHI(izl, iyh, ixh) - 4.0 * PHI(izl, iyh, ixh - 1) + 1.0 * PHI(izl, iyh, ixh - 2);
    hlh = +3.0 * PHI(izh, iyl, ixh) - 4.0 * PHI(izh, iyl, ixh - 1) + 1.0 * PHI(izh, iyl, ixh - 2);
    llh = +3.0 * PHI(izl, iyl, ixh) - 4.0 * PHI(izl, iyl, ixh - 1) + 1.0 * PHI(izl, iyl, ixh - 2);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[0] = Ewl + wzh * (Ewh - Ewl);

  /* y component, vector-math.c:659-691 */
  if (iyl >= 0) {
    hlh = PHI(izh, iyl + 1, ixh) - PHI(izh, iyl - 1, ixh);
    llh = PHI(izl, iyl + 1, ixh) - PHI(izl, iyl - 1, ixh);
    hll = PHI(izh, iyl + 1, ixl) - PHI(izh, iyl - 1, ixl);
    lll = PHI(izl, iyl + 1, ixl) - PHI(izl, iyl - 1, ixl);
  } else {
    hlh = -1.0 * PHI(izh, iyl + 2, ixh) + 4.0 * PHI(izh, iyl + 1, ixh) - 3.0 * PHI(izh, iyl, ixh);
    llh = -1.0 * PHI(izl, iyl + 2, ixh) + 4.0 * PHI(izl, iyl + 1, ixh) - 3.0 * PHI(izl, iyl, ixh);
    hll = -1.0 * PHI(izh, iyl + 2, ixl) + 4.0 * PHI(izh, iyl + 1, ixl) - 3.0 * PHI(izh, iyl, ixl);
    lll = -1.0 * PHI(izl, iyl + 2, ixl) + 4.0 * PHI(izl, iyl + 1, ixl) - 3.0 * PHI(izl, iyl, ixl);
  }
  if (iyh < Ny) {
    hhh = PHI(izh, iyh + 1, ixh) - PHI(izh, iyh - 1, ixh);
    lhh = PHI(izl, iyh + 1, ixh) - PHI(izl, iyh - 1, ixh);
    hhl = PHI(izh, iyh + 1, ixl) - PHI(izh, iyh - 1, ixl);
    lhl = PHI(izl, iyh + 1, ixl) - PHI(izl, iyh - 1, ixl);
  } else {
    hhh = +3.0 * PHI(izh, iyh, ixh) - 4.0 * PHI(izh, iyh - 1, ixh) + 1.0 * PHI(izh, iyh - 2, ixh);
    lhh = +3.0 * PHI(izl, iyh, ixh) - 4.0 * PHI(izl, iyh - 1, ixh) + 1.0 * PHI(izl, iyh - 2, ixh);
    hhl = +3.0 * PHI(izh, iyh, ixl) - 4.0 * PHI(izh, iyh - 1, ixl) + 1.0 * PHI(izh, iyh - 2, ixl);
    lhl = +3.0 * PHI(izl, iyh, ixl) - 4.0 * PHI(izl, iyh - 1, ixl) + 1.0 * PHI(izl, iyh - 2, ixl);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[1] = Ewl + wzh * (Ewh - Ewl);

  /* z component, vector-math.c:693-719 */
  if (izl >= 0) {
    lhh = PHI(izl + 1, iyh, ixh) - PHI(izl - 1, iyh, ixh);
    llh = PHI(izl + 1, iyl, ixh) - PHI(izl - 1, iyl, ixh);
    lhl = PHI(izl + 1, iyh, ixl) - PHI(izl - 1, iyh, ixl);
    lll = PHI(izl + 1, iyl, ixl) - PHI(izl - 1, iyl, ixl);
  } else {
    lhh = -1.0 * PHI(izl + 2, iyh, ixh) + 4.0 * PHI(izl + 1, iyh, ixh) - 3.0 * PHI(izl, iyh, ixh);
    llh = -1.0 * PHI(izl + 2, iyl, ixh) + 4.0 * PHI(izl + 1, iyl, ixh) - 3.0 * PHI(izl, iyl, ixh);
    lhl = -1.0 * PHI(izl + 2, iyh, ixl) + 4.0 * PHI(izl + 1, iyh, ixl) - 3.0 * PHI(izl, iyh, ixl);
    lll = -1.0 * PHI(izl + 2, iyl, ixl) + 4.0 * PHI(izl + 1, iyl, ixl) - 3.0 * PHI(izl, iyl, ixl);
  }
  if (izh < Nz) {
    hhh = PHI(izh + 1, iyh, ixh) - PHI(izh - 1, iyh, ixh);
    hlh = PHI(izh + 1, iyl, ixh) - PHI(izh - 1, iyl, ixh);
    hhl = PHI(izh + 1, iyh, ixl) - PHI(izh - 1, iyh, ixl);
    hll = PHI(izh + 1, iyl, ixl) - PHI(izh - 1, iyl, ixl);
  } else {
    hhh = +3.0 * PHI(izh, iyh, ixh) - 4.0 * PHI(izh - 1, iyh, ixh) + 1.0 * PHI(izh - 2, iyh, ixh);
    hlh = +3.0 * PHI(izh, iyl, ixh) - 4.0 * PHI(izh - 1, iyl, ixh) + 1.0 * PHI(izh - 2, iyl, ixh);
    hhl = +3.0 * PHI(izh, iyh, ixl) - 4.0 * PHI(izh - 1, iyh, ixl) + 1.0 * PHI(izh - 2, iyh, ixl);
    hll = +3.0 * PHI(izh, iyl, ixl) - 4.0 * PHI(izh - 1, iyl, ixl) + 1.0 * PHI(izh - 2, iyl, ixl);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[2] = Ewl + wzh * (Ewh - Ewl);
}
#undef PHI

/* velocities.c:281-371 (BorisMoverBz3D): first half E-kick, shear rotation, no
 * second half-kick.  Coordinates are divided by dx (velocities.c:342-344).    */
void oracle_boris_bz3d(const oracle_params *p, double dt, long np,
                       const double *pos, double *vel, const double *phi)
{
  double h[3] = {1.0 / p->dx, 1.0 / p->dy, 1.0 / p->dz};
  double tscale = 0.5 * (p->qi / p->mi) * dt;
  double t = tscale * p->B0;
  double s = 2.0 * t / (1 + t * t);
  double Escale[3];
  double E[3];
  long ip;
  int d;
  for (d = 0; d < 3; d++) Escale[d] = -0.5 * h[d] * tscale;
  for (ip = 0; ip < np; ip++) {
    double x = pos[ip * 3 + 0] / p->dx;
    double y = pos[ip * 3 + 1] / p->dy;
    double z = pos[ip * 3 + 2] / p->dz;
    double vxm, vym, vz, vxt, vxp, vyp;
    oracle_difference_vector3d(phi, p->Nx, p->Ny, p->Nz, x, y, z, E);
    vxm = vel[ip * 3 + 0] + Escale[0] * E[0];
    vym = vel[ip * 3 + 1] + Escale[1] * E[1];
    vz  = vel[ip * 3 + 2] + Escale[2] * E[2];
    vxt = vxm + vym * t;
    vyp = vym - vxt * s;
    vxp = vxt + vyp * t;
    vel[ip * 3 + 0] = vxp;
    vel[ip * 3 + 1] = vyp;
    vel[ip * 3 + 2] = vz;
  }
}

/* positions.c:455-459 (UpdatePositions drift). */
void oracle_update_positions(double dt, long np, double *pos, const double *vel)
{
  long ip;
  int d;
  for (ip = 0; ip < np; ip++)
    for (d = 0; d < 3; d++) pos[ip * 3 + d] += vel[ip * 3 + d] * dt;
}

/* boundaries.c:6-34 (ApplyBC, all-periodic): one wrap per axis. */
void oracle_apply_bc(const oracle_params *p, long np, double *pos)
{
  double L[3] = {p->Lx, p->Ly, p->Lz};
  double r0[3] = {p->x0, p->y0, p->z0};
  double r1[3] = {p->x1, p->y1, p->z1};
  long ip;
  int d;
  for (ip = 0; ip < np; ip++) {
    for (d = 0; d < 3; d++) {
      double r = pos[ip * 3 + d];
      if (r < r0[d]) r += L[d];
      if (r >= r1[d]) r -= L[d];
      pos[ip * 3 + d] = r;
    }
  }
}

/* Particles DMSwarmMigrate would drop (boundaries.c:92, 101-109): a coordinate
 * outside [x0, x1] or NaN.  A coordinate equal to x1 (a -ulp wrap) is kept: the
 * cell DM's point location treats the closed upper face as inside.            */
long oracle_count_outside(const oracle_params *p, long np, const double *pos)
{
  double r0[3] = {p->x0, p->y0, p->z0};
  double r1[3] = {p->x1, p->y1, p->z1};
  long ip, n = 0;
  int d;
  for (ip = 0; ip < np; ip++) {
    int out = 0;
    for (d = 0; d < 3; d++) {
      double r = pos[ip * 3 + d];
      if (!(r >= r0[d] && r <= r1[d])) out = 1;
    }
    n += out;
  }
  return n;
}

/* moments.c:108-226 (Collect3DFluidMoments).  Accumulation into the ghost-width-1
 * local array followed by the ADD_VALUES fold equals accumulation at the wrapped
 * node; a coordinate equal to L (node N+1, outside the reference's ghosted array)
 * is wrapped the same way.                                                     */
void oracle_collect_moments(const oracle_params *p, long np, const double *pos,
                            const double *vel, double *m)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  long ncell = (long)Nx * Ny * Nz;
  long ip;
  memset(m, 0, sizeof(double) * 4 * ncell);
  for (ip = 0; ip < np; ip++) {
    double x = pos[ip * 3 + 0] / p->dx;
    double y = pos[ip * 3 + 1] / p->dy;
    double z = pos[ip * 3 + 2] / p->dz;
    int ixl = (int)x, ixh = ixl + 1;
    int iyl = (int)y, iyh = iyl + 1;
    int izl = (int)z, izh = izl + 1;
    double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    double wzh = z - (double)izl, wzl = 1.0 - wzh;
    double w8[8];
    long n8[8];
    int c, d;
    w8[0] = wzh * wyh * wxh; n8[0] = nidx(ixh, iyh, izh, Nx, Ny, Nz);
    w8[1] = wzl * wyh * wxh; n8[1] = nidx(ixh, iyh, izl, Nx, Ny, Nz);
    w8[2] = wzh * wyl * wxh; n8[2] = nidx(ixh, iyl, izh, Nx, Ny, Nz);
    w8[3] = wzl * wyl * wxh; n8[3] = nidx(ixh, iyl, izl, Nx, Ny, Nz);
    w8[4] = wzh * wyh * wxl; n8[4] = nidx(ixl, iyh, izh, Nx, Ny, Nz);
    w8[5] = wzl * wyh * wxl; n8[5] = nidx(ixl, iyh, izl, Nx, Ny, Nz);
    w8[6] = wzh * wyl * wxl; n8[6] = nidx(ixl, iyl, izh, Nx, Ny, Nz);
    w8[7] = wzl * wyl * wxl; n8[7] = nidx(ixl, iyl, izl, Nx, Ny, Nz);
    for (c = 0; c < 8; c++) m[n8[c] * 4 + 0] += w8[c];
    for (d = 0; d < 3; d++) {
      double w = vel[ip * 3 + d];
      for (c = 0; c < 8; c++) m[n8[c] * 4 + 1 + d] += w * w8[c];
    }
  }
}

/* Per-cell particle counts from (int)(pos/d) per axis, the cell the reference's
 * deposit assigns (moments.c:154-174); coordinates equal to L count in cell 0. */
void oracle_cell_counts(const oracle_params *p, long np, const double *pos, int *counts)
{
  long ip;
  memset(counts, 0, sizeof(int) * (long)p->Nx * p->Ny * p->Nz);
  for (ip = 0; ip < np; ip++) {
    int i = (int)(pos[ip * 3 + 0] / p->dx);
    int j = (int)(pos[ip * 3 + 1] / p->dy);
    int k = (int)(pos[ip * 3 + 2] / p->dz);
    counts[nidx(i, j, k, p->Nx, p->Ny, p->Nz)]++;
  }
}

/* ------------------------------------------------------------------------- */
/* random.c: NR 2nd-edition generators with their static state made explicit. */
/* ------------------------------------------------------------------------- */
#define NR_IA 16807
#define NR_IM 2147483647
#define NR_AM (1.0 / NR_IM)
#define NR_IQ 127773
#define NR_IR 2836
#define NR_NTAB 32
#define NR_NDIV (1 + (NR_IM - 1) / NR_NTAB)
#define NR_RNMX (1.0 - DBL_EPSILON)
#define NR_MBIG 1000000000
#define NR_MSEED 161803398
#define NR_MZ 0
#define NR_FAC (1.0 / NR_MBIG)

void oracle_nr_init(oracle_nr_state *s) { memset(s, 0, sizeof(*s)); }

long oracle_getseed(int rank) { return (long)(-(rank + 1) * 12345); } /* random.c:292-295 */

/* random.c:233-289 (Sobseq).  The table initialisation loop of the reference runs
 * to j = MAXDIM*MAXBIT+1, one entry past both arrays; entries beyond the 24 listed
 * initial values are zero either way and the stray write is not restated.  The
 * recurrence for degrees above 2 XORs iu[j-1] for every set bit of the polynomial
 * where Numerical Recipes XORs iu[j-l]; the restatement follows the reference. */
void oracle_sobseq_init(oracle_sobol_state *s)
{
  static const unsigned long mdeg[7] = {0, 1, 2, 3, 3, 4, 4};
  static const unsigned long ip[7] = {0, 0, 1, 1, 2, 1, 4};
  static const unsigned long ic[25] = {0, 1, 1, 1, 1, 1, 1, 3, 1, 3, 3, 1, 1, 5, 7, 7, 3, 3, 5, 15, 11, 5, 15, 13, 9};
  const int MAXBIT = 30, MAXDIM = 6;
  int j, k, l;
  memset(s, 0, sizeof(*s));
  for (j = 1; j <= MAXDIM * MAXBIT; j++) s->iv[j] = j < 25 ? ic[j] : 0;
  /* iu[j] = &iv[(j-1)*MAXDIM], so iu[j][k] = iv[(j-1)*MAXDIM + k]. */
#define IU(J, K) s->iv[((J) - 1) * MAXDIM + (K)]
  for (k = 1; k <= MAXDIM; k++) {
    for (j = 1; j <= (int)mdeg[k]; j++) IU(j, k) <<= (MAXBIT - j);
    for (j = (int)mdeg[k] + 1; j <= MAXBIT; j++) {
      unsigned long ipp = ip[k];
      unsigned long i = IU(j - (int)mdeg[k], k);
      i ^= (i >> mdeg[k]);
      for (l = (int)mdeg[k] - 1; l >= 1; l--) {
        if (ipp & 1) i ^= IU(j - 1, k);
        ipp >>= 1;
      }
      IU(j, k) = i;
    }
  }
#undef IU
  s->fac = 1.0 / ((long)1 << MAXBIT);
  s->in = 0;
}

/* Next point: x[0..min(n,6)-1].  Returns -1 once 2^30 - 1 points have been drawn
 * (the reference's "MAXBIT too small" error). */
int oracle_sobseq_next(oracle_sobol_state *s, int n, double *x)
{
  const int MAXBIT = 30, MAXDIM = 6;
  unsigned long im = s->in++;
  int j, k;
  for (j = 1; j <= MAXBIT; j++) {
    if (!(im & 1)) break;
    im >>= 1;
  }
  if (j > MAXBIT) return -1;
  im = (unsigned long)(j - 1) * MAXDIM;
  for (k = 1; k <= (n < MAXDIM ? n : MAXDIM); k++) {
    s->ix[k] ^= s->iv[im + k];
    x[k - 1] = s->ix[k] * s->fac;
  }
  return 0;
}

/* PETSc's rand48 generators (PETSCRANDER48, and PETSCRAND48 over libc drand48): the
 * 48-bit LCG x <- (0x5DEECE66D x + 0xB) mod 2^48 seeded as srand48 seeds it,
 * x = (seed << 16) | 0x330E, each value x / 2^48 (exact in FP64). */
void oracle_rand48_seed(oracle_rand48 *r, long seed)
{
  r->x = (((unsigned long long)(unsigned int)seed) << 16) | 0x330EULL;
}

double oracle_rand48_next(oracle_rand48 *r)
{
  r->x = (0x5DEECE66DULL * r->x + 0xBULL) & 0xFFFFFFFFFFFFULL;
  return ldexp((double)r->x, -48);
}

/* positions.c:139-234 (SobolDistribution), one rank: Sobseq(-1) once, then one
 * 3-D point per ion, pos = s * L (no origin offset). */
int oracle_sobol_positions(const oracle_params *p, long np, double *pos)
{
  oracle_sobol_state st;
  const double L[3] = {p->Lx, p->Ly, p->Lz};
  double sv[3];
  long ip;
  int d;
  oracle_sobseq_init(&st);
  for (ip = 0; ip < np; ip++) {
    if (oracle_sobseq_next(&st, 3, sv)) return -1;
    for (d = 0; d < 3; d++) pos[ip * 3 + d] = sv[d] * L[d];
  }
  return 0;
}

/* positions.c:236-250 (SinusoidalDistribution): 1 + 0.25 sin(2 pi x), x in units of Lx. */
static double sinusoidal_density(const double r[3]) { return 1.0 + 0.25 * sin(2 * 3.14159265358979323846 * r[0]); }

/* positions.c:253-388 (Rejection), one rank: the maximum over grid nodes (i/Nx, j/Ny,
 * k/Nz) (the density depends on x alone), then per trial three values for the
 * position and one for the test, accepted when w/max > v.  Returns the trials. */
long oracle_sinusoidal_positions_nd(const oracle_params *p, int ndim, long np, double *pos)
{
  oracle_rand48 r;
  const double L[3] = {p->Lx, p->Ly, p->Lz};
  double maxVal = 0.0, normVal, sv[3], w, v;
  long ip = 0, it = 0;
  int i, d;
  oracle_rand48_seed(&r, 7);
  for (i = 0; i < p->Nx; i++) {
    sv[0] = (double)i / p->Nx;
    sv[1] = sv[2] = 0.0;
    w = sinusoidal_density(sv);
    maxVal = maxVal > w ? maxVal : w;
  }
  normVal = 1.0 / maxVal;
  while (ip < np) {
    for (d = 0; d < ndim; d++) sv[d] = oracle_rand48_next(&r);
    w = sinusoidal_density(sv);
    v = oracle_rand48_next(&r);
    if (w * normVal > v) {
      for (d = 0; d < ndim; d++) pos[ip * ndim + d] = sv[d] * L[d];
      ip++;
    }
    it++;
  }
  return it;
}

long oracle_sinusoidal_positions(const oracle_params *p, long np, double *pos)
{
  return oracle_sinusoidal_positions_nd(p, 3, np, pos);
}

/* random.c:24-72 (ran1). */
float oracle_ran1(oracle_nr_state *s, long *idum)
{
  int j;
  long k;
  float temp;
  if (*idum <= 0 || !s->r1_iy) {
    if (-(*idum) < 1) *idum = 1;
    else *idum = -(*idum);
    for (j = NR_NTAB + 7; j >= 0; j--) {
      k = (*idum) / NR_IQ;
      *idum = NR_IA * (*idum - k * NR_IQ) - NR_IR * k;
      if (*idum < 0) *idum += NR_IM;
      if (j < NR_NTAB) s->r1_iv[j] = *idum;
    }
    s->r1_iy = s->r1_iv[0];
  }
  k = (*idum) / NR_IQ;
  *idum = NR_IA * (*idum - k * NR_IQ) - NR_IR * k;
  if (*idum < 0) *idum += NR_IM;
  j = (int)(s->r1_iy / NR_NDIV);
  s->r1_iy = s->r1_iv[j];
  s->r1_iv[j] = *idum;
  if ((temp = NR_AM * s->r1_iy) > NR_RNMX) return NR_RNMX;
  return temp;
}

/* random.c:94-154 (ran3). */
float oracle_ran3(oracle_nr_state *s, long *idum)
{
  long mj, mk;
  int i, ii, k;
  if (*idum < 0 || s->r3_iff == 0) {
    s->r3_iff = 1;
    mj = labs(NR_MSEED - labs(*idum));
    mj %= NR_MBIG;
    s->r3_ma[55] = mj;
    mk = 1;
    for (i = 1; i <= 54; i++) {
      ii = (21 * i) % 55;
      s->r3_ma[ii] = mk;
      mk = mj - mk;
      if (mk < NR_MZ) mk += NR_MBIG;
      mj = s->r3_ma[ii];
    }
    for (k = 1; k <= 4; k++) {
      for (i = 1; i <= 55; i++) {
        s->r3_ma[i] -= s->r3_ma[1 + (i + 30) % 55];
        if (s->r3_ma[i] < NR_MZ) s->r3_ma[i] += NR_MBIG;
      }
    }
    s->r3_inext = 0;
    s->r3_inextp = 31;
    *idum = 1;
  }
  if (++s->r3_inext == 56) s->r3_inext = 1;
  if (++s->r3_inextp == 56) s->r3_inextp = 1;
  mj = s->r3_ma[s->r3_inext] - s->r3_ma[s->r3_inextp];
  if (mj < NR_MZ) mj += NR_MBIG;
  s->r3_ma[s->r3_inext] = mj;
  return mj * NR_FAC;
}

/* random.c:172-204 (gasdev, polar Box-Muller over ran1). */
float oracle_gasdev(oracle_nr_state *s, long *idum)
{
  float fac, rsq, v1, v2;
  if (*idum < 0) s->g_iset = 0;
  if (s->g_iset == 0) {
    do {
      v1 = 2.0 * oracle_ran1(s, idum) - 1.0;
      v2 = 2.0 * oracle_ran1(s, idum) - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (rsq >= 1.0 || rsq == 0.0);
    fac = sqrt(-2.0 * log(rsq) / rsq);
    s->g_gset = v1 * fac;
    s->g_iset = 1;
    return v2 * fac;
  }
  s->g_iset = 0;
  return s->g_gset;
}

/* n consecutive values of ran1 (which = 0), ran3 (1) or gasdev (2) from fresh static
 * state and seed `seed`: the bulk form of the three calls above. */
int oracle_nr_stream(int which, long seed, long n, float *out)
{
  oracle_nr_state s;
  long idum = seed, i;
  if (which < 0 || which > 2) return 1;
  oracle_nr_init(&s);
  for (i = 0; i < n; i++)
    out[i] = which == 0 ? oracle_ran1(&s, &idum) : which == 1 ? oracle_ran3(&s, &idum) : oracle_gasdev(&s, &idum);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* collisions.c:9-167 (ScatterElastic).                                      */
/* ------------------------------------------------------------------------- */

/* x86-64 double -> int32 conversion (cvttsd2si): out-of-range and NaN give INT_MIN. */
static int x86_double_to_int(double v)
{
  if (!(v > -2147483649.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)v;
}

static double vrm_of(const oracle_params *p)
{
  double vi0[3] = {p->vi0x, p->vi0y, p->vi0z};
  double vn0[3] = {p->vn0x, p->vn0y, p->vn0z};
  double tmp = 0.0;
  int d;
  for (d = 0; d < 3; d++) tmp += (vi0[d] - vn0[d]) * (vi0[d] - vn0[d]);
  return 4.0 * p->viT + sqrt(tmp); /* collisions.c:48-52 */
}

/* collisions.c:58-63 */
int oracle_collision_count(const oracle_params *p, long np, double dt)
{
  double fc = p->nui * dt;
  double vrm = vrm_of(p);
  int Nc = x86_double_to_int(((double)np * fc * vrm) / ((p->vnT + p->vn0) * sqrt(p->mn / p->mi)));
  if (Nc > np) Nc = (int)np;
  else if (Nc < 0) Nc = 0;
  return Nc;
}

/* One attempt of the collisions.c:70-157 loop body on velocity vi_in, in ndim (2 or 3)
 * dimensions: every loop runs over ndim components; vr[2] keeps its initial 0 in 2-D
 * (collisions.c:32), so cos(theta) = 0 there. */
void oracle_collide_one_nd(const oracle_params *p, int ndim, const oracle_attempt *a, const double vi_in[3],
                           double vf_out[3], int *accepted, int *failed)
{
  double vn0[3] = {p->vn0x, p->vn0y, p->vn0z};
  double vnT = p->vnT, mi = p->mi, mn = p->mn, M = mn + mi;
  double vrm = vrm_of(p);
  double vi[3] = {0, 0, 0}, vn[3] = {0, 0, 0}, vr[3] = {0.0, 0.0, 0.0}, vc[3] = {0, 0, 0}, u2[3], vf[3] = {0, 0, 0};
  double tmp = 0.0, vrr, vcr, u0x, u0y, u0z, uperp, uphi;
  double costht, sintht, cosphi, sinphi, u1x, u1y, u1z, vfr, ratio;
  float acc;
  int d;
  *accepted = 0;
  *failed = 0;
  for (d = 0; d < ndim; d++) {
    vi[d] = vi_in[d];
    vn[d] = a->g[d] * vnT + vn0[d];
    vr[d] = vi[d] - vn[d];
    tmp += vr[d] * vr[d];
  }
  vrr = sqrt(tmp);
  acc = a->u_acc1 * a->u_acc2; /* PetscSqr(ran3(&seed)): two draws, float product */
  if (!(acc < vrr * vrr / (vrm * vrm))) {
    for (d = 0; d < ndim; d++) vf_out[d] = vi_in[d];
    return;
  }
  *accepted = 1;
  tmp = 0.0;
  for (d = 0; d < ndim; d++) {
    vc[d] = (vi[d] * mi + vn[d] * mn) / M;
    tmp += (vi[d] - vc[d]) * (vi[d] - vc[d]);
  }
  vcr = sqrt(tmp);
  u0z = 2.0 * a->u_z - 1.0;
  uperp = sqrt(1.0 - u0z * u0z);
  uphi = 2 * 3.14159265358979323846 * a->u_phi;
  u0x = uperp * cos(uphi);
  u0y = uperp * sin(uphi);
  costht = vr[2] / vrr;
  sintht = sqrt(1.0 - costht * costht);
  cosphi = vr[0] / (vrr * sintht);
  sinphi = vr[1] / (vrr * sintht);
  u1x = u0z * sintht + u0x * costht;
  u1y = u0y;
  u1z = u0z * costht - u0x * sintht;
  u2[0] = u1x * cosphi - u1y * sinphi;
  u2[1] = u1x * sinphi + u1y * cosphi;
  u2[2] = u1z;
  tmp = 0.0;
  for (d = 0; d < ndim; d++) {
    vf[d] = vc[d] + vcr * u2[d];
    tmp += vf[d] * vf[d];
  }
  vfr = sqrt(tmp);
  ratio = vfr / p->viT;
  if (ratio > 10) {
    *failed = 1;
    for (d = 0; d < ndim; d++) vf_out[d] = vi_in[d];
    return;
  }
  for (d = 0; d < ndim; d++) vf_out[d] = vf[d];
}

void oracle_collide_one(const oracle_params *p, const oracle_attempt *a, const double vi_in[3],
                        double vf_out[3], int *accepted, int *failed)
{
  oracle_collide_one_nd(p, 3, a, vi_in, vf_out, accepted, failed);
}

static long pick_ip(long np, float u)
{
  /* (PetscInt)(np*ran3(&seed)): int*float is a float product (collisions.c:72).
   * A product that rounds up to np would index past the array in the reference;
   * both implementations clamp it to np-1. */
  long ip = (long)(int)((float)np * u);
  if (ip >= np) ip = np - 1;
  if (ip < 0) ip = 0;
  return ip;
}

/* Return codes: 0 ok, 2 collision failure (collisions.c:145-147 Abort), 3 records
 * exhausted before Nc successes.  vel is double[np][ndim]. */
int oracle_scatter_elastic_predrawn_nd(const oracle_params *p, int ndim, double dt, long np, double *vel,
                                       const oracle_attempt *att, long n_att, int Nc,
                                       long *n_used, int *Ns_out, int *Nf_out)
{
  int Ns = 0, Nf = 0, rc = 0;
  long a = 0;
  (void)dt;
  while (Ns < Nc) {
    long ip;
    int acc, fail, d;
    double vi[3] = {0, 0, 0}, vf[3];
    if (a >= n_att) { rc = 3; break; }
    ip = pick_ip(np, att[a].u_ip);
    for (d = 0; d < ndim; d++) vi[d] = vel[ip * ndim + d];
    oracle_collide_one_nd(p, ndim, &att[a], vi, vf, &acc, &fail);
    a++;
    if (!acc) continue;
    if (fail) {
      Nf++;
      if (Nf >= 10) { rc = 2; break; }
    } else {
      for (d = 0; d < ndim; d++) vel[ip * ndim + d] = vf[d];
      Ns++;
    }
  }
  if (n_used) *n_used = a;
  if (Ns_out) *Ns_out = Ns;
  if (Nf_out) *Nf_out = Nf;
  return rc;
}

int oracle_scatter_elastic_predrawn(const oracle_params *p, double dt, long np, double *vel,
                                    const oracle_attempt *att, long n_att, int Nc,
                                    long *n_used, int *Ns_out, int *Nf_out)
{
  return oracle_scatter_elastic_predrawn_nd(p, 3, dt, np, vel, att, n_att, Nc, n_used, Ns_out, Nf_out);
}

/* The reference stream: ScatterElastic with ran3/gasdev drawn in the reference's
 * order from the caller-held static state, starting from a local seed (getseed):
 * per attempt ran3 for the ion, ndim gasdev for the neutral, two ran3 for the
 * acceptance and, when accepted, two more for the angles.  Each consumed attempt is
 * appended to log (g[d] = 0 for d >= ndim; u_z, u_phi = -1 when not drawn).
 * vel is double[np][ndim]. */
int oracle_scatter_elastic_nr_nd(const oracle_params *p, int ndim, double dt, long np, double *vel,
                                 oracle_nr_state *st, long seed, oracle_attempt *log,
                                 long log_cap, long *n_log, int *Ns_out, int *Nf_out)
{
  int Nc = oracle_collision_count(p, np, dt);
  double vrm = vrm_of(p);
  double vn0[3] = {p->vn0x, p->vn0y, p->vn0z};
  int Ns = 0, Nf = 0, rc = 0;
  long nl = 0;
  while (Ns < Nc) {
    oracle_attempt a;
    long ip;
    int d, acc, fail;
    double vi[3] = {0, 0, 0}, vf[3], tmp = 0.0, vrr;
    a.u_ip = oracle_ran3(st, &seed);
    ip = pick_ip(np, a.u_ip);
    a.g[0] = a.g[1] = a.g[2] = 0.0f;
    for (d = 0; d < ndim; d++) {
      a.g[d] = oracle_gasdev(st, &seed);
      vi[d] = vel[ip * ndim + d];
      tmp += (vi[d] - (a.g[d] * p->vnT + vn0[d])) * (vi[d] - (a.g[d] * p->vnT + vn0[d]));
    }
    vrr = sqrt(tmp);
    a.u_acc1 = oracle_ran3(st, &seed);
    a.u_acc2 = oracle_ran3(st, &seed);
    a.u_z = -1.0f;
    a.u_phi = -1.0f;
    /* The acceptance decision needs vrr before the angle draws. */
    if ((float)(a.u_acc1 * a.u_acc2) < vrr * vrr / (vrm * vrm)) {
      a.u_z = oracle_ran3(st, &seed);
      a.u_phi = oracle_ran3(st, &seed);
    }
    if (log && nl < log_cap) log[nl] = a;
    nl++;
    oracle_collide_one_nd(p, ndim, &a, vi, vf, &acc, &fail);
    if (!acc) continue;
    if (fail) {
      Nf++;
      if (Nf >= 10) { rc = 2; break; }
    } else {
      for (d = 0; d < ndim; d++) vel[ip * ndim + d] = vf[d];
      Ns++;
    }
  }
  if (n_log) *n_log = nl;
  if (Ns_out) *Ns_out = Ns;
  if (Nf_out) *Nf_out = Nf;
  return rc;
}

int oracle_scatter_elastic_nr(const oracle_params *p, double dt, long np, double *vel,
                              oracle_nr_state *st, long seed, oracle_attempt *log,
                              long log_cap, long *n_log, int *Ns_out, int *Nf_out)
{
  return oracle_scatter_elastic_nr_nd(p, 3, dt, np, vel, st, seed, log, log_cap, n_log, Ns_out, Nf_out);
}

/* ------------------------------------------------------------------------- */
/* Two-dimensional variants (cli.ndim = 2: src/setup.c:99-104, Nz = 1).        */
/* Particles are double[np][2]; moments FluidNode[Ny][Nx] with Gz = 0;        */
/* phi double[Ny][Nx].                                                         */
/* ------------------------------------------------------------------------- */
#define PHI2(jj, ii) F[(long)wrapi((jj), Ny) * Nx + wrapi((ii), Nx)]

/* vector-math.c:529-592 (DifferenceVector2D): plain bilinear weights (no swap). */
void oracle_difference_vector2d(const double *F, int Nx, int Ny, double x0, double y0, double f[2])
{
  int ixl = (int)x0, ixh = ixl + 1, iyl = (int)y0, iyh = iyl + 1;
  double wxh = x0 - (double)ixl, wyh = y0 - (double)iyl;
  double hh, lh, hl, ll, Ewh, Ewl;
  if (ixl >= 0) {
    hl = PHI2(iyh, ixl + 1) - PHI2(iyh, ixl - 1);
    ll = PHI2(iyl, ixl + 1) - PHI2(iyl, ixl - 1);
  } else {
    hl = -1.0 * PHI2(iyh, ixl + 2) + 4.0 * PHI2(iyh, ixl + 1) - 3.0 * PHI2(iyh, ixl);
    ll = -1.0 * PHI2(iyl, ixl + 2) + 4.0 * PHI2(iyl, ixl + 1) - 3.0 * PHI2(iyl, ixl);
  }
  if (ixh < Nx) {
    hh = PHI2(iyh, ixh + 1) - PHI2(iyh, ixh - 1);
    lh = PHI2(iyl, ixh + 1) - PHI2(iyl, ixh - 1);
  } else {
    hh = +3.0 * PHI2(iyh, ixh) - 4.0 * PHI2(iyh, ixh - 1) + 1.0 * PHI2(iyh, ixh - 2);
    lh = +3.0 * PHI2(iyl, ixh) - 4.0 * PHI2(iyl, ixh - 1) + 1.0 * PHI2(iyl, ixh - 2);
  }
  Ewh = lh + wyh * (hh - lh);
  Ewl = ll + wyh * (hl - ll);
  f[0] = Ewl + wxh * (Ewh - Ewl);
  if (iyl >= 0) {
    lh = PHI2(iyl + 1, ixh) - PHI2(iyl - 1, ixh);
    ll = PHI2(iyl + 1, ixl) - PHI2(iyl - 1, ixl);
  } else {
    lh = -1.0 * PHI2(iyl + 2, ixh) + 4.0 * PHI2(iyl + 1, ixh) - 3.0 * PHI2(iyl, ixh);
    ll = -1.0 * PHI2(iyl + 2, ixl) + 4.0 * PHI2(iyl + 1, ixl) - 3.0 * PHI2(iyl, ixl);
  }
  if (iyh < Ny) {
    hh = PHI2(iyh + 1, ixh) - PHI2(iyh - 1, ixh);
    hl = PHI2(iyh + 1, ixl) - PHI2(iyh - 1, ixl);
  } else {
    hh = +3.0 * PHI2(iyh, ixh) - 4.0 * PHI2(iyh - 1, ixh) + 1.0 * PHI2(iyh - 2, ixh);
    hl = +3.0 * PHI2(iyh, ixl) - 4.0 * PHI2(iyh - 1, ixl) + 1.0 * PHI2(iyh - 2, ixl);
  }
  Ewh = lh + wyh * (hh - lh);
  Ewl = ll + wyh * (hl - ll);
  f[1] = Ewl + wxh * (Ewh - Ewl);
}
#undef PHI2

/* velocities.c:186-273 (BorisMoverBz2D). */
void oracle_boris_bz2d(const oracle_params *p, double dt, long np, const double *pos, double *vel,
                       const double *phi)
{
  double h[2] = {1.0 / p->dx, 1.0 / p->dy};
  double tscale = 0.5 * (p->qi / p->mi) * dt;
  double t = tscale * p->B0;
  double s = 2.0 * t / (1 + t * t);
  double Escale[2], E[2];
  long ip;
  int d;
  for (d = 0; d < 2; d++) Escale[d] = -0.5 * h[d] * tscale;
  for (ip = 0; ip < np; ip++) {
    double x = pos[ip * 2 + 0] / p->dx;
    double y = pos[ip * 2 + 1] / p->dy;
    double vxm, vym, vxt, vxp, vyp;
    oracle_difference_vector2d(phi, p->Nx, p->Ny, x, y, E);
    vxm = vel[ip * 2 + 0] + Escale[0] * E[0];
    vym = vel[ip * 2 + 1] + Escale[1] * E[1];
    vxt = vxm + vym * t;
    vyp = vym - vxt * s;
    vxp = vxt + vyp * t;
    vel[ip * 2 + 0] = vxp;
    vel[ip * 2 + 1] = vyp;
  }
}

/* positions.c:436-470 (UpdatePositions) and boundaries.c:6-42 (Apply2DBC), ndim = 2. */
void oracle_update_positions2d(double dt, long np, double *pos, const double *vel)
{
  long ip;
  int d;
  for (ip = 0; ip < np; ip++)
    for (d = 0; d < 2; d++) pos[ip * 2 + d] += vel[ip * 2 + d] * dt;
}

void oracle_apply_bc2d(const oracle_params *p, long np, double *pos)
{
  double L[2] = {p->Lx, p->Ly}, r0[2] = {p->x0, p->y0}, r1[2] = {p->x1, p->y1};
  long ip;
  int d;
  for (ip = 0; ip < np; ip++)
    for (d = 0; d < 2; d++) {
      double r = pos[ip * 2 + d];
      if (r < r0[d]) r += L[d];
      if (r >= r1[d]) r -= L[d];
      pos[ip * 2 + d] = r;
    }
}

/* moments.c:5-104 (Collect2DFluidMoments): m[Ny][Nx][4], Gz left at 0. */
void oracle_collect_moments2d(const oracle_params *p, long np, const double *pos, const double *vel, double *m)
{
  int Nx = p->Nx, Ny = p->Ny;
  long ip;
  memset(m, 0, sizeof(double) * 4 * (long)Nx * Ny);
  for (ip = 0; ip < np; ip++) {
    double x = pos[ip * 2 + 0] / p->dx, y = pos[ip * 2 + 1] / p->dy;
    int ixl = (int)x, ixh = ixl + 1, iyl = (int)y, iyh = iyl + 1;
    double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    double w4[4] = {wyh * wxh, wyl * wxh, wyh * wxl, wyl * wxl};
    long n4[4];
    int c, d;
    n4[0] = (long)wrapi(iyh, Ny) * Nx + wrapi(ixh, Nx);
    n4[1] = (long)wrapi(iyl, Ny) * Nx + wrapi(ixh, Nx);
    n4[2] = (long)wrapi(iyh, Ny) * Nx + wrapi(ixl, Nx);
    n4[3] = (long)wrapi(iyl, Ny) * Nx + wrapi(ixl, Nx);
    for (c = 0; c < 4; c++) m[n4[c] * 4 + 0] += w4[c];
    for (d = 0; d < 2; d++) {
      double w = vel[ip * 2 + d];
      for (c = 0; c < 4; c++) m[n4[c] * 4 + 1 + d] += w * w4[c];
    }
  }
}

/* lhs.c:194-334 (ComputeFullLHS2D): 9 coefficients per row in the reference's
 * column order (i+1), (i-1), (j+1), (j-1), (i+1,j+1), (i+1,j-1), (i-1,j+1), (i-1,j-1),
 * (i,j) (lhs.c:268-302). */
static const int lhs2_off[9][2] = {{+1, 0}, {-1, 0}, {0, +1}, {0, -1}, {+1, +1}, {+1, -1}, {-1, +1}, {-1, -1}, {0, 0}};

void oracle_full_lhs2d_coeffs(const oracle_params *p, const double *m, double *coef)
{
  int Nx = p->Nx, Ny = p->Ny, i, j;
  double kappa = p->kappa_e;
  double hxx = 1.0 / (2.0 * p->dx * p->dx), hyy = 1.0 / (2.0 * p->dy * p->dy), hod = 1.0 / (8.0 * p->dy * p->dx);
  double sxx = p->scale * hxx, syy = p->scale * hyy, sod = p->scale * kappa * hod;
#define N2(ii, jj) m[((long)wrapi((jj), Ny) * Nx + wrapi((ii), Nx)) * 4]
  for (j = 0; j < Ny; j++)
    for (i = 0; i < Nx; i++) {
      double nijk = N2(i, j), nmjk = N2(i - 1, j), npjk = N2(i + 1, j), nimk = N2(i, j - 1), nipk = N2(i, j + 1);
      double *c = &coef[((long)j * Nx + i) * 9];
      c[4] = sod * (nipk - npjk);
      c[5] = sod * (npjk - nimk);
      c[6] = sod * (nmjk - nipk);
      c[7] = sod * (nimk - nmjk);
      c[0] = sxx * (npjk + nijk) + sod * (nipk - nimk);
      c[1] = sxx * (nijk + nmjk) - sod * (nipk - nimk);
      c[2] = syy * (nipk + nijk) - sod * (npjk - nmjk);
      c[3] = syy * (nijk + nimk) + sod * (npjk - nmjk);
      c[8] = -(sxx * (npjk + 2 * nijk + nmjk) + syy * (nipk + 2 * nijk + nimk));
    }
#undef N2
}

void oracle_full_lhs2d_csr(const oracle_params *p, const double *m, int *rowptr, int *cols, double *vals)
{
  int Nx = p->Nx, Ny = p->Ny, i, j, e;
  long n = (long)Nx * Ny;
  oracle_full_lhs2d_coeffs(p, m, vals);
  for (j = 0; j < Ny; j++)
    for (i = 0; i < Nx; i++) {
      long r = (long)j * Nx + i;
      rowptr[r] = (int)(r * 9);
      for (e = 0; e < 9; e++)
        cols[r * 9 + e] = (int)((long)wrapi(j + lhs2_off[e][1], Ny) * Nx + wrapi(i + lhs2_off[e][0], Nx));
    }
  rowptr[n] = (int)(n * 9);
}

void oracle_full_lhs2d_apply(const oracle_params *p, const double *m, const double *x, double *y)
{
  int Nx = p->Nx, Ny = p->Ny, i, j, e;
  long n = (long)Nx * Ny;
  double *coef = (double *)malloc(sizeof(double) * 9 * n);
  oracle_full_lhs2d_coeffs(p, m, coef);
  for (j = 0; j < Ny; j++)
    for (i = 0; i < Nx; i++) {
      long r = (long)j * Nx + i;
      double acc = 0.0;
      for (e = 0; e < 9; e++)
        acc += coef[r * 9 + e] * x[(long)wrapi(j + lhs2_off[e][1], Ny) * Nx + wrapi(i + lhs2_off[e][0], Nx)];
      y[r] = acc;
    }
  free(coef);
}

/* rhs.c:151-256 (ComputeFullRHS2D), then the null-space (mean) removal. */
void oracle_full_rhs2d(const oracle_params *p, const double *m, double *b)
{
  int Nx = p->Nx, Ny = p->Ny, i, j;
  double kappa = p->kappa_e, E0 = p->E0, scale = p->scale;
  double cth = p->gammaT * OR_KB * p->Te / OR_Q;
  double cG = (1 + kappa * kappa) * p->me * p->nue / OR_Q;
  double hx = 1.0 / (2.0 * p->dx), hy = 1.0 / (2.0 * p->dy);
  double hxx = 1.0 / (p->dx * p->dx), hyy = 1.0 / (p->dy * p->dy);
#define M2(ii, jj, dof) m[((long)wrapi((jj), Ny) * Nx + wrapi((ii), Nx)) * 4 + (dof)]
  for (j = 0; j < Ny; j++)
    for (i = 0; i < Nx; i++) {
      double nijk = M2(i, j, 0), npjk = M2(i + 1, j, 0), nmjk = M2(i - 1, j, 0);
      double nipk = M2(i, j + 1, 0), nimk = M2(i, j - 1, 0);
      double dndx = hx * (npjk - nmjk);
      double dndy = hy * (nipk - nimk);
      double d2ndxx = hxx * (npjk - 2 * nijk + nmjk);
      double d2ndyy = hyy * (nipk - 2 * nijk + nimk);
      double dGxdx = hx * (M2(i + 1, j, 1) - M2(i - 1, j, 1));
      double dGydy = hy * (M2(i, j + 1, 2) - M2(i, j - 1, 2));
      double Eterm = (dndy - kappa * dndx) * E0;
      double Pterm = cth * (d2ndxx + d2ndyy);
      double Gterm = cG * (dGxdx + dGydy);
      b[(long)j * Nx + i] = scale * (Eterm + Pterm + Gterm);
    }
#undef M2
  oracle_remove_mean((long)Nx * Ny, b);
}

/* ------------------------------------------------------------------------- */
/* lhs.c:337-505 (ComputeFullLHS3D): 11 coefficients per row in the reference's */
/* column order (i+1), (i-1), (j+1), (j-1), (k+1), (k-1), (i+1,j+1), (i+1,j-1), */
/* (i-1,j+1), (i-1,j-1), (i,j,k) (lhs.c:426-479).                              */
/* ------------------------------------------------------------------------- */
static const int lhs_off[11][3] = {
  {+1, 0, 0}, {-1, 0, 0}, {0, +1, 0}, {0, -1, 0}, {0, 0, +1}, {0, 0, -1},
  {+1, +1, 0}, {+1, -1, 0}, {-1, +1, 0}, {-1, -1, 0}, {0, 0, 0}};

void oracle_full_lhs3d_coeffs(const oracle_params *p, const double *m, double *coef)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double kappa = p->kappa_e;
  double hxx = 1.0 / (2.0 * p->dx * p->dx);
  double hyy = 1.0 / (2.0 * p->dy * p->dy);
  double hzz = 1.0 / (2.0 * p->dz * p->dz);
  double hod = 1.0 / (8.0 * p->dy * p->dx);
  double sxx = p->scale * hxx;
  double syy = p->scale * hyy;
  double szz = p->scale * (1 + kappa * kappa) * hzz;
  double sod = p->scale * kappa * hod;
  int i, j, k;
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        long r = ((long)k * Ny + j) * Nx + i;
        double nijk = m[nidx(i, j, k, Nx, Ny, Nz) * 4];
        double nmjk = m[nidx(i - 1, j, k, Nx, Ny, Nz) * 4];
        double npjk = m[nidx(i + 1, j, k, Nx, Ny, Nz) * 4];
        double nimk = m[nidx(i, j - 1, k, Nx, Ny, Nz) * 4];
        double nipk = m[nidx(i, j + 1, k, Nx, Ny, Nz) * 4];
        double nijm = m[nidx(i, j, k - 1, Nx, Ny, Nz) * 4];
        double nijp = m[nidx(i, j, k + 1, Nx, Ny, Nz) * 4];
        double *c = &coef[r * 11];
        c[6] = sod * (nipk - npjk);
        c[7] = sod * (npjk - nimk);
        c[8] = sod * (nmjk - nipk);
        c[9] = sod * (nimk - nmjk);
        c[0] = sxx * (npjk + nijk) + sod * (nipk - nimk);
        c[1] = sxx * (nijk + nmjk) - sod * (nipk - nimk);
        c[2] = syy * (nipk + nijk) - sod * (npjk - nmjk);
        c[3] = syy * (nijk + nimk) + sod * (npjk - nmjk);
        c[4] = szz * (nijp + nijk);
        c[5] = szz * (nijk + nijm);
        c[10] = -(sxx * (npjk + 2 * nijk + nmjk) + syy * (nipk + 2 * nijk + nimk) +
                  szz * (nijp + 2 * nijk + nijm));
      }
}

/* CSR with 11 entries per row, columns in the reference order (not sorted). */
void oracle_full_lhs3d_csr(const oracle_params *p, const double *m,
                           int *rowptr, int *cols, double *vals)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  long n = (long)Nx * Ny * Nz, r;
  int i, j, k, e;
  oracle_full_lhs3d_coeffs(p, m, vals);
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        r = ((long)k * Ny + j) * Nx + i;
        rowptr[r] = (int)(r * 11);
        for (e = 0; e < 11; e++)
          cols[r * 11 + e] = (int)nidx(i + lhs_off[e][0], j + lhs_off[e][1], k + lhs_off[e][2], Nx, Ny, Nz);
      }
  rowptr[n] = (int)(n * 11);
}

void oracle_full_lhs3d_apply(const oracle_params *p, const double *m, const double *x, double *y)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  long n = (long)Nx * Ny * Nz;
  double *coef = (double *)malloc(sizeof(double) * 11 * n);
  int i, j, k, e;
  oracle_full_lhs3d_coeffs(p, m, coef);
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        long r = ((long)k * Ny + j) * Nx + i;
        double acc = 0.0;
        for (e = 0; e < 11; e++)
          acc += coef[r * 11 + e] * x[nidx(i + lhs_off[e][0], j + lhs_off[e][1], k + lhs_off[e][2], Nx, Ny, Nz)];
        y[r] = acc;
      }
  free(coef);
}

/* lhs.c:100-191 (ComputeLaplacianLHS3D), the 7-point test operator. */
void oracle_laplacian_lhs3d_apply(const oracle_params *p, const double *x, double *y)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double vpjk = 1.0 / (p->dx * p->dx), vmjk = vpjk;
  double vipk = 1.0 / (p->dy * p->dy), vimk = vipk;
  double vijp = 1.0 / (p->dz * p->dz), vijm = vijp;
  double vijk = -(vpjk + vipk + vijp + vmjk + vimk + vijm);
  int i, j, k;
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        long r = ((long)k * Ny + j) * Nx + i;
        y[r] = vpjk * x[nidx(i + 1, j, k, Nx, Ny, Nz)] + vmjk * x[nidx(i - 1, j, k, Nx, Ny, Nz)] +
               vipk * x[nidx(i, j + 1, k, Nx, Ny, Nz)] + vimk * x[nidx(i, j - 1, k, Nx, Ny, Nz)] +
               vijp * x[nidx(i, j, k + 1, Nx, Ny, Nz)] + vijm * x[nidx(i, j, k - 1, Nx, Ny, Nz)] +
               vijk * x[r];
      }
}

/* MatNullSpaceRemove with the constant null space: subtract the mean. */
void oracle_remove_mean(long n, double *v)
{
  double s = 0.0;
  long i;
  for (i = 0; i < n; i++) s += v[i];
  s /= (double)n;
  for (i = 0; i < n; i++) v[i] -= s;
}

/* rhs.c:259-373 (ComputeFullRHS3D). */
void oracle_full_rhs3d(const oracle_params *p, const double *m, double *b)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
  double kappa = p->kappa_e, E0 = p->E0, scale = p->scale;
  double cth = p->gammaT * OR_KB * p->Te / OR_Q;
  double cG = (1 + kappa * kappa) * p->me * p->nue / OR_Q;
  double hx = 1.0 / (2.0 * p->dx), hy = 1.0 / (2.0 * p->dy), hz = 1.0 / (2.0 * p->dz);
  double hxx = 1.0 / (p->dx * p->dx), hyy = 1.0 / (p->dy * p->dy), hzz = 1.0 / (p->dz * p->dz);
  int i, j, k;
#define MN(ii, jj, kk, dof) m[nidx((ii), (jj), (kk), Nx, Ny, Nz) * 4 + (dof)]
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        double nijk = MN(i, j, k, 0), npjk = MN(i + 1, j, k, 0), nmjk = MN(i - 1, j, k, 0);
        double nipk = MN(i, j + 1, k, 0), nimk = MN(i, j - 1, k, 0);
        double nijp = MN(i, j, k + 1, 0), nijm = MN(i, j, k - 1, 0);
        double dndx = hx * (npjk - nmjk);
        double dndy = hy * (nipk - nimk);
        double d2ndxx = hxx * (npjk - 2 * nijk + nmjk);
        double d2ndyy = hyy * (nipk - 2 * nijk + nimk);
        double d2ndzz = hzz * (nijp - 2 * nijk + nijm);
        double dGxdx = hx * (MN(i + 1, j, k, 1) - MN(i - 1, j, k, 1));
        double dGydy = hy * (MN(i, j + 1, k, 2) - MN(i, j - 1, k, 2));
        double dGzdz = hz * (MN(i, j, k + 1, 3) - MN(i, j, k - 1, 3));
        double Eterm = (dndy - kappa * dndx) * E0;
        double Pterm = cth * (d2ndxx + d2ndyy + (1 + kappa * kappa) * d2ndzz);
        double Gterm = cG * (dGxdx + dGydy + dGzdz);
        b[((long)k * Ny + j) * Nx + i] = scale * (Eterm + Pterm + Gterm);
      }
#undef MN
  oracle_remove_mean((long)Nx * Ny * Nz, b);
}

/* rhs.c:86-148 (ComputeSinusoidalRHS3D). */
void oracle_sinusoidal_rhs3d(const oracle_params *p, double *b)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, i, j, k;
  const double PI = 3.14159265358979323846;
  for (k = 0; k < Nz; k++)
    for (j = 0; j < Ny; j++)
      for (i = 0; i < Nx; i++) {
        double x = ((double)i + 0.5) * p->dx;
        double y = ((double)j + 0.5) * p->dy;
        double z = ((double)k + 0.5) * p->dz;
        double Cx = cos(2 * PI * x), Cy = cos(2 * PI * y), Cz = cos(2 * PI * z);
        double val = Cx * Cy * Cz;
        b[((long)k * Ny + j) * Nx + i] = p->n0 * val * p->scale;
      }
  oracle_remove_mean((long)Nx * Ny * Nz, b);
}

/* ------------------------------------------------------------------------- */
/* CPU baseline loop: src/pic.c:246-254 order with the potential held fixed.   */
/* ------------------------------------------------------------------------- */
static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int oracle_particle_steps(const oracle_params *p, double dt, int nt, long np, double *pos,
                          double *vel, const double *phi, double *moments,
                          oracle_nr_state *st, int collide, double t[4])
{
  int it, rc = 0;
  t[0] = t[1] = t[2] = t[3] = 0.0;
  for (it = 0; it < nt; it++) {
    double a = now_s(), b, c, d, e;
    int Ns, Nf;
    long nl;
    oracle_boris_bz3d(p, dt, np, pos, vel, phi);
    b = now_s();
    if (collide && p->mn > 0.0) {
      rc = oracle_scatter_elastic_nr(p, dt, np, vel, st, oracle_getseed(0), NULL, 0, &nl, &Ns, &Nf);
      if (rc) return rc;
    }
    c = now_s();
    oracle_update_positions(dt, np, pos, vel);
    oracle_apply_bc(p, np, pos);
    d = now_s();
    oracle_collect_moments(p, np, pos, vel, moments);
    e = now_s();
    t[0] += b - a;
    t[1] += c - b;
    t[2] += d - c;
    t[3] += e - d;
  }
  return rc;
}
