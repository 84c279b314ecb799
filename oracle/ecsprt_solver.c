/*
 * ecsprt_solver.c -- CPU FP64 solve of the potential equation (TEST INFRASTRUCTURE and the
 * timed CPU baseline of the solve).
 *
 * The system is the reference's: the 11-point operator of ComputeFullLHS3D
 * (src/lhs.c:337-505, coefficients lhs.c:409-423) built from the density, the forcing of
 * ComputeFullRHS3D (src/rhs.c:259-373) with the constant null space removed, solved as
 * KSPSolve does in src/potential.c:11 -- here with the GMRES(30) + geometric-multigrid
 * method the GPU path uses, restated independently on one core:
 *   - right-preconditioned GMRES(m), classical Gram-Schmidt (PETSc's default
 *     KSP_GMRES_CGS_REFINE_NEVER), converged on the true residual ||b - A x|| <= rtol ||b||;
 *   - one V-cycle per Arnoldi step: red-black Gauss-Seidel (2 + 2 sweeps), full-weighting
 *     restriction, linear prolongation, coarse operators rediscretised from the restricted
 *     density, an axis coarsened when it halves and its coupling is within a factor 2 of
 *     the strongest, 50 sweeps on the coarsest level (at most 128 nodes);
 *   - the mean removed from the solution.
 * Nothing here is loaded by the product library.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ecsprt_oracle.h"

typedef struct {
  int Nx, Ny, Nz;
  long n;
  double sxx, syy, szz, sod;
  double *dn, *x, *b, *r;
  int *ip, *im, *jp, *jm, *kp, *km; /* periodic neighbour indices per axis */
  int cx, cy, cz;                   /* axes halved towards the next level */
} olevel;

#define OS_MAXLEV 24

typedef struct {
  olevel L[OS_MAXLEV];
  int nl;
} omg;

static int *wrap_table(int N, int d)
{
  int *t = (int *)malloc(sizeof(int) * N), i;
  for (i = 0; i < N; i++) t[i] = ((i + d) % N + N) % N;
  return t;
}

/* Scale factors of ComputeFullLHS3D (lhs.c:366-375) on a level of spacing dx, dy, dz. */
static void level_consts(olevel *L, const oracle_params *p, double dx, double dy, double dz)
{
  double k = p->kappa_e;
  L->sxx = p->scale / (2.0 * dx * dx);
  L->syy = p->scale / (2.0 * dy * dy);
  L->szz = p->scale * (1 + k * k) / (2.0 * dz * dz);
  L->sod = p->scale * k / (8.0 * dy * dx);
}

static void level_alloc(olevel *L, int Nx, int Ny, int Nz)
{
  L->Nx = Nx;
  L->Ny = Ny;
  L->Nz = Nz;
  L->n = (long)Nx * Ny * Nz;
  L->dn = (double *)calloc(L->n, sizeof(double));
  L->x = (double *)calloc(L->n, sizeof(double));
  L->b = (double *)calloc(L->n, sizeof(double));
  L->r = (double *)calloc(L->n, sizeof(double));
  L->ip = wrap_table(Nx, 1);
  L->im = wrap_table(Nx, -1);
  L->jp = wrap_table(Ny, 1);
  L->jm = wrap_table(Ny, -1);
  L->kp = wrap_table(Nz, 1);
  L->km = wrap_table(Nz, -1);
  L->cx = L->cy = L->cz = 0;
}

static void level_free(olevel *L)
{
  free(L->dn);
  free(L->x);
  free(L->b);
  free(L->r);
  free(L->ip);
  free(L->im);
  free(L->jp);
  free(L->jm);
  free(L->kp);
  free(L->km);
}

#define IDX(L, i, j, k) (((long)(k) * (L)->Ny + (j)) * (L)->Nx + (i))

/* Row (i,j,k): coefficients in the order of lhs.c:426-479 and the off-diagonal sum. */
static double row_apply(const olevel *L, const double *x, int i, int j, int k, double *diag)
{
  const double *n = L->dn;
  int ip = L->ip[i], im = L->im[i], jp = L->jp[j], jm = L->jm[j], kp = L->kp[k], km = L->km[k];
  double nijk = n[IDX(L, i, j, k)], nmjk = n[IDX(L, im, j, k)], npjk = n[IDX(L, ip, j, k)];
  double nimk = n[IDX(L, i, jm, k)], nipk = n[IDX(L, i, jp, k)];
  double nijm = n[IDX(L, i, j, km)], nijp = n[IDX(L, i, j, kp)];
  double sxx = L->sxx, syy = L->syy, szz = L->szz, sod = L->sod;
  double cE = sxx * (npjk + nijk) + sod * (nipk - nimk);
  double cW = sxx * (nijk + nmjk) - sod * (nipk - nimk);
  double cN = syy * (nipk + nijk) - sod * (npjk - nmjk);
  double cS = syy * (nijk + nimk) + sod * (npjk - nmjk);
  double cT = szz * (nijp + nijk), cB = szz * (nijk + nijm);
  double cNE = sod * (nipk - npjk), cSE = sod * (npjk - nimk);
  double cNW = sod * (nmjk - nipk), cSW = sod * (nimk - nmjk);
  *diag = -(sxx * (npjk + 2 * nijk + nmjk) + syy * (nipk + 2 * nijk + nimk) + szz * (nijp + 2 * nijk + nijm));
  return cE * x[IDX(L, ip, j, k)] + cW * x[IDX(L, im, j, k)] + cN * x[IDX(L, i, jp, k)] + cS * x[IDX(L, i, jm, k)] +
         cT * x[IDX(L, i, j, kp)] + cB * x[IDX(L, i, j, km)] + cNE * x[IDX(L, ip, jp, k)] +
         cSE * x[IDX(L, ip, jm, k)] + cNW * x[IDX(L, im, jp, k)] + cSW * x[IDX(L, im, jm, k)];
}

static void level_apply(const olevel *L, const double *x, double *y)
{
  int i, j, k;
  double d;
  for (k = 0; k < L->Nz; k++)
    for (j = 0; j < L->Ny; j++)
      for (i = 0; i < L->Nx; i++) {
        double off = row_apply(L, x, i, j, k, &d);
        y[IDX(L, i, j, k)] = off + d * x[IDX(L, i, j, k)];
      }
}

static void level_residual(const olevel *L, const double *x, const double *b, double *r)
{
  int i, j, k;
  double d;
  for (k = 0; k < L->Nz; k++)
    for (j = 0; j < L->Ny; j++)
      for (i = 0; i < L->Nx; i++) {
        long q = IDX(L, i, j, k);
        double off = row_apply(L, x, i, j, k, &d);
        r[q] = b[q] - (off + d * x[q]);
      }
}

/* Red-black Gauss-Seidel: colour (i+j+k)&1 == first, then the other, in place. */
static void level_sweep(const olevel *L, double *x, const double *b)
{
  int c, i, j, k;
  double d;
  for (c = 0; c < 2; c++)
    for (k = 0; k < L->Nz; k++)
      for (j = 0; j < L->Ny; j++)
        for (i = (c + j + k) & 1; i < L->Nx; i += 2) {
          long q = IDX(L, i, j, k);
          double off = row_apply(L, x, i, j, k, &d);
          if (d != 0.0) x[q] = (b[q] - off) / d;
        }
}

/* Full weighting along the coarsened axes (1/4, 1/2, 1/4 per axis, periodic). */
static void restrict_fw(const olevel *F, const olevel *C, const double *f, double *c)
{
  int i, j, k, a, bb, e;
  for (k = 0; k < C->Nz; k++)
    for (j = 0; j < C->Ny; j++)
      for (i = 0; i < C->Nx; i++) {
        int fi = F->cx ? 2 * i : i, fj = F->cy ? 2 * j : j, fk = F->cz ? 2 * k : k;
        double acc = 0.0;
        for (e = -F->cz; e <= F->cz; e++)
          for (bb = -F->cy; bb <= F->cy; bb++)
            for (a = -F->cx; a <= F->cx; a++) {
              double w = (F->cx ? (a ? 0.25 : 0.5) : 1.0) * (F->cy ? (bb ? 0.25 : 0.5) : 1.0) *
                         (F->cz ? (e ? 0.25 : 0.5) : 1.0);
              int ii = (fi + a + F->Nx) % F->Nx, jj = (fj + bb + F->Ny) % F->Ny, kk = (fk + e + F->Nz) % F->Nz;
              acc += w * f[IDX(F, ii, jj, kk)];
            }
        c[IDX(C, i, j, k)] = acc;
      }
}

/* x_f += linear interpolation of x_c along the coarsened axes. */
static void prolong_add(const olevel *F, const olevel *C, const double *xc, double *xf)
{
  int i, j, k;
  for (k = 0; k < F->Nz; k++)
    for (j = 0; j < F->Ny; j++)
      for (i = 0; i < F->Nx; i++) {
        int ia[2], ja[2], ka[2], na = 1, nb = 1, nc = 1, a, bb, e;
        double wa = 1.0, wb = 1.0, wc = 1.0, acc = 0.0;
        ia[0] = F->cx ? i >> 1 : i;
        ja[0] = F->cy ? j >> 1 : j;
        ka[0] = F->cz ? k >> 1 : k;
        if (F->cx && (i & 1)) { ia[1] = ((i + 1) >> 1) % C->Nx; na = 2; wa = 0.5; }
        if (F->cy && (j & 1)) { ja[1] = ((j + 1) >> 1) % C->Ny; nb = 2; wb = 0.5; }
        if (F->cz && (k & 1)) { ka[1] = ((k + 1) >> 1) % C->Nz; nc = 2; wc = 0.5; }
        for (e = 0; e < nc; e++)
          for (bb = 0; bb < nb; bb++)
            for (a = 0; a < na; a++) acc += wa * wb * wc * xc[IDX(C, ia[a], ja[bb], ka[e])];
        xf[IDX(F, i, j, k)] += acc;
      }
}

static void mg_build(omg *mg, const oracle_params *p, const double *moments)
{
  int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, l;
  double dx = p->dx, dy = p->dy, dz = p->dz, k2 = 1 + p->kappa_e * p->kappa_e;
  long q;
  mg->nl = 0;
  for (l = 0; l < OS_MAXLEV; l++) {
    olevel *L = &mg->L[l];
    double a[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), k2 / (dz * dz)}, amax = 0.0;
    int N[3] = {Nx, Ny, Nz}, can[3], cf[3] = {0, 0, 0}, any = 0, d;
    level_alloc(L, Nx, Ny, Nz);
    level_consts(L, p, dx, dy, dz);
    if (l == 0)
      for (q = 0; q < L->n; q++) L->dn[q] = moments[4 * q];
    else
      restrict_fw(&mg->L[l - 1], L, mg->L[l - 1].dn, L->dn);
    mg->nl = l + 1;
    for (d = 0; d < 3; d++) {
      can[d] = N[d] % 2 == 0 && N[d] >= 4;
      if (can[d] && a[d] > amax) amax = a[d];
    }
    for (d = 0; d < 3; d++)
      if (can[d] && a[d] >= 0.5 * amax) cf[d] = any = 1;
    if (!any || L->n <= 128) break;
    L->cx = cf[0];
    L->cy = cf[1];
    L->cz = cf[2];
    if (cf[0]) { Nx /= 2; dx *= 2; }
    if (cf[1]) { Ny /= 2; dy *= 2; }
    if (cf[2]) { Nz /= 2; dz *= 2; }
  }
}

static void mg_free(omg *mg)
{
  int l;
  for (l = 0; l < mg->nl; l++) level_free(&mg->L[l]);
  mg->nl = 0;
}

/* z = V-cycle(r) from a zero initial guess on every level. */
static void vcycle(omg *mg, const double *r, double *z)
{
  int l, s;
  olevel *L0 = &mg->L[0];
  memcpy(L0->b, r, sizeof(double) * L0->n);
  for (l = 0; l < mg->nl - 1; l++) {
    olevel *L = &mg->L[l];
    memset(L->x, 0, sizeof(double) * L->n);
    for (s = 0; s < 2; s++) level_sweep(L, L->x, L->b);
    level_residual(L, L->x, L->b, L->r);
    restrict_fw(L, &mg->L[l + 1], L->r, mg->L[l + 1].b);
  }
  {
    olevel *C = &mg->L[mg->nl - 1];
    memset(C->x, 0, sizeof(double) * C->n);
    for (s = 0; s < 50; s++) level_sweep(C, C->x, C->b);
  }
  for (l = mg->nl - 2; l >= 0; l--) {
    olevel *L = &mg->L[l];
    prolong_add(L, &mg->L[l + 1], mg->L[l + 1].x, L->x);
    for (s = 0; s < 2; s++) level_sweep(L, L->x, L->b);
  }
  memcpy(z, L0->x, sizeof(double) * L0->n);
}

static double dot(long n, const double *a, const double *b)
{
  double s = 0.0;
  long i;
  for (i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int oracle_solve_gmres_mg(const oracle_params *p, const double *m_op, const double *b, double *x, double rtol,
                          int restart, int maxit, int *iters_out, double *rnorm_out)
{
  omg mg;
  const int m = restart > 0 ? restart : 30;
  const long n = (long)p->Nx * p->Ny * p->Nz;
  double *V = (double *)malloc(sizeof(double) * n * (m + 1)), *Z = (double *)malloc(sizeof(double) * n * m);
  double *w = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double)), *cs = (double *)calloc(m, sizeof(double));
  double *sn = (double *)calloc(m, sizeof(double)), *g = (double *)calloc(m + 1, sizeof(double));
  double *y = (double *)calloc(m, sizeof(double));
  double bnorm = sqrt(dot(n, b, b)), tol = rtol * bnorm, beta = 0.0;
  int it = 0, conv = 0, j, q;
  long i;
  mg_build(&mg, p, m_op);
  memset(x, 0, sizeof(double) * n);
  while (1) {
    level_residual(&mg.L[0], x, b, r);
    beta = sqrt(dot(n, r, r));
    if (beta <= tol) { conv = 1; break; }
    if (it >= maxit || !isfinite(beta)) break;
    for (i = 0; i < n; i++) V[i] = r[i] / beta;
    memset(g, 0, sizeof(double) * (m + 1));
    g[0] = beta;
    for (j = 0; j < m && it < maxit; j++, it++) {
      double hn, den, a0, b0;
      vcycle(&mg, V + (long)j * n, Z + (long)j * n);
      level_apply(&mg.L[0], Z + (long)j * n, w);
      for (q = 0; q <= j; q++) H[q * m + j] = dot(n, V + (long)q * n, w);
      for (q = 0; q <= j; q++) {
        const double h = H[q * m + j], *v = V + (long)q * n;
        for (i = 0; i < n; i++) w[i] -= h * v[i];
      }
      hn = sqrt(dot(n, w, w));
      H[(j + 1) * m + j] = hn;
      if (hn > 0)
        for (i = 0; i < n; i++) V[(long)(j + 1) * n + i] = w[i] / hn;
      for (q = 0; q < j; q++) {
        a0 = H[q * m + j];
        b0 = H[(q + 1) * m + j];
        H[q * m + j] = cs[q] * a0 + sn[q] * b0;
        H[(q + 1) * m + j] = -sn[q] * a0 + cs[q] * b0;
      }
      a0 = H[j * m + j];
      b0 = H[(j + 1) * m + j];
      den = hypot(a0, b0);
      cs[j] = den > 0 ? a0 / den : 1.0;
      sn[j] = den > 0 ? b0 / den : 0.0;
      H[j * m + j] = den;
      H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      if (fabs(g[j + 1]) <= tol || hn == 0.0) { j++; it++; break; }
    }
    for (q = j - 1; q >= 0; q--) {
      double s = g[q];
      int rr;
      for (rr = q + 1; rr < j; rr++) s -= H[q * m + rr] * y[rr];
      y[q] = H[q * m + q] != 0 ? s / H[q * m + q] : 0.0;
    }
    for (q = 0; q < j; q++) {
      const double *zq = Z + (long)q * n;
      for (i = 0; i < n; i++) x[i] += y[q] * zq[i];
    }
  }
  oracle_remove_mean(n, x);
  if (iters_out) *iters_out = it;
  if (rnorm_out) *rnorm_out = beta;
  mg_free(&mg);
  free(V);
  free(Z);
  free(w);
  free(r);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(y);
  return conv ? 0 : 1;
}
