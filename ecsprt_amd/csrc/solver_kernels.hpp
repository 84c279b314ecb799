// Device code of the potential solve, shared by the single-problem solver (solver.hip)
// and the batched one (bsolve.hip): the 11-point operator of ComputeFullLHS3D
// (src/lhs.c:337-505) from the density, red-black Gauss-Seidel sweeps, multigrid
// transfers, the single-workgroup coarse tail, Krylov vector kernels and the forcing of
// ComputeFullRHS3D (src/rhs.c:259-373).
#pragma once
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace ecsprt {

// Physical constants, src/constants.h:4-7.
constexpr double KB_ = 1.3807e-23;
constexpr double Q_ = 1.6022e-19;

struct StencilConsts {
  int Nx, Ny, Nz;
  double sxx, syy, szz, sod;
};

// 11 coefficients of a row in the reference order (lhs.c:409-423) from the density at
// the node (nijk) and its six axis neighbours (m/p: minus/plus along the axis).
template <typename T>
__host__ __device__ inline void coeffs7(const StencilConsts &s, T nijk, T nmjk, T npjk, T nimk, T nipk, T nijm, T nijp,
                                        T c[11]) {
  const T sxx = (T)s.sxx, syy = (T)s.syy, szz = (T)s.szz, sod = (T)s.sod;
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
  c[10] = -(sxx * (npjk + 2 * nijk + nmjk) + syy * (nipk + 2 * nijk + nimk) + szz * (nijp + 2 * nijk + nijm));
}

// 11 coefficients of row (i,j,k) of the periodic grid.
template <typename T>
__host__ __device__ inline void row_coeffs(const T *__restrict__ dn, const StencilConsts &s, int i, int j, int k,
                                           T c[11]) {
  const int Nx = s.Nx, Ny = s.Ny, Nz = s.Nz;
  const int ip = i + 1 == Nx ? 0 : i + 1, im = i == 0 ? Nx - 1 : i - 1;
  const int jp = j + 1 == Ny ? 0 : j + 1, jm = j == 0 ? Ny - 1 : j - 1;
  const int kp = k + 1 == Nz ? 0 : k + 1, km = k == 0 ? Nz - 1 : k - 1;
  const long row = (long)Ny * Nx;
  const T nijk = dn[k * row + (long)j * Nx + i];
  const T nmjk = dn[k * row + (long)j * Nx + im];
  const T npjk = dn[k * row + (long)j * Nx + ip];
  const T nimk = dn[k * row + (long)jm * Nx + i];
  const T nipk = dn[k * row + (long)jp * Nx + i];
  const T nijm = dn[km * row + (long)j * Nx + i];
  const T nijp = dn[kp * row + (long)j * Nx + i];
  coeffs7(s, nijk, nmjk, npjk, nimk, nipk, nijm, nijp, c);
}

// Off-diagonal part of row (i,j,k) applied to x.
template <typename T>
__host__ __device__ inline T offdiag(const T *__restrict__ x, const StencilConsts &s, int i, int j, int k, const T c[11]) {
  const int Nx = s.Nx, Ny = s.Ny, Nz = s.Nz;
  const int ip = i + 1 == Nx ? 0 : i + 1, im = i == 0 ? Nx - 1 : i - 1;
  const int jp = j + 1 == Ny ? 0 : j + 1, jm = j == 0 ? Ny - 1 : j - 1;
  const int kp = k + 1 == Nz ? 0 : k + 1, km = k == 0 ? Nz - 1 : k - 1;
  const long row = (long)Ny * Nx;
  const long zc = k * row, zp = kp * row, zm = km * row;
  const long yc = (long)j * Nx, yp = (long)jp * Nx, ym = (long)jm * Nx;
  return c[0] * x[zc + yc + ip] + c[1] * x[zc + yc + im] + c[2] * x[zc + yp + i] + c[3] * x[zc + ym + i] +
         c[4] * x[zp + yc + i] + c[5] * x[zm + yc + i] + c[6] * x[zc + yp + ip] + c[7] * x[zc + ym + ip] +
         c[8] * x[zc + yp + im] + c[9] * x[zc + ym + im];
}

// Node-wise kernels run on a 3D grid of 64 x 4 thread blocks, one thread per node:
// block (x, y) tiles an xy plane, blockIdx.z is the plane.
struct Launch3 {
  dim3 grid, block;
};
static Launch3 launch3(int Nx, int Ny, int Nz) {
  return {dim3((Nx + 63) / 64, (Ny + 3) / 4, Nz), dim3(64, 4, 1)};
}
#define NODE_IJK(NX, NY)                                              \
  const int i = blockIdx.x * blockDim.x + threadIdx.x;                \
  const int j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z; \
  if (i >= (NX) || j >= (NY)) return;                                 \
  const long q = ((long)k * (NY) + j) * (NX) + i;

template <typename T>
__global__ void k_apply(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ x, T *__restrict__ y) {
  NODE_IJK(s.Nx, s.Ny)
  T c[11];
  row_coeffs(dn, s, i, j, k, c);
  y[q] = offdiag(x, s, i, j, k, c) + c[10] * x[q];
}

// r = b - A x
template <typename T>
__global__ void k_residual(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ x, const T *__restrict__ b,
                           T *__restrict__ r) {
  NODE_IJK(s.Nx, s.Ny)
  T c[11];
  row_coeffs(dn, s, i, j, k, c);
  r[q] = b[q] - (offdiag(x, s, i, j, k, c) + c[10] * x[q]);
}

// Periodic index for a within one period of [0, n).
__host__ __device__ inline int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
// z index of a halo plane: one plane (the two-dimensional run) is its own periodic image.
__host__ __device__ inline int wrapz(int a, int n) { return n == 1 ? 0 : wrap1(a, n); }

struct CoarsenMap {
  int Nxf, Nyf, Nzf, Nxc, Nyc, Nzc;
  int cx, cy, cz;
};

// Linear interpolation along the coarsened axes of the coarse vector at fine node (i,j,k).
template <typename T>
__device__ inline T prolong_point(const CoarsenMap &m, const T *xc, int i, int j, int k) {
  int ia[2], ja[2], ka[2];
  T wa[2], wb[2], wc[2];
  int na = 1, nb = 1, nc = 1;
  if (m.cx && (i & 1)) { ia[0] = i >> 1; ia[1] = wrap1((i + 1) >> 1, m.Nxc); wa[0] = wa[1] = (T)0.5; na = 2; }
  else { ia[0] = m.cx ? i >> 1 : i; wa[0] = (T)1; }
  if (m.cy && (j & 1)) { ja[0] = j >> 1; ja[1] = wrap1((j + 1) >> 1, m.Nyc); wb[0] = wb[1] = (T)0.5; nb = 2; }
  else { ja[0] = m.cy ? j >> 1 : j; wb[0] = (T)1; }
  if (m.cz && (k & 1)) { ka[0] = k >> 1; ka[1] = wrap1((k + 1) >> 1, m.Nzc); wc[0] = wc[1] = (T)0.5; nc = 2; }
  else { ka[0] = m.cz ? k >> 1 : k; wc[0] = (T)1; }
  T acc = 0;
  for (int c = 0; c < nc; ++c)
    for (int b = 0; b < nb; ++b)
      for (int a = 0; a < na; ++a) acc += wc[c] * wb[b] * wa[a] * xc[((long)ka[c] * m.Nyc + ja[b]) * m.Nxc + ia[a]];
  return acc;
}

// 1/x from the hardware reciprocal refined by Newton steps (two in FP64, one in
// FP32): within an ulp of the quotient, for the smoother's diagonal scaling.
__device__ inline double smooth_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
__device__ inline float smooth_rcp(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

// Gauss-Seidel value of LDS node p (a box of HX x HY nodes per plane) from the
// density sd and the neighbour values xs; `old` where the diagonal vanishes.
template <typename T, int HX, int HY>
__device__ inline T relax_lds(const StencilConsts &s, const T *sd, const T *xs, int p, T bq, T old) {
  constexpr int DY = HX, DZ = HX * HY;
  T c[11];
  coeffs7(s, sd[p], sd[p - 1], sd[p + 1], sd[p - DY], sd[p + DY], sd[p - DZ], sd[p + DZ], c);
  if (c[10] == T(0)) return old;
  const T off = c[0] * xs[p + 1] + c[1] * xs[p - 1] + c[2] * xs[p + DY] + c[3] * xs[p - DY] + c[4] * xs[p + DZ] +
                c[5] * xs[p - DZ] + c[6] * xs[p + DY + 1] + c[7] * xs[p - DY + 1] + c[8] * xs[p + DY - 1] +
                c[9] * xs[p - DY - 1];
  return (bq - off) * smooth_rcp(c[10]);
}

// One red-black Gauss-Seidel sweep (colour `first`, then the other; colour of node
// (i,j,k) = (i+j+k)&1) from xin (zero when null) into xout, on tiles of TX x TY x TZ
// nodes.  Density and potential are staged in LDS with a halo of 2.  The first colour
// is updated on the tile plus a halo of 1 into a second LDS array, so the second
// colour's update on the tile reads LDS only.  The 11-point stencil's xy-diagonal
// couplings join nodes of one colour; a colour's update reads those neighbours from
// before its half-sweep, so the result does not depend on thread order.
// Tile t of the sweep (the body of k_sweep; the batched solver maps its blocks to tiles).
template <typename T, int TX, int TY, int TZ>
__device__ inline void sweep_box(const StencilConsts &s, const T *__restrict__ dn, const T *__restrict__ xin,
                                 const T *__restrict__ b, T *__restrict__ xout, int first, int t) {
  constexpr int HX = TX + 4, HY = TY + 4, HZ = TZ + 4, H3 = HX * HY * HZ;
  constexpr int RX = TX + 2, RY = TY + 2, RZ = TZ + 2, R3 = RX * RY * RZ;
  __shared__ T sx[H3], sn[H3], sd[H3];
  const int Nx = s.Nx, Ny = s.Ny, Nz = s.Nz;
  const int ntx = (Nx + TX - 1) / TX, nty = (Ny + TY - 1) / TY;
  const int i0 = (t % ntx) * TX, j0 = ((t / ntx) % nty) * TY, k0 = (t / (ntx * nty)) * TZ;
  const int ex = min(TX, Nx - i0), ey = min(TY, Ny - j0), ez = min(TZ, Nz - k0);
  for (int q = threadIdx.x; q < H3; q += blockDim.x) {
    const int li = q % HX - 2, lj = (q / HX) % HY - 2, lk = q / (HX * HY) - 2;
    if (li > ex + 1 || lj > ey + 1 || lk > ez + 1) continue;
    const long g = ((long)wrapz(k0 + lk, Nz) * Ny + wrap1(j0 + lj, Ny)) * Nx + wrap1(i0 + li, Nx);
    sd[q] = dn[g];
    sx[q] = xin ? xin[g] : T(0);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < R3; q += blockDim.x) {
    const int li = q % RX - 1, lj = (q / RX) % RY - 1, lk = q / (RX * RY) - 1;
    if (li > ex || lj > ey || lk > ez) continue;
    const int gi = wrap1(i0 + li, Nx), gj = wrap1(j0 + lj, Ny), gk = wrapz(k0 + lk, Nz);
    const int p = ((lk + 2) * HY + (lj + 2)) * HX + (li + 2);
    T v = sx[p];
    if (((gi + gj + gk) & 1) == first)
      v = relax_lds<T, HX, HY>(s, sd, sx, p, b[((long)gk * Ny + gj) * Nx + gi], v);
    sn[p] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < TX * TY * TZ; q += blockDim.x) {
    const int li = q % TX, lj = (q / TX) % TY, lk = q / (TX * TY);
    if (li >= ex || lj >= ey || lk >= ez) continue;
    const int gi = i0 + li, gj = j0 + lj, gk = k0 + lk;
    const int p = ((lk + 2) * HY + (lj + 2)) * HX + (li + 2);
    const long g = ((long)gk * Ny + gj) * Nx + gi;
    T v = sn[p];
    if (((gi + gj + gk) & 1) != first) v = relax_lds<T, HX, HY>(s, sd, sn, p, b[g], v);
    xout[g] = v;
  }
}

template <typename T, int TX, int TY, int TZ>
__global__ void __launch_bounds__(512) k_sweep(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ xin,
                                              const T *__restrict__ b, T *__restrict__ xout, int first) {
  sweep_box<T, TX, TY, TZ>(s, dn, xin, b, xout, first, xcd_order(blockIdx.x, gridDim.x));
}

// Gauss-Seidel value of a node from three planes of density (d0 below, d1 at, d2 above
// the node; row stride DX, node index pd) and of potential (x0, x1, x2; row stride XX,
// node index px); `old` where the diagonal vanishes.  Same arithmetic as relax_lds.
template <typename T, int DX, int XX>
__device__ inline T relax_planes(const StencilConsts &s, const T *d0, const T *d1, const T *d2, int pd, const T *x0,
                                 const T *x1, const T *x2, int px, T bq, T old) {
  T c[11];
  coeffs7(s, d1[pd], d1[pd - 1], d1[pd + 1], d1[pd - DX], d1[pd + DX], d0[pd], d2[pd], c);
  if (c[10] == T(0)) return old;
  const T off = c[0] * x1[px + 1] + c[1] * x1[px - 1] + c[2] * x1[px + XX] + c[3] * x1[px - XX] + c[4] * x2[px] +
                c[5] * x0[px] + c[6] * x1[px + XX + 1] + c[7] * x1[px - XX + 1] + c[8] * x1[px + XX - 1] +
                c[9] * x1[px - XX - 1];
  return (bq - off) * smooth_rcp(c[10]);
}

// The sweep of k_sweep on a single plane (Nz = 1, the two-dimensional run): the
// periodic z neighbours of a node are the node itself, so one plane of density and
// potential with an xy halo of 2 is staged, and each relaxation is relax_planes with
// the same plane below, at and above (the arithmetic of the box sweep on that grid).
template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_sweep2d(StencilConsts s, const T *__restrict__ dn,
                                                    const T *__restrict__ xin, const T *__restrict__ b,
                                                    T *__restrict__ xout, int first, const T *__restrict__ xc,
                                                    CoarsenMap pm) {
  constexpr int HX = TX + 4, HY = TY + 4, H2 = HX * HY;
  constexpr int RX = TX + 2, RY = TY + 2, R2 = RX * RY;
  __shared__ T sx[H2], sn[H2], sd[H2];
  const int Nx = s.Nx, Ny = s.Ny;
  const int ntx = (Nx + TX - 1) / TX;
  const int i0 = (blockIdx.x % ntx) * TX, j0 = (blockIdx.x / ntx) * TY;
  const int ex = min(TX, Nx - i0), ey = min(TY, Ny - j0);
  for (int q = threadIdx.x; q < H2; q += blockDim.x) {
    const int li = q % HX - 2, lj = q / HX - 2;
    if (li > ex + 1 || lj > ey + 1) continue;
    const int gi = wrap1(i0 + li, Nx), gj = wrap1(j0 + lj, Ny);
    const long g = (long)gj * Nx + gi;
    sd[q] = dn[g];
    T v = xin ? xin[g] : T(0);
    if (xc) v += prolong_point(pm, xc, gi, gj, 0);  // the coarse correction, as k_prolong_add adds it
    sx[q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < R2; q += blockDim.x) {
    const int li = q % RX - 1, lj = q / RX - 1;
    if (li > ex || lj > ey) continue;
    const int gi = wrap1(i0 + li, Nx), gj = wrap1(j0 + lj, Ny);
    const int p = (lj + 2) * HX + (li + 2);
    T v = sx[p];
    if (((gi + gj) & 1) == first)
      v = relax_planes<T, HX, HX>(s, sd, sd, sd, p, sx, sx, sx, p, b[(long)gj * Nx + gi], v);
    sn[p] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < TX * TY; q += blockDim.x) {
    const int li = q % TX, lj = q / TX;
    if (li >= ex || lj >= ey) continue;
    const int gi = i0 + li, gj = j0 + lj;
    const int p = (lj + 2) * HX + (li + 2);
    const long g = (long)gj * Nx + gi;
    T v = sn[p];
    if (((gi + gj) & 1) != first) v = relax_planes<T, HX, HX>(s, sd, sd, sd, p, sn, sn, sn, p, b[g], v);
    xout[g] = v;
  }
}

// One red-black Gauss-Seidel sweep (colour `first`, then the other; colour of node
// (i,j,k) = (i+j+k)&1) from xin (zero when null) into xout, marching along z.  A
// workgroup owns a TX x TY column of nodes over KZ planes.  Planes of the old potential
// and of the density enter LDS rings one at a time with an xy halo of 2 (the next
// plane's loads are in flight while the current one is relaxed).  The first colour is
// relaxed on the plane below the newest, over the tile plus a halo of 1, into a third
// ring; the second colour three planes below the newest, over the tile, from the third
// ring (its three first-colour planes were all finished before this plane's barrier).
// The first colour reads old values everywhere (the 11-point stencil's xy diagonals
// join nodes of one colour), the second reads the first colour's new values and its
// own colour's old ones, so the result does not depend on thread order.  One barrier
// per plane: rings of 4, 6 and 4 planes keep every plane until no thread can still
// read it (a thread storing plane L+1 before its barrier overlaps threads still
// relaxing after plane L's, which read potential planes L-2..L, density L-4..L and
// first-colour planes L-4..L-2).
template <typename T, int TX, int TY>
__device__ inline void sweep_z(const StencilConsts &s, const T *__restrict__ dn, const T *__restrict__ xin,
                               const T *__restrict__ b, T *__restrict__ xout, int first, int KZ, int bid) {
  constexpr int NT = TX * TY;                            // one thread per tile column
  constexpr int HX = TX + 4, HY = TY + 4, HP = HX * HY;  // potential and density planes
  constexpr int RX = TX + 2, RY = TY + 2, RP = RX * RY;  // first-colour planes
  constexpr int NXO = 4, ND = 6, NSN = 4;
  constexpr int LPT = (HP + NT - 1) / NT;
  __shared__ T sxo[NXO * HP], sd[ND * HP], ssn[NSN * RP];
  const int Nx = s.Nx, Ny = s.Ny, Nz = s.Nz;
  const int ntx = (Nx + TX - 1) / TX, nty = (Ny + TY - 1) / TY;
  const int t = bid % (ntx * nty), kc = bid / (ntx * nty);
  const int i0 = (t % ntx) * TX, j0 = (t / ntx) * TY, k0 = kc * KZ;
  const int ex = min(TX, Nx - i0), ey = min(TY, Ny - j0), ez = min(KZ, Nz - k0);
  const int tid = threadIdx.x;
  const long plane = (long)Nx * Ny;
  int lq[LPT];
  long goff[LPT];
#pragma unroll
  for (int u = 0; u < LPT; ++u) {
    const int q = tid + u * NT;
    const int li = q % HX - 2, lj = q / HX - 2;
    lq[u] = (q < HP && li <= ex + 1 && lj <= ey + 1) ? q : -1;
    goff[u] = (long)wrap1(j0 + lj, Ny) * Nx + wrap1(i0 + li, Nx);
  }
  const int Lbeg = k0 - 2, Lend = k0 + ez + 1;
  auto fetch = [&](int L, T *rx, T *rd) {
    const long zo = (long)wrapz(L, Nz) * plane;
#pragma unroll
    for (int u = 0; u < LPT; ++u)
      if (lq[u] >= 0) {
        rd[u] = dn[zo + goff[u]];
        rx[u] = xin ? xin[zo + goff[u]] : T(0);
      }
  };
  // One plane step: plane L enters LDS from registers, which then take plane L+2, so
  // two planes of loads are in flight while the relaxations run.  Step Lend + 1 stores
  // nothing and only finishes the last plane's second colour.
  auto step = [&](int L, T *rx, T *rd) {
    const int r = L - Lbeg;
    if (L <= Lend) {
      T *xo = sxo + (r % NXO) * HP, *dd = sd + (r % ND) * HP;
#pragma unroll
      for (int u = 0; u < LPT; ++u)
        if (lq[u] >= 0) {
          xo[lq[u]] = rx[u];
          dd[lq[u]] = rd[u];
        }
    }
    __syncthreads();
    if (L + 2 <= Lend) fetch(L + 2, rx, rd);
    if (r >= 2 && L <= Lend) {  // first colour on plane L-1
      const int r1 = r - 1;
      const T *x0 = sxo + ((r1 - 1) % NXO) * HP, *x1 = sxo + (r1 % NXO) * HP, *x2 = sxo + ((r1 + 1) % NXO) * HP;
      const T *d0 = sd + ((r1 - 1) % ND) * HP, *d1 = sd + (r1 % ND) * HP, *d2 = sd + ((r1 + 1) % ND) * HP;
      T *sn = ssn + (r1 % NSN) * RP;
      const int gk = wrapz(L - 1, Nz);
      for (int q = tid; q < RP; q += NT) {
        const int li = q % RX - 1, lj = q / RX - 1;
        if (li > ex || lj > ey) continue;
        const int p = (lj + 2) * HX + (li + 2);
        const int gi = wrap1(i0 + li, Nx), gj = wrap1(j0 + lj, Ny);
        T v = x1[p];
        if (((gi + gj + gk) & 1) == first)
          v = relax_planes<T, HX, HX>(s, d0, d1, d2, p, x0, x1, x2, p, b[gk * plane + (long)gj * Nx + gi], v);
        sn[q] = v;
      }
    }
    if (r >= 5) {  // second colour on plane L-3, written out
      const int r3 = r - 3;
      const T *s0 = ssn + ((r3 - 1) % NSN) * RP, *s1 = ssn + (r3 % NSN) * RP, *s2 = ssn + ((r3 + 1) % NSN) * RP;
      const T *d0 = sd + ((r3 - 1) % ND) * HP, *d1 = sd + (r3 % ND) * HP, *d2 = sd + ((r3 + 1) % ND) * HP;
      const int gk = L - 3;
      const int li = tid % TX, lj = tid / TX;  // TX * TY == NT: one node per thread
      if (li < ex && lj < ey) {
        const int pd = (lj + 2) * HX + (li + 2), ps = (lj + 1) * RX + (li + 1);
        const int gi = i0 + li, gj = j0 + lj;
        const long g = gk * plane + (long)gj * Nx + gi;
        T v = s1[ps];
        if (((gi + gj + gk) & 1) != first) v = relax_planes<T, HX, RX>(s, d0, d1, d2, pd, s0, s1, s2, ps, b[g], v);
        xout[g] = v;
      }
    }
  };
  T rxa[LPT], rda[LPT], rxb[LPT], rdb[LPT];
  fetch(Lbeg, rxa, rda);
  fetch(Lbeg + 1, rxb, rdb);  // Lend >= Lbeg + 4
  for (int L = Lbeg; L <= Lend + 1; L += 2) {
    step(L, rxa, rda);
    if (L + 1 <= Lend + 1) step(L + 1, rxb, rdb);
  }
}

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_sweep_z(StencilConsts s, const T *__restrict__ dn,
                                                    const T *__restrict__ xin, const T *__restrict__ b,
                                                    T *__restrict__ xout, int first, int KZ) {
  sweep_z<T, TX, TY>(s, dn, xin, b, xout, first, KZ, xcd_order(blockIdx.x, gridDim.x));
}

constexpr int SW_TX = 16, SW_TY = 8, SW_TZ = 4;  // box of the one-shot sweep
constexpr int S2_TX = 32, S2_TY = 16;             // tile of the single-plane sweep
constexpr int SZ_TX = 32, SZ_TY = 16;            // xy tile of the marching sweep
constexpr long SZ_MIN_POINTS = 1L << 21;         // levels at least this large march
// In the lockstep batch every launch spans all problems, so the marching sweep's lower
// load redundancy pays from much smaller levels (given 8 planes or more).
constexpr long SZ_MIN_POINTS_BATCH = 1L << 18;
constexpr int DENSE_MAX = 1024;                // largest coarsest level solved by a dense inverse
constexpr size_t TAIL_LDS_MAX = 140 * 1024;    // LDS of the single-workgroup tail

template <typename T>
__global__ void k_jacobi(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ r, T *__restrict__ z) {
  NODE_IJK(s.Nx, s.Ny)
  T c[11];
  row_coeffs(dn, s, i, j, k, c);
  z[q] = c[10] != (T)0 ? r[q] / c[10] : (T)0;
}

// Full weighting [1/4 1/2 1/4] along the coarsened axes, injection along the others.
template <typename T>
__device__ inline T restrict_point(const CoarsenMap &m, const T *f, int I, int J, int K) {
  const int fi = m.cx ? 2 * I : I, fj = m.cy ? 2 * J : J, fk = m.cz ? 2 * K : K;
  double acc = 0.0;
  for (int c = -m.cz; c <= m.cz; ++c) {
    const double wz = m.cz ? (c == 0 ? 0.5 : 0.25) : 1.0;
    const int kk = wrap1(fk + c, m.Nzf);
    for (int b = -m.cy; b <= m.cy; ++b) {
      const double wy = m.cy ? (b == 0 ? 0.5 : 0.25) : 1.0;
      const int jj = wrap1(fj + b, m.Nyf);
      for (int a = -m.cx; a <= m.cx; ++a) {
        const double wx = m.cx ? (a == 0 ? 0.5 : 0.25) : 1.0;
        acc += wz * wy * wx * (double)f[((long)kk * m.Nyf + jj) * m.Nxf + wrap1(fi + a, m.Nxf)];
      }
    }
  }
  return (T)acc;
}


template <typename T>
__global__ void k_restrict(CoarsenMap m, const T *__restrict__ f, T *__restrict__ cvec) {
  NODE_IJK(m.Nxc, m.Nyc)
  cvec[q] = restrict_point(m, f, i, j, k);
}

// x_f += P x_c
template <typename T>
__global__ void k_prolong_add(CoarsenMap m, const T *__restrict__ xc, T *__restrict__ xf) {
  NODE_IJK(m.Nxf, m.Nyf)
  xf[q] += prolong_point(m, xc, i, j, k);
}

// ---- coarse end of the V-cycle in one workgroup -----------------------------------
// Levels whose density, two potential buffers, right-hand side and residual fit in
// LDS run in one launch: smoothing, residual, restriction and prolongation are
// workgroup-wide passes between barriers, and the coarsest level is solved with a
// dense inverse of (A + alpha 1 1^T), exact for the mean-free right-hand sides the
// cycle produces (A's rows and columns sum to zero), followed by mean removal.
constexpr int TAIL_MAX = 8;
constexpr int TAIL_THREADS = 1024;

template <typename T>
struct TailArgs {
  int nl;  // tail levels, the last being the coarsest
  int pre, post, nc;
  StencilConsts s[TAIL_MAX];
  CoarsenMap m[TAIL_MAX];  // level l -> l + 1
  int off[TAIL_MAX];       // LDS block of level l: dn, x, x2, b, r (n_l elements each)
  const T *dn[TAIL_MAX];
  const T *ainv;           // nc x nc, row-major
  const T *bin;            // right-hand side of the first tail level
  T *xout;                 // its solution
};

// One workgroup-wide red-black sweep on LDS arrays: X -> X2 (first colour) -> X.
template <typename T>
__device__ inline void tail_sweep(const StencilConsts &s, const T *D, T *X, T *X2, const T *B, int first, bool zero) {
  const int n = s.Nx * s.Ny * s.Nz;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / (s.Nx * s.Ny);
    T v = zero ? T(0) : X[q];
    if (((i + j + k) & 1) == first) {
      T c[11];
      row_coeffs(D, s, i, j, k, c);
      if (c[10] != T(0)) v = (B[q] - (zero ? T(0) : offdiag(X, s, i, j, k, c))) / c[10];
    }
    X2[q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / (s.Nx * s.Ny);
    T v = X2[q];
    if (((i + j + k) & 1) != first) {
      T c[11];
      row_coeffs(D, s, i, j, k, c);
      if (c[10] != T(0)) v = (B[q] - offdiag(X2, s, i, j, k, c)) / c[10];
    }
    X[q] = v;
  }
  __syncthreads();
}

// The V-cycle's coarse end for one problem, in one workgroup (the body of k_tail; the
// batched solver runs one workgroup per problem).
template <typename T>
__device__ inline void tail_body(const TailArgs<T> &a) {
  extern __shared__ double tail_lds[];
  T *sm = reinterpret_cast<T *>(tail_lds);
  __shared__ double red[TAIL_THREADS];
  for (int l = 0; l < a.nl; ++l) {
    const int n = a.s[l].Nx * a.s[l].Ny * a.s[l].Nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) sm[a.off[l] + q] = a.dn[l][q];
  }
  {
    const int n = a.s[0].Nx * a.s[0].Ny * a.s[0].Nz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) sm[a.off[0] + 3 * n + q] = a.bin[q];
  }
  __syncthreads();
  for (int l = 0; l + 1 < a.nl; ++l) {
    const StencilConsts &s = a.s[l];
    const int n = s.Nx * s.Ny * s.Nz;
    T *D = sm + a.off[l], *X = D + n, *X2 = X + n, *B = X2 + n, *R = B + n;
    for (int it = 0; it < a.pre; ++it) tail_sweep(s, D, X, X2, B, 0, it == 0);
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / (s.Nx * s.Ny);
      T c[11];
      row_coeffs(D, s, i, j, k, c);
      R[q] = B[q] - (offdiag(X, s, i, j, k, c) + c[10] * X[q]);
    }
    __syncthreads();
    const CoarsenMap &m = a.m[l];
    const int nc = m.Nxc * m.Nyc * m.Nzc;
    T *Bc = sm + a.off[l + 1] + 3 * nc;
    for (int q = threadIdx.x; q < nc; q += blockDim.x)
      Bc[q] = restrict_point(m, R, q % m.Nxc, (q / m.Nxc) % m.Nyc, q / (m.Nxc * m.Nyc));
    __syncthreads();
  }
  {
    const int l = a.nl - 1, n = a.nc;
    T *X = sm + a.off[l] + n, *B = X + 2 * n;
    double part = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      T acc = 0;
      const T *row = a.ainv + (long)q * n;
      for (int c = 0; c < n; ++c) acc += row[c] * B[c];
      X[q] = acc;
      part += (double)acc;
    }
    red[threadIdx.x] = part;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    const T mean = (T)(red[0] / n);
    for (int q = threadIdx.x; q < n; q += blockDim.x) X[q] -= mean;
    __syncthreads();
  }
  for (int l = a.nl - 2; l >= 0; --l) {
    const StencilConsts &s = a.s[l];
    const int n = s.Nx * s.Ny * s.Nz;
    T *D = sm + a.off[l], *X = D + n, *X2 = X + n, *B = X2 + n;
    const CoarsenMap &m = a.m[l];
    const T *Xc = sm + a.off[l + 1] + m.Nxc * m.Nyc * m.Nzc;
    for (int q = threadIdx.x; q < n; q += blockDim.x)
      X[q] += prolong_point(m, Xc, q % s.Nx, (q / s.Nx) % s.Ny, q / (s.Nx * s.Ny));
    __syncthreads();
    for (int it = 0; it < a.post; ++it) tail_sweep(s, D, X, X2, B, 1, false);
  }
  const int n0 = a.s[0].Nx * a.s[0].Ny * a.s[0].Nz;
  for (int q = threadIdx.x; q < n0; q += blockDim.x) a.xout[q] = sm[a.off[0] + n0 + q];
}

template <typename T>
__global__ void __launch_bounds__(TAIL_THREADS) k_tail(TailArgs<T> a) {
  tail_body(a);
}

// ---- vector kernels ------------------------------------------------------------
constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

// partial[blk*m + q] = sum over the block's stripe of V[q][i] * w[i], q < m.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_dot_multi(int m, T *const *__restrict__ V, const T *__restrict__ w,
                                                           long n, double *__restrict__ partial) {
  __shared__ double red[RED_THREADS];
  for (int q0 = 0; q0 < m; q0 += 8) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int mq = min(8, m - q0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
      const double wi = (double)w[i];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < mq) acc[q] += (double)V[q0 + q][i] * wi;
    }
    for (int q = 0; q < mq; ++q) {
      red[threadIdx.x] = acc[q];
      __syncthreads();
      for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
      }
      if (threadIdx.x == 0) partial[(long)blockIdx.x * m + q0 + q] = red[0];
      __syncthreads();
    }
  }
}

// out[q] = scale * sum_b partial[b*m + q] (+ out[q] if accumulate), fixed order.
static __global__ void k_reduce_partials(int m, int nb, const double *__restrict__ partial, double *__restrict__ out,
                                  int accumulate) {
  __shared__ double red[256];
  for (int q = 0; q < m; ++q) {
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) a += partial[(long)b * m + q];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = accumulate ? out[q] + red[0] : red[0];
    __syncthreads();
  }
}

// w -= sum_q h[q] V[q]
template <typename T>
__global__ void k_sub_multi(int m, T *const *__restrict__ V, const double *__restrict__ h, T *__restrict__ w, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = (double)w[i];
    for (int q = 0; q < m; ++q) a -= h[q] * (double)V[q][i];
    w[i] = (T)a;
  }
}

// Fused Gram-Schmidt step for m <= 8 basis vectors: w -= sum_q h[q] V[q] (as
// k_sub_multi), then per block either the products V[q].w (norm == 0; partial[b*m+q]) or
// w.w (norm == 1; partial[b]) of the updated, stored w, reduced as k_dot_multi does.
// One pass over V and w instead of two.
template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_sub_dot_multi(int m, T *const *__restrict__ V,
                                                               const double *__restrict__ h, T *__restrict__ w, long n,
                                                               double *__restrict__ partial, int norm) {
  __shared__ double red[RED_THREADS];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double hq[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) hq[q] = q < m ? h[q] : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double v[8];
    double a = (double)w[i];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < m) {
        v[q] = (double)V[q][i];
        a -= hq[q] * v[q];
      }
    const T wt = (T)a;
    w[i] = wt;
    const double wn = (double)wt;
    if (norm) {
      acc[0] += wn * wn;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < m) acc[q] += v[q] * wn;
    }
  }
  const int mq = norm ? 1 : m;
  for (int q = 0; q < mq; ++q) {
    red[threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(long)blockIdx.x * mq + q] = red[0];
    __syncthreads();
  }
}

// x += sum_q y[q] Z[q]
template <typename T>
__global__ void k_add_multi(int m, T *const *__restrict__ Z, const double *__restrict__ y, T *__restrict__ x, long n,
                            double scale) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int q = 0; q < m; ++q) a += y[q] * (double)Z[q][i];
    x[i] = (T)((double)x[i] + scale * a);
  }
}

// y = x / *nrm (nrm on the device)
template <typename T>
__global__ void k_scale_by_inv(const T *__restrict__ x, T *__restrict__ y, const double *__restrict__ nrm2, long n,
                               int is_sq) {
  const double d = is_sq ? sqrt(*nrm2) : *nrm2;
  const double inv = d > 0 ? 1.0 / d : 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (T)((double)x[i] * inv);
}

template <typename T>
__global__ void k_sub_scalar(T *__restrict__ v, const double *__restrict__ sum, long n) {
  const double mean = *sum / (double)n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[i] = (T)((double)v[i] - mean);
}

template <typename T, typename U>
__global__ void k_convert(const T *__restrict__ a, U *__restrict__ b, long n, double scale) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (U)((double)a[i] * scale);
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, long n, T v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = v;
}

template <typename T>
__global__ void k_axpy_from(const T *__restrict__ x, T *__restrict__ y, long n, double a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    y[i] = (T)((double)y[i] + a * (double)x[i]);
}

// Density of the FluidNode moments -> level-0 density.
template <typename T>
__global__ void k_density(const double *__restrict__ m, T *__restrict__ dn, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dn[i] = (T)m[i * 4];
}

// ComputeFullRHS3D body, src/rhs.c:321-351.
struct RhsConsts {
  int Nx, Ny, Nz;
  double kappa, E0, scale, cth, cG, hx, hy, hz, hxx, hyy, hzz;
};

// The forcing at node (i, j, k).
__device__ inline double rhs_point(const RhsConsts &c, const double *__restrict__ m, int i, int j, int k) {
  const int ip = i + 1 == c.Nx ? 0 : i + 1, im = i == 0 ? c.Nx - 1 : i - 1;
  const int jp = j + 1 == c.Ny ? 0 : j + 1, jm = j == 0 ? c.Ny - 1 : j - 1;
  const int kp = k + 1 == c.Nz ? 0 : k + 1, km = k == 0 ? c.Nz - 1 : k - 1;
  const long row = (long)c.Ny * c.Nx;
#define MM(ii, jj, kk, d) m[(((long)(kk)) * row + (long)(jj) * c.Nx + (ii)) * 4 + (d)]
  const double nijk = MM(i, j, k, 0), npjk = MM(ip, j, k, 0), nmjk = MM(im, j, k, 0);
  const double nipk = MM(i, jp, k, 0), nimk = MM(i, jm, k, 0);
  const double nijp = MM(i, j, kp, 0), nijm = MM(i, j, km, 0);
  const double dndx = c.hx * (npjk - nmjk);
  const double dndy = c.hy * (nipk - nimk);
  const double d2ndxx = c.hxx * (npjk - 2 * nijk + nmjk);
  const double d2ndyy = c.hyy * (nipk - 2 * nijk + nimk);
  const double d2ndzz = c.hzz * (nijp - 2 * nijk + nijm);
  const double dGxdx = c.hx * (MM(ip, j, k, 1) - MM(im, j, k, 1));
  const double dGydy = c.hy * (MM(i, jp, k, 2) - MM(i, jm, k, 2));
  const double dGzdz = c.hz * (MM(i, j, kp, 3) - MM(i, j, km, 3));
#undef MM
  const double Eterm = (dndy - c.kappa * dndx) * c.E0;
  const double Pterm = c.cth * (d2ndxx + d2ndyy + (1 + c.kappa * c.kappa) * d2ndzz);
  const double Gterm = c.cG * (dGxdx + dGydy + dGzdz);
  return c.scale * (Eterm + Pterm + Gterm);
}

static __global__ void k_rhs(RhsConsts c, const double *__restrict__ m, double *__restrict__ b) {
  NODE_IJK(c.Nx, c.Ny)
  b[q] = rhs_point(c, m, i, j, k);
}

// ComputeSinusoidalRHS3D body (src/rhs.c:120-133): n0 scale cos(2 pi x) cos(2 pi y)
// cos(2 pi z) at the cell centres x = (i + 1/2) dx, ...
static __global__ void k_rhs_sinusoidal(int Nx, int Ny, double dx, double dy, double dz, double n0, double scale,
                                        double *__restrict__ b) {
  NODE_IJK(Nx, Ny)
  const double PI = 3.14159265358979323846;
  const double x = ((double)i + 0.5) * dx, y = ((double)j + 0.5) * dy, z = ((double)k + 0.5) * dz;
  const double val = cos(2 * PI * x) * cos(2 * PI * y) * cos(2 * PI * z);
  b[q] = n0 * val * scale;
}

// Host: the stencil constants of a level, the forcing constants, and the dense inverse
// of the coarsest operator plus alpha 1 1^T (Gauss-Jordan with partial pivoting in FP64)
// from the level's density dnh.
inline StencilConsts stencil_consts(const ecsprt_params &p, int Nx, int Ny, int Nz, double dx, double dy, double dz) {
  // Scale factors of ComputeFullLHS3D, src/lhs.c:366-375.
  StencilConsts s;
  s.Nx = Nx;
  s.Ny = Ny;
  s.Nz = Nz;
  const double kappa = p.kappa_e;
  const double hxx = 1.0 / (2.0 * dx * dx), hyy = 1.0 / (2.0 * dy * dy), hzz = 1.0 / (2.0 * dz * dz);
  const double hod = 1.0 / (8.0 * dy * dx);
  s.sxx = p.scale * hxx;
  s.syy = p.scale * hyy;
  s.szz = p.ndim == 2 ? 0.0 : p.scale * (1 + kappa * kappa) * hzz;  // ComputeFullLHS2D: no z coupling
  s.sod = p.scale * kappa * hod;
  return s;
}

inline RhsConsts rhs_consts(const ecsprt_params &p) {
  RhsConsts k;
  k.Nx = p.Nx;
  k.Ny = p.Ny;
  k.Nz = p.Nz;
  k.kappa = p.kappa_e;
  k.E0 = p.E0;
  k.scale = p.scale;
  // src/rhs.c:291-300
  k.cth = p.gammaT * KB_ * p.Te / Q_;
  k.cG = (1 + p.kappa_e * p.kappa_e) * p.me * p.nue / Q_;
  k.hx = 1.0 / (2.0 * p.dx);
  k.hy = 1.0 / (2.0 * p.dy);
  k.hz = 1.0 / (2.0 * p.dz);
  k.hxx = 1.0 / (p.dx * p.dx);
  k.hyy = 1.0 / (p.dy * p.dy);
  k.hzz = 1.0 / (p.dz * p.dz);
  if (p.ndim == 2) k.hz = k.hzz = 0.0;  // ComputeFullRHS2D (src/rhs.c:151-256) has no z terms
  return k;
}

// z-chunk of the marching sweep: about 512 workgroups (two per CU), at least 6 planes
// per chunk so the z halo of 4 planes stays a modest overhead.
inline int sweep_kz(int Nx, int Ny, int Nz) {
  const int nxy = ((Nx + SZ_TX - 1) / SZ_TX) * ((Ny + SZ_TY - 1) / SZ_TY);
  const int chunks = std::max(1, 512 / nxy);
  return std::min(Nz, std::max(6, (Nz + chunks - 1) / chunks));
}

template <typename T>
inline std::vector<double> coarse_inverse(const StencilConsts &s, const std::vector<T> &dnh) {
  const int n = s.Nx * s.Ny * s.Nz;
  std::vector<double> A((size_t)n * n, 0.0), Inv((size_t)n * n, 0.0);
  double diag = 0.0;
  for (int k = 0; k < s.Nz; ++k)
    for (int j = 0; j < s.Ny; ++j)
      for (int i = 0; i < s.Nx; ++i) {
        T cf[11];
        row_coeffs(dnh.data(), s, i, j, k, cf);
        const int ip = (i + 1) % s.Nx, im = (i + s.Nx - 1) % s.Nx, jp = (j + 1) % s.Ny, jm = (j + s.Ny - 1) % s.Ny;
        const int kp = (k + 1) % s.Nz, km = (k + s.Nz - 1) % s.Nz;
        auto id = [&](int a, int b2, int d) { return ((long)d * s.Ny + b2) * s.Nx + a; };
        const long q = id(i, j, k);
        double *row = &A[(size_t)q * n];
        row[id(ip, j, k)] += cf[0];
        row[id(im, j, k)] += cf[1];
        row[id(i, jp, k)] += cf[2];
        row[id(i, jm, k)] += cf[3];
        row[id(i, j, kp)] += cf[4];
        row[id(i, j, km)] += cf[5];
        row[id(ip, jp, k)] += cf[6];
        row[id(ip, jm, k)] += cf[7];
        row[id(im, jp, k)] += cf[8];
        row[id(im, jm, k)] += cf[9];
        row[q] += cf[10];
        diag += std::fabs((double)cf[10]);
      }
  const double alpha = diag > 0 ? diag / n / n : 1.0;
  for (auto &v : A) v += alpha;
  for (int q = 0; q < n; ++q) Inv[(size_t)q * n + q] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
    ECS_CHECK(A[(size_t)piv * n + col] != 0.0, ECSPRT_ERR_NOT_CONVERGED, "singular coarse operator");
    if (piv != col)
      for (int q = 0; q < n; ++q) {
        std::swap(A[(size_t)piv * n + q], A[(size_t)col * n + q]);
        std::swap(Inv[(size_t)piv * n + q], Inv[(size_t)col * n + q]);
      }
    const double d = 1.0 / A[(size_t)col * n + col];
    for (int q = 0; q < n; ++q) {
      A[(size_t)col * n + q] *= d;
      Inv[(size_t)col * n + q] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      const double f = A[(size_t)r * n + col];
      if (f == 0.0) continue;
      for (int q = 0; q < n; ++q) {
        A[(size_t)r * n + q] -= f * A[(size_t)col * n + q];
        Inv[(size_t)r * n + q] -= f * Inv[(size_t)col * n + q];
      }
    }
  }
  return Inv;
}

}  // namespace ecsprt
