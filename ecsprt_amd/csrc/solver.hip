// Potential solve of the hot path on gfx950: the variable-coefficient 11-point
// operator of ComputeFullLHS3D (src/lhs.c:337-505) applied matrix-free from the
// density, the forcing of ComputeFullRHS3D (src/rhs.c:259-373), a geometric
// multigrid preconditioner and right-preconditioned GMRES(m) replacing PETSc's
// KSPSolve (src/potential.c:11).
//
// Multigrid: coarse operators are the same 11-point formula rediscretised on the
// coarse spacing from the full-weighting restriction of the density.  An axis is
// coarsened when its coupling is within a factor 2 of the strongest one, so the
// (1+kappa_e^2) z-anisotropy is removed by z-semicoarsening before full
// coarsening; red-black point Gauss-Seidel smooths; the coarsest level is solved by
// one workgroup in LDS.  All reductions are two-stage with a fixed order, so every
// solve is bitwise reproducible.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace ecsprt {

// Physical constants, src/constants.h:4-7.
constexpr double KB_ = 1.3807e-23;
constexpr double Q_ = 1.6022e-19;

struct StencilConsts {
  int Nx, Ny, Nz;
  double sxx, syy, szz, sod;
};

// 11 coefficients of row (i,j,k) in the reference order (lhs.c:409-423).
template <typename T>
__device__ inline void row_coeffs(const T *__restrict__ dn, const StencilConsts &s, int i, int j, int k, T c[11]) {
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

// Off-diagonal part of row (i,j,k) applied to x.
template <typename T>
__device__ inline T offdiag(const T *__restrict__ x, const StencilConsts &s, int i, int j, int k, const T c[11]) {
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

template <typename T>
__global__ void k_apply(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ x, T *__restrict__ y) {
  const long n = (long)s.Nx * s.Ny * s.Nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / ((long)s.Nx * s.Ny);
    T c[11];
    row_coeffs(dn, s, i, j, k, c);
    y[q] = offdiag(x, s, i, j, k, c) + c[10] * x[q];
  }
}

// r = b - A x
template <typename T>
__global__ void k_residual(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ x, const T *__restrict__ b,
                           T *__restrict__ r) {
  const long n = (long)s.Nx * s.Ny * s.Nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / ((long)s.Nx * s.Ny);
    T c[11];
    row_coeffs(dn, s, i, j, k, c);
    r[q] = b[q] - (offdiag(x, s, i, j, k, c) + c[10] * x[q]);
  }
}

// One colour of red-black Gauss-Seidel: nodes with (i+j+k)&1 == colour.
template <typename T>
__global__ void k_smooth_rb(StencilConsts s, const T *__restrict__ dn, T *__restrict__ x, const T *__restrict__ b,
                            int colour) {
  const long n = (long)s.Nx * s.Ny * s.Nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / ((long)s.Nx * s.Ny);
    if (((i + j + k) & 1) != colour) continue;
    T c[11];
    row_coeffs(dn, s, i, j, k, c);
    if (c[10] != (T)0) x[q] = (b[q] - offdiag(x, s, i, j, k, c)) / c[10];
  }
}

template <typename T>
__global__ void k_jacobi(StencilConsts s, const T *__restrict__ dn, const T *__restrict__ r, T *__restrict__ z) {
  const long n = (long)s.Nx * s.Ny * s.Nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / ((long)s.Nx * s.Ny);
    T c[11];
    row_coeffs(dn, s, i, j, k, c);
    z[q] = c[10] != (T)0 ? r[q] / c[10] : (T)0;
  }
}

struct CoarsenMap {
  int Nxf, Nyf, Nzf, Nxc, Nyc, Nzc;
  int cx, cy, cz;
};

// Full weighting [1/4 1/2 1/4] along the coarsened axes, injection along the others.
template <typename T>
__global__ void k_restrict(CoarsenMap m, const T *__restrict__ f, T *__restrict__ cvec) {
  const long n = (long)m.Nxc * m.Nyc * m.Nzc;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int I = q % m.Nxc, J = (q / m.Nxc) % m.Nyc, K = q / ((long)m.Nxc * m.Nyc);
    const int fi = m.cx ? 2 * I : I, fj = m.cy ? 2 * J : J, fk = m.cz ? 2 * K : K;
    double acc = 0.0;
    for (int c = -m.cz; c <= m.cz; ++c) {
      const double wz = m.cz ? (c == 0 ? 0.5 : 0.25) : 1.0;
      const int kk = (fk + c + m.Nzf) % m.Nzf;
      for (int b = -m.cy; b <= m.cy; ++b) {
        const double wy = m.cy ? (b == 0 ? 0.5 : 0.25) : 1.0;
        const int jj = (fj + b + m.Nyf) % m.Nyf;
        for (int a = -m.cx; a <= m.cx; ++a) {
          const double wx = m.cx ? (a == 0 ? 0.5 : 0.25) : 1.0;
          const int ii = (fi + a + m.Nxf) % m.Nxf;
          acc += wz * wy * wx * (double)f[((long)kk * m.Nyf + jj) * m.Nxf + ii];
        }
      }
    }
    cvec[q] = (T)acc;
  }
}

// x_f += P x_c, linear interpolation along the coarsened axes.
template <typename T>
__global__ void k_prolong_add(CoarsenMap m, const T *__restrict__ xc, T *__restrict__ xf) {
  const long n = (long)m.Nxf * m.Nyf * m.Nzf;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % m.Nxf, j = (q / m.Nxf) % m.Nyf, k = q / ((long)m.Nxf * m.Nyf);
    int ia[2], ja[2], ka[2];
    T wa[2], wb[2], wc[2];
    int na = 1, nb = 1, nc = 1;
    if (m.cx) {
      if (i & 1) { ia[0] = i >> 1; ia[1] = ((i + 1) >> 1) % m.Nxc; wa[0] = wa[1] = (T)0.5; na = 2; }
      else { ia[0] = i >> 1; wa[0] = (T)1; }
    } else { ia[0] = i; wa[0] = (T)1; }
    if (m.cy) {
      if (j & 1) { ja[0] = j >> 1; ja[1] = ((j + 1) >> 1) % m.Nyc; wb[0] = wb[1] = (T)0.5; nb = 2; }
      else { ja[0] = j >> 1; wb[0] = (T)1; }
    } else { ja[0] = j; wb[0] = (T)1; }
    if (m.cz) {
      if (k & 1) { ka[0] = k >> 1; ka[1] = ((k + 1) >> 1) % m.Nzc; wc[0] = wc[1] = (T)0.5; nc = 2; }
      else { ka[0] = k >> 1; wc[0] = (T)1; }
    } else { ka[0] = k; wc[0] = (T)1; }
    T acc = 0;
    for (int c = 0; c < nc; ++c)
      for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a)
          acc += wc[c] * wb[b] * wa[a] * xc[((long)ka[c] * m.Nyc + ja[b]) * m.Nxc + ia[a]];
    xf[q] += acc;
  }
}

// Coarsest level: zero guess, `sweeps` red-black sweeps and mean removal, all in
// LDS by one workgroup (n <= 4096).
template <typename T>
__global__ void __launch_bounds__(1024) k_coarse_solve(StencilConsts s, const T *__restrict__ dn,
                                                       const T *__restrict__ b, T *__restrict__ x, int sweeps) {
  __shared__ T sx[4096];
  __shared__ double red[1024];
  const int n = s.Nx * s.Ny * s.Nz;
  for (int q = threadIdx.x; q < n; q += blockDim.x) sx[q] = 0;
  __syncthreads();
  for (int it = 0; it < sweeps; ++it) {
    for (int colour = 0; colour < 2; ++colour) {
      for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % s.Nx, j = (q / s.Nx) % s.Ny, k = q / (s.Nx * s.Ny);
        if (((i + j + k) & 1) != colour) continue;
        T c[11];
        row_coeffs(dn, s, i, j, k, c);
        if (c[10] != (T)0) sx[q] = (b[q] - offdiag(sx, s, i, j, k, c)) / c[10];
      }
      __syncthreads();
    }
  }
  double part = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) part += (double)sx[q];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const T mean = (T)(red[0] / n);
  for (int q = threadIdx.x; q < n; q += blockDim.x) x[q] = sx[q] - mean;
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
__global__ void k_reduce_partials(int m, int nb, const double *__restrict__ partial, double *__restrict__ out,
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

__global__ void k_rhs(RhsConsts c, const double *__restrict__ m, double *__restrict__ b) {
  const long n = (long)c.Nx * c.Ny * c.Nz;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    const int i = q % c.Nx, j = (q / c.Nx) % c.Ny, k = q / ((long)c.Nx * c.Ny);
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
    b[q] = c.scale * (Eterm + Pterm + Gterm);
  }
}

// ---- host structures ------------------------------------------------------------
static int gsz(long n) {
  long b = (n + 255) / 256;
  return (int)std::max(1L, std::min(b, 8192L));
}

template <typename T>
struct Level {
  StencilConsts s;
  long n = 0;
  double dx = 1, dy = 1, dz = 1;
  CoarsenMap to_coarse{};  // map from this level to the next
  T *dn = nullptr, *x = nullptr, *b = nullptr, *r = nullptr;
};

template <typename T>
struct MG {
  std::vector<Level<T>> L;
};

template <typename T>
struct Gmres {
  int m = 0;
  long n = 0;
  std::vector<T *> V, Z;
  T **dV = nullptr, **dZ = nullptr;
  T *w = nullptr, *r = nullptr;
  double *partial = nullptr, *dh = nullptr, *hpin = nullptr, *dy = nullptr, *ypin = nullptr;
};

struct Solver {
  MG<double> mg64;
  MG<float> mg32;
  Gmres<double> g64;
  Gmres<float> g32;
  double *r64 = nullptr, *sum = nullptr;
  float *b32 = nullptr, *x32 = nullptr;
  bool have32 = false;
};

static StencilConsts stencil_consts(const ecsprt_params &p, int Nx, int Ny, int Nz, double dx, double dy, double dz) {
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
  s.szz = p.scale * (1 + kappa * kappa) * hzz;
  s.sod = p.scale * kappa * hod;
  return s;
}

template <typename T>
static void mg_create(ecsprt_hip_ctx *c, MG<T> &mg) {
  const ecsprt_params &p = c->p;
  int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
  double dx = p.dx, dy = p.dy, dz = p.dz;
  const double k2 = 1 + p.kappa_e * p.kappa_e;
  for (int lev = 0; lev < 16; ++lev) {
    Level<T> L;
    L.s = stencil_consts(p, Nx, Ny, Nz, dx, dy, dz);
    L.n = (long)Nx * Ny * Nz;
    L.dx = dx;
    L.dy = dy;
    L.dz = dz;
    ECS_HIP(hipMalloc(&L.dn, sizeof(T) * L.n));
    ECS_HIP(hipMalloc(&L.x, sizeof(T) * L.n));
    ECS_HIP(hipMalloc(&L.b, sizeof(T) * L.n));
    ECS_HIP(hipMalloc(&L.r, sizeof(T) * L.n));
    // Coarsening rule: strong axes (coupling >= half the strongest) that halve cleanly.
    const double a[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), k2 / (dz * dz)};
    const int N[3] = {Nx, Ny, Nz};
    const double amax = std::max(a[0], std::max(a[1], a[2]));
    int cf[3] = {0, 0, 0};
    int any = 0;
    for (int d = 0; d < 3; ++d)
      if (a[d] >= 0.5 * amax && N[d] % 2 == 0 && N[d] >= 4) { cf[d] = 1; any = 1; }
    if (!any || L.n <= 512) {
      mg.L.push_back(L);
      break;
    }
    L.to_coarse = CoarsenMap{Nx, Ny, Nz, cf[0] ? Nx / 2 : Nx, cf[1] ? Ny / 2 : Ny, cf[2] ? Nz / 2 : Nz,
                             cf[0], cf[1], cf[2]};
    mg.L.push_back(L);
    Nx = L.to_coarse.Nxc;
    Ny = L.to_coarse.Nyc;
    Nz = L.to_coarse.Nzc;
    if (cf[0]) dx *= 2;
    if (cf[1]) dy *= 2;
    if (cf[2]) dz *= 2;
  }
}

template <typename T>
static void mg_free(MG<T> &mg) {
  for (auto &L : mg.L) {
    (void)hipFree(L.dn);
    (void)hipFree(L.x);
    (void)hipFree(L.b);
    (void)hipFree(L.r);
  }
  mg.L.clear();
}

// Level-0 density from the moments, restricted down the hierarchy.
template <typename T>
static void mg_set_density(ecsprt_hip_ctx *c, MG<T> &mg) {
  hipStream_t st = c->stream;
  k_density<T><<<gsz(mg.L[0].n), 256, 0, st>>>(c->d_moments, mg.L[0].dn, mg.L[0].n);
  ECS_HIP(hipGetLastError());
  for (size_t l = 0; l + 1 < mg.L.size(); ++l) {
    k_restrict<T><<<gsz(mg.L[l + 1].n), 256, 0, st>>>(mg.L[l].to_coarse, mg.L[l].dn, mg.L[l + 1].dn);
    ECS_HIP(hipGetLastError());
  }
}

// One V-cycle: z = M^-1 r (r, z on level 0).
template <typename T>
static void mg_vcycle(ecsprt_hip_ctx *c, MG<T> &mg, const T *r, T *z) {
  hipStream_t st = c->stream;
  const int pre = c->o.mg_pre > 0 ? c->o.mg_pre : 2;
  const int post = c->o.mg_post > 0 ? c->o.mg_post : 2;
  const int csw = c->o.mg_coarse_sweeps > 0 ? c->o.mg_coarse_sweeps : 50;
  const int nl = (int)mg.L.size();
  const T *b = r;
  T *x = z;
  for (int l = 0; l < nl; ++l) {
    Level<T> &L = mg.L[l];
    if (l > 0) {
      b = L.b;
      x = L.x;
    }
    if (l == nl - 1) {
      if (L.n <= 4096) {
        k_coarse_solve<T><<<1, 1024, 0, st>>>(L.s, L.dn, b, x, csw);
      } else {
        k_fill<T><<<gsz(L.n), 256, 0, st>>>(x, L.n, (T)0);
        for (int it = 0; it < csw; ++it)
          for (int col = 0; col < 2; ++col) k_smooth_rb<T><<<gsz(L.n), 256, 0, st>>>(L.s, L.dn, x, b, col);
      }
      ECS_HIP(hipGetLastError());
      break;
    }
    k_fill<T><<<gsz(L.n), 256, 0, st>>>(x, L.n, (T)0);
    for (int it = 0; it < pre; ++it)
      for (int col = 0; col < 2; ++col) k_smooth_rb<T><<<gsz(L.n), 256, 0, st>>>(L.s, L.dn, x, b, col);
    k_residual<T><<<gsz(L.n), 256, 0, st>>>(L.s, L.dn, x, b, L.r);
    k_restrict<T><<<gsz(mg.L[l + 1].n), 256, 0, st>>>(L.to_coarse, L.r, mg.L[l + 1].b);
    ECS_HIP(hipGetLastError());
  }
  for (int l = nl - 2; l >= 0; --l) {
    Level<T> &L = mg.L[l];
    T *xl = l == 0 ? z : L.x;
    const T *bl = l == 0 ? r : L.b;
    k_prolong_add<T><<<gsz(L.n), 256, 0, st>>>(L.to_coarse, mg.L[l + 1].x, xl);
    for (int it = 0; it < post; ++it)
      for (int col = 1; col >= 0; --col) k_smooth_rb<T><<<gsz(L.n), 256, 0, st>>>(L.s, L.dn, xl, bl, col);
    ECS_HIP(hipGetLastError());
  }
}

template <typename T>
static void precond(ecsprt_hip_ctx *c, MG<T> &mg, const T *r, T *z) {
  const int pc = c->o.pc_type;
  Level<T> &L = mg.L[0];
  if (pc == ECSPRT_PC_NONE) {
    ECS_HIP(hipMemcpyAsync(z, r, sizeof(T) * L.n, hipMemcpyDeviceToDevice, c->stream));
  } else if (pc == ECSPRT_PC_JACOBI) {
    k_jacobi<T><<<gsz(L.n), 256, 0, c->stream>>>(L.s, L.dn, r, z);
    ECS_HIP(hipGetLastError());
  } else {
    mg_vcycle(c, mg, r, z);
  }
}

template <typename T>
static void gmres_create(Gmres<T> &g, int m, long n) {
  g.m = m;
  g.n = n;
  g.V.resize(m + 1);
  g.Z.resize(m);
  for (int q = 0; q <= m; ++q) ECS_HIP(hipMalloc(&g.V[q], sizeof(T) * n));
  for (int q = 0; q < m; ++q) ECS_HIP(hipMalloc(&g.Z[q], sizeof(T) * n));
  ECS_HIP(hipMalloc(&g.dV, sizeof(T *) * (m + 1)));
  ECS_HIP(hipMalloc(&g.dZ, sizeof(T *) * m));
  ECS_HIP(hipMemcpy(g.dV, g.V.data(), sizeof(T *) * (m + 1), hipMemcpyHostToDevice));
  ECS_HIP(hipMemcpy(g.dZ, g.Z.data(), sizeof(T *) * m, hipMemcpyHostToDevice));
  ECS_HIP(hipMalloc(&g.w, sizeof(T) * n));
  ECS_HIP(hipMalloc(&g.r, sizeof(T) * n));
  ECS_HIP(hipMalloc(&g.partial, sizeof(double) * RED_BLOCKS * (m + 2)));
  ECS_HIP(hipMalloc(&g.dh, sizeof(double) * 2 * (m + 2)));
  ECS_HIP(hipMalloc(&g.dy, sizeof(double) * (m + 1)));
  ECS_HIP(hipHostMalloc((void **)&g.hpin, sizeof(double) * 2 * (m + 2), hipHostMallocDefault));
  ECS_HIP(hipHostMalloc((void **)&g.ypin, sizeof(double) * (m + 1), hipHostMallocDefault));
}

template <typename T>
static void gmres_free(Gmres<T> &g) {
  for (T *v : g.V) (void)hipFree(v);
  for (T *v : g.Z) (void)hipFree(v);
  g.V.clear();
  g.Z.clear();
  (void)hipFree(g.dV);
  (void)hipFree(g.dZ);
  (void)hipFree(g.w);
  (void)hipFree(g.r);
  (void)hipFree(g.partial);
  (void)hipFree(g.dh);
  (void)hipFree(g.dy);
  if (g.hpin) (void)hipHostFree(g.hpin);
  if (g.ypin) (void)hipHostFree(g.ypin);
  g = Gmres<T>{};
}

// dots[q] = V[q] . w for q < m (into dst on device); accumulate adds to dst.
template <typename T>
static void dots(ecsprt_hip_ctx *c, Gmres<T> &g, int m, T *const *dV, const T *w, double *dst, int accumulate) {
  const int nb = (int)std::min<long>(RED_BLOCKS, (g.n + RED_THREADS - 1) / RED_THREADS);
  k_dot_multi<T><<<nb, RED_THREADS, 0, c->stream>>>(m, dV, w, g.n, g.partial);
  k_reduce_partials<<<1, 256, 0, c->stream>>>(m, nb, g.partial, dst, accumulate);
  ECS_HIP(hipGetLastError());
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_dot1(const T *__restrict__ a, const T *__restrict__ b, long n,
                                                      double *__restrict__ partial) {
  __shared__ double red[RED_THREADS];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(RED_THREADS) k_sum1(const T *__restrict__ a, long n, double *__restrict__ partial) {
  __shared__ double red[RED_THREADS];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc += (double)a[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
static void dot_single(ecsprt_hip_ctx *c, double *partial, const T *a, const T *b, long n, double *dst) {
  const int nb = (int)std::min<long>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
  k_dot1<T><<<nb, RED_THREADS, 0, c->stream>>>(a, b, n, partial);
  k_reduce_partials<<<1, 256, 0, c->stream>>>(1, nb, partial, dst, 0);
  ECS_HIP(hipGetLastError());
}

template <typename T>
static void remove_mean_dev(ecsprt_hip_ctx *c, double *partial, double *sumdst, T *v, long n) {
  const int nb = (int)std::min<long>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
  k_sum1<T><<<nb, RED_THREADS, 0, c->stream>>>(v, n, partial);
  k_reduce_partials<<<1, 256, 0, c->stream>>>(1, nb, partial, sumdst, 0);
  k_sub_scalar<T><<<gsz(n), 256, 0, c->stream>>>(v, sumdst, n);
  ECS_HIP(hipGetLastError());
}

static double fetch1(ecsprt_hip_ctx *c, const double *d, double *pin) {
  ECS_HIP(hipMemcpyAsync(pin, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
  return pin[0];
}

// Right-preconditioned GMRES(m) with classical Gram-Schmidt and one
// re-orthogonalisation; convergence on the true residual ||b - A x|| <= tol.
// Returns iterations; rnorm receives the final true residual norm.
template <typename T>
static int gmres(ecsprt_hip_ctx *c, MG<T> &mg, Gmres<T> &g, const T *b, T *x, double tol, int maxit,
                 double &rnorm, bool *converged) {
  hipStream_t st = c->stream;
  Level<T> &L0 = mg.L[0];
  const long n = L0.n;
  const int m = g.m;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1);
  int it = 0;
  *converged = false;
  double beta = 0.0;
  while (true) {
    k_residual<T><<<gsz(n), 256, 0, st>>>(L0.s, L0.dn, x, b, g.r);
    ECS_HIP(hipGetLastError());
    dot_single<T>(c, g.partial, g.r, g.r, n, g.dh);
    beta = std::sqrt(fetch1(c, g.dh, g.hpin));
    rnorm = beta;
    if (beta <= tol) {
      *converged = true;
      break;
    }
    if (!std::isfinite(beta)) break;
    if (it >= maxit) break;
    k_scale_by_inv<T><<<gsz(n), 256, 0, st>>>(g.r, g.V[0], g.dh, n, 1);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int j = 0;
    bool done = false;
    for (; j < m && it < maxit; ++j, ++it) {
      precond<T>(c, mg, g.V[j], g.Z[j]);
      k_apply<T><<<gsz(n), 256, 0, st>>>(L0.s, L0.dn, g.Z[j], g.w);
      ECS_HIP(hipGetLastError());
      // CGS2: h = V^T w; w -= V h; h2 = V^T w; w -= V h2.
      dots<T>(c, g, j + 1, g.dV, g.w, g.dh, 0);
      k_sub_multi<T><<<gsz(n), 256, 0, st>>>(j + 1, g.dV, g.dh, g.w, n);
      dots<T>(c, g, j + 1, g.dV, g.w, g.dh + (m + 2), 0);
      k_sub_multi<T><<<gsz(n), 256, 0, st>>>(j + 1, g.dV, g.dh + (m + 2), g.w, n);
      dot_single<T>(c, g.partial, g.w, g.w, n, g.dh + (j + 1));
      ECS_HIP(hipGetLastError());
      ECS_HIP(hipMemcpyAsync(g.hpin, g.dh, sizeof(double) * 2 * (m + 2), hipMemcpyDeviceToHost, st));
      ECS_HIP(hipStreamSynchronize(st));
      for (int q = 0; q <= j; ++q) H[(size_t)q * m + j] = g.hpin[q] + g.hpin[(m + 2) + q];
      const double hn = std::sqrt(std::max(g.hpin[j + 1], 0.0));
      if (!std::isfinite(hn)) {
        it = maxit;
        break;
      }
      H[(size_t)(j + 1) * m + j] = hn;
      // w/hn -> V[j+1]
      if (hn > 0) k_scale_by_inv<T><<<gsz(n), 256, 0, st>>>(g.w, g.V[j + 1], g.dh + (j + 1), n, 1);
      for (int q = 0; q < j; ++q) {
        const double a = H[(size_t)q * m + j], bb = H[(size_t)(q + 1) * m + j];
        H[(size_t)q * m + j] = cs[q] * a + sn[q] * bb;
        H[(size_t)(q + 1) * m + j] = -sn[q] * a + cs[q] * bb;
      }
      const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
      const double den = std::hypot(a, bb);
      cs[j] = den > 0 ? a / den : 1.0;
      sn[j] = den > 0 ? bb / den : 0.0;
      H[(size_t)j * m + j] = den;
      H[(size_t)(j + 1) * m + j] = 0.0;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      if (std::fabs(gv[j + 1]) <= tol || hn == 0.0) {
        ++j;
        ++it;
        done = true;
        break;
      }
    }
    // Back substitution for y (j x j) and x += Z y.
    std::vector<double> y(j, 0.0);
    for (int q = j - 1; q >= 0; --q) {
      double s = gv[q];
      for (int r = q + 1; r < j; ++r) s -= H[(size_t)q * m + r] * y[r];
      y[q] = H[(size_t)q * m + q] != 0 ? s / H[(size_t)q * m + q] : 0.0;
    }
    for (int q = 0; q < j; ++q) g.ypin[q] = y[q];
    ECS_HIP(hipMemcpyAsync(g.dy, g.ypin, sizeof(double) * std::max(j, 1), hipMemcpyHostToDevice, st));
    if (j > 0) k_add_multi<T><<<gsz(n), 256, 0, st>>>(j, g.dZ, g.dy, x, n, 1.0);
    ECS_HIP(hipGetLastError());
    ECS_HIP(hipStreamSynchronize(st));
    (void)done;
    if (it >= maxit) {
      k_residual<T><<<gsz(n), 256, 0, st>>>(L0.s, L0.dn, x, b, g.r);
      dot_single<T>(c, g.partial, g.r, g.r, n, g.dh);
      rnorm = std::sqrt(fetch1(c, g.dh, g.hpin));
      *converged = rnorm <= tol;
      break;
    }
  }
  return it;
}

// ---- entry points used by api.hip ------------------------------------------------
void solver_create(ecsprt_hip_ctx *c) {
  Solver *s = new Solver();
  c->solver = s;
  mg_create<double>(c, s->mg64);
  const int m = c->o.ksp_restart > 0 ? c->o.ksp_restart : 30;
  gmres_create<double>(s->g64, m, c->nnode);
  ECS_HIP(hipMalloc(&s->r64, sizeof(double) * c->nnode));
  ECS_HIP(hipMalloc(&s->sum, sizeof(double) * 4));
  if (c->o.solve_precision == ECSPRT_SOLVE_MIXED) {
    mg_create<float>(c, s->mg32);
    gmres_create<float>(s->g32, m, c->nnode);
    ECS_HIP(hipMalloc(&s->b32, sizeof(float) * c->nnode));
    ECS_HIP(hipMalloc(&s->x32, sizeof(float) * c->nnode));
    s->have32 = true;
  }
}

void solver_destroy(ecsprt_hip_ctx *c) {
  Solver *s = c->solver;
  if (!s) return;
  mg_free(s->mg64);
  mg_free(s->mg32);
  gmres_free(s->g64);
  gmres_free(s->g32);
  (void)hipFree(s->r64);
  (void)hipFree(s->sum);
  (void)hipFree(s->b32);
  (void)hipFree(s->x32);
  delete s;
  c->solver = nullptr;
}

void solver_build_operator(ecsprt_hip_ctx *c) {
  Solver *s = c->solver;
  mg_set_density<double>(c, s->mg64);
  if (s->have32) mg_set_density<float>(c, s->mg32);
  c->operator_built = true;
}

void solver_rhs(ecsprt_hip_ctx *c, double *b) {
  const ecsprt_params &p = c->p;
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
  k_rhs<<<gsz(c->nnode), 256, 0, c->stream>>>(k, c->d_moments, b);
  ECS_HIP(hipGetLastError());
  solver_remove_mean(c, b);
}

void solver_matmult(ecsprt_hip_ctx *c, const double *x, double *y) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  Level<double> &L = c->solver->mg64.L[0];
  k_apply<double><<<gsz(L.n), 256, 0, c->stream>>>(L.s, L.dn, x, y);
  ECS_HIP(hipGetLastError());
}

void solver_pcapply(ecsprt_hip_ctx *c, const double *r, double *z) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  precond<double>(c, c->solver->mg64, r, z);
}

void solver_remove_mean(ecsprt_hip_ctx *c, double *v) {
  remove_mean_dev<double>(c, c->solver->g64.partial, c->solver->sum, v, c->nnode);
}

// KSPSolve: A x = b.  Returns ECSPRT_OK or ECSPRT_ERR_NOT_CONVERGED.
int solver_solve(ecsprt_hip_ctx *c, const double *b, double *x) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  Solver *s = c->solver;
  const long n = c->nnode;
  const double rtol = c->o.ksp_rtol > 0 ? c->o.ksp_rtol : 1e-5;
  const double atol = c->o.ksp_atol > 0 ? c->o.ksp_atol : 1e-50;
  const int maxit = c->o.ksp_max_it > 0 ? c->o.ksp_max_it : 10000;
  if (c->o.initial_guess == 0) {
    k_fill<double><<<gsz(n), 256, 0, c->stream>>>(x, n, 0.0);
    ECS_HIP(hipGetLastError());
  }
  dot_single<double>(c, s->g64.partial, b, b, n, s->g64.dh);
  const double bnorm = std::sqrt(fetch1(c, s->g64.dh, s->g64.hpin));
  const double tol = std::max(rtol * bnorm, atol);
  c->last_bnorm = bnorm;
  double rnorm = 0.0;
  bool conv = false;
  int iters = 0;
  if (c->o.solve_precision == ECSPRT_SOLVE_MIXED && s->have32) {
    // FP64 iterative refinement around an FP32 GMRES+GMG inner solve.
    const double inner = c->o.inner_rtol > 0 ? c->o.inner_rtol : 1e-4;
    Level<double> &L0 = s->mg64.L[0];
    for (int outer = 0; outer < 100; ++outer) {
      k_residual<double><<<gsz(n), 256, 0, c->stream>>>(L0.s, L0.dn, x, b, s->r64);
      ECS_HIP(hipGetLastError());
      dot_single<double>(c, s->g64.partial, s->r64, s->r64, n, s->g64.dh);
      rnorm = std::sqrt(fetch1(c, s->g64.dh, s->g64.hpin));
      if (rnorm <= tol) {
        conv = true;
        break;
      }
      if (!std::isfinite(rnorm)) break;
      if (iters >= maxit) break;
      k_convert<double, float><<<gsz(n), 256, 0, c->stream>>>(s->r64, s->b32, n, 1.0 / rnorm);
      k_fill<float><<<gsz(n), 256, 0, c->stream>>>(s->x32, n, 0.0f);
      double rn32 = 0.0;
      bool c32 = false;
      iters += gmres<float>(c, s->mg32, s->g32, s->b32, s->x32, inner, maxit - iters, rn32, &c32);
      // x += rnorm * e, accumulated in FP64.
      k_convert<float, double><<<gsz(n), 256, 0, c->stream>>>(s->x32, s->r64, n, rnorm);
      k_axpy_from<double><<<gsz(n), 256, 0, c->stream>>>(s->r64, x, n, 1.0);
      ECS_HIP(hipGetLastError());
    }
  } else {
    iters = gmres<double>(c, s->mg64, s->g64, b, x, tol, maxit, rnorm, &conv);
  }
  solver_remove_mean(c, x);
  ECS_HIP(hipStreamSynchronize(c->stream));
  c->last_iters = iters;
  c->last_rnorm = rnorm;
  return conv ? ECSPRT_OK : ECSPRT_ERR_NOT_CONVERGED;
}

}  // namespace ecsprt
