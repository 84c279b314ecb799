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
// coarsening.  Red-black Gauss-Seidel smooths, one LDS-tiled launch per sweep.  The
// levels small enough for one workgroup's LDS (the coarse end of the V-cycle) run in
// a single launch that ends in a dense inverse of the coarsest operator.  Smoothing
// and all reductions have a fixed order, so every solve is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include <mutex>

#include "common.hpp"
#include "solver_kernels.hpp"

namespace ecsprt {

// ---- host structures ------------------------------------------------------------
// Algorithmic HBM bytes of the solve: every launch adds its compulsory reads and writes
// (each distinct array read or written once per point).
static inline void tally(ecsprt_hip_ctx *c, double bytes) { c->solve_bytes += bytes; }

static int gsz(long n) {
  long b = (n + 255) / 256;
  return (int)std::max(1L, std::min(b, 8192L));
}

// Solver workspace: from the context's arena when it has one, else hipMalloc.
template <typename P>
static void sv_alloc(ecsprt_hip_ctx *c, P **p, size_t bytes) {
  if (Arena *a = c->arena) {
    const size_t off = (a->dev_used + 255) & ~(size_t)255;
    ECS_CHECK(off + bytes <= a->dev_cap, ECSPRT_ERR_NOMEM, "solver arena exhausted");
    *p = reinterpret_cast<P *>(a->dev + off);
    a->dev_used = off + bytes;
    return;
  }
  ECS_HIP(hipMalloc(p, bytes));
}
template <typename P>
static void sv_host_alloc(ecsprt_hip_ctx *c, P **p, size_t bytes) {
  if (Arena *a = c->arena) {
    const size_t off = (a->host_used + 63) & ~(size_t)63;
    ECS_CHECK(off + bytes <= a->host_cap, ECSPRT_ERR_NOMEM, "solver host arena exhausted");
    *p = reinterpret_cast<P *>(a->host + off);
    a->host_used = off + bytes;
    return;
  }
  ECS_HIP(hipHostMalloc((void **)p, bytes, hipHostMallocDefault));
}
static void sv_free(ecsprt_hip_ctx *c, void *p) {
  if (!c->arena && p) (void)hipFree(p);
}
static void sv_host_free(ecsprt_hip_ctx *c, void *p) {
  if (!c->arena && p) (void)hipHostFree(p);
}

template <typename T>
struct Level {
  StencilConsts s;
  long n = 0;
  double dx = 1, dy = 1, dz = 1;
  CoarsenMap to_coarse{};  // map from this level to the next
  T *dn = nullptr, *x = nullptr, *x2 = nullptr, *b = nullptr, *r = nullptr;
};

template <typename T>
struct MG {
  std::vector<Level<T>> L;
  T *tmp0 = nullptr;  // second level-0 buffer of the ping-pong sweeps
  T *ainv = nullptr;  // dense inverse of the coarsest operator (+ alpha 1 1^T), or null
  int tail = 0;       // first level of the single-workgroup tail (== L.size(): none)
  size_t tail_lds = 0;
  TailArgs<T> targs{};
};

template <typename T>
struct Gmres {
  int m = 0;
  long n = 0;
  std::vector<T *> V, Z;
  T **dV = nullptr, **dZ = nullptr;
  T *w = nullptr, *r = nullptr;
  double *partial = nullptr, *dh = nullptr, *dy = nullptr, *ypin = nullptr;
  // Pinned: two slots of one Hessenberg column (2 (m+2) doubles each, alternating
  // between Arnoldi steps), then one double for a norm read back late.
  double *hpin = nullptr, *bpin = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};  // each column slot's copy done
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
    sv_alloc(c, &L.dn, sizeof(T) * L.n);
    sv_alloc(c, &L.x, sizeof(T) * L.n);
    sv_alloc(c, &L.x2, sizeof(T) * L.n);
    sv_alloc(c, &L.b, sizeof(T) * L.n);
    sv_alloc(c, &L.r, sizeof(T) * L.n);
    // Coarsening rule: among the axes that halve cleanly (even, at least 4 nodes),
    // those whose coupling is within a factor 2 of the strongest of them.
    const double a[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), k2 / (dz * dz)};
    const int N[3] = {Nx, Ny, Nz};
    bool can[3];
    double amax = 0.0;
    for (int d = 0; d < 3; ++d) {
      can[d] = N[d] % 2 == 0 && N[d] >= 4;
      if (can[d]) amax = std::max(amax, a[d]);
    }
    int cf[3] = {0, 0, 0};
    int any = 0;
    for (int d = 0; d < 3; ++d)
      if (can[d] && a[d] >= 0.5 * amax) { cf[d] = 1; any = 1; }
    if (!any || L.n <= 128) {
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
  sv_alloc(c, &mg.tmp0, sizeof(T) * mg.L[0].n);
  // The tail: the coarsest levels whose five LDS arrays fit in one workgroup, when
  // the coarsest is small enough for a dense inverse.
  const int nl = (int)mg.L.size();
  mg.tail = nl;
  if (mg.L.back().n <= DENSE_MAX) {
    size_t bytes = 0;
    int t = nl;
    while (t > 0 && nl - (t - 1) <= TAIL_MAX) {
      const size_t add = 5 * sizeof(T) * (size_t)mg.L[t - 1].n;
      if (bytes + add > TAIL_LDS_MAX) break;
      bytes += add;
      --t;
    }
    if (t < nl) {
      mg.tail = t;
      mg.tail_lds = bytes;
      TailArgs<T> &a = mg.targs;
      a.nl = nl - t;
      a.nc = (int)mg.L.back().n;
      size_t off = 0;
      for (int l = t; l < nl; ++l) {
        a.s[l - t] = mg.L[l].s;
        a.m[l - t] = mg.L[l].to_coarse;
        a.off[l - t] = (int)off;
        a.dn[l - t] = mg.L[l].dn;
        off += 5 * (size_t)mg.L[l].n;
      }
      // Once per process: slot-pool batches create solvers on several threads at once.
      static std::once_flag once;
      std::call_once(once, [] {
        ECS_HIP(hipFuncSetAttribute((const void *)k_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)TAIL_LDS_MAX));
      });
    }
  }
}

template <typename T>
static void mg_free(ecsprt_hip_ctx *c, MG<T> &mg) {
  for (auto &L : mg.L) {
    sv_free(c, L.dn);
    sv_free(c, L.x);
    sv_free(c, L.x2);
    sv_free(c, L.b);
    sv_free(c, L.r);
  }
  sv_free(c, mg.tmp0);
  sv_free(c, mg.ainv);
  mg = MG<T>{};
}

// Dense inverse of the coarsest operator plus alpha 1 1^T (Gauss-Jordan with partial
// pivoting in FP64 on the host), built whenever the operator is.
template <typename T>
static void mg_dense_coarse(ecsprt_hip_ctx *c, MG<T> &mg) {
  Level<T> &L = mg.L.back();
  const int n = (int)L.n;
  if (n > DENSE_MAX) return;
  std::vector<T> dnh(n);
  ECS_HIP(hipMemcpyAsync(dnh.data(), L.dn, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
  const std::vector<double> Inv = coarse_inverse(L.s, dnh);
  std::vector<T> h((size_t)n * n);
  for (size_t q = 0; q < h.size(); ++q) h[q] = (T)Inv[q];
  if (!mg.ainv) sv_alloc(c, &mg.ainv, sizeof(T) * h.size());
  ECS_HIP(hipMemcpyAsync(mg.ainv, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
  mg.targs.ainv = mg.ainv;
}

// Level-0 density from the moments, restricted down the hierarchy.
template <typename T>
static void mg_set_density(ecsprt_hip_ctx *c, MG<T> &mg) {
  hipStream_t st = c->stream;
  k_density<T><<<gsz(mg.L[0].n), 256, 0, st>>>(c->d_moments, mg.L[0].dn, mg.L[0].n);
  ECS_HIP(hipGetLastError());
  for (size_t l = 0; l + 1 < mg.L.size(); ++l) {
    const Launch3 g3 = launch3(mg.L[l + 1].s.Nx, mg.L[l + 1].s.Ny, mg.L[l + 1].s.Nz);
    k_restrict<T><<<g3.grid, g3.block, 0, st>>>(mg.L[l].to_coarse, mg.L[l].dn, mg.L[l + 1].dn);
    ECS_HIP(hipGetLastError());
  }
  mg_dense_coarse(c, mg);
}

template <typename T>
static void remove_mean_dev(ecsprt_hip_ctx *c, double *partial, double *sumdst, T *v, long n);

// The marching sweep loads each node about 2.1 times at 128^3 (the box sweep 3.75
// times) but runs its planes in sequence; it is faster only where the level has
// enough points to keep every CU streaming (measured: 33 against 43 us at 2^21 FP64
// points, slower from 2^20 points down).
template <typename T>
static bool marching_level(const Level<T> &L, int mode) {
  return L.s.Nz != 1 && (mode == 2 || (mode == 0 && (long)L.s.Nx * L.s.Ny * L.s.Nz >= SZ_MIN_POINTS));
}

// xc (with its map pm) adds the prolonged coarse correction to xin as the sweep loads
// it; single-plane levels only (in the box sweep, whose halo is 3.75 times its tile,
// the repeated interpolation cost more than the pass it saved: 4.53 against 4.37 ms
// per 128^3 solve; on single planes 2.14 against 2.22 ms per 1024^2 solve).
template <typename T>
static void launch_sweep(hipStream_t st, const Level<T> &L, const T *xin, const T *b, T *xout, int first, int mode,
                         const T *xc = nullptr, const CoarsenMap &pm = CoarsenMap{}) {
  if (L.s.Nz == 1) {
    const int nt = ((L.s.Nx + S2_TX - 1) / S2_TX) * ((L.s.Ny + S2_TY - 1) / S2_TY);
    k_sweep2d<T, S2_TX, S2_TY><<<nt, S2_TX * S2_TY, 0, st>>>(L.s, L.dn, xin, b, xout, first, xc, pm);
  } else if (marching_level(L, mode)) {
    const int kz = sweep_kz(L.s.Nx, L.s.Ny, L.s.Nz);
    const int nt = ((L.s.Nx + SZ_TX - 1) / SZ_TX) * ((L.s.Ny + SZ_TY - 1) / SZ_TY) * ((L.s.Nz + kz - 1) / kz);
    k_sweep_z<T, SZ_TX, SZ_TY><<<nt, SZ_TX * SZ_TY, 0, st>>>(L.s, L.dn, xin, b, xout, first, kz);
  } else {
    const int nt = ((L.s.Nx + SW_TX - 1) / SW_TX) * ((L.s.Ny + SW_TY - 1) / SW_TY) * ((L.s.Nz + SW_TZ - 1) / SW_TZ);
    k_sweep<T, SW_TX, SW_TY, SW_TZ><<<nt, 512, 0, st>>>(L.s, L.dn, xin, b, xout, first);
  }
}

// One V-cycle: z = M^-1 r (r, z on level 0).  Each level ping-pongs between two
// buffers, starting on the one that leaves the last post-smoothing sweep in the
// level's home buffer (z on level 0).
template <typename T>
static void mg_vcycle(ecsprt_hip_ctx *c, MG<T> &mg, const T *r, T *z) {
  hipStream_t st = c->stream;
  const int pre = c->o.mg_pre > 0 ? c->o.mg_pre : 2;
  const int post = c->o.mg_post > 0 ? c->o.mg_post : 2;
  const int csw = c->o.mg_coarse_sweeps > 0 ? c->o.mg_coarse_sweeps : 50;
  const int nl = (int)mg.L.size();
  const int lt = mg.tail;
  std::vector<T *> cur(nl, nullptr);
  auto home = [&](int l) { return l == 0 ? z : mg.L[l].x; };
  auto tmp = [&](int l) { return l == 0 ? mg.tmp0 : mg.L[l].x2; };
  auto rhs = [&](int l) { return l == 0 ? r : (const T *)mg.L[l].b; };
  for (int l = 0; l < std::min(lt, nl); ++l) {
    Level<T> &L = mg.L[l];
    const int total = l == nl - 1 ? csw : pre + post;
    const int nsw = l == nl - 1 ? csw : pre;
    const T *in = nullptr;
    for (int k = 1; k <= nsw; ++k) {
      T *out = ((total - k) % 2 == 0) ? home(l) : tmp(l);
      launch_sweep(st, L, in, rhs(l), out, 0, c->o.mg_sweep);
      tally(c, (double)sizeof(T) * L.n * (in ? 4 : 3));
      in = out;
    }
    cur[l] = const_cast<T *>(in);
    if (l == nl - 1) {  // coarsest level without a tail: smoothed, then made mean-free
      double *partial = std::is_same<T, double>::value ? c->solver->g64.partial : c->solver->g32.partial;
      remove_mean_dev<T>(c, partial, c->solver->sum, cur[l], L.n);
      break;
    }
    const Launch3 gf = launch3(L.s.Nx, L.s.Ny, L.s.Nz);
    const Launch3 gc = launch3(mg.L[l + 1].s.Nx, mg.L[l + 1].s.Ny, mg.L[l + 1].s.Nz);
    k_residual<T><<<gf.grid, gf.block, 0, st>>>(L.s, L.dn, in, rhs(l), L.r);
    k_restrict<T><<<gc.grid, gc.block, 0, st>>>(L.to_coarse, L.r, mg.L[l + 1].b);
    ECS_HIP(hipGetLastError());
    tally(c, (double)sizeof(T) * (5 * L.n + mg.L[l + 1].n));
  }
  if (lt < nl) {
    TailArgs<T> a = mg.targs;
    a.pre = pre;
    a.post = post;
    a.bin = rhs(lt);
    a.xout = home(lt);
    k_tail<T><<<1, TAIL_THREADS, mg.tail_lds, st>>>(a);
    ECS_HIP(hipGetLastError());
    cur[lt] = a.xout;
    double tb = 2.0 * mg.L[lt].n + (double)mg.L.back().n * mg.L.back().n;
    for (int l = lt; l < nl; ++l) tb += (double)mg.L[l].n;
    tally(c, sizeof(T) * tb);
  }
  for (int l = std::min(lt, nl - 1) - 1; l >= 0; --l) {
    Level<T> &L = mg.L[l];
    const T *in = cur[l];
    const int total = pre + post;
    int k = pre + 1;
    if (post > 0 && L.s.Nz == 1) {
      // the first post-smoothing sweep adds the coarse correction as it loads x
      T *out = ((total - k) % 2 == 0) ? home(l) : tmp(l);
      launch_sweep(st, L, in, rhs(l), out, 1, c->o.mg_sweep, (const T *)cur[l + 1], L.to_coarse);
      tally(c, (double)sizeof(T) * (4 * L.n + mg.L[l + 1].n));
      in = out;
      ++k;
    } else {
      const Launch3 gf = launch3(L.s.Nx, L.s.Ny, L.s.Nz);
      k_prolong_add<T><<<gf.grid, gf.block, 0, st>>>(L.to_coarse, cur[l + 1], cur[l]);
      tally(c, (double)sizeof(T) * (2 * L.n + mg.L[l + 1].n));
    }
    for (; k <= total; ++k) {
      T *out = ((total - k) % 2 == 0) ? home(l) : tmp(l);
      launch_sweep(st, L, in, rhs(l), out, 1, c->o.mg_sweep);
      tally(c, (double)sizeof(T) * 4 * L.n);
      in = out;
    }
    ECS_HIP(hipGetLastError());
  }
}

template <typename T>
static void precond(ecsprt_hip_ctx *c, MG<T> &mg, const T *r, T *z) {
  const int pc = c->o.pc_type;
  Level<T> &L = mg.L[0];
  if (pc == ECSPRT_PC_NONE) {
    ECS_HIP(hipMemcpyAsync(z, r, sizeof(T) * L.n, hipMemcpyDeviceToDevice, c->stream));
    tally(c, 2.0 * sizeof(T) * L.n);
  } else if (pc == ECSPRT_PC_JACOBI) {
    const Launch3 g3 = launch3(L.s.Nx, L.s.Ny, L.s.Nz);
    k_jacobi<T><<<g3.grid, g3.block, 0, c->stream>>>(L.s, L.dn, r, z);
    ECS_HIP(hipGetLastError());
    tally(c, 3.0 * sizeof(T) * L.n);
  } else {
    mg_vcycle(c, mg, r, z);
  }
}

// GMRES(m) workspace.  m = 0 keeps only the reduction buffers (the FP64 side of the
// mixed-precision solve needs no Krylov basis).
template <typename T>
static void gmres_create(ecsprt_hip_ctx *c, Gmres<T> &g, int m, long n) {
  g.m = m;
  g.n = n;
  if (m > 0) {
    g.V.resize(m + 1);
    g.Z.resize(m);
    for (int q = 0; q <= m; ++q) sv_alloc(c, &g.V[q], sizeof(T) * n);
    for (int q = 0; q < m; ++q) sv_alloc(c, &g.Z[q], sizeof(T) * n);
    sv_alloc(c, &g.dV, sizeof(T *) * (m + 1));
    sv_alloc(c, &g.dZ, sizeof(T *) * m);
    ECS_HIP(hipMemcpyAsync(g.dV, g.V.data(), sizeof(T *) * (m + 1), hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipMemcpyAsync(g.dZ, g.Z.data(), sizeof(T *) * m, hipMemcpyHostToDevice, c->stream));
    ECS_HIP(hipStreamSynchronize(c->stream));
    sv_alloc(c, &g.w, sizeof(T) * n);
    sv_alloc(c, &g.r, sizeof(T) * n);
  }
  sv_alloc(c, &g.partial, sizeof(double) * RED_BLOCKS * (m + 2));
  sv_alloc(c, &g.dh, sizeof(double) * 2 * (m + 2));
  sv_alloc(c, &g.dy, sizeof(double) * (m + 1));
  sv_host_alloc(c, &g.hpin, sizeof(double) * (4 * (m + 2) + 1));
  g.bpin = g.hpin + 4 * (m + 2);
  sv_host_alloc(c, &g.ypin, sizeof(double) * (m + 1));
  if (m > 0)
    for (hipEvent_t &e : g.ev) ECS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

template <typename T>
static void gmres_free(ecsprt_hip_ctx *c, Gmres<T> &g) {
  for (T *v : g.V) sv_free(c, v);
  for (T *v : g.Z) sv_free(c, v);
  sv_free(c, g.dV);
  sv_free(c, g.dZ);
  sv_free(c, g.w);
  sv_free(c, g.r);
  sv_free(c, g.partial);
  sv_free(c, g.dh);
  sv_free(c, g.dy);
  sv_host_free(c, g.hpin);
  sv_host_free(c, g.ypin);
  for (hipEvent_t e : g.ev)
    if (e) (void)hipEventDestroy(e);
  g = Gmres<T>{};
}

// dots[q] = V[q] . w for q < m (into dst on device); accumulate adds to dst.
template <typename T>
static void dots(ecsprt_hip_ctx *c, Gmres<T> &g, int m, T *const *dV, const T *w, double *dst, int accumulate) {
  const int nb = (int)std::min<long>(RED_BLOCKS, (g.n + RED_THREADS - 1) / RED_THREADS);
  k_dot_multi<T><<<nb, RED_THREADS, 0, c->stream>>>(m, dV, w, g.n, g.partial);
  k_reduce_partials<<<1, 256, 0, c->stream>>>(m, nb, g.partial, dst, accumulate);
  tally(c, (double)sizeof(T) * (m + 1) * g.n);
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
  tally(c, (double)sizeof(T) * (a == b ? 1 : 2) * n);
  ECS_HIP(hipGetLastError());
}

template <typename T>
static void remove_mean_dev(ecsprt_hip_ctx *c, double *partial, double *sumdst, T *v, long n) {
  const int nb = (int)std::min<long>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
  k_sum1<T><<<nb, RED_THREADS, 0, c->stream>>>(v, n, partial);
  k_reduce_partials<<<1, 256, 0, c->stream>>>(1, nb, partial, sumdst, 0);
  k_sub_scalar<T><<<gsz(n), 256, 0, c->stream>>>(v, sumdst, n);
  tally(c, 3.0 * sizeof(T) * n);
  ECS_HIP(hipGetLastError());
}

static double fetch1(ecsprt_hip_ctx *c, const double *d, double *pin) {
  ECS_HIP(hipMemcpyAsync(pin, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
  return pin[0];
}

// Right-preconditioned GMRES(m) with classical Gram-Schmidt (optionally one
// re-orthogonalisation); convergence on the true residual ||b - A x|| <= tol.
// Returns iterations; rnorm receives the final true residual norm.  single_cycle: one
// restart cycle stopped by the Arnoldi residual estimate (the inner solve of the
// mixed-precision refinement, whose FP32 true residual has a rounding floor).
// x_zero: x is zero on entry, so the first residual is b itself; beta0 >= 0 is then
// ||b||, with b.b already in g.dh[0], else ||b|| is read back with the first column.
//
// The Arnoldi steps are pipelined one deep: step j+1 is queued before the host waits
// for step j's Hessenberg column whenever the residual estimates so far say that step
// j cannot reach tol (the last estimate times the fastest reduction seen in this
// cycle, or 0.1 before one is seen, still 10 times above tol).  A step queued this way
// after all is never used: the solution only takes the columns the host processed, so
// the iterates do not depend on the guess.
template <typename T>
static int gmres(ecsprt_hip_ctx *c, MG<T> &mg, Gmres<T> &g, const T *b, T *x, double tol, int maxit,
                 double &rnorm, bool *converged, bool single_cycle = false, bool x_zero = false,
                 double beta0 = -1.0) {
  hipStream_t st = c->stream;
  Level<T> &L0 = mg.L[0];
  const long n = L0.n;
  const int m = g.m;
  const bool refine = c->o.gmres_refine != 0;
  const Launch3 g0 = launch3(L0.s.Nx, L0.s.Ny, L0.s.Nz);
  const int nb = (int)std::min<long>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
  const size_t col = 2 * (size_t)(m + 2);  // doubles of one column slot
  constexpr double SPEC_MARGIN = 10.0, SPEC_RATIO0 = 0.1;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gv(m + 1);
  // Arnoldi step q: Z[q] = M^-1 V[q], w = A Z[q], w orthogonalised against V[0..q]
  // with its column in g.dh, the column copied to slot q & 1 (event ev[q & 1]), then
  // V[q+1] = w / |w| on the device (zero when |w| = 0, and then never used).
  auto enqueue = [&](int q) {
    precond<T>(c, mg, g.V[q], g.Z[q]);
    k_apply<T><<<g0.grid, g0.block, 0, st>>>(L0.s, L0.dn, g.Z[q], g.w);
    ECS_HIP(hipGetLastError());
    tally(c, 3.0 * sizeof(T) * n);
    // Classical Gram-Schmidt, h = V^T w; w -= V h; |w|^2 (PETSc's default), or with
    // gmres_refine CGS2: h2 = V^T w; w -= V h2 before the norm.  Up to 8 basis vectors
    // the subtractions carry the following products in the same pass.
    dots<T>(c, g, q + 1, g.dV, g.w, g.dh, 0);
    if (!refine) {
      if (q + 1 <= 8) {
        k_sub_dot_multi<T><<<nb, RED_THREADS, 0, st>>>(q + 1, g.dV, g.dh, g.w, n, g.partial, 1);
        k_reduce_partials<<<1, 256, 0, st>>>(1, nb, g.partial, g.dh + (q + 1), 0);
        tally(c, (double)sizeof(T) * (q + 3) * n);
      } else {
        k_sub_multi<T><<<gsz(n), 256, 0, st>>>(q + 1, g.dV, g.dh, g.w, n);
        tally(c, (double)sizeof(T) * (q + 3) * n);
        dot_single<T>(c, g.partial, g.w, g.w, n, g.dh + (q + 1));
      }
    } else if (q + 1 <= 8) {
      k_sub_dot_multi<T><<<nb, RED_THREADS, 0, st>>>(q + 1, g.dV, g.dh, g.w, n, g.partial, 0);
      k_reduce_partials<<<1, 256, 0, st>>>(q + 1, nb, g.partial, g.dh + (m + 2), 0);
      k_sub_dot_multi<T><<<nb, RED_THREADS, 0, st>>>(q + 1, g.dV, g.dh + (m + 2), g.w, n, g.partial, 1);
      k_reduce_partials<<<1, 256, 0, st>>>(1, nb, g.partial, g.dh + (q + 1), 0);
      tally(c, 2.0 * sizeof(T) * (q + 3) * n);
    } else {
      k_sub_multi<T><<<gsz(n), 256, 0, st>>>(q + 1, g.dV, g.dh, g.w, n);
      dots<T>(c, g, q + 1, g.dV, g.w, g.dh + (m + 2), 0);
      k_sub_multi<T><<<gsz(n), 256, 0, st>>>(q + 1, g.dV, g.dh + (m + 2), g.w, n);
      tally(c, 2.0 * sizeof(T) * (q + 3) * n);
      dot_single<T>(c, g.partial, g.w, g.w, n, g.dh + (q + 1));
    }
    ECS_HIP(hipGetLastError());
    ECS_HIP(hipMemcpyAsync(g.hpin + (q & 1) * col, g.dh, sizeof(double) * col, hipMemcpyDeviceToHost, st));
    ECS_HIP(hipEventRecord(g.ev[q & 1], st));
    k_scale_by_inv<T><<<gsz(n), 256, 0, st>>>(g.w, g.V[q + 1], g.dh + (q + 1), n, 1);
    ECS_HIP(hipGetLastError());
    tally(c, 2.0 * sizeof(T) * n);
  };
  int it = 0;
  *converged = false;
  double beta = 0.0;
  bool first = true;
  while (true) {
    const T *r0 = g.r;
    bool beta_late = false;
    if (first && x_zero) {
      r0 = b;
      if (beta0 >= 0) {
        beta = beta0;
      } else {
        dot_single<T>(c, g.partial, b, b, n, g.dh);
        ECS_HIP(hipMemcpyAsync(g.bpin, g.dh, sizeof(double), hipMemcpyDeviceToHost, st));
        beta_late = true;
      }
    } else {
      k_residual<T><<<g0.grid, g0.block, 0, st>>>(L0.s, L0.dn, x, b, g.r);
      ECS_HIP(hipGetLastError());
      tally(c, 4.0 * sizeof(T) * n);
      dot_single<T>(c, g.partial, g.r, g.r, n, g.dh);
      beta = std::sqrt(fetch1(c, g.dh, g.hpin));
    }
    first = false;
    if (!beta_late) {
      rnorm = beta;
      if (beta <= tol) {
        *converged = true;
        break;
      }
      if (!std::isfinite(beta)) break;
    }
    if (it >= maxit) break;
    k_scale_by_inv<T><<<gsz(n), 256, 0, st>>>(r0, g.V[0], g.dh, n, 1);
    tally(c, 2.0 * sizeof(T) * n);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int j = 0, queued = 1;
    bool done = false, stop = false;
    double ratio = SPEC_RATIO0;  // fastest reduction of the residual estimate per step
    enqueue(0);
    for (; j < m && it < maxit; ++j, ++it) {
      const bool more = j + 1 < m && it + 1 < maxit;
      // the residual estimate before step j (unknown, so far above tol, while the
      // norm of b is still on its way)
      const double est = beta_late ? HUGE_VAL : std::fabs(gv[j]);
      if (more && queued == j + 1 && est * ratio > SPEC_MARGIN * tol) {
        enqueue(j + 1);
        ++queued;
      }
      ECS_HIP(hipEventSynchronize(g.ev[j & 1]));
      if (beta_late) {  // step 0: the norm of b arrived before its column
        beta_late = false;
        beta = std::sqrt(g.bpin[0]);
        rnorm = beta;
        gv[0] = beta;
        if (beta <= tol) {
          *converged = true;
          stop = true;
          break;
        }
        if (!std::isfinite(beta)) {
          stop = true;
          break;
        }
      }
      const double *hp = g.hpin + (j & 1) * col;
      for (int q = 0; q <= j; ++q) H[(size_t)q * m + j] = hp[q] + (refine ? hp[(m + 2) + q] : 0.0);
      const double hn = std::sqrt(std::max(hp[j + 1], 0.0));
      if (!std::isfinite(hn)) {
        it = maxit;
        break;
      }
      H[(size_t)(j + 1) * m + j] = hn;
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
      ratio = std::min(ratio, std::fabs(sn[j]));  // the estimate fell by |sn[j]| in step j
      if (more && queued == j + 1) {
        enqueue(j + 1);
        ++queued;
      }
    }
    if (stop) break;
    // Back substitution for y (j x j) and x += Z y.
    std::vector<double> y(j, 0.0);
    for (int q = j - 1; q >= 0; --q) {
      double s = gv[q];
      for (int r = q + 1; r < j; ++r) s -= H[(size_t)q * m + r] * y[r];
      y[q] = H[(size_t)q * m + q] != 0 ? s / H[(size_t)q * m + q] : 0.0;
    }
    for (int q = 0; q < j; ++q) g.ypin[q] = y[q];
    if (j > 0) {
      // ypin is next written after a later synchronisation of this stream
      ECS_HIP(hipMemcpyAsync(g.dy, g.ypin, sizeof(double) * j, hipMemcpyHostToDevice, st));
      k_add_multi<T><<<gsz(n), 256, 0, st>>>(j, g.dZ, g.dy, x, n, 1.0);
      tally(c, (double)sizeof(T) * (j + 2) * n);
    }
    ECS_HIP(hipGetLastError());
    if (single_cycle) {
      rnorm = std::fabs(gv[j]);
      *converged = done;
      break;
    }
    if (it >= maxit) {
      k_residual<T><<<g0.grid, g0.block, 0, st>>>(L0.s, L0.dn, x, b, g.r);
      tally(c, 4.0 * sizeof(T) * n);
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
  const bool mixed = c->o.solve_precision == ECSPRT_SOLVE_MIXED;
  gmres_create<double>(c, s->g64, mixed ? 0 : m, c->nnode);
  sv_alloc(c, &s->r64, sizeof(double) * c->nnode);
  sv_alloc(c, &s->sum, sizeof(double) * 4);
  if (mixed) {
    mg_create<float>(c, s->mg32);
    gmres_create<float>(c, s->g32, m, c->nnode);
    sv_alloc(c, &s->b32, sizeof(float) * c->nnode);
    sv_alloc(c, &s->x32, sizeof(float) * c->nnode);
    s->have32 = true;
  }
}

void solver_destroy(ecsprt_hip_ctx *c) {
  Solver *s = c->solver;
  if (!s) return;
  mg_free(c, s->mg64);
  mg_free(c, s->mg32);
  gmres_free(c, s->g64);
  gmres_free(c, s->g32);
  sv_free(c, s->r64);
  sv_free(c, s->sum);
  sv_free(c, s->b32);
  sv_free(c, s->x32);
  delete s;
  c->solver = nullptr;
}

void solver_build_operator(ecsprt_hip_ctx *c) {
  Solver *s = c->solver;
  mg_set_density<double>(c, s->mg64);
  if (s->have32) mg_set_density<float>(c, s->mg32);
  c->operator_built = true;
}

// The FP64 level-0 density of the operator as last built.
const double *solver_density(const ecsprt_hip_ctx *c) { return c->solver->mg64.L[0].dn; }

void solver_rhs(ecsprt_hip_ctx *c, double *b) {
  const ecsprt_params &p = c->p;
  const RhsConsts k = rhs_consts(p);
  const Launch3 g3 = launch3(p.Nx, p.Ny, p.Nz);
  if (c->rhs_sinusoidal) {
    k_rhs_sinusoidal<<<g3.grid, g3.block, 0, c->stream>>>(p.Nx, p.Ny, p.dx, p.dy, p.dz, p.n0, p.scale, b);
    tally(c, 8.0 * c->nnode);
  } else {
    k_rhs<<<g3.grid, g3.block, 0, c->stream>>>(k, c->d_moments, b);
    tally(c, 40.0 * c->nnode);
  }
  ECS_HIP(hipGetLastError());
  solver_remove_mean(c, b);
}

void solver_matmult(ecsprt_hip_ctx *c, const double *x, double *y) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  Level<double> &L = c->solver->mg64.L[0];
  const Launch3 g3 = launch3(L.s.Nx, L.s.Ny, L.s.Nz);
  k_apply<double><<<g3.grid, g3.block, 0, c->stream>>>(L.s, L.dn, x, y);
  ECS_HIP(hipGetLastError());
}

void solver_pcapply(ecsprt_hip_ctx *c, const double *r, double *z) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  precond<double>(c, c->solver->mg64, r, z);
}

void solver_remove_mean(ecsprt_hip_ctx *c, double *v) {
  remove_mean_dev<double>(c, c->solver->g64.partial, c->solver->sum, v, c->nnode);
}

// Upper bound of the solver workspace of an n-node grid (batch arenas are sized with
// it): level buffers sum to at most 2n nodes per array, GMRES keeps 2m+3 vectors.
size_t solver_arena_bytes(const ecsprt_options &o, long n) {
  const int m = o.ksp_restart > 0 ? o.ksp_restart : 30;
  const bool mixed = o.solve_precision == ECSPRT_SOLVE_MIXED;
  auto mg = [&](size_t t) { return t * (11 * (size_t)n + (size_t)DENSE_MAX * DENSE_MAX); };
  auto gm = [&](size_t t, int mm) {
    return (mm > 0 ? t * (size_t)(2 * mm + 3) * n + 16 * (size_t)(2 * mm + 1) : 0) + 8 * (size_t)RED_BLOCKS * (mm + 2) +
           8 * (size_t)(3 * mm + 5);
  };
  size_t b = mg(8) + gm(8, mixed ? 0 : m) + 8 * (size_t)n + 32 + 8 * (size_t)n;  // + r64, sum, rhs
  if (mixed) b += mg(4) + gm(4, m) + 8 * (size_t)n;
  return b + 256 * 512;  // alignment of every buffer
}

// One standalone solve on a context without particles: operator from c->d_moments,
// the RHS of ComputeFullRHS3D, solution into c->d_phi.  The solver is built and
// released around the solve (from the arena when the context has one).
int solver_standalone(ecsprt_hip_ctx *c) {
  c->solve_bytes = 0.0;
  double *rhs = nullptr;
  sv_alloc(c, &rhs, sizeof(double) * c->nnode);
  solver_create(c);
  int rc = ECSPRT_OK;
  try {
    solver_build_operator(c);
    solver_rhs(c, rhs);
    rc = solver_solve(c, rhs, c->d_phi);
  } catch (...) {
    solver_destroy(c);
    sv_free(c, rhs);
    throw;
  }
  solver_destroy(c);
  sv_free(c, rhs);
  return rc;
}

// KSPSolve: A x = b.  Returns ECSPRT_OK or ECSPRT_ERR_NOT_CONVERGED.
int solver_solve(ecsprt_hip_ctx *c, const double *b, double *x) {
  ECS_CHECK(c->operator_built, ECSPRT_ERR_STATE, "operator not built (call ecsprt_hip_lhs)");
  Solver *s = c->solver;
  const long n = c->nnode;
  const double rtol = c->o.ksp_rtol > 0 ? c->o.ksp_rtol : 1e-5;
  const double atol = c->o.ksp_atol > 0 ? c->o.ksp_atol : 1e-50;
  const int maxit = c->o.ksp_max_it > 0 ? c->o.ksp_max_it : 10000;
  const bool x_zero = c->o.initial_guess == 0;
  if (x_zero) {
    k_fill<double><<<gsz(n), 256, 0, c->stream>>>(x, n, 0.0);
    ECS_HIP(hipGetLastError());
    tally(c, 8.0 * n);
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
    // Inner tolerance: the outer target split evenly over the fewest refinement cycles
    // that each ask for at most 1e-5 (FP32 resolves that comfortably): 1e-5 twice for
    // rtol 1e-10, 1e-4 twice for 1e-8.
    double inner = c->o.inner_rtol;
    if (inner <= 0) {
      const double lr = -std::log10(std::max(rtol, 1e-15));
      const int cycles = std::max(1, (int)std::ceil(lr / 5.0 - 1e-9));
      inner = std::pow(10.0, -lr / cycles);
    }
    Level<double> &L0 = s->mg64.L[0];
    for (int outer = 0; outer < 100; ++outer) {
      const double *r = s->r64;
      if (outer == 0 && x_zero) {
        r = b;  // x is zero: the residual is b itself
        rnorm = bnorm;
      } else {
        const Launch3 g3 = launch3(L0.s.Nx, L0.s.Ny, L0.s.Nz);
        k_residual<double><<<g3.grid, g3.block, 0, c->stream>>>(L0.s, L0.dn, x, b, s->r64);
        ECS_HIP(hipGetLastError());
        tally(c, 32.0 * n);
        dot_single<double>(c, s->g64.partial, s->r64, s->r64, n, s->g64.dh);
        rnorm = std::sqrt(fetch1(c, s->g64.dh, s->g64.hpin));
      }
      if (rnorm <= tol) {
        conv = true;
        break;
      }
      if (!std::isfinite(rnorm)) break;
      if (iters >= maxit) break;
      k_convert<double, float><<<gsz(n), 256, 0, c->stream>>>(r, s->b32, n, 1.0 / rnorm);
      k_fill<float><<<gsz(n), 256, 0, c->stream>>>(s->x32, n, 0.0f);
      tally(c, 16.0 * n);
      double rn32 = 0.0;
      bool c32 = false;
      iters += gmres<float>(c, s->mg32, s->g32, s->b32, s->x32, inner, maxit - iters, rn32, &c32, true, true);
      // x += rnorm * e, accumulated in FP64.
      k_convert<float, double><<<gsz(n), 256, 0, c->stream>>>(s->x32, s->r64, n, rnorm);
      k_axpy_from<double><<<gsz(n), 256, 0, c->stream>>>(s->r64, x, n, 1.0);
      tally(c, 36.0 * n);
      ECS_HIP(hipGetLastError());
    }
  } else {
    iters = gmres<double>(c, s->mg64, s->g64, b, x, tol, maxit, rnorm, &conv, false, x_zero, bnorm);
  }
  solver_remove_mean(c, x);
  ECS_HIP(hipStreamSynchronize(c->stream));
  c->last_iters = iters;
  c->last_rnorm = rnorm;
  return conv ? ECSPRT_OK : ECSPRT_ERR_NOT_CONVERGED;
}

}  // namespace ecsprt
