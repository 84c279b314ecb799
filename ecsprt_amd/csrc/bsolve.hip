// Batched standalone potential solves with one launch per phase for the whole batch.
//
// Every problem of the batch is the solve of src/solver.c:107-203 on its own grid: the
// operator of ComputeFullLHS3D (src/lhs.c:337-505) from its density, the forcing of
// ComputeFullRHS3D (src/rhs.c:259-373) with the mean removed, right-preconditioned
// GMRES(m) with one multigrid V-cycle per step (the method of solver.hip), in FP64 or as
// FP32 GMRES/multigrid inside FP64 iterative refinement.  Instead of one host thread and
// stream per problem, one host thread drives all problems in lockstep: every kernel of
// the V-cycle, the operator and the Krylov updates is launched once for all problems
// that take part, each launch carrying a table of segments (one problem's level, its
// pointers and the first block of its share); a block finds its segment by binary search
// and skips it when its problem has converged.  The Krylov scalars of all problems cross
// to the host in one copy per Arnoldi step.  The launch tables of the V-cycle are built
// once per batch, so a solve allocates nothing.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "solver_kernels.hpp"

namespace ecsprt {

constexpr int BT = 256;          // threads of the node-wise and vector kernels
constexpr int BVEC_PER_BLOCK = 8;  // elements per thread and block pass of the vector kernels

// One problem's share of a V-cycle launch: its level's stencil and transfer map, its
// buffers (pointers into the batch arena; the level-0 right-hand side and home buffer
// advance by one vector per Krylov index j, the *_js strides), the marching chunk, the
// problem and the first block.
template <typename T>
struct BSeg {
  StencilConsts s;
  CoarsenMap m;
  const T *dn;
  const T *in;
  const T *b;
  T *out;
  long in_js, b_js, out_js;
  int kz, prob, start, nodes;
};

// One problem's level-0 vectors for the Krylov kernels.
template <typename T>
struct BVec {
  StencilConsts s;
  const T *dn;
  long n;
  T *V, *Z, *w, *r, *x;
  const T *b;
  double *partial, *dh, *dy;
  int prob, start, nb;
};

template <typename S>
__device__ inline int seg_index(const S *sg, int nseg) {
  int lo = 0, hi = nseg - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sg[mid].start <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int TX, int TY, int TZ>
__global__ void __launch_bounds__(512) kb_sweep_box(const BSeg<T> *sg, int nseg, int j, const int *act, int first) {
  const BSeg<T> g = sg[seg_index(sg, nseg)];
  if (!act[g.prob]) return;
  sweep_box<T, TX, TY, TZ>(g.s, g.dn, g.in ? g.in + j * g.in_js : nullptr, g.b + j * g.b_js, g.out + j * g.out_js,
                           first, blockIdx.x - g.start);
}

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) kb_sweep_z(const BSeg<T> *sg, int nseg, int j, const int *act, int first) {
  const BSeg<T> g = sg[seg_index(sg, nseg)];
  if (!act[g.prob]) return;
  sweep_z<T, TX, TY>(g.s, g.dn, g.in ? g.in + j * g.in_js : nullptr, g.b + j * g.b_js, g.out + j * g.out_js, first,
                     g.kz, blockIdx.x - g.start);
}

// r = b - A x on the segment's level.
template <typename T>
__global__ void __launch_bounds__(BT) kb_residual(const BSeg<T> *sg, int nseg, int j, const int *act) {
  const BSeg<T> g = sg[seg_index(sg, nseg)];
  if (!act[g.prob]) return;
  const int q = (blockIdx.x - g.start) * BT + threadIdx.x;
  if (q >= g.nodes) return;
  const int i = q % g.s.Nx, jj = (q / g.s.Nx) % g.s.Ny, k = q / (g.s.Nx * g.s.Ny);
  const T *x = g.in + j * g.in_js, *b = g.b + j * g.b_js;
  T c[11];
  row_coeffs(g.dn, g.s, i, jj, k, c);
  g.out[j * g.out_js + q] = b[q] - (offdiag(x, g.s, i, jj, k, c) + c[10] * x[q]);
}

// Coarse node q of the segment: out = restriction of in.
template <typename T>
__global__ void __launch_bounds__(BT) kb_restrict(const BSeg<T> *sg, int nseg, const int *act) {
  const BSeg<T> g = sg[seg_index(sg, nseg)];
  if (!act[g.prob]) return;
  const int q = (blockIdx.x - g.start) * BT + threadIdx.x;
  if (q >= g.nodes) return;
  const CoarsenMap &m = g.m;
  g.out[q] = restrict_point(m, g.in, q % m.Nxc, (q / m.Nxc) % m.Nyc, q / (m.Nxc * m.Nyc));
}

// Fine node q of the segment: out += prolongation of in.
template <typename T>
__global__ void __launch_bounds__(BT) kb_prolong_add(const BSeg<T> *sg, int nseg, int j, const int *act) {
  const BSeg<T> g = sg[seg_index(sg, nseg)];
  if (!act[g.prob]) return;
  const int q = (blockIdx.x - g.start) * BT + threadIdx.x;
  if (q >= g.nodes) return;
  const CoarsenMap &m = g.m;
  g.out[j * g.out_js + q] += prolong_point(m, g.in, q % m.Nxf, (q / m.Nxf) % m.Nyf, q / (m.Nxf * m.Nyf));
}

// One workgroup per problem.  A problem whose tail starts at level 0 (bin == null) takes
// the V-cycle's right-hand side V[j] and writes Z[j].
template <typename T>
__global__ void __launch_bounds__(TAIL_THREADS) kb_tail(const TailArgs<T> *ta, const BVec<T> *vs, int j,
                                                        const int *act) {
  if (!act[blockIdx.x]) return;
  TailArgs<T> a = ta[blockIdx.x];
  if (!a.bin) {
    const BVec<T> &v = vs[blockIdx.x];
    a.bin = v.V + j * v.n;
    a.xout = v.Z + j * v.n;
  }
  tail_body(a);
}

// ---- Krylov kernels: each problem's blocks stride over its n elements -------------
#define BVEC_PROLOGUE                                   \
  const BVec<T> v = vs[seg_index(vs, nseg)];            \
  if (!act[v.prob]) return;                             \
  const long i0 = (long)(blockIdx.x - v.start) * BT + threadIdx.x, istep = (long)v.nb * BT;

// w = A Z[j]
template <typename T>
__global__ void __launch_bounds__(BT) kb_apply(const BVec<T> *vs, int nseg, int j, const int *act) {
  BVEC_PROLOGUE
  const T *x = v.Z + j * v.n;
  for (long q = i0; q < v.n; q += istep) {
    const int i = (int)(q % v.s.Nx), jj = (int)((q / v.s.Nx) % v.s.Ny), k = (int)(q / ((long)v.s.Nx * v.s.Ny));
    T c[11];
    row_coeffs(v.dn, v.s, i, jj, k, c);
    v.w[q] = offdiag(x, v.s, i, jj, k, c) + c[10] * x[q];
  }
}

// r = b - A x
template <typename T>
__global__ void __launch_bounds__(BT) kb_residual0(const BVec<T> *vs, int nseg, const int *act) {
  BVEC_PROLOGUE
  for (long q = i0; q < v.n; q += istep) {
    const int i = (int)(q % v.s.Nx), jj = (int)((q / v.s.Nx) % v.s.Ny), k = (int)(q / ((long)v.s.Nx * v.s.Ny));
    T c[11];
    row_coeffs(v.dn, v.s, i, jj, k, c);
    v.r[q] = v.b[q] - (offdiag(v.x, v.s, i, jj, k, c) + c[10] * v.x[q]);
  }
}

__device__ inline void block_sum_store(double a, double *dst) {
  __shared__ double red[BT];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int o = BT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst = red[0];
  __syncthreads();
}

// partial[b * m + q] over the block's stride: which = 0: V[q] . w for q < m; 1: r . r;
// 2: w . w (m = 1 for the last two).
template <typename T>
__global__ void __launch_bounds__(BT) kb_dots(const BVec<T> *vs, int nseg, const int *act, int m, int which) {
  BVEC_PROLOGUE
  const T *w = which == 1 ? v.r : v.w;
  const int blk = blockIdx.x - v.start;
  for (int q0 = 0; q0 < m; q0 += 8) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int mq = min(8, m - q0);
    for (long q = i0; q < v.n; q += istep) {
      const double wi = (double)w[q];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < mq) acc[u] += (double)(which ? w[q] : v.V[(q0 + u) * v.n + q]) * wi;
    }
    for (int u = 0; u < mq; ++u) block_sum_store(acc[u], &v.partial[(long)blk * m + q0 + u]);
  }
}

// w -= sum_q h[q] V[q] (h = dh[hoff..]), then partial sums of w.w (norm) or of V[q].w.
template <typename T>
__global__ void __launch_bounds__(BT) kb_sub_dots(const BVec<T> *vs, int nseg, const int *act, int m, int hoff,
                                                  int norm) {
  BVEC_PROLOGUE
  const int blk = blockIdx.x - v.start;
  double hq[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) hq[u] = u < m ? v.dh[hoff + u] : 0.0;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long q = i0; q < v.n; q += istep) {
    double a = (double)v.w[q], vv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < m) {
        vv[u] = (double)v.V[u * v.n + q];
        a -= hq[u] * vv[u];
      }
    const T wt = (T)a;
    v.w[q] = wt;
    const double wn = (double)wt;
    if (norm) {
      acc[0] += wn * wn;
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < m) acc[u] += vv[u] * wn;
    }
  }
  const int mq = norm ? 1 : m;
  for (int u = 0; u < mq; ++u) block_sum_store(acc[u], &v.partial[(long)blk * mq + u]);
}

// w -= sum_q h[q] V[q] for any m (h = dh[hoff..]).
template <typename T>
__global__ void __launch_bounds__(BT) kb_sub(const BVec<T> *vs, int nseg, const int *act, int m, int hoff) {
  BVEC_PROLOGUE
  for (long q = i0; q < v.n; q += istep) {
    double a = (double)v.w[q];
    for (int u = 0; u < m; ++u) a -= v.dh[hoff + u] * (double)v.V[u * v.n + q];
    v.w[q] = (T)a;
  }
}

// dh[off + q] = sum of the problem's nb partials of entry q (fixed order), q < m.
template <typename T>
__global__ void __launch_bounds__(BT) kb_reduce(const BVec<T> *vs, int nprob, const int *act, int m, int off) {
  const BVec<T> v = vs[blockIdx.x];
  if (!act[v.prob]) return;
  for (int q = 0; q < m; ++q) {
    double a = 0.0;
    for (int b = threadIdx.x; b < v.nb; b += BT) a += v.partial[(long)b * m + q];
    block_sum_store(a, &v.dh[off + q]);
  }
}

// V[j+1] = src / sqrt(dh[k]) (src = w, or r for the first vector).
template <typename T>
__global__ void __launch_bounds__(BT) kb_scale(const BVec<T> *vs, int nseg, const int *act, int j, int k,
                                               int from_r) {
  BVEC_PROLOGUE
  const double d = sqrt(v.dh[k]);
  const double inv = d > 0 ? 1.0 / d : 0.0;
  const T *src = from_r ? v.r : v.w;
  T *dst = v.V + (long)j * v.n;
  for (long q = i0; q < v.n; q += istep) dst[q] = (T)((double)src[q] * inv);
}

// x += sum_{q < count[prob]} dy[q] Z[q]
template <typename T>
__global__ void __launch_bounds__(BT) kb_add(const BVec<T> *vs, int nseg, const int *act, const int *count) {
  BVEC_PROLOGUE
  const int m = count[v.prob];
  for (long q = i0; q < v.n; q += istep) {
    double a = 0.0;
    for (int u = 0; u < m; ++u) a += v.dy[u] * (double)v.Z[u * v.n + q];
    v.x[q] = (T)((double)v.x[q] + a);
  }
}

template <typename T>
__global__ void __launch_bounds__(BT) kb_fill_x(const BVec<T> *vs, int nseg, const int *act) {
  BVEC_PROLOGUE
  for (long q = i0; q < v.n; q += istep) v.x[q] = T(0);
}

// Mixed refinement: FP64 residual r64 -> FP32 right-hand side b32 = r64 / ||r64||
// (scale[prob] = 1/||r64||), and x64 += ||r64|| x32 (scale[prob] = ||r64||).
__global__ void __launch_bounds__(BT) kb_to32(const BVec<double> *v64, const BVec<float> *v32, int nseg,
                                              const int *act, const double *scale) {
  const int si = seg_index(v64, nseg);
  const BVec<double> a = v64[si];
  if (!act[a.prob]) return;
  const BVec<float> f = v32[si];
  const long i0 = (long)(blockIdx.x - a.start) * BT + threadIdx.x, istep = (long)a.nb * BT;
  const double sc = scale[a.prob];
  float *b32 = const_cast<float *>(f.b);
  for (long q = i0; q < a.n; q += istep) {
    b32[q] = (float)(a.r[q] * sc);
    f.x[q] = 0.0f;
  }
}

__global__ void __launch_bounds__(BT) kb_from32(const BVec<double> *v64, const BVec<float> *v32, int nseg,
                                                const int *act, const double *scale) {
  const int si = seg_index(v64, nseg);
  const BVec<double> a = v64[si];
  if (!act[a.prob]) return;
  const BVec<float> f = v32[si];
  const long i0 = (long)(blockIdx.x - a.start) * BT + threadIdx.x, istep = (long)a.nb * BT;
  const double sc = scale[a.prob];
  for (long q = i0; q < a.n; q += istep) a.x[q] = a.x[q] + sc * (double)f.x[q];
}

// Mean removal (MatNullSpaceRemove, src/rhs.c:364-366) in three launches: per-block sums
// into partial, their fixed-order sum into dh[0] (kb_reduce), then the subtraction; of the
// solution x (rhs = 0) or of the forcing b (rhs = 1).
__global__ void __launch_bounds__(BT) kb_sum(const BVec<double> *vs, int nseg, const int *act, int rhs) {
  const BVec<double> v = vs[seg_index(vs, nseg)];
  if (!act[v.prob]) return;
  const long i0 = (long)(blockIdx.x - v.start) * BT + threadIdx.x, istep = (long)v.nb * BT;
  const double *x = rhs ? v.b : v.x;
  double a = 0.0;
  for (long q = i0; q < v.n; q += istep) a += x[q];
  block_sum_store(a, &v.partial[blockIdx.x - v.start]);
}

__global__ void __launch_bounds__(BT) kb_sub_mean(const BVec<double> *vs, int nseg, const int *act, int rhs) {
  const BVec<double> v = vs[seg_index(vs, nseg)];
  if (!act[v.prob]) return;
  const long i0 = (long)(blockIdx.x - v.start) * BT + threadIdx.x, istep = (long)v.nb * BT;
  const double mean = v.dh[0] / (double)v.n;
  double *x = rhs ? const_cast<double *>(v.b) : v.x;
  for (long q = i0; q < v.n; q += istep) x[q] -= mean;
}

// Density of each problem's moments into its level-0 density; the forcing into b64.
// With dens set instead (the frozen operators of an ensemble), that density is copied.
struct BProbIn {
  const double *moments;
  const double *dens;
  double *rhs;
  RhsConsts rc;
  float *dn32;
  double *dn64;
  long n;
  int start;
};

__global__ void __launch_bounds__(BT) kb_density_rhs(const BProbIn *pin, int nseg) {
  const BProbIn p = pin[seg_index(pin, nseg)];
  const long q = (long)(blockIdx.x - p.start) * BT + threadIdx.x;
  if (q >= p.n) return;
  const double d = p.moments[4 * q];
  if (p.dn64) p.dn64[q] = d;
  if (p.dn32) p.dn32[q] = (float)d;
  const int Nx = p.rc.Nx, Ny = p.rc.Ny;
  const int i = (int)(q % Nx), j = (int)((q / Nx) % Ny), k = (int)(q / ((long)Nx * Ny));
  p.rhs[q] = rhs_point(p.rc, p.moments, i, j, k);
}

__global__ void __launch_bounds__(BT) kb_density_copy(const BProbIn *pin, int nseg) {
  const BProbIn p = pin[seg_index(pin, nseg)];
  const long q = (long)(blockIdx.x - p.start) * BT + threadIdx.x;
  if (!p.dens || q >= p.n) return;
  const double d = p.dens[q];
  if (p.dn64) p.dn64[q] = d;
  if (p.dn32) p.dn32[q] = (float)d;
}

}  // namespace ecsprt

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
namespace ecsprt {

namespace {

int bnb(long n) { return (int)std::max(1L, std::min(1024L, (n + (long)BT * BVEC_PER_BLOCK - 1) / ((long)BT * BVEC_PER_BLOCK))); }
int bnodes(long n) { return (int)((n + BT - 1) / BT); }

// Multigrid level of one problem in precision T.
template <typename T>
struct BLevel {
  StencilConsts s;
  CoarsenMap m{};
  long n = 0;
  T *dn = nullptr, *x = nullptr, *x2 = nullptr, *b = nullptr, *r = nullptr;
};

// One problem's multigrid hierarchy and Krylov vectors in precision T.
template <typename T>
struct BProb {
  std::vector<BLevel<T>> L;
  int tail = 0;
  size_t tail_lds = 0;
  T *tmp0 = nullptr, *ainv = nullptr;
  T *V = nullptr, *Z = nullptr, *w = nullptr;
  double *partial = nullptr, *dh = nullptr, *dy = nullptr;
};

// A launch of the V-cycle program: its kernel, segment table and block count, the
// problem of each segment and the segment's compulsory bytes.
enum BKind { BK_SWEEP_BOX, BK_SWEEP_Z, BK_RESIDUAL, BK_RESTRICT, BK_PROLONG, BK_TAIL };
struct BLaunch {
  BKind kind;
  int first = 0;
  void *segs = nullptr;
  int nseg = 0, blocks = 0;
  std::vector<int> prob;
  std::vector<double> bytes;
};

// Bump allocator over one device block.
struct BArena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  template <typename P>
  P *take(size_t count) {
    const size_t off = (used + 255) & ~(size_t)255;
    if (base) {
      ECS_CHECK(off + sizeof(P) * count <= cap, ECSPRT_ERR_NOMEM, "batch arena exhausted");
      used = off + sizeof(P) * count;
      return reinterpret_cast<P *>(base + off);
    }
    used = off + sizeof(P) * count;  // sizing pass
    return nullptr;
  }
};

}  // namespace

// The whole batch in lockstep: per precision the problems' hierarchies, the V-cycle
// program, the Krylov segment tables; the FP64 level-0 data of every problem.
struct LockstepBatch {
  std::vector<ecsprt_params> pp;  // per problem (grid and physics)
  ecsprt_options o;
  hipStream_t st = nullptr;
  int P = 0, m = 30, pre = 2, post = 2;
  bool mixed = false;
  std::vector<int> dims;
  BArena arena;
  // FP64 level 0 of every problem: density (when the cycles are FP32), forcing, residual
  std::vector<double *> dn64, rhs64, r64;
  std::vector<long> n;
  BProb<double> *p64 = nullptr;  // per problem when the cycles are FP64
  BProb<float> *p32 = nullptr;   // per problem when they are FP32
  std::vector<BLaunch> prog;     // the V-cycle
  std::vector<BLaunch> dens;     // restriction of the densities down the hierarchies
  void *vsegs = nullptr;         // BVec<T> per problem (the cycle precision)
  BVec<double> *vsegs64 = nullptr;  // the FP64 outer vectors (mixed) or the cycle's (FP64)
  BVec<float> *vsegs32 = nullptr;
  BProbIn *pin = nullptr;
  void *tails = nullptr;
  int vblocks = 0, nblocks_in = 0;
  size_t tail_lds = 0;
  int *d_act = nullptr, *d_count = nullptr, *h_count = nullptr;
  // The flags are copied from a ring of pinned rows: a row is rewritten only after the
  // copies of the rows before it have run (every Arnoldi step synchronises, and a step
  // sets the flags at most four times).
  static constexpr int ACT_RING = 32;
  int *h_act_ring = nullptr, act_slot = 0;
  std::vector<char> act_now;
  double *h_dh = nullptr, *h_dy = nullptr, *d_scale = nullptr, *h_scale = nullptr;
  double *dh_all = nullptr, *dy_all = nullptr;  // device: P x DH, P x (m + 1)
  int DH = 0;
  double bytes = 0.0;  // algorithmic bytes of the last solve
  std::vector<int> iters;
  std::vector<double> rnorm, bnorm;
  std::vector<double *> phi_ptr;
  std::vector<const double *> mom_ptr;
  ~LockstepBatch();
};

namespace {

template <typename T>
BProb<T> *&probs(LockstepBatch &B);
template <>
BProb<double> *&probs<double>(LockstepBatch &B) {
  return B.p64;
}
template <>
BProb<float> *&probs<float>(LockstepBatch &B) {
  return B.p32;
}

// Hierarchy of problem q (the coarsening and tail rules of the single-problem solver).
template <typename T>
void build_levels(LockstepBatch &B, int q, BProb<T> &P) {
  const ecsprt_params &p = B.pp[q];
  int Nx = B.dims[3 * q], Ny = B.dims[3 * q + 1], Nz = B.dims[3 * q + 2];
  double dx = p.dx, dy = p.dy, dz = p.dz;
  const double k2 = 1 + p.kappa_e * p.kappa_e;
  P.L.clear();
  for (int lev = 0; lev < 16; ++lev) {
    BLevel<T> L;
    L.s = stencil_consts(p, Nx, Ny, Nz, dx, dy, dz);
    L.n = (long)Nx * Ny * Nz;
    const double a[3] = {1.0 / (dx * dx), 1.0 / (dy * dy), k2 / (dz * dz)};
    const int N[3] = {Nx, Ny, Nz};
    bool can[3];
    double amax = 0.0;
    for (int d = 0; d < 3; ++d) {
      can[d] = N[d] % 2 == 0 && N[d] >= 4;
      if (can[d]) amax = std::max(amax, a[d]);
    }
    int cf[3] = {0, 0, 0}, any = 0;
    for (int d = 0; d < 3; ++d)
      if (can[d] && a[d] >= 0.5 * amax) {
        cf[d] = 1;
        any = 1;
      }
    if (!any || L.n <= 128) {
      P.L.push_back(L);
      break;
    }
    L.m = CoarsenMap{Nx, Ny, Nz, cf[0] ? Nx / 2 : Nx, cf[1] ? Ny / 2 : Ny, cf[2] ? Nz / 2 : Nz, cf[0], cf[1], cf[2]};
    P.L.push_back(L);
    Nx = L.m.Nxc;
    Ny = L.m.Nyc;
    Nz = L.m.Nzc;
    if (cf[0]) dx *= 2;
    if (cf[1]) dy *= 2;
    if (cf[2]) dz *= 2;
  }
  const int nl = (int)P.L.size();
  ECS_CHECK(P.L.back().n <= DENSE_MAX, ECSPRT_ERR_ARG,
            "lockstep batch: the coarsest level of every problem must have at most 1024 nodes");
  size_t bytes = 0;
  int t = nl;
  while (t > 0 && nl - (t - 1) <= TAIL_MAX) {
    const size_t add = 5 * sizeof(T) * (size_t)P.L[t - 1].n;
    if (bytes + add > TAIL_LDS_MAX) break;
    bytes += add;
    --t;
  }
  ECS_CHECK(t < nl, ECSPRT_ERR_ARG, "lockstep batch: a problem's coarse levels do not fit the tail");
  P.tail = t;
  P.tail_lds = bytes;
}

template <typename T>
void carve(LockstepBatch &B, BProb<T> &P, long n0) {
  BArena &A = B.arena;
  for (auto &L : P.L) {
    L.dn = A.take<T>(L.n);
    L.x = A.take<T>(L.n);
    L.x2 = A.take<T>(L.n);
    L.b = A.take<T>(L.n);
    L.r = A.take<T>(L.n);
  }
  P.tmp0 = A.take<T>(n0);
  P.ainv = A.take<T>((size_t)P.L.back().n * P.L.back().n);
  P.V = A.take<T>((size_t)(B.m + 1) * n0);
  P.Z = A.take<T>((size_t)B.m * n0);
  P.w = A.take<T>(n0);
  P.partial = A.take<double>((size_t)bnb(n0) * (B.m + 2));
  P.dh = nullptr;  // rows of dh_all
  P.dy = nullptr;
}

template <typename S>
void *upload(LockstepBatch &B, const std::vector<S> &v) {
  S *d = B.arena.take<S>(std::max<size_t>(v.size(), 1));
  if (d && !v.empty()) ECS_HIP(hipMemcpy(d, v.data(), sizeof(S) * v.size(), hipMemcpyHostToDevice));
  return d;
}

// The V-cycle program of precision T for every problem (see mg_vcycle in solver.hip: the
// same sweeps, buffers and order, level by level across the batch).
template <typename T>
void build_program(LockstepBatch &B) {
  BProb<T> *PP = probs<T>(B);
  B.prog.clear();
  const int total = B.pre + B.post;
  int maxl = 0;
  for (int q = 0; q < B.P; ++q) maxl = std::max(maxl, PP[q].tail);
  struct Buf {
    T *p;
    long js;
  };
  auto home = [&](int q, int l) { return l == 0 ? Buf{PP[q].Z, B.n[q]} : Buf{PP[q].L[l].x, 0}; };
  auto tmp = [&](int q, int l) { return l == 0 ? Buf{PP[q].tmp0, 0} : Buf{PP[q].L[l].x2, 0}; };
  auto rhs = [&](int q, int l) { return l == 0 ? Buf{PP[q].V, B.n[q]} : Buf{PP[q].L[l].b, 0}; };
  auto sel = [&](int q, int l, int k) { return ((total - k) % 2 == 0) ? home(q, l) : tmp(q, l); };
  auto sweeps = [&](int l, int k, int first) {
    std::vector<BSeg<T>> box, mz;
    BLaunch lb{BK_SWEEP_BOX}, lz{BK_SWEEP_Z};
    lb.first = lz.first = first;
    int bb = 0, bz = 0;
    for (int q = 0; q < B.P; ++q) {
      if (l >= PP[q].tail) continue;
      const BLevel<T> &L = PP[q].L[l];
      const Buf in = k == 1 ? Buf{nullptr, 0} : sel(q, l, k - 1), b = rhs(q, l), out = sel(q, l, k);
      BSeg<T> g{};
      g.s = L.s;
      g.dn = L.dn;
      g.in = in.p;
      g.in_js = in.js;
      g.b = b.p;
      g.b_js = b.js;
      g.out = out.p;
      g.out_js = out.js;
      g.prob = q;
      const double by = (double)sizeof(T) * L.n * (k == 1 ? 3 : 4);
      if (L.n >= SZ_MIN_POINTS_BATCH && L.s.Nz >= 8) {
        g.kz = sweep_kz(L.s.Nx, L.s.Ny, L.s.Nz);
        g.start = bz;
        bz += ((L.s.Nx + SZ_TX - 1) / SZ_TX) * ((L.s.Ny + SZ_TY - 1) / SZ_TY) * ((L.s.Nz + g.kz - 1) / g.kz);
        mz.push_back(g);
        lz.prob.push_back(q);
        lz.bytes.push_back(by);
      } else {
        g.start = bb;
        bb += ((L.s.Nx + SW_TX - 1) / SW_TX) * ((L.s.Ny + SW_TY - 1) / SW_TY) * ((L.s.Nz + SW_TZ - 1) / SW_TZ);
        box.push_back(g);
        lb.prob.push_back(q);
        lb.bytes.push_back(by);
      }
    }
    if (!box.empty()) {
      lb.segs = upload(B, box);
      lb.nseg = (int)box.size();
      lb.blocks = bb;
      B.prog.push_back(lb);
    }
    if (!mz.empty()) {
      lz.segs = upload(B, mz);
      lz.nseg = (int)mz.size();
      lz.blocks = bz;
      B.prog.push_back(lz);
    }
  };
  for (int l = 0; l < maxl; ++l) {
    for (int k = 1; k <= B.pre; ++k) sweeps(l, k, 0);
    std::vector<BSeg<T>> rs, rr;
    BLaunch lr{BK_RESIDUAL}, lt{BK_RESTRICT};
    int br = 0, bt = 0;
    for (int q = 0; q < B.P; ++q) {
      if (l >= PP[q].tail) continue;
      const BLevel<T> &L = PP[q].L[l], &C = PP[q].L[l + 1];
      const Buf in = sel(q, l, B.pre), b = rhs(q, l);
      BSeg<T> g{};
      g.s = L.s;
      g.dn = L.dn;
      g.in = in.p;
      g.in_js = in.js;
      g.b = b.p;
      g.b_js = b.js;
      g.out = L.r;
      g.prob = q;
      g.nodes = (int)L.n;
      g.start = br;
      br += bnodes(L.n);
      rs.push_back(g);
      lr.prob.push_back(q);
      lr.bytes.push_back(4.0 * sizeof(T) * L.n);
      BSeg<T> h{};
      h.m = L.m;
      h.in = L.r;
      h.out = C.b;
      h.prob = q;
      h.nodes = (int)C.n;
      h.start = bt;
      bt += bnodes(C.n);
      rr.push_back(h);
      lt.prob.push_back(q);
      lt.bytes.push_back((double)sizeof(T) * (L.n + C.n));
    }
    lr.segs = upload(B, rs);
    lr.nseg = (int)rs.size();
    lr.blocks = br;
    B.prog.push_back(lr);
    lt.segs = upload(B, rr);
    lt.nseg = (int)rr.size();
    lt.blocks = bt;
    B.prog.push_back(lt);
  }
  {
    BLaunch ta{BK_TAIL};
    ta.blocks = B.P;
    for (int q = 0; q < B.P; ++q) {
      ta.prob.push_back(q);
      double by = 0;
      for (int l = PP[q].tail; l < (int)PP[q].L.size(); ++l) by += (double)sizeof(T) * PP[q].L[l].n;
      by += 2.0 * sizeof(T) * PP[q].L[PP[q].tail].n;
      ta.bytes.push_back(by);
    }
    B.prog.push_back(ta);
  }
  for (int l = maxl - 1; l >= 0; --l) {
    std::vector<BSeg<T>> ps;
    BLaunch lp{BK_PROLONG};
    int bp = 0;
    for (int q = 0; q < B.P; ++q) {
      if (l >= PP[q].tail) continue;
      const BLevel<T> &L = PP[q].L[l], &C = PP[q].L[l + 1];
      const Buf out = sel(q, l, B.pre);
      BSeg<T> g{};
      g.m = L.m;
      g.in = C.x;  // the coarse level's home buffer (the tail's output when l + 1 is the tail)
      g.out = out.p;
      g.out_js = out.js;
      g.prob = q;
      g.nodes = (int)L.n;
      g.start = bp;
      bp += bnodes(L.n);
      ps.push_back(g);
      lp.prob.push_back(q);
      lp.bytes.push_back((double)sizeof(T) * (2 * L.n + C.n));
    }
    lp.segs = upload(B, ps);
    lp.nseg = (int)ps.size();
    lp.blocks = bp;
    B.prog.push_back(lp);
    for (int k = B.pre + 1; k <= total; ++k) sweeps(l, k, 1);
  }
}

// Density restriction down every hierarchy (the operator set-up of each solve).
template <typename T>
void build_density(LockstepBatch &B) {
  BProb<T> *PP = probs<T>(B);
  B.dens.clear();
  int maxl = 0;
  for (int q = 0; q < B.P; ++q) maxl = std::max(maxl, (int)PP[q].L.size() - 1);
  for (int l = 0; l < maxl; ++l) {
    std::vector<BSeg<T>> rr;
    BLaunch lt{BK_RESTRICT};
    int bt = 0;
    for (int q = 0; q < B.P; ++q) {
      if (l + 1 >= (int)PP[q].L.size()) continue;
      const BLevel<T> &L = PP[q].L[l], &C = PP[q].L[l + 1];
      BSeg<T> h{};
      h.m = L.m;
      h.in = L.dn;
      h.out = C.dn;
      h.prob = q;
      h.nodes = (int)C.n;
      h.start = bt;
      bt += bnodes(C.n);
      rr.push_back(h);
      lt.prob.push_back(q);
      lt.bytes.push_back(0.0);
    }
    lt.segs = upload(B, rr);
    lt.nseg = (int)rr.size();
    lt.blocks = bt;
    B.dens.push_back(lt);
  }
}

void set_act(LockstepBatch &B, const std::vector<char> &a) {
  int *row = B.h_act_ring + (size_t)B.act_slot * B.P;
  B.act_slot = (B.act_slot + 1) % LockstepBatch::ACT_RING;
  for (int q = 0; q < B.P; ++q) row[q] = a[q] ? 1 : 0;
  B.act_now.assign(a.begin(), a.end());
  ECS_HIP(hipMemcpyAsync(B.d_act, row, sizeof(int) * B.P, hipMemcpyHostToDevice, B.st));
}

double active_bytes(const LockstepBatch &B, const BLaunch &L) {
  double s = 0;
  for (size_t i = 0; i < L.prob.size(); ++i)
    if (B.act_now[L.prob[i]]) s += L.bytes[i];
  return s;
}

template <typename T>
BVec<T> *cycle_vecs(LockstepBatch &B);
template <>
BVec<double> *cycle_vecs<double>(LockstepBatch &B) {
  return B.vsegs64;
}
template <>
BVec<float> *cycle_vecs<float>(LockstepBatch &B) {
  return B.vsegs32;
}

template <typename T>
void run_launch(LockstepBatch &B, const BLaunch &L, int j) {
  if (L.blocks == 0) return;
  const BSeg<T> *sg = static_cast<const BSeg<T> *>(L.segs);
  switch (L.kind) {
    case BK_SWEEP_BOX:
      kb_sweep_box<T, SW_TX, SW_TY, SW_TZ><<<L.blocks, 512, 0, B.st>>>(sg, L.nseg, j, B.d_act, L.first);
      break;
    case BK_SWEEP_Z:
      kb_sweep_z<T, SZ_TX, SZ_TY><<<L.blocks, SZ_TX * SZ_TY, 0, B.st>>>(sg, L.nseg, j, B.d_act, L.first);
      break;
    case BK_RESIDUAL: kb_residual<T><<<L.blocks, BT, 0, B.st>>>(sg, L.nseg, j, B.d_act); break;
    case BK_RESTRICT: kb_restrict<T><<<L.blocks, BT, 0, B.st>>>(sg, L.nseg, B.d_act); break;
    case BK_PROLONG: kb_prolong_add<T><<<L.blocks, BT, 0, B.st>>>(sg, L.nseg, j, B.d_act); break;
    case BK_TAIL:
      kb_tail<T><<<B.P, TAIL_THREADS, B.tail_lds, B.st>>>(static_cast<const TailArgs<T> *>(B.tails),
                                                          cycle_vecs<T>(B), j, B.d_act);
      break;
  }
  ECS_HIP(hipGetLastError());
  B.bytes += active_bytes(B, L);
}

// Vector bytes of the active problems: `per` values of T per point.
template <typename T>
double vec_bytes(const LockstepBatch &B, double per) {
  double s = 0;
  for (int q = 0; q < B.P; ++q)
    if (B.act_now[q]) s += per * sizeof(T) * B.n[q];
  return s;
}

// One GMRES cycle (at most m Arnoldi steps) in precision T for the problems with run[q],
// from the residual in their BVec r (x is updated by x += Z y): right preconditioning by
// the V-cycle program, classical Gram-Schmidt, Givens rotations on the host.  tol[q]
// stops a problem's Arnoldi process when its residual estimate falls to it; steps[q]
// receives its steps, est[q] the estimate, beta[q] the initial norm.
template <typename T>
void gmres_cycle(LockstepBatch &B, const std::vector<char> &run, const std::vector<double> &tol,
                 std::vector<int> &steps, std::vector<double> &est, std::vector<double> &beta) {
  const int P = B.P, m = B.m, DH = B.DH;
  BVec<T> *vs = cycle_vecs<T>(B);
  const int nv = P, vb = B.vblocks;
  std::vector<char> act(run);
  set_act(B, act);
  // beta = ||r||, V0 = r / beta
  kb_dots<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, 1, 1);
  kb_reduce<T><<<P, BT, 0, B.st>>>(vs, P, B.d_act, 1, 0);
  B.bytes += vec_bytes<T>(B, 1);
  ECS_HIP(hipMemcpyAsync(B.h_dh, B.dh_all, sizeof(double) * P * DH, hipMemcpyDeviceToHost, B.st));
  ECS_HIP(hipStreamSynchronize(B.st));
  std::vector<std::vector<double>> H(P), g(P), cs(P), sn(P);
  std::vector<char> done(P, 1);
  for (int q = 0; q < P; ++q) {
    steps[q] = 0;
    beta[q] = 0.0;
    if (!run[q]) continue;
    beta[q] = std::sqrt(std::max(B.h_dh[(size_t)q * DH], 0.0));
    est[q] = beta[q];
    if (!(beta[q] > 0.0) || !std::isfinite(beta[q])) {
      act[q] = 0;
      continue;
    }
    done[q] = 0;
    H[q].assign((size_t)(m + 1) * m, 0.0);
    g[q].assign(m + 1, 0.0);
    g[q][0] = beta[q];
    cs[q].assign(m, 0.0);
    sn[q].assign(m, 0.0);
  }
  set_act(B, act);
  kb_scale<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, 0, 0, 1);
  B.bytes += vec_bytes<T>(B, 2);
  for (int j = 0; j < m; ++j) {
    bool any = false;
    for (int q = 0; q < P; ++q) {
      act[q] = run[q] && !done[q];
      any = any || act[q];
    }
    if (!any) break;
    set_act(B, act);
    for (const BLaunch &L : B.prog) run_launch<T>(B, L, j);
    kb_apply<T><<<vb, BT, 0, B.st>>>(vs, nv, j, B.d_act);
    B.bytes += vec_bytes<T>(B, 3);
    // classical Gram-Schmidt: h = V^T w, then w -= V h with its norm
    kb_dots<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, j + 1, 0);
    kb_reduce<T><<<P, BT, 0, B.st>>>(vs, P, B.d_act, j + 1, 0);
    B.bytes += vec_bytes<T>(B, j + 2);
    if (j + 1 <= 8) {
      kb_sub_dots<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, j + 1, 0, 1);
    } else {
      kb_sub<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, j + 1, 0);
      kb_dots<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, 1, 2);
      B.bytes += vec_bytes<T>(B, 1);
    }
    kb_reduce<T><<<P, BT, 0, B.st>>>(vs, P, B.d_act, 1, j + 1);
    B.bytes += vec_bytes<T>(B, j + 3);
    ECS_HIP(hipGetLastError());
    ECS_HIP(hipMemcpyAsync(B.h_dh, B.dh_all, sizeof(double) * P * DH, hipMemcpyDeviceToHost, B.st));
    ECS_HIP(hipStreamSynchronize(B.st));
    for (int q = 0; q < P; ++q) {
      if (!act[q]) continue;
      const double *dh = B.h_dh + (size_t)q * DH;
      std::vector<double> &Hq = H[q];
      for (int u = 0; u <= j; ++u) Hq[(size_t)u * m + j] = dh[u];
      const double hn = std::sqrt(std::max(dh[j + 1], 0.0));
      steps[q] = j + 1;
      if (!std::isfinite(hn)) {
        done[q] = 1;
        continue;
      }
      Hq[(size_t)(j + 1) * m + j] = hn;
      for (int u = 0; u < j; ++u) {
        const double a = Hq[(size_t)u * m + j], bb = Hq[(size_t)(u + 1) * m + j];
        Hq[(size_t)u * m + j] = cs[q][u] * a + sn[q][u] * bb;
        Hq[(size_t)(u + 1) * m + j] = -sn[q][u] * a + cs[q][u] * bb;
      }
      const double a = Hq[(size_t)j * m + j], bb = Hq[(size_t)(j + 1) * m + j];
      const double den = std::hypot(a, bb);
      cs[q][j] = den > 0 ? a / den : 1.0;
      sn[q][j] = den > 0 ? bb / den : 0.0;
      Hq[(size_t)j * m + j] = den;
      Hq[(size_t)(j + 1) * m + j] = 0.0;
      g[q][j + 1] = -sn[q][j] * g[q][j];
      g[q][j] = cs[q][j] * g[q][j];
      est[q] = std::fabs(g[q][j + 1]);
      if (est[q] <= tol[q] || hn == 0.0) done[q] = 1;
    }
    // V[j+1] = w / ||w|| for the problems that go on
    if (j + 1 < m) {
      bool more = false;
      for (int q = 0; q < P; ++q) {
        act[q] = run[q] && !done[q];
        more = more || act[q];
      }
      if (more) {
        set_act(B, act);
        kb_scale<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, j + 1, j + 1, 0);
        B.bytes += vec_bytes<T>(B, 2);
      }
    }
  }
  // y from the triangular system, x += Z y
  for (int q = 0; q < P; ++q) {
    double *y = B.h_dy + (size_t)q * (m + 1);
    B.h_count[q] = 0;
    if (!run[q] || steps[q] == 0 || H[q].empty()) continue;
    const int jq = steps[q];
    for (int u = jq - 1; u >= 0; --u) {
      double s = g[q][u];
      for (int r = u + 1; r < jq; ++r) s -= H[q][(size_t)u * m + r] * y[r];
      y[u] = H[q][(size_t)u * m + u] != 0 ? s / H[q][(size_t)u * m + u] : 0.0;
    }
    B.h_count[q] = jq;
  }
  for (int q = 0; q < P; ++q) act[q] = run[q] && B.h_count[q] > 0;
  ECS_HIP(hipMemcpyAsync(B.dy_all, B.h_dy, sizeof(double) * P * (m + 1), hipMemcpyHostToDevice, B.st));
  ECS_HIP(hipMemcpyAsync(B.d_count, B.h_count, sizeof(int) * P, hipMemcpyHostToDevice, B.st));
  set_act(B, act);
  kb_add<T><<<vb, BT, 0, B.st>>>(vs, nv, B.d_act, B.d_count);
  ECS_HIP(hipGetLastError());
  for (int q = 0; q < P; ++q)
    if (act[q]) B.bytes += (double)sizeof(T) * B.n[q] * (B.h_count[q] + 2);
  ECS_HIP(hipStreamSynchronize(B.st));
}

// ||r64|| of the FP64 residual r64 = b - A x of every running problem into norms.
void outer_residual(LockstepBatch &B, const std::vector<char> &run, std::vector<double> &norms) {
  set_act(B, run);
  kb_residual0<double><<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, B.P, B.d_act);
  kb_dots<double><<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, B.P, B.d_act, 1, 1);
  kb_reduce<double><<<B.P, BT, 0, B.st>>>(B.vsegs64, B.P, B.d_act, 1, 0);
  ECS_HIP(hipGetLastError());
  B.bytes += vec_bytes<double>(B, 4) + vec_bytes<double>(B, 1) + 8.0 * 0;
  ECS_HIP(hipMemcpyAsync(B.h_dh, B.dh_all, sizeof(double) * B.P * B.DH, hipMemcpyDeviceToHost, B.st));
  ECS_HIP(hipStreamSynchronize(B.st));
  for (int q = 0; q < B.P; ++q) norms[q] = run[q] ? std::sqrt(std::max(B.h_dh[(size_t)q * B.DH], 0.0)) : 0.0;
}

// The hierarchies' densities and coarsest inverses of the problems whose flags are set.
template <typename T>
void setup_operators(LockstepBatch &B) {
  BProb<T> *PP = probs<T>(B);
  for (const BLaunch &L : B.dens) run_launch<T>(B, L, 0);
  // dense inverses of the coarsest operators
  std::vector<std::vector<T>> dn(B.P);
  for (int q = 0; q < B.P; ++q) {
    if (!B.act_now[q]) continue;
    const BLevel<T> &C = PP[q].L.back();
    dn[q].resize(C.n);
    ECS_HIP(hipMemcpyAsync(dn[q].data(), C.dn, sizeof(T) * C.n, hipMemcpyDeviceToHost, B.st));
  }
  ECS_HIP(hipStreamSynchronize(B.st));
  std::vector<T> inv;
  for (int q = 0; q < B.P; ++q) {
    if (!B.act_now[q]) continue;
    const BLevel<T> &C = PP[q].L.back();
    const std::vector<double> I = coarse_inverse(C.s, dn[q]);
    inv.resize(I.size());
    for (size_t u = 0; u < I.size(); ++u) inv[u] = (T)I[u];
    ECS_HIP(hipMemcpy(PP[q].ainv, inv.data(), sizeof(T) * inv.size(), hipMemcpyHostToDevice));
  }
}

template <typename T>
void build_tails(LockstepBatch &B) {
  BProb<T> *PP = probs<T>(B);
  std::vector<TailArgs<T>> ta(B.P);
  B.tail_lds = 0;
  for (int q = 0; q < B.P; ++q) {
    BProb<T> &P = PP[q];
    TailArgs<T> &a = ta[q];
    const int nl = (int)P.L.size(), t = P.tail;
    a = TailArgs<T>{};
    a.nl = nl - t;
    a.nc = (int)P.L.back().n;
    a.pre = B.pre;
    a.post = B.post;
    size_t off = 0;
    for (int l = t; l < nl; ++l) {
      a.s[l - t] = P.L[l].s;
      a.m[l - t] = P.L[l].m;
      a.off[l - t] = (int)off;
      a.dn[l - t] = P.L[l].dn;
      off += 5 * (size_t)P.L[l].n;
    }
    a.ainv = P.ainv;
    a.bin = t > 0 ? P.L[t].b : nullptr;  // level 0: V[j] -> Z[j], set by the kernel
    a.xout = t > 0 ? P.L[t].x : nullptr;
    B.tail_lds = std::max(B.tail_lds, P.tail_lds);
  }
  B.tails = upload(B, ta);
  ECS_HIP(hipFuncSetAttribute((const void *)kb_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)TAIL_LDS_MAX));
}

// Vector segment tables of precision T.
template <typename T>
std::vector<BVec<T>> make_vecs(LockstepBatch &B, bool outer) {
  std::vector<BVec<T>> v(B.P);
  int start = 0;
  for (int q = 0; q < B.P; ++q) {
    BVec<T> &s = v[q];
    s = BVec<T>{};
    s.n = B.n[q];
    s.prob = q;
    s.start = start;
    s.nb = bnb(B.n[q]);
    start += s.nb;
    s.dh = B.dh_all + (size_t)q * B.DH;
    s.dy = B.dy_all + (size_t)q * (B.m + 1);
  }
  B.vblocks = start;
  return v;
}

}  // namespace

LockstepBatch::~LockstepBatch() {
  if (st) (void)hipStreamSynchronize(st);
  (void)hipFree(arena.base);
  if (h_act_ring) (void)hipHostFree(h_act_ring);
  if (h_count) (void)hipHostFree(h_count);
  if (h_dh) (void)hipHostFree(h_dh);
  if (h_dy) (void)hipHostFree(h_dy);
  if (h_scale) (void)hipHostFree(h_scale);
  delete[] p64;
  delete[] p32;
  if (st) (void)hipStreamDestroy(st);
}

// Layout of the whole batch in the arena (sizing pass with base == null, then the real
// pass): per problem the FP64 level-0 vectors, the cycle precision's hierarchy and
// Krylov vectors (and in mixed mode the FP32 right-hand side and solution), the
// segment tables.
static void lockstep_layout(LockstepBatch &B) {
  BArena &A = B.arena;
  A.used = 0;
  B.dn64.assign(B.P, nullptr);
  B.rhs64.assign(B.P, nullptr);
  B.r64.assign(B.P, nullptr);
  B.d_act = A.take<int>(B.P);
  B.d_count = A.take<int>(B.P);
  B.d_scale = A.take<double>(B.P);
  B.dh_all = A.take<double>((size_t)B.P * B.DH);
  B.dy_all = A.take<double>((size_t)B.P * (B.m + 1));
  std::vector<float *> b32(B.P, nullptr), x32(B.P, nullptr);
  for (int q = 0; q < B.P; ++q) {
    const long n = B.n[q];
    B.rhs64[q] = A.take<double>(n);
    B.r64[q] = A.take<double>(n);
    if (B.mixed) {
      B.dn64[q] = A.take<double>(n);
      b32[q] = A.take<float>(n);
      x32[q] = A.take<float>(n);
      carve(B, B.p32[q], n);
    } else {
      carve(B, B.p64[q], n);
    }
  }
  // segment tables
  auto v64 = make_vecs<double>(B, true);
  for (int q = 0; q < B.P; ++q) {
    BVec<double> &s = v64[q];
    s.r = B.r64[q];
    s.b = B.rhs64[q];
    s.partial = B.mixed ? nullptr : (B.p64 ? B.p64[q].partial : nullptr);
    if (!B.mixed && B.p64) {
      const BProb<double> &P = B.p64[q];
      s.s = P.L[0].s;
      s.dn = P.L[0].dn;
      s.V = P.V;
      s.Z = P.Z;
      s.w = P.w;
    } else {
      const ecsprt_params &pq = B.pp[q];
      s.s = stencil_consts(pq, B.dims[3 * q], B.dims[3 * q + 1], B.dims[3 * q + 2], pq.dx, pq.dy, pq.dz);
      s.dn = B.dn64[q];
    }
  }
  if (B.mixed) {
    // FP64 outer partial sums: their own rows after the FP32 problems' arrays
    for (int q = 0; q < B.P; ++q) v64[q].partial = A.take<double>((size_t)bnb(B.n[q]) * 2);
    auto v32 = make_vecs<float>(B, false);
    for (int q = 0; q < B.P; ++q) {
      BVec<float> &s = v32[q];
      const BProb<float> &P = B.p32[q];
      s.s = P.L[0].s;
      s.dn = P.L[0].dn;
      s.V = P.V;
      s.Z = P.Z;
      s.w = P.w;
      s.r = b32[q];
      s.b = b32[q];
      s.x = x32[q];
      s.partial = P.partial;
    }
    B.vsegs32 = static_cast<BVec<float> *>(upload(B, v32));
  }
  B.vsegs64 = static_cast<BVec<double> *>(upload(B, v64));  // x (phi) filled per solve
  // density and forcing inputs (moments and phi per solve; the table is rewritten then)
  std::vector<BProbIn> pin(B.P);
  int start = 0;
  for (int q = 0; q < B.P; ++q) {
    pin[q].rc = rhs_consts(B.pp[q]);
    pin[q].rhs = B.rhs64[q];
    pin[q].n = B.n[q];
    pin[q].start = start;
    start += bnodes(B.n[q]);
    if (B.mixed) {
      pin[q].dn32 = B.p32[q].L[0].dn;
      pin[q].dn64 = B.dn64[q];
    } else {
      pin[q].dn64 = B.p64[q].L[0].dn;
    }
  }
  B.nblocks_in = start;
  B.pin = static_cast<BProbIn *>(upload(B, pin));
  if (B.mixed) {
    build_program<float>(B);
    build_density<float>(B);
    build_tails<float>(B);
  } else {
    build_program<double>(B);
    build_density<double>(B);
    build_tails<double>(B);
  }
}

LockstepBatch *lockstep_create_runs(const std::vector<ecsprt_params> &pp, const ecsprt_options &o) {
  auto *B = new LockstepBatch();
  try {
    const int nprob = (int)pp.size();
    B->pp = pp;
    B->o = o;
    B->P = nprob;
    B->dims.resize(3 * nprob);
    for (int q = 0; q < nprob; ++q) {
      B->dims[3 * q] = pp[q].Nx;
      B->dims[3 * q + 1] = pp[q].Ny;
      B->dims[3 * q + 2] = pp[q].Nz;
    }
    B->m = o.ksp_restart > 0 ? o.ksp_restart : 30;
    B->pre = o.mg_pre > 0 ? o.mg_pre : 2;
    B->post = o.mg_post > 0 ? o.mg_post : 2;
    B->mixed = o.solve_precision == ECSPRT_SOLVE_MIXED;
    ECS_CHECK(o.pc_type == ECSPRT_PC_GMG, ECSPRT_ERR_ARG, "lockstep batch: multigrid preconditioning only");
    B->DH = 2 * (B->m + 2);
    B->n.resize(nprob);
    for (int q = 0; q < nprob; ++q) B->n[q] = (long)pp[q].Nx * pp[q].Ny * pp[q].Nz;
    if (B->mixed) {
      B->p32 = new BProb<float>[nprob];
      for (int q = 0; q < nprob; ++q) build_levels<float>(*B, q, B->p32[q]);
    } else {
      B->p64 = new BProb<double>[nprob];
      for (int q = 0; q < nprob; ++q) build_levels<double>(*B, q, B->p64[q]);
    }
    // sizing pass, then allocation and the real layout
    B->arena.base = nullptr;
    lockstep_layout(*B);
    const size_t need = B->arena.used + 4096;
    if (hipMalloc(&B->arena.base, need) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "lockstep batch arena"};
    B->arena.cap = need;
    ECS_HIP(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking));
    ECS_HIP(hipHostMalloc((void **)&B->h_act_ring, sizeof(int) * nprob * LockstepBatch::ACT_RING,
                          hipHostMallocDefault));
    ECS_HIP(hipHostMalloc((void **)&B->h_count, sizeof(int) * nprob, hipHostMallocDefault));
    ECS_HIP(hipHostMalloc((void **)&B->h_dh, sizeof(double) * nprob * B->DH, hipHostMallocDefault));
    ECS_HIP(hipHostMalloc((void **)&B->h_dy, sizeof(double) * nprob * (B->m + 1), hipHostMallocDefault));
    ECS_HIP(hipHostMalloc((void **)&B->h_scale, sizeof(double) * nprob, hipHostMallocDefault));
    ECS_HIP(hipMemset(B->arena.base, 0, need));  // one-time clear of the workspace
    lockstep_layout(*B);
    ECS_HIP(hipDeviceSynchronize());
  } catch (...) {
    delete B;
    throw;
  }
  return B;
}

LockstepBatch *lockstep_create(const ecsprt_params &p, const ecsprt_options &o, int nprob, const int *dims) {
  std::vector<ecsprt_params> pp(nprob, p);
  for (int q = 0; q < nprob; ++q) {
    pp[q].Nx = dims[3 * q];
    pp[q].Ny = dims[3 * q + 1];
    pp[q].Nz = dims[3 * q + 2];
  }
  return lockstep_create_runs(pp, o);
}

void lockstep_destroy(LockstepBatch *B) { delete B; }

hipStream_t lockstep_stream(const LockstepBatch &B) { return B.st; }

// GMRES with mixed refinement or in FP64 for the problems in take, from the x already in
// place, each to rtol times its initial residual; then the solutions' means removed.
static int lockstep_iterate(LockstepBatch &B, const std::vector<char> &take) {
  const int P = B.P;
  std::vector<char> run(take);
  const double rtol = B.o.ksp_rtol > 0 ? B.o.ksp_rtol : 1e-5, atol = B.o.ksp_atol > 0 ? B.o.ksp_atol : 1e-50;
  const int maxit = B.o.ksp_max_it > 0 ? B.o.ksp_max_it : 10000;
  std::vector<double> norms(P), tol(P), tol_in(P), beta(P), est(P);
  std::vector<int> steps(P);
  std::vector<char> conv(P, 0);
  double inner = B.o.inner_rtol;
  if (inner <= 0) {
    const double lr = -std::log10(std::max(rtol, 1e-15));
    const int cycles = std::max(1, (int)std::ceil(lr / 5.0 - 1e-9));
    inner = std::pow(10.0, -lr / cycles);
  }
  for (int outer = 0; outer < 200; ++outer) {
    for (int q = 0; q < P; ++q) run[q] = take[q] && !conv[q];
    outer_residual(B, run, norms);
    bool any = false;
    for (int q = 0; q < P; ++q) {
      if (!run[q]) continue;
      if (outer == 0) {
        tol[q] = std::max(rtol * norms[q], atol);
        B.bnorm[q] = norms[q];
      }
      B.rnorm[q] = norms[q];
      if (norms[q] <= tol[q]) {
        conv[q] = 1;
        run[q] = 0;
      } else if (!std::isfinite(norms[q]) || B.iters[q] >= maxit) {
        run[q] = 0;
      } else {
        any = true;
      }
    }
    if (!any) break;
    if (B.mixed) {
      for (int q = 0; q < P; ++q) B.h_scale[q] = run[q] ? 1.0 / norms[q] : 0.0;
      ECS_HIP(hipMemcpyAsync(B.d_scale, B.h_scale, sizeof(double) * P, hipMemcpyHostToDevice, B.st));
      set_act(B, run);
      kb_to32<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, B.vsegs32, P, B.d_act, B.d_scale);
      ECS_HIP(hipGetLastError());
      B.bytes += vec_bytes<double>(B, 1) + vec_bytes<float>(B, 2);
      for (int q = 0; q < P; ++q) tol_in[q] = inner;
      gmres_cycle<float>(B, run, tol_in, steps, est, beta);
      for (int q = 0; q < P; ++q) B.h_scale[q] = run[q] ? norms[q] : 0.0;
      ECS_HIP(hipMemcpyAsync(B.d_scale, B.h_scale, sizeof(double) * P, hipMemcpyHostToDevice, B.st));
      set_act(B, run);
      kb_from32<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, B.vsegs32, P, B.d_act, B.d_scale);
      ECS_HIP(hipGetLastError());
      B.bytes += vec_bytes<double>(B, 2) + vec_bytes<float>(B, 1);
    } else {
      gmres_cycle<double>(B, run, tol, steps, est, beta);
    }
    for (int q = 0; q < P; ++q)
      if (run[q]) B.iters[q] += steps[q];
  }
  // mean removal of every solution
  set_act(B, take);
  kb_sum<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 0);
  kb_reduce<double><<<P, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 1, 0);
  kb_sub_mean<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 0);
  ECS_HIP(hipGetLastError());
  B.bytes += vec_bytes<double>(B, 3);
  ECS_HIP(hipStreamSynchronize(B.st));
  int rc = ECSPRT_OK;
  for (int q = 0; q < P; ++q)
    if (take[q] && !conv[q]) rc = ECSPRT_ERR_NOT_CONVERGED;
  return rc;
}

// Solution pointers x = phi[q] and forcing pointers b = rhs[q] (rhs null: the batch's own
// forcing buffers) into the FP64 vector table, densities into the input table.
static void set_pointers(LockstepBatch &B, const double *const *moments, const double *const *dens,
                         const double *const *rhs, double *const *phi) {
  const int P = B.P;
  std::vector<BProbIn> pin(P);
  ECS_HIP(hipMemcpy(pin.data(), B.pin, sizeof(BProbIn) * P, hipMemcpyDeviceToHost));
  std::vector<BVec<double>> v64(P);
  ECS_HIP(hipMemcpy(v64.data(), B.vsegs64, sizeof(BVec<double>) * P, hipMemcpyDeviceToHost));
  for (int q = 0; q < P; ++q) {
    pin[q].moments = moments ? moments[q] : nullptr;
    pin[q].dens = dens ? dens[q] : nullptr;
    v64[q].x = phi[q];
    v64[q].b = rhs ? rhs[q] : B.rhs64[q];
  }
  ECS_HIP(hipMemcpy(B.pin, pin.data(), sizeof(BProbIn) * P, hipMemcpyHostToDevice));
  ECS_HIP(hipMemcpy(B.vsegs64, v64.data(), sizeof(BVec<double>) * P, hipMemcpyHostToDevice));
}

// All problems: operator and forcing from moments[q], solution into phi[q].
int lockstep_solve(LockstepBatch &B, double *const *moments, double *const *phi) {
  const int P = B.P;
  B.bytes = 0.0;
  B.iters.assign(P, 0);
  B.rnorm.assign(P, 0.0);
  B.bnorm.assign(P, 0.0);
  set_pointers(B, moments, nullptr, nullptr, phi);
  std::vector<char> run(P, 1);
  set_act(B, run);
  kb_density_rhs<<<B.nblocks_in, BT, 0, B.st>>>(B.pin, P);
  ECS_HIP(hipGetLastError());
  for (int q = 0; q < P; ++q) B.bytes += 48.0 * B.n[q];
  if (B.mixed) setup_operators<float>(B);
  else setup_operators<double>(B);
  // the forcing's mean removed
  kb_sum<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 1);
  kb_reduce<double><<<P, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 1, 0);
  kb_sub_mean<<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act, 1);
  ECS_HIP(hipGetLastError());
  B.bytes += vec_bytes<double>(B, 3);
  kb_fill_x<double><<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act);
  ECS_HIP(hipGetLastError());
  return lockstep_iterate(B, run);
}

// The problems in take: forcing rhs[q] (mean already removed), solution into phi[q]
// from a zero start.  A problem with dens[q] set first takes its operator from that
// level-0 density; the others keep theirs.
int lockstep_solve_given(LockstepBatch &B, const double *const *dens, const double *const *rhs, double *const *phi,
                         const std::vector<char> &take) {
  const int P = B.P;
  B.bytes = 0.0;
  B.iters.assign(P, 0);
  B.rnorm.assign(P, 0.0);
  B.bnorm.assign(P, 0.0);
  set_pointers(B, nullptr, dens, rhs, phi);
  std::vector<char> ops(P, 0);
  bool any = false;
  for (int q = 0; q < P; ++q) any |= (ops[q] = dens[q] != nullptr);
  if (any) {
    set_act(B, ops);
    kb_density_copy<<<B.nblocks_in, BT, 0, B.st>>>(B.pin, P);
    ECS_HIP(hipGetLastError());
    if (B.mixed) setup_operators<float>(B);
    else setup_operators<double>(B);
  }
  set_act(B, take);
  kb_fill_x<double><<<B.vblocks, BT, 0, B.st>>>(B.vsegs64, P, B.d_act);
  ECS_HIP(hipGetLastError());
  return lockstep_iterate(B, take);
}

void lockstep_bnorm(const LockstepBatch &B, double *bnorm) {
  for (int q = 0; q < B.P; ++q) bnorm[q] = q < (int)B.bnorm.size() ? B.bnorm[q] : 0.0;
}

void lockstep_stats(const LockstepBatch &B, int *iters, double *rnorm, double *bytes) {
  for (int q = 0; q < B.P; ++q) {
    if (iters) iters[q] = q < (int)B.iters.size() ? B.iters[q] : 0;
    if (rnorm) rnorm[q] = q < (int)B.rnorm.size() ? B.rnorm[q] : 0.0;
  }
  if (bytes) *bytes = B.bytes;
}

}  // namespace ecsprt
