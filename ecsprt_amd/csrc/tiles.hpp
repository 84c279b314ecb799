// Device and host helpers shared by the FP64 and mixed-precision tile kernels.
#pragma once
#include <cmath>

#include "common.hpp"
#include "pic_math.hpp"

namespace ecsprt {

enum : int { M_PUSH = 1, M_DRIFT = 2, M_DEPOSIT = 4, M_REBIN = 8, M_COLLIDE = 16 };

// Argument structs read afresh.  A by-value argument struct is loaded into scalar
// registers at kernel entry and each field stays live to its last use; in the fused
// tile kernels that is about twice the scalar register budget, and the compiler spills
// the excess to vector-register lanes, one VALU instruction per read inside the particle
// loops.  Read through a constant-address-space pointer that an empty asm statement
// redefines at each call, the fields are not carried between loop iterations: they
// become scalar loads, which the scalar cache serves.  The struct must stay unchanged
// while the kernel runs.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ inline const T &fresh(ConstPtr<T> p) {
  asm volatile("" : "+s"(p));
  return *(const T *)p;
}

// The kernel's argument T, for a kernel whose sole explicit argument is a T (the
// code-object ABI places it at offset 0 of the kernel-argument segment).
template <typename T>
__device__ inline ConstPtr<T> kernarg() {
  return (ConstPtr<T>)__builtin_amdgcn_kernarg_segment_ptr();
}

// Workgroups of one launch over the tiles of several runs: run r owns workgroups
// [start[r], start[r + 1]) and its argument struct is args[r] (device memory).
constexpr int RUNS_MAX = 64;
struct RunTable {
  const void *args;
  int nrun;
  int start[RUNS_MAX + 1];
};

// Run of workgroup gi (binary search over the starts, read from the kernel argument).
__device__ inline int run_of(ConstPtr<RunTable> t, int gi) {
  int lo = 0, hi = t->nrun - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t->start[mid] <= gi) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}


__device__ inline int tile_of(const TileGeom &g, double xg, double yg, double zg) {
  int ci = (int)xg, cj = (int)yg, ck = (int)zg;
  if (ci >= g.Nx) ci = g.Nx - 1;
  if (cj >= g.Ny) cj = g.Ny - 1;
  if (ck >= g.Nz) ck = g.Nz - 1;
  return ((ck / g.Tz) * g.nty + cj / g.Ty) * g.ntx + ci / g.Tx;
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  return v;
}

// Adds the workgroup's sums of a, b, c to ctr[0..2] with one device atomic each (wave sums
// through red, one row per wave; at most 16 waves): thousands of workgroups ending on
// per-wave atomics to the same three addresses would serialise there.  Every thread calls it.
__device__ inline void block_add3(unsigned long long (*red)[3], unsigned a, unsigned b, unsigned c,
                                  unsigned long long *ctr) {
  const unsigned long long s0 = wave_sum(a), s1 = wave_sum(b), s2 = wave_sum(c);
  if (__lane_id() == 0) {
    red[threadIdx.x >> 6][0] = s0;
    red[threadIdx.x >> 6][1] = s1;
    red[threadIdx.x >> 6][2] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += red[w][threadIdx.x];
    if (t) atomicAdd(&ctr[threadIdx.x], t);
  }
}

// The per-tile slot counters (next) and late-list lengths (late_cnt) take device
// atomics from the 27 neighbouring workgroups every batch.  Packed 32 to a 128-byte
// line, the atomics on one line serialise as if on one counter; each tile's counter
// sits on a line of its own.
constexpr int TILE_CTR_STRIDE = 32;
__host__ __device__ inline long tile_ctr(int t) { return (long)t * TILE_CTR_STRIDE; }

// Tile offset d = b - a on a periodic axis of n tiles folded to -1, 0, +1; 2 when
// the destination is further than one tile.
__device__ inline int tile_step(int d, int n) {
  if (d > 1) d -= n;
  else if (d < -1) d += n;
  return (d >= -1 && d <= 1) ? d : 2;
}

// Exact fixed-point accumulation of an FP64 value already scaled to integer units
// (|v| < 2^51): v + 1.5*2^52 rounds v to the nearest integer into the low mantissa bits,
// whose difference from the bit pattern of 1.5*2^52 is that integer as a signed 64-bit
// value; one integer LDS atomic adds it.  The sum is exact and independent of order.
__device__ inline long long fx_round(double scaled) {
  const double y = scaled + 6755399441055744.0;
  return __double_as_longlong(y) - 0x4338000000000000LL;
}
__device__ inline void lds_addfx(long long *p, double scaled) {
  __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p), (unsigned long long)fx_round(scaled),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Slot of this ion in the list of its destination tile, code 0..26 among the 27
// neighbours nbr (codes 27 and 28 take none), reserved by its wave alone: the lanes
// bound for one destination take consecutive slots after one device atomic on that
// tile's counter, made by the lowest of them, so no workgroup barrier is needed.  One
// pass per distinct destination in the wave (about ten).  Called by every lane of the
// wave.
__device__ inline int wave_reserve(int code, const int *nbr, int *next) {
  unsigned long long todo = __ballot(code < 27);
  unsigned long long mine = 0;
  while (todo) {
    const int c = __builtin_amdgcn_readlane(code, (int)__builtin_ctzll(todo));
    const unsigned long long m = __ballot(code == c);
    if (code == c) mine = m;
    todo &= ~m;
  }
  const int below =
      (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mine >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mine, 0u));
  int base = 0;
  if (code < 27 && below == 0) base = atomicAdd(&next[tile_ctr(nbr[code])], (int)__popcll(mine));
  base = __shfl(base, mine ? (int)__builtin_ctzll(mine) : 0);
  return base + below;
}

// Periodic index for i within one period of [0, n) (no integer division).
__device__ inline int wrap_near(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }

// Deposit of an ion outside the fixed-point LDS path (beyond the tile's region without
// a late list, or with a velocity component at or above vmax): global FP64 atomics into
// the overflow moments.
__device__ inline void deposit_overflow(double *ovf, const TileGeom &g, int ixl, int iyl, int izl,
                                              const double w8[8], double vx, double vy, double vz) {
  const int xs[2] = {wrap_near(ixl, g.Nx), wrap_near(ixl + 1, g.Nx)};
  const int ys[2] = {wrap_near(iyl, g.Ny), wrap_near(iyl + 1, g.Ny)};
  const int zs[2] = {wrap_near(izl, g.Nz), wrap_near(izl + 1, g.Nz)};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ax = (c < 4) ? 1 : 0, ay = (c & 2) ? 0 : 1, az = (c & 1) ? 0 : 1;
    const long node = ((long)zs[az] * g.Ny + ys[ay]) * g.Nx + xs[ax];
    const double wt = w8[c];
    unsafeAtomicAdd(&ovf[node * 4 + 0], wt);
    unsafeAtomicAdd(&ovf[node * 4 + 1], vx * wt);
    unsafeAtomicAdd(&ovf[node * 4 + 2], vy * wt);
    unsafeAtomicAdd(&ovf[node * 4 + 3], vz * wt);
  }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

inline int grid_for(long n, int block) {
  long b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > 65536) b = 65536;
  return (int)b;
}

CollideConsts collide_consts_of(const ecsprt_params &p);

// Mixed-precision storage (particles_mixed.hip).
void mx_free(ParticlesMix &P);
void mx_alloc(ParticlesMix &P, long n);
void mx_set_lds_attrs();
void mx_launch(ecsprt_hip_ctx *c, int mode, double dt, const PoissonParams *pt, uint32_t key0, uint32_t key1,
               uint32_t step);
void mx_rebin(ecsprt_hip_ctx *c);
void mx_bin_scatter(ecsprt_hip_ctx *c, long np);
void mx_unbin(ecsprt_hip_ctx *c);
void mx_unbin_list(ecsprt_hip_ctx *c, long n);
void mx_cell_counts(ecsprt_hip_ctx *c, int *d_counts);
void mx_late(ecsprt_hip_ctx *c);
void mx_collide_philox(ecsprt_hip_ctx *c, const PoissonParams &pt, const CollideConsts &cc, uint32_t key0,
                       uint32_t key1, uint32_t step);
double mx_vmax(const ecsprt_params &p);
int fx_bits(const ecsprt_hip_ctx *c);
void mx_launch_runs(hipStream_t st, const std::vector<ecsprt_hip_ctx *> &cs, double dt,
                    const std::vector<const PoissonParams *> &pt, const std::vector<uint32_t> &key0,
                    const std::vector<uint32_t> &key1, const std::vector<uint32_t> &step, void *h_args, void *d_args);
size_t mx_args_bytes();

// Fill kernels launched on the context's stream (particles.hip).
void fill_f64(ecsprt_hip_ctx *c, double *p, long n, double v);
void fill_i32(ecsprt_hip_ctx *c, int *p, long n, int v);
void zero_ctr(ecsprt_hip_ctx *c);

inline DriftConsts drift_consts(const ecsprt_params &p, double dt) {
  DriftConsts d;
  d.dt = dt;
  d.L[0] = p.Lx;
  d.L[1] = p.Ly;
  d.L[2] = p.Lz;
  d.r0[0] = p.x0;
  d.r0[1] = p.y0;
  d.r0[2] = p.z0;
  d.r1[0] = p.x1;
  d.r1[1] = p.y1;
  d.r1[2] = p.z1;
  return d;
}

// BorisMoverBz3D scalars, src/velocities.c:307-320.
inline PushConsts push_consts(const ecsprt_params &p, double dt) {
  PushConsts k;
  double h[3] = {1.0 / p.dx, 1.0 / p.dy, 1.0 / p.dz};
  double tscale = 0.5 * (p.qi / p.mi) * dt;
  k.t = tscale * p.B0;
  k.s = 2.0 * k.t / (1 + k.t * k.t);
  for (int d = 0; d < 3; ++d) k.Escale[d] = -0.5 * h[d] * tscale;
  k.dx = p.dx;
  k.dy = p.dy;
  k.dz = p.dz;
  const double d3[3] = {p.dx, p.dy, p.dz};
  k.pow2 = 1;
  for (int d = 0; d < 3; ++d) {
    int e;
    k.pow2 &= std::frexp(d3[d], &e) == 0.5 ? 1 : 0;
    k.inv[d] = 1.0 / d3[d];
  }
  return k;
}

constexpr int CTR_N = 16;  // device counters per context (and their pinned mirror)

inline void read_ctr(ecsprt_hip_ctx *c) {
  ECS_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(unsigned long long) * CTR_N, hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace ecsprt
