// Two-dimensional ion stages (cli.ndim = 2, src/setup.c:99-104): the push
// (BorisMoverBz2D, src/velocities.c:186-273), the drift and periodic wrap
// (UpdatePositions and Apply2DBC with ndim = 2, src/positions.c:436-470,
// src/boundaries.c:6-42), the deposit (Collect2DFluidMoments, src/moments.c:5-104)
// and the two-dimensional ScatterElastic (src/collisions.c:9-170 with ndim = 2).
//
// Ions are kept in the reference's own layout, double[np][2] positions and
// velocities, with an id (the index they were set with).  Each deposit bins their
// indices by tile and, while it reads them in that order, writes them out again in tile
// order (a second buffer), so the next push, collisions and drift stream them with a
// local potential.  Collision draws are keyed by id, and get_particles returns the ions
// in id order.  The deposit is exact and order-independent: every weight
// (and every flux over vmax) is rounded once to a fixed-point grid of 2^-fx and added
// with 64-bit integer atomics, fx chosen from the fullest tile (or cell) so that no node
// sum can overflow; ions with a velocity component at or above vmax take FP64 atomics.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "philox.hpp"
#include "pic_math.hpp"
#include "tiles.hpp"

namespace ecsprt {

int collision_count(const ecsprt_params *p, long np, double dt);
double mx_vmax(const ecsprt_params &p);

namespace {

// DifferenceVector2D (src/vector-math.c:529-592) on the unghosted periodic potential:
// plain bilinear weights, one-sided differences at an upper node index N.
// Periodic index of base + d for a base already in [0, n) and |d| <= 2 < n.
__device__ inline int near_wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// wrap_index without the integer division for an index already in [0, n] (every
// in-domain coordinate, a coordinate at L giving n).
__device__ inline int wrap_fast(int i, int n) {
  return (unsigned)i < (unsigned)n ? i : (i == n ? 0 : wrap_index(i, n));
}

// P(j, i) is the potential at node (i, j), always with i - ixl and j - iyl in [-1, 2].
template <class Acc>
__device__ inline void diff_vec2d_at(const Acc &P, int Nx, int Ny, double x0, double y0, double f[2]) {
  const int ixl = (int)x0, ixh = ixl + 1, iyl = (int)y0, iyh = iyl + 1;
  const double wxh = x0 - (double)ixl, wyh = y0 - (double)iyl;
  double hh, lh, hl, ll;
  if (ixl >= 0) {
    hl = P(iyh, ixl + 1) - P(iyh, ixl - 1);
    ll = P(iyl, ixl + 1) - P(iyl, ixl - 1);
  } else {
    hl = -1.0 * P(iyh, ixl + 2) + 4.0 * P(iyh, ixl + 1) - 3.0 * P(iyh, ixl);
    ll = -1.0 * P(iyl, ixl + 2) + 4.0 * P(iyl, ixl + 1) - 3.0 * P(iyl, ixl);
  }
  if (ixh < Nx) {
    hh = P(iyh, ixh + 1) - P(iyh, ixh - 1);
    lh = P(iyl, ixh + 1) - P(iyl, ixh - 1);
  } else {
    hh = +3.0 * P(iyh, ixh) - 4.0 * P(iyh, ixh - 1) + 1.0 * P(iyh, ixh - 2);
    lh = +3.0 * P(iyl, ixh) - 4.0 * P(iyl, ixh - 1) + 1.0 * P(iyl, ixh - 2);
  }
  double Ewh = lh + wyh * (hh - lh), Ewl = ll + wyh * (hl - ll);
  f[0] = Ewl + wxh * (Ewh - Ewl);
  if (iyl >= 0) {
    lh = P(iyl + 1, ixh) - P(iyl - 1, ixh);
    ll = P(iyl + 1, ixl) - P(iyl - 1, ixl);
  } else {
    lh = -1.0 * P(iyl + 2, ixh) + 4.0 * P(iyl + 1, ixh) - 3.0 * P(iyl, ixh);
    ll = -1.0 * P(iyl + 2, ixl) + 4.0 * P(iyl + 1, ixl) - 3.0 * P(iyl, ixl);
  }
  if (iyh < Ny) {
    hh = P(iyh + 1, ixh) - P(iyh - 1, ixh);
    hl = P(iyh + 1, ixl) - P(iyh - 1, ixl);
  } else {
    hh = +3.0 * P(iyh, ixh) - 4.0 * P(iyh - 1, ixh) + 1.0 * P(iyh - 2, ixh);
    hl = +3.0 * P(iyh, ixl) - 4.0 * P(iyh - 1, ixl) + 1.0 * P(iyh - 2, ixl);
  }
  Ewh = lh + wyh * (hh - lh);
  Ewl = ll + wyh * (hl - ll);
  f[1] = Ewl + wxh * (Ewh - Ewl);
}

__device__ inline void diff_vec2d(const double *__restrict__ F, int Nx, int Ny, double x0, double y0, double f[2]) {
  const int ixl = (int)x0, iyl = (int)y0;
  // Every node read is (ixl + a, iyl + b) with a, b in [-1, 2]: one modulo per axis.
  const int bx = wrap_fast(ixl, Nx), by = wrap_fast(iyl, Ny);
  auto P = [&](int j, int i) { return F[(long)near_wrap(by + (j - iyl), Ny) * Nx + near_wrap(bx + (i - ixl), Nx)]; };
  diff_vec2d_at(P, Nx, Ny, x0, y0, f);
}

// BorisMoverBz2D body (src/velocities.c:254-262).
__global__ void k_push2d(long np, const double *__restrict__ phi, int Nx, int Ny, PushConsts k,
                         const double *__restrict__ pos, double *__restrict__ vel) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    const double x = cell_coord(pos[2 * p], k.dx, k.inv[0], k.pow2);
    const double y = cell_coord(pos[2 * p + 1], k.dy, k.inv[1], k.pow2);
    double E[2];
    diff_vec2d(phi, Nx, Ny, x, y, E);
    const double vxm = vel[2 * p] + k.Escale[0] * E[0];
    const double vym = vel[2 * p + 1] + k.Escale[1] * E[1];
    const double vxt = vxm + vym * k.t;
    const double vyp = vym - vxt * k.s;
    const double vxp = vxt + vyp * k.t;
    vel[2 * p] = vxp;
    vel[2 * p + 1] = vyp;
  }
}

// UpdatePositions (src/positions.c:455-459) and ApplyBC (src/boundaries.c:19-28), one
// wrap per axis; ions still outside [r0, r1] (or not finite) are counted in lost.
__global__ void k_drift2d(long np, DriftConsts d, double *__restrict__ pos, const double *__restrict__ vel,
                          unsigned long long *lost) {
  unsigned long long nl = 0;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    bool out = false;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      double r = pos[2 * p + a] + vel[2 * p + a] * d.dt;
      if (r < d.r0[a]) r += d.L[a];
      if (r >= d.r1[a]) r -= d.L[a];
      pos[2 * p + a] = r;
      out |= !(r >= d.r0[a] && r <= d.r1[a]);
    }
    nl += out ? 1ull : 0ull;
  }
  nl = wave_sum(nl);
  if (__lane_id() == 0 && nl) atomicAdd(lost, nl);
}

__global__ void k_apply_bc2d(long np, DriftConsts d, double *__restrict__ pos) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      double r = pos[2 * p + a];
      if (r < d.r0[a]) r += d.L[a];
      if (r >= d.r1[a]) r -= d.L[a];
      pos[2 * p + a] = r;
    }
  }
}

// Ions per cell ((int)(pos/d) per axis, a coordinate at L in cell 0) and the largest
// count, which bounds every node sum of weights by four times it.
__global__ void k_cell_counts2d(long np, int Nx, int Ny, double dx, double dy, const double *__restrict__ pos,
                                int *__restrict__ counts) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    const int i = wrap_index((int)(pos[2 * p] / dx), Nx), j = wrap_index((int)(pos[2 * p + 1] / dy), Ny);
    atomicAdd(&counts[(long)j * Nx + i], 1);
  }
}

__global__ void k_max_i32(const int *__restrict__ a, long n, int *out) {
  int m = 0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    m = max(m, a[q]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_down(m, o));
  if (__lane_id() == 0) atomicMax(out, m);
}

// Fixed-point bits for node sums bounded by 4 * maxcell (weights <= 1, |flux/vmax| < 1);
// maxcell is the fullest cell, or on the tiled path the fullest tile.
__device__ inline int fx2d_bits(int maxcell) {
  int b = 0;
  while (b < 62 && (1ll << b) < 4ll * (long long)maxcell + 1) ++b;
  return min(50, 62 - b);
}

// Collect2DFluidMoments (src/moments.c:44-82): bilinear weights hh, lh, hl, ll of the
// four corner nodes; moments n, Gx, Gy as fixed-point integers acc[node][3].
__global__ void k_deposit2d(long np, int Nx, int Ny, PushConsts k, double inv_vmax, const int *__restrict__ maxcell,
                            const double *__restrict__ pos, const double *__restrict__ vel,
                            unsigned long long *__restrict__ acc, double *__restrict__ ovf) {
  const double sw = ldexp(1.0, fx2d_bits(*maxcell));
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    const double x = cell_coord(pos[2 * p], k.dx, k.inv[0], k.pow2);
    const double y = cell_coord(pos[2 * p + 1], k.dy, k.inv[1], k.pow2);
    const int ixl = (int)x, iyl = (int)y;
    const double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    const double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    const double w4[4] = {wyh * wxh, wyl * wxh, wyh * wxl, wyl * wxl};
    const int bx = wrap_fast(ixl, Nx), by = wrap_fast(iyl, Ny);
    const int bx1 = near_wrap(bx + 1, Nx), by1 = near_wrap(by + 1, Ny);
    const int ci[4] = {bx1, bx1, bx, bx}, cj[4] = {by1, by, by1, by};
    const double vx = vel[2 * p], vy = vel[2 * p + 1];
    const double ux = vx * inv_vmax, uy = vy * inv_vmax;
    const bool slow = fabs(ux) < 1.0 && fabs(uy) < 1.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const long node = (long)cj[c] * Nx + ci[c];
      if (slow) {
        atomicAdd(&acc[node * 3 + 0], (unsigned long long)fx_round(w4[c] * sw));
        atomicAdd(&acc[node * 3 + 1], (unsigned long long)fx_round(ux * w4[c] * sw));
        atomicAdd(&acc[node * 3 + 2], (unsigned long long)fx_round(uy * w4[c] * sw));
      } else {
        unsafeAtomicAdd(&ovf[node * 4 + 0], w4[c]);
        unsafeAtomicAdd(&ovf[node * 4 + 1], vx * w4[c]);
        unsafeAtomicAdd(&ovf[node * 4 + 2], vy * w4[c]);
      }
    }
  }
}

// moments[node] = {n, Gx, Gy, 0} from the fixed-point sums and the overflow moments.
__global__ void k_fold2d(long n, const unsigned long long *__restrict__ acc, const double *__restrict__ ovf,
                         const int *__restrict__ maxcell, double vmax, double *__restrict__ m) {
  const double sn = ldexp(1.0, -fx2d_bits(*maxcell)), sg = vmax * sn;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    m[q * 4 + 0] = (double)(long long)acc[q * 3 + 0] * sn + ovf[q * 4 + 0];
    m[q * 4 + 1] = (double)(long long)acc[q * 3 + 1] * sg + ovf[q * 4 + 1];
    m[q * 4 + 2] = (double)(long long)acc[q * 3 + 2] * sg + ovf[q * 4 + 2];
    m[q * 4 + 3] = 0.0;
  }
}

// ---- binned deposit --------------------------------------------------------------
// Ion indices are binned by tile of T2 x T2 cells (a counting sort of indices: the ions
// themselves stay in place, in the reference's order), then one workgroup per tile sums
// its ions into fixed-point LDS accumulators over the tile's (T2+1)^2 nodes and adds the
// region to the global sums: 12 LDS atomics per ion instead of 12 global ones.
constexpr int T2 = 32;
constexpr int MAX_TILES2 = 4096;
constexpr int BIN_THREADS2 = 256;

struct Bin2 {
  int Nx, Ny, ntx, nty;
};

// The ion's wrapped cell (ix, iy) from its cell coordinates.
__device__ inline void cell_of_xy(const PushConsts &k, const Bin2 &b, double px, double py, int &ix, int &iy) {
  ix = wrap_fast((int)cell_coord(px, k.dx, k.inv[0], k.pow2), b.Nx);
  iy = wrap_fast((int)cell_coord(py, k.dy, k.inv[1], k.pow2), b.Ny);
}

__device__ inline void cell_of(const PushConsts &k, const Bin2 &b, const double *__restrict__ pos, long p, int &ix,
                               int &iy) {
  cell_of_xy(k, b, pos[2 * p], pos[2 * p + 1], ix, iy);
}

// LDS atomicAdd(&h[key], 1) of every active lane, returning each lane's old value;
// the lanes whose key is that of the first active lane (most of them: ions are stored
// in tile order) take theirs from one atomic, numbered by lane.
__device__ inline int wave_count(int *h, int key) {
  const int k0 = __builtin_amdgcn_readfirstlane(key);
  const unsigned long long same = __ballot(key == k0);
  int r = 0;
  if (key == k0) {
    const int below =
        (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)same, 0u));
    if (below == 0) r = atomicAdd(&h[k0], (int)__popcll(same));
    r = __builtin_amdgcn_readlane(r, __ffsll((long long)same) - 1) + below;
  } else {
    r = atomicAdd(&h[key], 1);
  }
  return r;
}

// Per-tile counts (LDS histogram per workgroup, added to tile_counts once per non-empty
// bin).  Workgroup w takes ions [w*chunk, (w+1)*chunk).
__global__ void __launch_bounds__(BIN_THREADS2) k_bin_hist2d(long np, long chunk, Bin2 b, PushConsts k,
                                                            const double *__restrict__ pos,
                                                            int *__restrict__ tile_counts) {
  __shared__ int h[MAX_TILES2];
  const int nt = b.ntx * b.nty;
  for (int q = threadIdx.x; q < nt; q += blockDim.x) h[q] = 0;
  __syncthreads();
  const long c0 = blockIdx.x * chunk, c1 = min(np, c0 + chunk);
  for (long p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
    int ix, iy;
    cell_of(k, b, pos, p, ix, iy);
    wave_count(h, (iy / T2) * b.ntx + ix / T2);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nt; q += blockDim.x)
    if (h[q]) atomicAdd(&tile_counts[q], h[q]);
}

// Exclusive scan of the tile counts into off[] and fill[] (one workgroup), and the
// largest tile count into *maxc: a node lies in at most four tile regions, so its sums
// are bounded by four times it (the fixed-point scale is chosen from that bound).
__global__ void __launch_bounds__(BIN_THREADS2) k_scan_tiles2d(int nt, const int *counts, int *counts_out,
                                                              int *__restrict__ off, int *__restrict__ fill,
                                                              int *__restrict__ maxc) {
  __shared__ long long part[BIN_THREADS2];
  __shared__ int pmax[BIN_THREADS2];
  const int per = (nt + BIN_THREADS2 - 1) / BIN_THREADS2;
  const int q0 = threadIdx.x * per, q1 = min(nt, q0 + per);
  long long s = 0;
  int mx = 0;
  for (int q = q0; q < q1; ++q) {
    s += counts[q];
    mx = max(mx, counts[q]);
  }
  part[threadIdx.x] = s;
  pmax[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long a = 0;
    int m = 0;
    for (int t = 0; t < BIN_THREADS2; ++t) {
      const long long v = part[t];
      part[t] = a;
      a += v;
      m = max(m, pmax[t]);
    }
    *maxc = m;
  }
  __syncthreads();
  long long a = part[threadIdx.x];
  for (int q = q0; q < q1; ++q) {
    const int n = counts[q];
    off[q] = (int)a;
    fill[q] = (int)a;
    counts_out[q] = n;  // the counts the deposit and the next push read
    a += n;
  }
}

// Scatter of ion indices to their tile's range of perm: each workgroup counts its
// chunk per tile in LDS, reserves one range per non-empty tile, then places its ions.
__global__ void __launch_bounds__(BIN_THREADS2) k_bin_scatter2d(long np, long chunk, Bin2 b, PushConsts k,
                                                               const double *__restrict__ pos,
                                                               const int *__restrict__ tile_in,
                                                               int *__restrict__ fill, int *__restrict__ perm) {
  // The tile of ion p: recorded by the fused push (tile_in), or from its position.
  auto tile_of = [&](long p) {
    if (tile_in) return tile_in[p];
    int ix, iy;
    cell_of(k, b, pos, p, ix, iy);
    return (iy / T2) * b.ntx + ix / T2;
  };
  __shared__ int h[MAX_TILES2];
  const int nt = b.ntx * b.nty;
  for (int q = threadIdx.x; q < nt; q += blockDim.x) h[q] = 0;
  __syncthreads();
  const long c0 = blockIdx.x * chunk, c1 = min(np, c0 + chunk);
  for (long p = c0 + threadIdx.x; p < c1; p += blockDim.x) wave_count(h, tile_of(p));
  __syncthreads();
  for (int q = threadIdx.x; q < nt; q += blockDim.x) {
    const int n = h[q];
    h[q] = n ? atomicAdd(&fill[q], n) : 0;  // the workgroup's first slot in tile q
  }
  __syncthreads();
  for (long p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
    const int slot = wave_count(h, tile_of(p));
    if (slot < np) perm[slot] = (int)p;  // always, when the counts are those of these positions
  }
}

// One workgroup per tile: fixed-point LDS sums of n, Gx/vmax, Gy/vmax over the tile's
// (T2+1)^2 nodes from its listed ions, then added (wrapped) into the global sums.
__global__ void __launch_bounds__(BIN_THREADS2) k_deposit2d_tiles(Bin2 b, PushConsts k, double inv_vmax,
                                                                 const int *__restrict__ maxcell,
                                                                 const int *__restrict__ off,
                                                                 const int *__restrict__ counts,
                                                                 const int *__restrict__ perm,
                                                                 const double *__restrict__ pos,
                                                                 const double *__restrict__ vel,
                                                                 const int *__restrict__ id,
                                                                 double *__restrict__ pos_out,
                                                                 double *__restrict__ vel_out,
                                                                 int *__restrict__ id_out,
                                                                 unsigned long long *__restrict__ acc,
                                                                 double *__restrict__ ovf) {
  constexpr int R = T2 + 1, R2 = R * R;
  __shared__ unsigned long long s[3 * R2];
  for (int q = threadIdx.x; q < 3 * R2; q += blockDim.x) s[q] = 0;
  __syncthreads();
  const int t = blockIdx.x;
  const int i0 = (t % b.ntx) * T2, j0 = (t / b.ntx) * T2;
  const double sw = ldexp(1.0, fx2d_bits(*maxcell));
  const int beg = off[t], n = counts[t];
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const long p = perm[beg + q];
    const double px = pos[2 * p], py = pos[2 * p + 1];
    const double x = cell_coord(px, k.dx, k.inv[0], k.pow2);
    const double y = cell_coord(py, k.dy, k.inv[1], k.pow2);
    const int ixl = (int)x, iyl = (int)y;
    const double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    const double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    const double w4[4] = {wyh * wxh, wyl * wxh, wyh * wxl, wyl * wxl};
    const double vx = vel[2 * p], vy = vel[2 * p + 1];
    // the ion, copied in tile order: the following stages stream it with a local potential
    const long o = beg + q;
    pos_out[2 * o] = px;
    pos_out[2 * o + 1] = py;
    vel_out[2 * o] = vx;
    vel_out[2 * o + 1] = vy;
    id_out[o] = id[p];
    const double ux = vx * inv_vmax, uy = vy * inv_vmax;
    const int la = wrap_fast(ixl, b.Nx) - i0, lb = wrap_fast(iyl, b.Ny) - j0;  // in [0, T2)
    const int na[4] = {la + 1, la + 1, la, la}, nb[4] = {lb + 1, lb, lb + 1, lb};
    if (fabs(ux) < 1.0 && fabs(uy) < 1.0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = nb[c] * R + na[c];
        lds_addfx(reinterpret_cast<long long *>(&s[r]), w4[c] * sw);
        lds_addfx(reinterpret_cast<long long *>(&s[R2 + r]), ux * w4[c] * sw);
        lds_addfx(reinterpret_cast<long long *>(&s[2 * R2 + r]), uy * w4[c] * sw);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const long node = (long)near_wrap(j0 + nb[c], b.Ny) * b.Nx + near_wrap(i0 + na[c], b.Nx);
        unsafeAtomicAdd(&ovf[node * 4 + 0], w4[c]);
        unsafeAtomicAdd(&ovf[node * 4 + 1], vx * w4[c]);
        unsafeAtomicAdd(&ovf[node * 4 + 2], vy * w4[c]);
      }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < R2; r += blockDim.x) {
    const int a = r % R, bb = r / R;
    if (i0 + a > b.Nx || j0 + bb > b.Ny) continue;  // beyond the grid's last node (partial tile)
    const unsigned long long v0 = s[r], v1 = s[R2 + r], v2 = s[2 * R2 + r];
    if (!(v0 | v1 | v2)) continue;
    const long node = (long)near_wrap(j0 + bb, b.Ny) * b.Nx + near_wrap(i0 + a, b.Nx);
    atomicAdd(&acc[node * 3 + 0], v0);
    atomicAdd(&acc[node * 3 + 1], v1);
    atomicAdd(&acc[node * 3 + 2], v2);
  }
}

// Statistical collisions (Philox mode) in two dimensions: a Poisson number of attempts
// per ion keyed by (ion index, step), each the two-dimensional scattering formula.  The
// draw runs over every ion; workgroup w lists the ions of its chunk that have attempts
// (a few in a hundred) in its own segment of the list (LDS ranks, no global atomics),
// and a second kernel runs the attempts of each segment with every lane busy instead
// of leaving most of each wave idle through the scattering arithmetic.
constexpr int COLL_CHUNK2 = 4096;  // ions per selecting workgroup (16 per thread)

__global__ void __launch_bounds__(256) k_collide2d_select(long np, PoissonParams pt, uint32_t key0, uint32_t key1,
                                                          uint32_t step, const int *__restrict__ id,
                                                          int *__restrict__ list, int *__restrict__ seg) {
  __shared__ int s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const long c0 = (long)blockIdx.x * COLL_CHUNK2, c1 = min(np, c0 + COLL_CHUNK2);
  for (long p0 = c0; p0 < c1; p0 += blockDim.x) {  // uniform per wave
    const long p = p0 + threadIdx.x;
    const int k = p < c1 && poisson_any(pt, key0, key1, step, id[p]) ? 1 : 0;
    const unsigned long long m = __ballot(k > 0);
    if (!m) continue;
    const int below = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    int base = 0;
    if (k > 0 && below == 0) base = atomicAdd(&s_n, (int)__popcll(m));
    base = __builtin_amdgcn_readlane(base, __ffsll((long long)m) - 1);
    if (k > 0) list[c0 + base + below] = (int)p;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    seg[2 * blockIdx.x] = (int)c0;
    seg[2 * blockIdx.x + 1] = s_n;
  }
}

constexpr int COLL_SEGS2 = 8;  // most list segments per applying workgroup

// Segments per applying workgroup: about 8192 workgroups of a few hundred listed ions.
int apply_segs(long nseg) { return (int)std::max(1L, std::min((long)COLL_SEGS2, nseg / 8192)); }

// Apply the attempts of the ions listed in G consecutive segments (segment g holds
// seg[2g + 1] ion slots from list[seg[2g]] on), so every thread has work.  The
// workgroup's attempt, success and failure counts go to the counters with one atomic
// each.
__global__ void __launch_bounds__(256) k_collide2d_apply(long nseg, int G, const int *__restrict__ list,
                                                         const int *__restrict__ seg, PoissonParams pt,
                                                         CollideConsts cc, uint32_t key0, uint32_t key1,
                                                         uint32_t step, int write, const int *__restrict__ id,
                                                         double *__restrict__ vel, unsigned long long *ctr) {
  __shared__ int s_off[COLL_SEGS2 + 1], s_beg[COLL_SEGS2];
  __shared__ unsigned long long s_red[3][4];
  const long g0 = (long)blockIdx.x * G;
  if (threadIdx.x == 0) {
    int t = 0;
    for (int g = 0; g < COLL_SEGS2; ++g) {
      s_off[g] = t;
      s_beg[g] = 0;
      if (g < G && g0 + g < nseg) {
        s_beg[g] = seg[2 * (g0 + g)];
        t += seg[2 * (g0 + g) + 1];
      }
    }
    s_off[COLL_SEGS2] = t;
  }
  __syncthreads();
  const int total = s_off[COLL_SEGS2];
  unsigned natt = 0, nsucc = 0, nfail = 0;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    int g = 0;
    while (q >= s_off[g + 1]) ++g;
    const long p = list[(long)s_beg[g] + (q - s_off[g])];
    const int pid = id[p];
    const int k = poisson_attempts(pt, key0, key1, step, pid);
    double vx = vel[2 * p], vy = vel[2 * p + 1], vz = 0.0;
    collide_attempts<double, 2>(cc, key0, key1, step, pid, 0, k, vx, vy, vz, natt, nsucc, nfail);
    if (write) {
      vel[2 * p] = vx;
      vel[2 * p + 1] = vy;
    }
  }
  const unsigned long long s0 = wave_sum(natt), s1 = wave_sum(nsucc), s2 = wave_sum(nfail);
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    s_red[0][w] = s0;
    s_red[1][w] = s1;
    s_red[2][w] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const unsigned long long v = s_red[threadIdx.x][0] + s_red[threadIdx.x][1] + s_red[threadIdx.x][2] +
                                 s_red[threadIdx.x][3];
    if (v) atomicAdd(&ctr[3 + threadIdx.x], v);
  }
}

constexpr int HALO2 = T2 + 3;  // a tile's nodes -1 .. T2 + 1 per axis

constexpr int QCAP2 = 512;  // LDS collision queue of k_push2d_tiles: two workgroups' worth

// UpdatePositions (src/positions.c:455-459) and the wrap of one ion, stored and
// returned in q; true if it is still outside [r0, r1] (or not finite).
__device__ inline bool drift_one2d(const DriftConsts &d, long p, double px, double py, double vx, double vy,
                                   double *__restrict__ pos, double q[2]) {
  const double r0v[2] = {px, py}, v[2] = {vx, vy};
  bool out = false;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    double r = r0v[a] + v[a] * d.dt;
    if (r < d.r0[a]) r += d.L[a];
    if (r >= d.r1[a]) r -= d.L[a];
    pos[2 * p + a] = r;
    q[a] = r;
    out |= !(r >= d.r0[a] && r <= d.r1[a]);
  }
  return out;
}

// BorisMoverBz2D (src/velocities.c:254-262) over ions stored in tile order.  Workgroup
// (t, s) takes the s-th contiguous share of tile t's ions and reads the potential from
// the tile's nodes staged in LDS; an ion whose cell is outside the tile (not where the
// last deposit put it) reads it from memory.  With fused set it also draws each ion's
// collision attempt count and moves it (UpdatePositions and the wrap): an ion with no
// attempt at once, the others through an LDS queue that the whole workgroup works off
// 256 at a time (attempts, then the move), so collisions run on full waves while the
// ions' lines are still in the cache.  The moved ions are counted per tile into
// tnext (the deposit's binning then needs no counting pass): the tile and its eight
// neighbours in LDS, any other tile with a global atomic.
__global__ void __launch_bounds__(256, 5) k_push2d_tiles(Bin2 b, int split, const int *__restrict__ toff,
                                                      const int *__restrict__ tcnt, const double *__restrict__ phi,
                                                      PushConsts k, int fused, int collide, DriftConsts d,
                                                      PoissonParams pt, CollideConsts cc, uint32_t key0,
                                                      uint32_t key1, uint32_t step, const int *__restrict__ id,
                                                      double *__restrict__ pos, double *__restrict__ vel,
                                                      int *__restrict__ tnext, int *__restrict__ tile_out,
                                                      unsigned long long *ctr) {
  __shared__ double s_phi[HALO2 * HALO2];
  __shared__ int s_h[9];
  __shared__ double s_qd[QCAP2][4];  // position, then pushed velocity
  __shared__ int s_qp[QCAP2], s_qid[QCAP2];
  __shared__ int s_qn;
  __shared__ unsigned long long s_red[4][4];
  const int t = blockIdx.x / split, sp = blockIdx.x - t * split;
  const int i0 = (t % b.ntx) * T2, j0 = (t / b.ntx) * T2;
  for (int q = threadIdx.x; q < HALO2 * HALO2; q += blockDim.x) {
    const int ly = q / HALO2, lx = q - ly * HALO2;
    s_phi[q] = phi[(long)wrap_index(j0 - 1 + ly, b.Ny) * b.Nx + wrap_index(i0 - 1 + lx, b.Nx)];
  }
  if (threadIdx.x < 9) s_h[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_qn = 0;
  __syncthreads();
  const int tx = t % b.ntx, ty = t / b.ntx;
  // Tile of a moved ion: its own (bin 4, one LDS atomic per wave), a neighbour's bin, or
  // a global atomic.  Called by every lane of the wave; an ion still outside the domain
  // (the step fails) is not counted.
  auto count_tile = [&](bool moved, long p, double qx, double qy) {
    int bin = -1, tile = 0;
    if (moved) {
      int ix, iy;
      cell_of_xy(k, b, qx, qy, ix, iy);
      const int ux = ix / T2, uy = iy / T2;
      tile = uy * b.ntx + ux;
      tile_out[p] = tile;
      const int dx = tile_step(ux - tx, b.ntx), dy = tile_step(uy - ty, b.nty);
      bin = (dx < 2 && dy < 2) ? (dy + 1) * 3 + (dx + 1) : 9;
    }
    const unsigned long long own = __ballot(bin == 4);
    if (own && bin == 4 &&
        (int)__builtin_amdgcn_mbcnt_hi((unsigned)(own >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)own, 0u)) == 0)
      atomicAdd(&s_h[4], (int)__popcll(own));
    if (bin >= 0 && bin < 9 && bin != 4) atomicAdd(&s_h[bin], 1);
    if (bin == 9) atomicAdd(&tnext[tile], 1);
  };
  const int n = tcnt[t], m = (n + split - 1) / split;
  const long beg = (long)toff[t] + (long)sp * m;
  const int cnt = max(0, min(m, n - sp * m));
  unsigned natt = 0, nsucc = 0, nfail = 0, nl = 0;
  // Entries from .. from + nw - 1 of the queue, one per thread.
  auto work_off = [&](int from, int nw) {
    bool moved = false;
    double qx = 0.0, qy = 0.0;
    long p = 0;
    if ((int)threadIdx.x < nw) {
      const int e = from + threadIdx.x;
      p = s_qp[e];
      const int pid = s_qid[e];
      const double px = s_qd[e][0], py = s_qd[e][1];
      double vx = s_qd[e][2], vy = s_qd[e][3], vz = 0.0;
      const int ka = poisson_attempts(pt, key0, key1, step, pid);
      collide_attempts<double, 2>(cc, key0, key1, step, pid, 0, ka, vx, vy, vz, natt, nsucc, nfail);
      vel[2 * p] = vx;
      vel[2 * p + 1] = vy;
      double qn[2];
      const bool out = drift_one2d(d, p, px, py, vx, vy, pos, qn);
      nl += out ? 1u : 0u;
      moved = !out;
      qx = qn[0];
      qy = qn[1];
    }
    count_tile(moved, p, qx, qy);
  };
  for (int q0 = 0; q0 < cnt; q0 += blockDim.x) {  // uniform per workgroup
    const int q = q0 + threadIdx.x;
    bool listed = false, moved = false;
    double qx = 0.0, qy = 0.0;
    long p = 0;
    int pid = 0;
    double px = 0, py = 0, vxp = 0, vyp = 0;
    if (q < cnt) {
      p = beg + q;
      px = pos[2 * p];
      py = pos[2 * p + 1];
      const double x = cell_coord(px, k.dx, k.inv[0], k.pow2);
      const double y = cell_coord(py, k.dy, k.inv[1], k.pow2);
      const int ixl = (int)x, iyl = (int)y;
      const int lx = wrap_fast(ixl, b.Nx) - i0, ly = wrap_fast(iyl, b.Ny) - j0;
      double E[2];
      if (lx >= 0 && lx < T2 && ly >= 0 && ly < T2) {
        auto P = [&](int j, int i) { return s_phi[(ly + 1 + (j - iyl)) * HALO2 + lx + 1 + (i - ixl)]; };
        diff_vec2d_at(P, b.Nx, b.Ny, x, y, E);
      } else {
        diff_vec2d(phi, b.Nx, b.Ny, x, y, E);
      }
      const double vxm = vel[2 * p] + k.Escale[0] * E[0];
      const double vym = vel[2 * p + 1] + k.Escale[1] * E[1];
      const double vxt = vxm + vym * k.t;
      vyp = vym - vxt * k.s;
      vxp = vxt + vyp * k.t;
      if (fused && collide) {
        pid = id[p];
        listed = poisson_any(pt, key0, key1, step, pid);
      }
      if (!listed) {
        vel[2 * p] = vxp;
        vel[2 * p + 1] = vyp;
        if (fused) {
          double qn[2];
          const bool out = drift_one2d(d, p, px, py, vxp, vyp, pos, qn);
          nl += out ? 1u : 0u;
          moved = !out;
          qx = qn[0];
          qy = qn[1];
        }
      }
    }
    if (fused) count_tile(moved, p, qx, qy);
    if (fused && collide) {
      const unsigned long long mk = __ballot(listed);
      if (mk) {
        const int below =
            (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mk, 0u));
        int base = 0;
        if (listed && below == 0) base = atomicAdd(&s_qn, (int)__popcll(mk));
        base = __builtin_amdgcn_readlane(base, __ffsll((long long)mk) - 1);
        if (listed) {
          const int e = base + below;  // < QCAP2: the queue holds < 256 entries before this append
          s_qp[e] = (int)p;
          s_qid[e] = pid;
          s_qd[e][0] = px;
          s_qd[e][1] = py;
          s_qd[e][2] = vxp;
          s_qd[e][3] = vyp;
        }
      }
      __syncthreads();
      const int nq = s_qn;
      __syncthreads();  // every thread has read s_qn before it changes
      if (nq >= (int)blockDim.x) {
        work_off(nq - blockDim.x, blockDim.x);
        if (threadIdx.x == 0) s_qn = nq - blockDim.x;
        __syncthreads();
      }
    }
  }
  if (fused) {
    if (collide) {
      __syncthreads();
      work_off(0, s_qn);
    }
    const unsigned long long r0 = wave_sum(natt), r1 = wave_sum(nsucc), r2 = wave_sum(nfail), r3 = wave_sum(nl);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) {
      s_red[0][w] = r0;
      s_red[1][w] = r1;
      s_red[2][w] = r2;
      s_red[3][w] = r3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      const unsigned long long v = s_red[threadIdx.x][0] + s_red[threadIdx.x][1] + s_red[threadIdx.x][2] +
                                   s_red[threadIdx.x][3];
      if (v) atomicAdd(&ctr[threadIdx.x < 3 ? 3 + threadIdx.x : 2], v);
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + 9) {
      const int q = threadIdx.x - 64, h = s_h[q];
      const int dx = q % 3 - 1, dy = q / 3 - 1;
      if (h) atomicAdd(&tnext[wrap_index(ty + dy, b.nty) * b.ntx + wrap_index(tx + dx, b.ntx)], h);
    }
  }
}

// Device initial conditions: positions uniform in the domain (x optionally displaced
// to a density 1 + amp cos(2 pi k x / Lx)), velocities Maxwellian per component.
__global__ void k_init2d(long np, double Lx, double Ly, double x0, double y0, double amp, int kk, double vtx,
                         double vty, double v0x, double v0y, uint32_t key0, uint32_t key1, double *__restrict__ pos,
                         double *__restrict__ vel) {
  const double twopi = 6.283185307179586;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    u32x4 c0 = {{(uint32_t)p, (uint32_t)(p >> 32), 0x2D2Du, 0u}};
    u32x4 c1 = {{(uint32_t)p, (uint32_t)(p >> 32), 0x2D2Du, 1u}};
    u32x4 r0 = philox4x32_10(c0, key0, key1);
    u32x4 r1 = philox4x32_10(c1, key0, key1);
    const double ux = ((double)(r0.v[0] >> 5) * 67108864.0 + (double)(r0.v[1] >> 6)) * (1.0 / 9007199254740992.0);
    const double uy = ((double)(r0.v[2] >> 5) * 67108864.0 + (double)(r0.v[3] >> 6)) * (1.0 / 9007199254740992.0);
    double xs = ux;
    if (amp != 0.0 && kk != 0) {
      const double kw = twopi * kk;
      for (int it = 0; it < 8; ++it) {
        const double F = xs - amp / kw * (cos(kw * xs) - 1.0) - ux;
        const double dF = 1.0 + amp * sin(kw * xs);
        xs -= F / dF;
      }
      xs = xs - floor(xs);
    }
    double x = x0 + xs * Lx, y = y0 + uy * Ly;
    if (x >= x0 + Lx) x = x0;
    if (y >= y0 + Ly) y = y0;
    pos[2 * p] = x;
    pos[2 * p + 1] = y;
    const double u1 = ((double)r1.v[0] + 0.5) * (1.0 / 4294967296.0);
    const double u2 = ((double)r1.v[1] + 0.5) * (1.0 / 4294967296.0);
    const double ra = sqrt(-2.0 * log(u1));
    vel[2 * p] = vtx * ra * cos(twopi * u2) + v0x;
    vel[2 * p + 1] = vty * ra * sin(twopi * u2) + v0y;
  }
}

__global__ void k_iota2d(long np, int *__restrict__ id) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) id[p] = (int)p;
}

__global__ void k_unsort2d(long np, const int *__restrict__ id, const double *__restrict__ pos,
                           const double *__restrict__ vel, double *__restrict__ pos_out, double *__restrict__ vel_out) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    const long q = id[p];
    pos_out[2 * q] = pos[2 * p];
    pos_out[2 * q + 1] = pos[2 * p + 1];
    vel_out[2 * q] = vel[2 * p];
    vel_out[2 * q + 1] = vel[2 * p + 1];
  }
}

__global__ void k_slot_of_id2d(long np, const int *__restrict__ id, int *__restrict__ slot) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x)
    slot[id[p]] = (int)p;
}

int grid2(long n) { return (int)std::max(1L, std::min(65536L, (n + 255) / 256)); }

}  // namespace

// ---- host side -------------------------------------------------------------------

namespace {

Bin2 bin2_of(const ecsprt_params &p) { return Bin2{p.Nx, p.Ny, (p.Nx + T2 - 1) / T2, (p.Ny + T2 - 1) / T2}; }

// Workgroups per tile in k_push2d_tiles: about 8192 over the grid.
int push_split(int nt) { return std::max(1, std::min(64, (8192 + nt - 1) / nt)); }

void ensure_perm(ecsprt_hip_ctx *c, long np) {
  if (c->d_perm2 && c->perm_cap2 >= np) return;
  (void)hipFree(c->d_perm2);
  c->d_perm2 = nullptr;
  c->perm_cap2 = 0;
  if (hipMalloc(&c->d_perm2, sizeof(int) * std::max(np, 1L)) != hipSuccess)
    throw Error{ECSPRT_ERR_NOMEM, "two-dimensional ion lists"};
  c->perm_cap2 = np;
}

// (start, count) of each list segment.
void ensure_segments(ecsprt_hip_ctx *c, long nseg) {
  if (c->d_seg2 && c->seg_cap2 >= nseg) return;
  (void)hipFree(c->d_seg2);
  c->d_seg2 = nullptr;
  c->seg_cap2 = 0;
  ECS_HIP(hipMalloc(&c->d_seg2, sizeof(int) * 2 * std::max(nseg, 1L)));
  c->seg_cap2 = nseg;
}

}  // namespace

void p2d_reserve(ecsprt_hip_ctx *c, long np) {
  ECS_CHECK(np > 0 && np < 2147483647L, ECSPRT_ERR_ARG, "np must be in [1, 2^31-1)");
  if (!c->d_acc2) {
    ECS_HIP(hipMalloc(&c->d_acc2, sizeof(unsigned long long) * 3 * c->nnode));
    ECS_HIP(hipMalloc(&c->d_cnt2, sizeof(int) * (c->nnode + 1)));
  }
  if (!c->d_ovf) ECS_HIP(hipMalloc(&c->d_ovf, sizeof(double) * 4 * c->nnode));
  if (!c->d_ctr) {
    ECS_HIP(hipMalloc(&c->d_ctr, sizeof(unsigned long long) * CTR_N));
    ECS_HIP(hipHostMalloc(&c->h_ctr, sizeof(unsigned long long) * CTR_N));
  }
  if (c->cap2 >= np && c->d_pos2) return;
  double **dv[4] = {&c->d_pos2, &c->d_vel2, &c->d_pos2b, &c->d_vel2b};
  int **iv[3] = {&c->d_id2, &c->d_id2b, &c->d_tile2};
  for (double **a : dv) {
    (void)hipFree(*a);
    *a = nullptr;
  }
  for (int **a : iv) {
    (void)hipFree(*a);
    *a = nullptr;
  }
  c->cap2 = 0;
  for (double **a : dv)
    if (hipMalloc(a, sizeof(double) * 2 * std::max(np, 1L)) != hipSuccess)
      throw Error{ECSPRT_ERR_NOMEM, "two-dimensional particle arrays"};
  for (int **a : iv)
    if (hipMalloc(a, sizeof(int) * std::max(np, 1L)) != hipSuccess)
      throw Error{ECSPRT_ERR_NOMEM, "two-dimensional particle ids"};
  c->cap2 = np;
}

void p2d_free(ecsprt_hip_ctx *c) {
  (void)hipFree(c->d_pos2);
  (void)hipFree(c->d_vel2);
  (void)hipFree(c->d_pos2b);
  (void)hipFree(c->d_vel2b);
  (void)hipFree(c->d_id2);
  (void)hipFree(c->d_id2b);
  (void)hipFree(c->d_tile2);
  c->d_pos2b = c->d_vel2b = nullptr;
  c->d_id2 = c->d_id2b = c->d_tile2 = nullptr;
  (void)hipFree(c->d_acc2);
  (void)hipFree(c->d_cnt2);
  (void)hipFree(c->d_perm2);
  (void)hipFree(c->d_tiles2);
  (void)hipFree(c->d_seg2);
  c->d_perm2 = c->d_tiles2 = c->d_seg2 = nullptr;
  c->seg_cap2 = 0;
  c->perm_cap2 = 0;
  c->d_pos2 = c->d_vel2 = nullptr;
  c->d_acc2 = nullptr;
  c->d_cnt2 = nullptr;
}

void p2d_set(ecsprt_hip_ctx *c, long np, const double *pos, const double *vel, hipMemcpyKind kind) {
  p2d_reserve(c, np);
  ECS_HIP(hipMemcpyAsync(c->d_pos2, pos, sizeof(double) * 2 * np, kind, c->stream));
  ECS_HIP(hipMemcpyAsync(c->d_vel2, vel, sizeof(double) * 2 * np, kind, c->stream));
  k_iota2d<<<grid2(np), 256, 0, c->stream>>>(np, c->d_id2);
  ECS_HIP(hipGetLastError());
  c->np = np;
  c->tiled2 = c->counted2 = false;
  c->coll_lambda = -1.0;
  ECS_HIP(hipStreamSynchronize(c->stream));
}

// Ions back in the order they were set (by id).
void p2d_get(ecsprt_hip_ctx *c, double *pos, double *vel) {
  const long np = c->np;
  ensure_stage(c, np);
  k_unsort2d<<<grid2(np), 256, 0, c->stream>>>(np, c->d_id2, c->d_pos2, c->d_vel2, c->d_stage, c->d_stage + 2 * np);
  ECS_HIP(hipGetLastError());
  ECS_HIP(hipMemcpyAsync(pos, c->d_stage, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipMemcpyAsync(vel, c->d_stage + 2 * np, sizeof(double) * 2 * np, hipMemcpyDeviceToHost, c->stream));
  ECS_HIP(hipStreamSynchronize(c->stream));
}

// Slot of each ion id (for the pre-drawn replay, whose records name ions by id).
void p2d_slot_of_id(ecsprt_hip_ctx *c, int *slot) {
  k_slot_of_id2d<<<grid2(c->np), 256, 0, c->stream>>>(c->np, c->d_id2, slot);
  ECS_HIP(hipGetLastError());
}

void p2d_init_device(ecsprt_hip_ctx *c, long np, uint64_t seed, double amp, int k) {
  p2d_reserve(c, np);
  const ecsprt_params &p = c->p;
  k_init2d<<<grid2(np), 256, 0, c->stream>>>(np, p.Lx, p.Ly, p.x0, p.y0, amp, k, p.viTx, p.viTy, p.vi0x, p.vi0y,
                                             (uint32_t)seed, (uint32_t)(seed >> 32), c->d_pos2, c->d_vel2);
  k_iota2d<<<grid2(np), 256, 0, c->stream>>>(np, c->d_id2);
  ECS_HIP(hipGetLastError());
  c->np = np;
  c->tiled2 = c->counted2 = false;
  c->coll_lambda = -1.0;
  ECS_HIP(hipStreamSynchronize(c->stream));
}

void p2d_push(ecsprt_hip_ctx *c, double dt) {
  if (c->tiled2) {
    const Bin2 b = bin2_of(c->p);
    const int nt = b.ntx * b.nty, split = push_split(nt);
    int *toff = c->d_tiles2 + MAX_TILES2, *tcnt = c->d_tiles2;
    k_push2d_tiles<<<nt * split, 256, 0, c->stream>>>(b, split, toff, tcnt, c->d_phi, push_consts(c->p, dt), 0, 0,
                                                      drift_consts(c->p, dt), poisson_params(0.0),
                                                      collide_consts_of(c->p), 0, 0, 0, c->d_id2, c->d_pos2,
                                                      c->d_vel2, nullptr, nullptr, nullptr);
  } else {
    k_push2d<<<grid2(c->np), 256, 0, c->stream>>>(c->np, c->d_phi, c->p.Nx, c->p.Ny, push_consts(c->p, dt),
                                                  c->d_pos2, c->d_vel2);
  }
  ECS_HIP(hipGetLastError());
}

// Push, Philox collisions and drift of one step in one pass over the tile-ordered ions
// (k_push2d_tiles with fused set): the same ions, draws and arithmetic as p2d_push,
// p2d_collide_philox and p2d_drift in turn.  False (nothing done) when the ions are not
// in tile order, collisions are pre-drawn or the Philox attempt rate has not been set
// yet (its dry runs are done by p2d_collide_philox).
bool p2d_push_collide_drift(ecsprt_hip_ctx *c, double dt) {
  c->counted2 = false;
  if (!c->tiled2) return false;
  const bool philox = c->p.mn > 0.0 && c->o.collide_mode == ECSPRT_COLLIDE_PHILOX;
  if (c->p.mn > 0.0 && c->o.collide_mode == ECSPRT_COLLIDE_PREDRAWN) return false;
  const int Nc = philox ? collision_count(&c->p, c->np, dt) : 0;
  const bool collide = philox && Nc > 0;
  if (collide && c->coll_lambda <= 0.0) return false;
  if (philox) {
    c->last_Nc = Nc;
    c->last_Ns = c->last_Nf = c->last_attempts = 0;
  }
  const Bin2 b = bin2_of(c->p);
  const int nt = b.ntx * b.nty, split = push_split(nt);
  int *toff = c->d_tiles2 + MAX_TILES2, *tcnt = c->d_tiles2;
  const PoissonParams pt = poisson_params(collide ? c->coll_lambda : 0.0);
  const uint32_t k0 = (uint32_t)c->o.seed, k1 = (uint32_t)(c->o.seed >> 32), step = (uint32_t)c->step_index;
  int *tnext = c->d_tiles2 + 3 * MAX_TILES2;
  zero_ctr(c);
  fill_i32(c, tnext, nt, 0);
  k_push2d_tiles<<<nt * split, 256, 0, c->stream>>>(b, split, toff, tcnt, c->d_phi, push_consts(c->p, dt), 1,
                                                    collide ? 1 : 0, drift_consts(c->p, dt), pt,
                                                    collide_consts_of(c->p), k0, k1, step, c->d_id2, c->d_pos2,
                                                    c->d_vel2, tnext, c->d_tile2, c->d_ctr);
  ECS_HIP(hipGetLastError());
  c->tiled2 = false;
  read_ctr(c);
  if (collide) {
    const long natt = (long)c->h_ctr[3], ns = (long)c->h_ctr[4], nf = (long)c->h_ctr[5];
    c->last_attempts = natt;
    c->last_Ns = ns;
    c->last_Nf = nf;
    if (ns > 0) c->coll_lambda = steer_lambda(c->coll_lambda, Nc, ns);
    ECS_CHECK(nf < 10, ECSPRT_ERR_COLLISION, "failed to collide " + std::to_string(nf) + " ion-neutral pairs");
  }
  ECS_CHECK(c->h_ctr[2] == 0, ECSPRT_ERR_COUNT_CHANGED,
            std::to_string(c->h_ctr[2]) + " ions left the periodic domain (not finite or moved more than L)");
  c->counted2 = true;
  return true;
}

// Drift and wrap; a lost ion aborts as a changed ion count does in the reference
// (src/boundaries.c:101-109).
void p2d_drift(ecsprt_hip_ctx *c, double dt) {
  c->tiled2 = c->counted2 = false;
  zero_ctr(c);
  k_drift2d<<<grid2(c->np), 256, 0, c->stream>>>(c->np, drift_consts(c->p, dt), c->d_pos2, c->d_vel2, c->d_ctr + 2);
  ECS_HIP(hipGetLastError());
  read_ctr(c);
  ECS_CHECK(c->h_ctr[2] == 0, ECSPRT_ERR_COUNT_CHANGED,
            std::to_string(c->h_ctr[2]) + " ions left the periodic domain (not finite or moved more than L)");
}

void p2d_apply_bc(ecsprt_hip_ctx *c, long np, double *pos) {
  c->tiled2 = c->counted2 = false;
  k_apply_bc2d<<<grid2(np), 256, 0, c->stream>>>(np, drift_consts(c->p, 0.0), pos);
  ECS_HIP(hipGetLastError());
}

void p2d_cell_counts(ecsprt_hip_ctx *c, int *d_counts) {
  fill_i32(c, d_counts, c->nnode, 0);
  k_cell_counts2d<<<grid2(c->np), 256, 0, c->stream>>>(c->np, c->p.Nx, c->p.Ny, c->p.dx, c->p.dy, c->d_pos2,
                                                       d_counts);
  ECS_HIP(hipGetLastError());
}

void p2d_deposit(ecsprt_hip_ctx *c) {
  const long n = c->nnode;
  const int Nx = c->p.Nx, Ny = c->p.Ny;
  const Bin2 b{Nx, Ny, (Nx + T2 - 1) / T2, (Ny + T2 - 1) / T2};
  const int nt = b.ntx * b.nty;
  int *maxcell = c->d_cnt2 + n;
  const PushConsts k = push_consts(c->p, 0.0);
  const double vmax = mx_vmax(c->p);
  fill_f64(c, reinterpret_cast<double *>(c->d_acc2), 3 * n, 0.0);  // all-zero bits
  fill_f64(c, c->d_ovf, 4 * n, 0.0);
  const long np = c->np;
  const bool counted = c->counted2;  // tile counts of these positions left by the fused push
  c->tiled2 = c->counted2 = false;
  if (nt <= MAX_TILES2) {
    ensure_perm(c, np);
    if (!c->d_tiles2) ECS_HIP(hipMalloc(&c->d_tiles2, sizeof(int) * 4 * MAX_TILES2));
    int *tcnt = c->d_tiles2, *toff = tcnt + MAX_TILES2, *tfill = toff + MAX_TILES2, *tnext = tfill + MAX_TILES2;
    const long nwg = std::max(1L, std::min(4096L, (np + 65535) / 65536));
    const long chunk = (np + nwg - 1) / nwg;
    if (!counted) {
      fill_i32(c, tcnt, nt, 0);
      k_bin_hist2d<<<nwg, BIN_THREADS2, 0, c->stream>>>(np, chunk, b, k, c->d_pos2, tcnt);
    }
    k_scan_tiles2d<<<1, BIN_THREADS2, 0, c->stream>>>(nt, counted ? tnext : tcnt, tcnt, toff, tfill, maxcell);
    k_bin_scatter2d<<<nwg, BIN_THREADS2, 0, c->stream>>>(np, chunk, b, k, c->d_pos2, counted ? c->d_tile2 : nullptr,
                                                         tfill, c->d_perm2);
    k_deposit2d_tiles<<<nt, BIN_THREADS2, 0, c->stream>>>(b, k, 1.0 / vmax, maxcell, toff, tcnt, c->d_perm2,
                                                          c->d_pos2, c->d_vel2, c->d_id2, c->d_pos2b, c->d_vel2b,
                                                          c->d_id2b, c->d_acc2, c->d_ovf);
    ECS_HIP(hipGetLastError());
    std::swap(c->d_pos2, c->d_pos2b);
    std::swap(c->d_vel2, c->d_vel2b);
    std::swap(c->d_id2, c->d_id2b);
    c->tiled2 = true;
  } else {
    // Grids of more than 4096 tiles: global fixed-point atomics per ion, the scale bounded
    // by the fullest cell.
    p2d_cell_counts(c, c->d_cnt2);
    fill_i32(c, maxcell, 1, 0);
    k_max_i32<<<grid2(n), 256, 0, c->stream>>>(c->d_cnt2, n, maxcell);
    k_deposit2d<<<grid2(np), 256, 0, c->stream>>>(np, Nx, Ny, k, 1.0 / vmax, maxcell, c->d_pos2, c->d_vel2,
                                                  c->d_acc2, c->d_ovf);
    ECS_HIP(hipGetLastError());
  }
  k_fold2d<<<grid2(n), 256, 0, c->stream>>>(n, c->d_acc2, c->d_ovf, maxcell, vmax, c->d_moments);
  ECS_HIP(hipGetLastError());
}

// Philox-mode collisions: the attempt rate is set by dry runs the first time (as the
// three-dimensional pilot does) and steered towards Nc collisions per step after.
void p2d_collide_philox(ecsprt_hip_ctx *c, double dt) {
  const int Nc = collision_count(&c->p, c->np, dt);
  c->last_Nc = Nc;
  c->last_Ns = c->last_Nf = c->last_attempts = 0;
  if (Nc == 0 || c->p.mn <= 0.0) return;
  const CollideConsts cc = collide_consts_of(c->p);
  const uint32_t k0 = (uint32_t)c->o.seed, k1 = (uint32_t)(c->o.seed >> 32);
  ensure_perm(c, c->np);  // the collision list shares the tile-list buffer
  const long nseg = (c->np + COLL_CHUNK2 - 1) / COLL_CHUNK2;
  ensure_segments(c, nseg);
  auto pass = [&](double lam, uint32_t step, int write) {
    zero_ctr(c);
    const PoissonParams pt = poisson_params(lam);
    k_collide2d_select<<<nseg, 256, 0, c->stream>>>(c->np, pt, k0, k1, step, c->d_id2, c->d_perm2, c->d_seg2);
    const int G = apply_segs(nseg);
    k_collide2d_apply<<<(nseg + G - 1) / G, 256, 0, c->stream>>>(nseg, G, c->d_perm2, c->d_seg2, pt, cc, k0, k1, step,
                                                                 write, c->d_id2, c->d_vel2, c->d_ctr);
    ECS_HIP(hipGetLastError());
    read_ctr(c);
  };
  if (c->coll_lambda <= 0.0) {
    double lam = (double)Nc / (double)c->np;
    for (int it = 0; it < 4 && lam > 0; ++it) {
      pass(std::min(lam, POISSON_LAMBDA_MAX), 0xFFFFFFF0u + it, 0);
      lam = pilot_rescale(lam, Nc, (double)c->h_ctr[4]);
    }
    pilot_check(lam);
    c->coll_lambda = lam;
  }
  pass(c->coll_lambda, (uint32_t)c->step_index, 1);
  const long natt = (long)c->h_ctr[3], ns = (long)c->h_ctr[4], nf = (long)c->h_ctr[5];
  c->last_attempts = natt;
  c->last_Ns = ns;
  c->last_Nf = nf;
  if (ns > 0) c->coll_lambda = steer_lambda(c->coll_lambda, Nc, ns);
  ECS_CHECK(nf < 10, ECSPRT_ERR_COLLISION, "failed to collide " + std::to_string(nf) + " ion-neutral pairs");
}

}  // namespace ecsprt
