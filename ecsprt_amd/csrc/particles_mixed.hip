// Mixed-precision particle stages on gfx950.  Each ion is a packed int32 cell index,
// an FP32 offset within that cell and an FP32 velocity (32 bytes with its id, against
// 52 for FP64).  The push gathers from an FP32 copy of the potential in LDS, the drift
// moves the offset and carries whole cells into the integer index (so the periodic
// wrap is exact), and the deposit computes FP32 cloud-in-cell weights and adds them
// to 64-bit fixed-point LDS accumulators.  Integer sums are exact and independent of
// the order in which ions arrive, so the moments are deterministic and carry only the
// FP32 rounding of each weight (about 1e-9 of a unit weight), not the growth of an
// FP32 running sum over the ~800 contributions a node receives.
//
// The tile structure, re-binning protocol and collision deferral are those of the
// FP64 tile kernel (particles.hip); only the particle arithmetic differs.
#include <mutex>

#include "common.hpp"
#include "philox.hpp"
#include "pic_math.hpp"
#include "tiles.hpp"

namespace ecsprt {

// FP32 push and drift constants (BorisMoverBz3D scalars, src/velocities.c:307-320).
struct MxConsts {
  float t, s;
  float Escale[3];
  float dtc[3];  // dt / cell width: cells crossed per unit velocity in one step
  float inv_vmax;  // 1 / vmax: fixed-point flux units are vmax * 2^-30
};

// Fixed-point deposit: a weight w in [0, 1] is stored as trunc(w * 2^30) and a flux
// contribution v*w as rint(v/vmax * w * 2^30); |v| < vmax keeps both below 2^30.  Ions
// with a faster component take the FP64 overflow path.
constexpr float MX_WSCALE = 1073741824.0f;  // 2^30

struct TileArgsMx {
  TileGeom g;
  ParticlesMix src, dst, emi, spill, def;
  long scap;
  PoissonParams pt;
  const int *base, *count, *cap;
  int *next;
  unsigned long long *ctr;
  const double *phi;
  long long *fold;
  double *ovf;
  uint4 *late;    // per-tile lists of ions whose deposit waits for k_late_mx: two uint4 per ion
  int *late_cnt;  // list lengths
  MxConsts k;
  CollideConsts cc;
  uint32_t key0, key1, step;
};

__device__ inline uint32_t *mx_at(const ParticlesMix &P, long i) { return P.p + (i >> 6) * 512 + (i & 63); }

__device__ inline void store_mx(const ParticlesMix &d, long i, uint32_t cell, float ox, float oy, float oz, float vx,
                                float vy, float vz, int id) {
  uint32_t *b = mx_at(d, i);
  b[64 * MX_CELL] = cell;
  b[64 * MX_OX] = __float_as_uint(ox);
  b[64 * MX_OY] = __float_as_uint(oy);
  b[64 * MX_OZ] = __float_as_uint(oz);
  b[64 * MX_VX] = __float_as_uint(vx);
  b[64 * MX_VY] = __float_as_uint(vy);
  b[64 * MX_VZ] = __float_as_uint(vz);
  b[64 * MX_ID] = (uint32_t)id;
}

__device__ inline void load_mx(const ParticlesMix &P, long i, uint32_t &cell, float &ox, float &oy, float &oz,
                               float &vx, float &vy, float &vz, int &id) {
  const uint32_t *b = mx_at(P, i);
  cell = b[64 * MX_CELL];
  ox = __uint_as_float(b[64 * MX_OX]);
  oy = __uint_as_float(b[64 * MX_OY]);
  oz = __uint_as_float(b[64 * MX_OZ]);
  vx = __uint_as_float(b[64 * MX_VX]);
  vy = __uint_as_float(b[64 * MX_VY]);
  vz = __uint_as_float(b[64 * MX_VZ]);
  id = (int)b[64 * MX_ID];
}

// Fixed-point LDS accumulation; the conversion rule is fixed, so sums stay exact and
// independent of order.  Weights are non-negative: truncated (one instruction, a bias
// below 2^-31 per weight) and zero-extended.  Fluxes are signed: rounded to nearest, so
// their sums carry no bias toward zero.
__device__ inline void lds_addw(long long *p, float scaled) {
  const unsigned long long q = (unsigned long long)__float2uint_rz(scaled);
  __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p), q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_addq(long long *p, float scaled) {
  const long long q = (long long)__float2int_rn(scaled);
  __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(p), (unsigned long long)q, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Drift along one axis: offset o plus d cells, whole cells moved into the index i and
// wrapped once over the N periodic cells (ApplyBC, src/boundaries.c:25-26, in cell
// units).  Returns false for a particle that is still outside after the wrap (or
// whose displacement is not finite): the reference drops it.
__device__ inline bool drift_axis(float &o, int &i, float d, int N) {
  const float u = o + d;
  const float f = floorf(u);
  if (!(fabsf(f) < 1048576.0f)) return false;
  // u - f is below 1; it rounds up to 1 only for u just below an integer.
  o = fminf(u - f, 0x1.fffffep-1f);
  int n = i + (int)f;
  if (n < 0) n += N;
  if (n >= N) n -= N;
  i = n;
  return n >= 0 && n < N;
}

// Stores an ion in tile dt, or in the spill list when the tile is full.
__device__ inline void place_mx(const TileArgsMx &a, int base, int cap, int slot, uint32_t cell, float ox, float oy,
                                float oz, float vx, float vy, float vz, int id) {
  if (slot < cap) {
    store_mx(a.dst, (long)base + slot, cell, ox, oy, oz, vx, vy, vz, id);
    return;
  }
  const unsigned long long sp = atomicAdd(&a.ctr[1], 1ull);
  if ((long)sp < a.scap) store_mx(a.spill, (long)sp, cell, ox, oy, oz, vx, vy, vz, id);
  else atomicAdd(&a.ctr[2], 1ull);
}

// Collect3DFluidMoments weights (src/moments.c:164-184) from the in-cell offsets,
// scaled by 2^30 through the z factor (exact: a power of two).
__device__ inline void weights_mx(float ox, float oy, float oz, float w8[8]) {
  const float wxh = ox, wxl = 1.0f - ox;
  const float wyh = oy, wyl = 1.0f - oy;
  const float wzh = oz * MX_WSCALE, wzl = (1.0f - oz) * MX_WSCALE;
  w8[0] = wzh * wyh * wxh;
  w8[1] = wzl * wyh * wxh;
  w8[2] = wzh * wyl * wxh;
  w8[3] = wzl * wyl * wxh;
  w8[4] = wzh * wyh * wxl;
  w8[5] = wzl * wyh * wxl;
  w8[6] = wzh * wyl * wxl;
  w8[7] = wzl * wyl * wxl;
}

// Deposit through FP64 global atomics on the overflow moments.
__device__ inline void deposit_direct_mx(double *ovf, const TileGeom &g, int ix, int iy, int iz, float ox, float oy,
                                         float oz, float vx, float vy, float vz) {
  float w8[8];
  weights_mx(ox, oy, oz, w8);
  double wd[8];
  for (int c = 0; c < 8; ++c) wd[c] = (double)w8[c] * (1.0 / 1073741824.0);
  deposit_overflow(ovf, g, ix, iy, iz, wd, (double)vx, (double)vy, (double)vz);
}
__device__ inline void deposit_direct_mx(const TileArgsMx &a, int ix, int iy, int iz, float ox, float oy, float oz,
                                         float vx, float vy, float vz) {
  deposit_direct_mx(a.ovf, a.g, ix, iy, iz, ox, oy, oz, vx, vy, vz);
}

// An ion beyond its source tile's deposit region lies in its destination tile dt:
// append cell, offsets and velocity to dt's list (one atomic, 32 contiguous bytes) for
// k_late_mx, which adds it in the exact fixed point of the tile regions, or deposit it
// directly when the list is full.
__device__ inline void late_append_mx(uint4 *late, int *late_cnt, double *ovf, const TileGeom &g, int dt, int base,
                                      int cap, int ix, int iy, int iz, float ox, float oy, float oz, float vx, float vy,
                                      float vz) {
  const int li = atomicAdd(&late_cnt[tile_ctr(dt)], 1);
  if (li < (cap >> 2)) {
    uint4 *e = late + 2 * ((long)(base >> 2) + li);
    e[0] = make_uint4(mx_pack(ix, iy, iz), __float_as_uint(ox), __float_as_uint(oy), __float_as_uint(oz));
    e[1] = make_uint4(__float_as_uint(vx), __float_as_uint(vy), __float_as_uint(vz), 0u);
  } else {
    deposit_direct_mx(ovf, g, ix, iy, iz, ox, oy, oz, vx, vy, vz);
  }
}

// Bit 31 of an emigrant's id marks an ion whose deposit is still owed (ids are below
// 2^31 - 1): k_rebin_mx lists it in its destination tile.
constexpr uint32_t EMI_DEPOSIT_MX = 0x80000000u;

struct TileWGMx {
  int t, tx, ty, tz, i0, j0, k0;
  int RAx, RAy, RAz;
  long long *sacc;
  int *s_nbr;             // the 27 neighbour tiles
  int *s_nslot, *s_ncap;  // slot offset and capacity of each neighbour tile
};

// Drift + wrap, deposit and re-binning of one ion per thread, in uniform control
// flow across each wave (re-binning uses wave-wide ballots).
template <int MODE, int TX, int TY, int TZ, int HH>
__device__ inline void tail_mx(const TileArgsMx &a, TileWGMx &w, bool valid, long p, int ix, int iy, int iz, float ox,
                               float oy, float oz, float vx, float vy, float vz, int pid) {
  const TileGeom &g = a.g;
  bool lost = false, late = false;
  if ((MODE & M_DRIFT) && valid) {
    // UpdatePositions (src/positions.c:455-459) with the wrap of src/boundaries.c:19-28.
    const bool okx = drift_axis(ox, ix, vx * a.k.dtc[0], g.Nx);
    const bool oky = drift_axis(oy, iy, vy * a.k.dtc[1], g.Ny);
    const bool okz = drift_axis(oz, iz, vz * a.k.dtc[2], g.Nz);
    lost = !(okx && oky && okz);
  }
  if ((MODE & M_DEPOSIT) && valid && !lost) {
    float w8[8];
    weights_mx(ox, oy, oz, w8);
    constexpr int RAX = TX + 1 + 2 * HH, RAY = TY + 1 + 2 * HH, RAZ = TZ + 1 + 2 * HH, RA3 = RAX * RAY * RAZ;
    int la = ix - (w.i0 - HH), lb = iy - (w.j0 - HH), lc = iz - (w.k0 - HH);
    if (la < 0) la += g.Nx; else if (la > w.RAx - 2) la -= g.Nx;
    if (lb < 0) lb += g.Ny; else if (lb > w.RAy - 2) lb -= g.Ny;
    if (lc < 0) lc += g.Nz; else if (lc > w.RAz - 2) lc -= g.Nz;
    const bool inreg = la >= 0 && la <= w.RAx - 2 && lb >= 0 && lb <= w.RAy - 2 && lc >= 0 && lc <= w.RAz - 2;
    const float ux = vx * a.k.inv_vmax, uy = vy * a.k.inv_vmax, uz = vz * a.k.inv_vmax;
    const bool slow = fabsf(ux) < 1.0f && fabsf(uy) < 1.0f && fabsf(uz) < 1.0f;
    if (inreg && slow) {
      const int base = (lc * RAY + lb) * RAX + la;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int ax = (c < 4) ? 1 : 0, ay = (c & 2) ? 0 : 1, az = (c & 1) ? 0 : 1;
        const int idx = base + (az * RAY + ay) * RAX + ax;
        const float wt = w8[c];
        lds_addw(&w.sacc[idx], wt);
        lds_addq(&w.sacc[RA3 + idx], ux * wt);
        lds_addq(&w.sacc[2 * RA3 + idx], uy * wt);
        lds_addq(&w.sacc[3 * RA3 + idx], uz * wt);
      }
    } else if ((MODE & M_REBIN) && slow) {
      late = true;  // deposited by k_late_mx once re-binned into its new tile
    } else {
      deposit_direct_mx(a, ix, iy, iz, ox, oy, oz, vx, vy, vz);
    }
  }
  if (MODE & M_REBIN) {
    int code = 28;
    if (valid && !lost) {
      const int sx = tile_step(ix / TX - w.tx, g.ntx), sy = tile_step(iy / TY - w.ty, g.nty);
      const int sz = tile_step(iz / TZ - w.tz, g.ntz);
      code = (sx < 2 && sy < 2 && sz < 2) ? ((sz + 1) * 3 + (sy + 1)) * 3 + (sx + 1) : 27;
    }
    const int slot = wave_reserve(code, w.s_nbr, a.next);
    if (code < 27) {
      const int dt = w.s_nbr[code], nb = w.s_nslot[code], nc = w.s_ncap[code];
      place_mx(a, nb, nc, slot, mx_pack(ix, iy, iz), ox, oy, oz, vx, vy, vz, pid);
      if (late) late_append_mx(a.late, a.late_cnt, a.ovf, g, dt, nb, nc, ix, iy, iz, ox, oy, oz, vx, vy, vz);
    } else if (code == 27) {
      // further than one tile: k_rebin_mx places it, and lists its deposit when owed
      const unsigned long long e = atomicAdd(&a.ctr[0], 1ull);
      store_mx(a.emi, (long)e, mx_pack(ix, iy, iz), ox, oy, oz, vx, vy, vz,
               late ? (int)((uint32_t)pid | EMI_DEPOSIT_MX) : pid);
    }
    if (valid && lost) atomicAdd(&a.ctr[2], 1ull);
  } else if ((MODE & M_PUSH) && valid) {
    uint32_t *b = mx_at(a.src, p);
    b[64 * MX_VX] = __float_as_uint(vx);
    b[64 * MX_VY] = __float_as_uint(vy);
    b[64 * MX_VZ] = __float_as_uint(vz);
  }
}

// The fused pass of workgroup wi of one run (arguments at ap): phase 1 pushes, decides
// collisions and finishes the ions without attempts; phases 2 and 3 collide the deferred
// ones; the deposit region goes to the run's fold buffer.
template <int MODE, int TX, int TY, int TZ, int HH>
__device__ __attribute__((always_inline)) inline void tile_mx(ConstPtr<TileArgsMx> ap, int wi) {
  static_assert(!(MODE & M_COLLIDE) || (MODE & M_REBIN), "collisions run only in the re-binning pass");
  const TileArgsMx &a = fresh(ap);
  extern __shared__ long long smemq[];
  __shared__ int s_nbr[27], s_nslot[27], s_ncap[27], s_ndef, s_n3;
  const TileGeom g = a.g;
  constexpr int NEX = TX + 1, NEY = TY + 1, NE3 = NEX * NEY * (TZ + 1);
  constexpr int RAX = TX + 1 + 2 * HH, RAY = TY + 1 + 2 * HH, RAZ = TZ + 1 + 2 * HH, RA3 = RAX * RAY * RAZ;
  TileWGMx w;
  w.t = wi / g.split;
  const int s = wi % g.split;
  w.tx = w.t % g.ntx;
  w.ty = (w.t / g.ntx) % g.nty;
  w.tz = w.t / (g.ntx * g.nty);
  w.i0 = w.tx * TX;
  w.j0 = w.ty * TY;
  w.k0 = w.tz * TZ;
  w.RAx = min(TX, g.Nx - w.i0) + 1 + 2 * HH;
  w.RAy = min(TY, g.Ny - w.j0) + 1 + 2 * HH;
  w.RAz = min(TZ, g.Nz - w.k0) + 1 + 2 * HH;
  w.s_nbr = s_nbr;
  w.s_nslot = s_nslot;
  w.s_ncap = s_ncap;
  // LDS: the accumulators first (8-byte aligned), then the gather's difference
  // numerators at the tile's (T+1)^3 corner nodes (FP64 differences rounded to FP32).
  w.sacc = smemq;
  float *sE = reinterpret_cast<float *>(smemq + ((MODE & M_DEPOSIT) ? 4 * RA3 : 0));
  const int tid = threadIdx.x;

  if (MODE & M_PUSH) {
    const PhiGlobal gphi{a.phi, g.Nx, g.Ny, g.Nz};
    for (int q = tid; q < NE3; q += blockDim.x) {
      const int li = q % NEX, lj = (q / NEX) % NEY, lk = q / (NEX * NEY);
      float d[3];
      node_diff<float>(gphi, g.Nx, g.Ny, g.Nz, w.i0 + li, w.j0 + lj, w.k0 + lk, d);
      sE[q] = d[0];
      sE[NE3 + q] = d[1];
      sE[2 * NE3 + q] = d[2];
    }
  }
  if (MODE & M_DEPOSIT) {
    for (int q = tid; q < 4 * RA3; q += blockDim.x) w.sacc[q] = 0;
  }
  if ((MODE & M_REBIN) && tid < 27) {
    const int sx = tid % 3 - 1, sy = (tid / 3) % 3 - 1, sz = tid / 9 - 1;
    const int nb = (wrap_index(w.tz + sz, g.ntz) * g.nty + wrap_index(w.ty + sy, g.nty)) * g.ntx +
                   wrap_index(w.tx + sx, g.ntx);
    s_nbr[tid] = nb;
    s_nslot[tid] = a.base[nb];
    s_ncap[tid] = a.cap[nb];
  }
  if (tid == 0) {
    s_ndef = 0;
    s_n3 = 0;
  }
  __syncthreads();

  const long cnt = a.count[w.t];
  const long beg = a.base[w.t] + cnt * s / g.split;
  const long end = a.base[w.t] + cnt * (s + 1) / g.split;
  unsigned natt = 0, nsucc = 0, nfail = 0;

  // Phase 1.  Two workgroups share a CU (64 VGPRs), so one's loads overlap the
  // other's arithmetic without a software prefetch.
  for (long p0 = beg; p0 < end; p0 += blockDim.x) {
    const TileArgsMx &A = fresh(ap);
    const long p = p0 + tid;
    const bool valid = p < end;
    uint32_t cell = 0;
    float ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
    int pid = 0;
    if (valid) load_mx(A.src, p, cell, ox, oy, oz, vx, vy, vz, pid);
    const int ix = mx_i(cell), iy = mx_j(cell), iz = mx_k(cell);
    if ((MODE & M_PUSH) && valid) {
      // BorisMoverBz3D body (src/velocities.c:342-356) in FP32.
      // Cells of a tile's ions lie in the tile and below N, so every corner node is in
      // the staged (T+1)^3 block.
      float E[3];
      gather_nodes<float, NEX, NEY, NE3>(sE, ((iz - w.k0) * NEY + (iy - w.j0)) * NEX + (ix - w.i0), ox, oy, oz, E);
      const float vxm = vx + A.k.Escale[0] * E[0];
      const float vym = vy + A.k.Escale[1] * E[1];
      const float vzm = vz + A.k.Escale[2] * E[2];
      const float vxt = vxm + vym * A.k.t;
      const float vyp = vym - vxt * A.k.s;
      const float vxp = vxt + vyp * A.k.t;
      vx = vxp;
      vy = vyp;
      vz = vzm;
    }
    bool defer = false;
    if ((MODE & M_COLLIDE) && valid && poisson_any(A.pt, A.key0, A.key1, A.step, pid)) {
      defer = true;
      const int q = atomicAdd(&s_ndef, 1);
      store_mx(A.def, beg + q, cell, ox, oy, oz, vx, vy, vz, pid);
    }
    tail_mx<MODE, TX, TY, TZ, HH>(A, w, valid && !defer, p, ix, iy, iz, ox, oy, oz, vx, vy, vz, pid);
  }
  if (MODE & M_COLLIDE) {
    // Phase 2: deferred ions, first attempt each (every lane busy).  The scattering
    // formula runs in FP32 on the same Philox draws as the FP64 path.  Ions with more
    // attempts (about 7% of them) are parked after the deferred ones in this
    // workgroup's slots of D and finish in phase 3, so no wave loops for a few lanes.
    __syncthreads();
    const int ndef = s_ndef;
    const long room = (end - beg) - ndef;
    for (int q0 = 0; q0 < ndef; q0 += blockDim.x) {
      const TileArgsMx &A = fresh(ap);
      const int q = q0 + tid;
      const bool valid = q < ndef;
      bool done = valid;
      uint32_t cell = 0;
      float ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
      int pid = 0;
      if (valid) {
        load_mx(A.def, beg + q, cell, ox, oy, oz, vx, vy, vz, pid);
        const int k = poisson_attempts(A.pt, A.key0, A.key1, A.step, pid);
        collide_attempts<float>(A.cc, A.key0, A.key1, A.step, pid, 0, 1, vx, vy, vz, natt, nsucc, nfail);
        if (k > 1) {
          const int r = atomicAdd(&s_n3, 1);
          if (r < room) {
            store_mx(A.def, beg + ndef + r, cell, ox, oy, oz, vx, vy, vz, pid);
            done = false;
          } else {
            collide_attempts<float>(A.cc, A.key0, A.key1, A.step, pid, 1, k, vx, vy, vz, natt, nsucc, nfail);
          }
        }
      }
      tail_mx<MODE, TX, TY, TZ, HH>(A, w, done, 0, mx_i(cell), mx_j(cell), mx_k(cell), ox, oy, oz, vx, vy, vz,
                                    pid);
    }
    // Phase 3: the remaining attempts of the parked ions.
    __syncthreads();
    const int n3 = (int)min((long)s_n3, room);
    for (int q0 = 0; q0 < n3; q0 += blockDim.x) {
      const TileArgsMx &A = fresh(ap);
      const int q = q0 + tid;
      const bool valid = q < n3;
      uint32_t cell = 0;
      float ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
      int pid = 0;
      if (valid) {
        load_mx(A.def, beg + ndef + q, cell, ox, oy, oz, vx, vy, vz, pid);
        const int k = poisson_attempts(A.pt, A.key0, A.key1, A.step, pid);
        collide_attempts<float>(A.cc, A.key0, A.key1, A.step, pid, 1, k, vx, vy, vz, natt, nsucc, nfail);
      }
      tail_mx<MODE, TX, TY, TZ, HH>(A, w, valid, 0, mx_i(cell), mx_j(cell), mx_k(cell), ox, oy, oz, vx, vy, vz,
                                    pid);
    }
    __shared__ unsigned long long s_red[16][3];
    block_add3(s_red, natt, nsucc, nfail, a.ctr + 3);
  }
  if (MODE & M_DEPOSIT) {
    __syncthreads();
    long long *out = a.fold + (size_t)wi * 4 * RA3;
    for (int q = tid; q < 4 * RA3; q += blockDim.x) out[q] = w.sacc[q];
  }
}

template <int MODE, int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(1024, 8) k_tile_mx(TileArgsMx) {
  tile_mx<MODE, TX, TY, TZ, HH>(kernarg<TileArgsMx>(), xcd_order(blockIdx.x, gridDim.x));
}

// One launch over the tiles of several runs of the same tile shape and block size.
template <int MODE, int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(1024, 8) k_tile_mx_runs(RunTable) {
  const ConstPtr<RunTable> t = kernarg<RunTable>();
  const int gi = xcd_order(blockIdx.x, gridDim.x);
  const int r = run_of(t, gi);
  tile_mx<MODE, TX, TY, TZ, HH>((ConstPtr<TileArgsMx>)t->args + r, gi - t->start[r]);
}

// Emigrants further than one tile -> destination tiles.
// Deposit of the ions listed for tile t (they arrived from beyond their source tile's
// deposit region and now lie in tile t): fixed-point LDS sums over the tile's
// (T+1)^3 nodes, added exactly into the tile's own slot of the fold buffer (its deposit
// region contains those nodes; no other workgroup touches the slot now).  Resets the
// list.
template <int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(256) k_late_mx(TileGeom g, const uint4 *__restrict__ late,
                                                int *__restrict__ late_cnt, const int *__restrict__ base,
                                                const int *__restrict__ cap, float inv_vmax,
                                                long long *__restrict__ fold, double *__restrict__ ovf) {
  constexpr int RAX = TX + 1 + 2 * HH, RAY = TY + 1 + 2 * HH, RA3 = RAX * RAY * (TZ + 1 + 2 * HH);
  constexpr int NX = TX + 1, NY = TY + 1, NZ = TZ + 1, N3 = NX * NY * NZ;
  __shared__ long long acc[4 * N3];
  const int t = blockIdx.x;
  const int n = min(late_cnt[tile_ctr(t)], cap[t] >> 2);
  __syncthreads();
  if (threadIdx.x == 0) late_cnt[tile_ctr(t)] = 0;
  if (n == 0) return;
  const int i0 = (t % g.ntx) * TX, j0 = ((t / g.ntx) % g.nty) * TY, k0 = (t / (g.ntx * g.nty)) * TZ;
  for (int q = threadIdx.x; q < 4 * N3; q += blockDim.x) acc[q] = 0;
  __syncthreads();
  const uint4 *list = late + 2 * (long)(base[t] >> 2);
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const uint4 e0 = list[2 * q], e1 = list[2 * q + 1];
    const uint32_t cell = e0.x;
    const float ox = __uint_as_float(e0.y), oy = __uint_as_float(e0.z), oz = __uint_as_float(e0.w);
    const float vx = __uint_as_float(e1.x), vy = __uint_as_float(e1.y), vz = __uint_as_float(e1.z);
    float w8[8];
    weights_mx(ox, oy, oz, w8);
    const float ux = vx * inv_vmax, uy = vy * inv_vmax, uz = vz * inv_vmax;
    const int la = mx_i(cell) - i0, lb = mx_j(cell) - j0, lc = mx_k(cell) - k0;
    if (la < 0 || la >= TX || lb < 0 || lb >= TY || lc < 0 || lc >= TZ) {  // not expected: listed ions lie in t
      double wd[8];
      for (int c = 0; c < 8; ++c) wd[c] = (double)w8[c] * (1.0 / 1073741824.0);
      deposit_overflow(ovf, g, mx_i(cell), mx_j(cell), mx_k(cell), wd, (double)vx, (double)vy, (double)vz);
      continue;
    }
    const int base0 = (lc * NY + lb) * NX + la;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ax = (c < 4) ? 1 : 0, ay = (c & 2) ? 0 : 1, az = (c & 1) ? 0 : 1;
      const int idx = base0 + (az * NY + ay) * NX + ax;
      const float wt = w8[c];
      lds_addw(&acc[idx], wt);
      lds_addq(&acc[N3 + idx], ux * wt);
      lds_addq(&acc[2 * N3 + idx], uy * wt);
      lds_addq(&acc[3 * N3 + idx], uz * wt);
    }
  }
  __syncthreads();
  long long *slot = fold + (size_t)t * g.split * 4 * RA3;
  for (int q = threadIdx.x; q < N3; q += blockDim.x) {
    const int li = q % NX, lj = (q / NX) % NY, lk = q / (NX * NY);
    const int r = ((lk + HH) * RAY + (lj + HH)) * RAX + (li + HH);
    for (int d = 0; d < 4; ++d) {
      const long long v = acc[d * N3 + q];
      if (v != 0) slot[d * RA3 + r] += v;
    }
  }
}

__global__ void k_rebin_mx(TileGeom g, ParticlesMix emi, ParticlesMix dst, ParticlesMix spill, long scap,
                           const int *base, const int *cap, int *next, unsigned long long *ctr, uint4 *late,
                           int *late_cnt, double *ovf) {
  const unsigned long long n = ctr[0];
  for (unsigned long long e = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; e < n;
       e += (unsigned long long)gridDim.x * blockDim.x) {
    uint32_t cell;
    float ox, oy, oz, vx, vy, vz;
    int tagged;
    load_mx(emi, (long)e, cell, ox, oy, oz, vx, vy, vz, tagged);
    const int id = (int)((uint32_t)tagged & ~EMI_DEPOSIT_MX);
    const int t = ((mx_k(cell) / g.Tz) * g.nty + mx_j(cell) / g.Ty) * g.ntx + mx_i(cell) / g.Tx;
    const int slot = atomicAdd(&next[tile_ctr(t)], 1);
    if (slot < cap[t]) {
      store_mx(dst, (long)base[t] + slot, cell, ox, oy, oz, vx, vy, vz, id);
    } else {
      const unsigned long long sp = atomicAdd(&ctr[1], 1ull);
      if ((long)sp < scap) store_mx(spill, (long)sp, cell, ox, oy, oz, vx, vy, vz, id);
      else atomicAdd(&ctr[2], 1ull);
    }
    if ((uint32_t)tagged & EMI_DEPOSIT_MX)
      late_append_mx(late, late_cnt, ovf, g, t, base[t], cap[t], mx_i(cell), mx_j(cell), mx_k(cell), ox, oy, oz, vx,
                     vy, vz);
  }
}

// AoS FP64 staging (pos[np][3], vel[np][3]) -> tile slots.  The cell is (int)(pos/d),
// as in the binning count, and is kept even when the FP64 remainder rounds up to 1 in
// FP32: the offset is clamped to the largest float below 1.  The reference's gather
// is discontinuous across cell faces (its x weight blends along z), so an ion must
// stay in the cell the FP64 path puts it in.
__global__ void k_bin_scatter_mx(TileGeom g, const double *stage, long np, const int *key, const int *base,
                                 ParticlesMix A, double dx, double dy, double dz) {
  const double *vel = stage + 3 * np;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    const int t = key[2 * p];
    if (t < 0) continue;
    const long slot = (long)base[t] + key[2 * p + 1];
    const double x = stage[p * 3] / dx, y = stage[p * 3 + 1] / dy, z = stage[p * 3 + 2] / dz;
    const int i = min((int)x, g.Nx - 1), j = min((int)y, g.Ny - 1), k = min((int)z, g.Nz - 1);
    const float below1 = 0x1.fffffep-1f;
    store_mx(A, slot, mx_pack(i, j, k), fminf((float)(x - i), below1), fminf((float)(y - j), below1),
             fminf((float)(z - k), below1), (float)vel[p * 3], (float)vel[p * 3 + 1], (float)vel[p * 3 + 2], (int)p);
  }
}

__device__ inline void unstage_mx(const ParticlesMix &A, long s, double *stage, double *vel, double dx, double dy,
                                  double dz) {
  uint32_t cell;
  float ox, oy, oz, vx, vy, vz;
  int pid;
  load_mx(A, s, cell, ox, oy, oz, vx, vy, vz, pid);
  const long id = pid;
  stage[id * 3] = ((double)mx_i(cell) + (double)ox) * dx;
  stage[id * 3 + 1] = ((double)mx_j(cell) + (double)oy) * dy;
  stage[id * 3 + 2] = ((double)mx_k(cell) + (double)oz) * dz;
  vel[id * 3] = vx;
  vel[id * 3 + 1] = vy;
  vel[id * 3 + 2] = vz;
}

__global__ void k_unbin_mx(ParticlesMix A, const int *base, const int *count, double *stage, long np, double dx,
                           double dy, double dz) {
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  for (long q = threadIdx.x; q < n; q += blockDim.x) unstage_mx(A, b + q, stage, stage + 3 * np, dx, dy, dz);
}

__global__ void k_unbin_list_mx(ParticlesMix S, long n, double *stage, long np, double dx, double dy, double dz) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
    unstage_mx(S, q, stage, stage + 3 * np, dx, dy, dz);
}

__global__ void k_cell_counts_mx(TileGeom g, ParticlesMix A, const int *base, const int *count, int *counts) {
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const uint32_t cell = mx_at(A, b + q)[64 * MX_CELL];
    atomicAdd(&counts[((long)mx_k(cell) * g.Ny + mx_j(cell)) * g.Nx + mx_i(cell)], 1);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Flux scale of the fixed-point deposit: the power of two at or above 64 times the
// largest thermal or drift speed in the run.
double mx_vmax(const ecsprt_params &p) {
  double v = std::max({p.viT, p.viTx, p.viTy, p.viTz, std::fabs(p.vi0x), std::fabs(p.vi0y), std::fabs(p.vi0z),
                       p.vnT, std::fabs(p.vn0x), std::fabs(p.vn0y), std::fabs(p.vn0z)});
  if (!(v > 0.0) || !std::isfinite(v)) v = 1.0;
  return std::exp2(std::ceil(std::log2(64.0 * v)));
}

static MxConsts mx_consts(const ecsprt_params &p, double dt) {
  const PushConsts d = push_consts(p, dt);
  MxConsts k;
  k.t = (float)d.t;
  k.s = (float)d.s;
  for (int i = 0; i < 3; ++i) k.Escale[i] = (float)d.Escale[i];
  k.dtc[0] = (float)(dt / p.dx);
  k.dtc[1] = (float)(dt / p.dy);
  k.dtc[2] = (float)(dt / p.dz);
  k.inv_vmax = (float)(1.0 / mx_vmax(p));
  return k;
}

void mx_free(ParticlesMix &P) {
  (void)hipFree(P.p);
  P = ParticlesMix{};
}

void mx_alloc(ParticlesMix &P, long n) {
  mx_free(P);
  const long blocks = (std::max(n, 1L) + 63) / 64;
  hipError_t e = hipMalloc(&P.p, sizeof(uint32_t) * 512 * blocks);
  if (e != hipSuccess) {
    P = ParticlesMix{};
    throw Error{ECSPRT_ERR_NOMEM, "particle arrays: " + std::string(hipGetErrorString(e))};
  }
}

template <int MODE, int TX, int TY, int TZ, int HH>
static void set_attr_mx() {
  // Once per kernel instance and process, also when contexts on several host threads
  // launch their first step at the same time.
  static std::once_flag once;
  std::call_once(once, [] {
    hipFuncAttributes fa;
    ECS_HIP(hipFuncGetAttributes(&fa, (const void *)k_tile_mx<MODE, TX, TY, TZ, HH>));
    const int dyn = 160 * 1024 - (int)fa.sharedSizeBytes;
    ECS_HIP(hipFuncSetAttribute((const void *)k_tile_mx<MODE, TX, TY, TZ, HH>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    if constexpr ((MODE & (M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN)) == (M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN)) {
      ECS_HIP(hipFuncGetAttributes(&fa, (const void *)k_tile_mx_runs<MODE, TX, TY, TZ, HH>));
      ECS_HIP(hipFuncSetAttribute((const void *)k_tile_mx_runs<MODE, TX, TY, TZ, HH>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - (int)fa.sharedSizeBytes));
    }
  });
}

template <int MODE>
static void set_attrs_mx() {
  set_attr_mx<MODE, 4, 4, 4, 2>();
  set_attr_mx<MODE, 8, 8, 4, 2>();
  set_attr_mx<MODE, 8, 8, 8, 2>();
}

void mx_set_lds_attrs() {
  set_attrs_mx<M_PUSH>();
  set_attrs_mx<M_DEPOSIT>();
  set_attrs_mx<M_DRIFT | M_REBIN>();
  set_attrs_mx<M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN>();
  set_attrs_mx<M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN>();
}

static size_t lds_mx(const TileGeom &g, int mode) {
  size_t b = 0;
  if (mode & M_PUSH) b += sizeof(float) * 3 * (g.Tx + 1) * (g.Ty + 1) * (g.Tz + 1);
  if (mode & M_DEPOSIT) b += sizeof(long long) * 4 * g.RAx * g.RAy * g.RAz;
  return b;
}

template <int MODE>
static void launch_mx(ecsprt_hip_ctx *c, const TileArgsMx &a) {
  const int block = c->block;
  const size_t lds = lds_mx(c->g, MODE);
  const int nwg = c->ntiles * c->g.split;
  const int tx = c->g.Tx, tz = c->g.Tz;
  if (tx == 8 && tz == 8) k_tile_mx<MODE, 8, 8, 8, 2><<<nwg, block, lds, c->stream>>>(a);
  else if (tx == 8 && tz == 4) k_tile_mx<MODE, 8, 8, 4, 2><<<nwg, block, lds, c->stream>>>(a);
  else k_tile_mx<MODE, 4, 4, 4, 2><<<nwg, block, lds, c->stream>>>(a);
  ECS_HIP(hipGetLastError());
}

// Arguments of a tile pass of context c (pt null: no collisions).
static TileArgsMx mx_args(ecsprt_hip_ctx *c, double dt, const PoissonParams *pt, uint32_t key0, uint32_t key1,
                          uint32_t step) {
  TileArgsMx a{};
  a.g = c->g;
  a.src = c->Am;
  a.dst = c->Bm;
  a.emi = c->Em;
  a.spill = c->Sm;
  a.def = c->Dm;
  a.scap = c->scap;
  a.base = c->d_base;
  a.count = c->d_count;
  a.cap = c->d_cap;
  a.next = c->d_next;
  a.ctr = c->d_ctr;
  a.phi = c->d_phi;
  a.fold = reinterpret_cast<long long *>(c->d_fold);
  a.ovf = c->d_ovf;
  a.late = reinterpret_cast<uint4 *>(c->d_late);
  a.late_cnt = c->d_late_cnt;
  a.k = mx_consts(c->p, dt);
  if (pt) {
    a.pt = *pt;
    a.cc = collide_consts_of(c->p);
  }
  a.key0 = key0;
  a.key1 = key1;
  a.step = step;
  return a;
}

void mx_launch(ecsprt_hip_ctx *c, int mode, double dt, const PoissonParams *pt, uint32_t key0, uint32_t key1,
               uint32_t step) {
  const TileArgsMx a = mx_args(c, dt, pt, key0, key1, step);
  switch (mode) {
    case M_PUSH: launch_mx<M_PUSH>(c, a); break;
    case M_DEPOSIT: launch_mx<M_DEPOSIT>(c, a); break;
    case M_DRIFT | M_REBIN: launch_mx<M_DRIFT | M_REBIN>(c, a); break;
    case M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN: launch_mx<M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN>(c, a); break;
    case M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN:
      launch_mx<M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN>(c, a);
      break;
    default: throw Error{ECSPRT_ERR_ARG, "unsupported tile pass"};
  }
}

template <int MODE, int TX, int TY, int TZ>
static void launch_mx_runs(hipStream_t st, const RunTable &t, int block, size_t lds) {
  k_tile_mx_runs<MODE, TX, TY, TZ, 2><<<t.start[t.nrun], block, lds, st>>>(t);
  ECS_HIP(hipGetLastError());
}

// The fused pass (with collisions when pt[0] is non-null; then every run has them) over
// the runs cs, which share the tile shape and block size, in one launch on stream st.
// The arguments are staged at h_args (pinned) and copied to d_args first.
void mx_launch_runs(hipStream_t st, const std::vector<ecsprt_hip_ctx *> &cs, double dt,
                    const std::vector<const PoissonParams *> &pt, const std::vector<uint32_t> &key0,
                    const std::vector<uint32_t> &key1, const std::vector<uint32_t> &step, void *h_args, void *d_args) {
  const int n = (int)cs.size();
  ECS_CHECK(n >= 1 && n <= RUNS_MAX, ECSPRT_ERR_ARG, "multi-run tile pass: 1 to 64 runs");
  TileArgsMx *h = static_cast<TileArgsMx *>(h_args);
  RunTable t{};
  t.args = d_args;
  t.nrun = n;
  long total = 0;
  for (int r = 0; r < n; ++r) {
    ecsprt_hip_ctx *c = cs[r];
    ECS_CHECK(c->g.Tx == cs[0]->g.Tx && c->g.Tz == cs[0]->g.Tz && c->block == cs[0]->block, ECSPRT_ERR_ARG,
              "multi-run tile pass: tile shapes or block sizes differ");
    h[r] = mx_args(c, dt, pt[r], key0[r], key1[r], step[r]);
    t.start[r] = (int)total;
    total += (long)c->ntiles * c->g.split;
  }
  ECS_CHECK(total < (1L << 31), ECSPRT_ERR_ARG, "multi-run tile pass: too many workgroups");
  t.start[n] = (int)total;
  ECS_HIP(hipMemcpyAsync(d_args, h, sizeof(TileArgsMx) * n, hipMemcpyHostToDevice, st));
  const bool coll = pt[0] != nullptr;
  const int mode = coll ? (M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN) : (M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN);
  const size_t lds = lds_mx(cs[0]->g, mode);
  const int block = cs[0]->block, tx = cs[0]->g.Tx, tz = cs[0]->g.Tz;
  constexpr int F = M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN, FC = F | M_COLLIDE;
  if (tx == 8 && tz == 8) {
    if (coll) launch_mx_runs<FC, 8, 8, 8>(st, t, block, lds);
    else launch_mx_runs<F, 8, 8, 8>(st, t, block, lds);
  } else if (tx == 8 && tz == 4) {
    if (coll) launch_mx_runs<FC, 8, 8, 4>(st, t, block, lds);
    else launch_mx_runs<F, 8, 8, 4>(st, t, block, lds);
  } else {
    if (coll) launch_mx_runs<FC, 4, 4, 4>(st, t, block, lds);
    else launch_mx_runs<F, 4, 4, 4>(st, t, block, lds);
  }
}

size_t mx_args_bytes() { return sizeof(TileArgsMx); }

// The Philox collide pass of the CollideFunc slot on mixed ions: the fused pass's
// phases 2 and 3 (FP32 scattering on the same draws, in this file's arithmetic), applied
// in place to every ion of every tile.  One workgroup per tile.
__global__ void __launch_bounds__(256) k_collide_philox_mx(ParticlesMix A, const int *__restrict__ base,
                                                           const int *__restrict__ count, CollideConsts cc,
                                                           PoissonParams pt, uint32_t key0, uint32_t key1,
                                                           uint32_t step, unsigned long long *ctr) {
  __syncthreads();
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  unsigned natt = 0, nsucc = 0, nfail = 0;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    uint32_t *w = mx_at(A, b + q);
    const int pid = (int)w[64 * MX_ID];
    const int k = poisson_attempts(pt, key0, key1, step, pid);
    if (k == 0) continue;
    float vx = __uint_as_float(w[64 * MX_VX]), vy = __uint_as_float(w[64 * MX_VY]), vz = __uint_as_float(w[64 * MX_VZ]);
    collide_attempts<float>(cc, key0, key1, step, pid, 0, k, vx, vy, vz, natt, nsucc, nfail);
    w[64 * MX_VX] = __float_as_uint(vx);
    w[64 * MX_VY] = __float_as_uint(vy);
    w[64 * MX_VZ] = __float_as_uint(vz);
  }
  const unsigned long long s0 = wave_sum(natt), s1 = wave_sum(nsucc), s2 = wave_sum(nfail);
  if (__lane_id() == 0) {
    if (s0) atomicAdd(&ctr[3], s0);
    if (s1) atomicAdd(&ctr[4], s1);
    if (s2) atomicAdd(&ctr[5], s2);
  }
}

void mx_collide_philox(ecsprt_hip_ctx *c, const PoissonParams &pt, const CollideConsts &cc, uint32_t key0,
                       uint32_t key1, uint32_t step) {
  k_collide_philox_mx<<<c->ntiles, 256, 0, c->stream>>>(c->Am, c->d_base, c->d_count, cc, pt, key0, key1, step,
                                                        c->d_ctr);
  ECS_HIP(hipGetLastError());
}

void mx_rebin(ecsprt_hip_ctx *c) {
  k_rebin_mx<<<2048, 256, 0, c->stream>>>(c->g, c->Em, c->Bm, c->Sm, c->scap, c->d_base, c->d_cap, c->d_next,
                                          c->d_ctr, reinterpret_cast<uint4 *>(c->d_late), c->d_late_cnt, c->d_ovf);
  ECS_HIP(hipGetLastError());
}

void mx_bin_scatter(ecsprt_hip_ctx *c, long np) {
  k_bin_scatter_mx<<<grid_for(np, 256), 256, 0, c->stream>>>(c->g, c->d_stage, np, c->d_key, c->d_base, c->Am,
                                                               c->p.dx, c->p.dy, c->p.dz);
  ECS_HIP(hipGetLastError());
}

void mx_unbin(ecsprt_hip_ctx *c) {
  k_unbin_mx<<<c->ntiles, 256, 0, c->stream>>>(c->Am, c->d_base, c->d_count, c->d_stage, c->np, c->p.dx, c->p.dy,
                                                c->p.dz);
  ECS_HIP(hipGetLastError());
}

void mx_unbin_list(ecsprt_hip_ctx *c, long n) {
  k_unbin_list_mx<<<grid_for(n, 256), 256, 0, c->stream>>>(c->Sm, n, c->d_stage, c->np, c->p.dx, c->p.dy, c->p.dz);
  ECS_HIP(hipGetLastError());
}

void mx_cell_counts(ecsprt_hip_ctx *c, int *d_counts) {
  k_cell_counts_mx<<<c->ntiles, 256, 0, c->stream>>>(c->g, c->Am, c->d_base, c->d_count, d_counts);
  ECS_HIP(hipGetLastError());
}

void mx_late(ecsprt_hip_ctx *c) {
  const float iv = (float)(1.0 / mx_vmax(c->p));
  long long *fq = reinterpret_cast<long long *>(c->d_fold);
  const uint4 *late = reinterpret_cast<const uint4 *>(c->d_late);
  const int tx = c->g.Tx, tz = c->g.Tz;
  if (tx == 8 && tz == 8)
    k_late_mx<8, 8, 8, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base,
                                                            c->d_cap, iv, fq, c->d_ovf);
  else if (tx == 8 && tz == 4)
    k_late_mx<8, 8, 4, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base,
                                                            c->d_cap, iv, fq, c->d_ovf);
  else
    k_late_mx<4, 4, 4, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base,
                                                            c->d_cap, iv, fq, c->d_ovf);
  ECS_HIP(hipGetLastError());
}

}  // namespace ecsprt
