// Particle stages of the hot path on gfx950: Boris push with LDS-staged field
// gather, drift + periodic wrap + re-binning, cloud-in-cell deposit into per-tile
// LDS accumulators with a deterministic fold, and the fused single-pass timestep.
//
// Particles are binned by tile (a T^3 brick of cells) in structure-of-arrays FP64
// storage, so a workgroup streams one tile's particles with coalesced loads while
// the potential it gathers from and the moments it deposits into sit in LDS.
#include <algorithm>
#include <numeric>
#include <type_traits>

#include <mutex>

#include "common.hpp"
#include "philox.hpp"
#include "pic_math.hpp"
#include "tiles.hpp"

namespace ecsprt {

struct TileArgs {
  TileGeom g;
  ParticlesF64 src, dst, emi, spill, def;
  long scap;
  PoissonParams pt;
  const int *base, *count, *cap;
  int *next;
  unsigned long long *ctr;
  const double *phi;
  long long *fold;  // per-workgroup deposit regions, fixed point (weights at fx_w, fluxes at fx_w / vmax)
  double *ovf;
  double fx_w, inv_vmax;
  double *late;   // per-tile lists of ions whose deposit waits for k_late: x, y, z, vx, vy, vz each
  int *late_cnt;  // list lengths
  PushConsts pc;
  DriftConsts dc;
  CollideConsts cc;
  uint32_t key0, key1, step;
};

__device__ inline void store_particle(const ParticlesF64 &d, long i, double x, double y, double z,
                                      double vx, double vy, double vz, int id) {
  d.x(i) = x;
  d.y(i) = y;
  d.z(i) = z;
  d.vx(i) = vx;
  d.vy(i) = vy;
  d.vz(i) = vz;
  d.id(i) = id;
}

// Place a particle whose destination slot was reserved: into the destination tile
// when there is room, else into the spill list (the host re-bins after the step).
__device__ inline void place(const TileArgs &a, int base, int cap, int slot, double x, double y, double z, double vx,
                             double vy, double vz, int id) {
  if (slot < cap) {
    store_particle(a.dst, (long)base + slot, x, y, z, vx, vy, vz, id);
    return;
  }
  const unsigned long long sp = atomicAdd(&a.ctr[1], 1ull);
  if ((long)sp < a.scap) store_particle(a.spill, (long)sp, x, y, z, vx, vy, vz, id);
  else atomicAdd(&a.ctr[2], 1ull);
}

// Deposit of one ion with global FP64 atomics (Collect3DFluidMoments weights): the
// rare ion whose tile's late list is full.
__device__ inline void deposit_direct(double *ovf, const TileGeom &g, const PushConsts &pc, double px, double py,
                                      double pz, double vx, double vy, double vz) {
  const double x = cell_coord(px, pc.dx, pc.inv[0], pc.pow2);
  const double y = cell_coord(py, pc.dy, pc.inv[1], pc.pow2);
  const double z = cell_coord(pz, pc.dz, pc.inv[2], pc.pow2);
  const int ixl = (int)x, iyl = (int)y, izl = (int)z;
  const double wxh = x - (double)ixl, wxl = 1.0 - wxh;
  const double wyh = y - (double)iyl, wyl = 1.0 - wyh;
  const double wzh = z - (double)izl, wzl = 1.0 - wzh;
  const double w8[8] = {wzh * wyh * wxh, wzl * wyh * wxh, wzh * wyl * wxh, wzl * wyl * wxh,
                        wzh * wyh * wxl, wzl * wyh * wxl, wzh * wyl * wxl, wzl * wyl * wxl};
  deposit_overflow(ovf, g, ixl, iyl, izl, w8, vx, vy, vz);
}

// An ion that left its source tile's deposit region lies in its destination tile dt:
// append its position and velocity to dt's list (one atomic, 48 contiguous bytes) for
// k_late, which adds it in the exact fixed point of the tile regions, or deposit it now
// when the tile's list is full.
__device__ inline void late_append(double *late, int *late_cnt, double *ovf, const TileGeom &g, const PushConsts &pc,
                                   int dt, int base, int cap, double px, double py, double pz, double vx, double vy,
                                   double vz) {
  const int li = atomicAdd(&late_cnt[tile_ctr(dt)], 1);
  if (li < (cap >> 2)) {
    double2 *e = reinterpret_cast<double2 *>(late + 6 * ((long)(base >> 2) + li));
    e[0] = make_double2(px, py);
    e[1] = make_double2(pz, vx);
    e[2] = make_double2(vy, vz);
  } else {
    deposit_direct(ovf, g, pc, px, py, pz, vx, vy, vz);
  }
}

// Bit 31 of an emigrant's id marks an ion whose deposit is still owed: k_rebin lists it
// in its destination tile (ids are below 2^31 - 1).
constexpr int EMI_DEPOSIT = (int)0x80000000u;

// Per-workgroup tile constants and LDS pointers used by the particle tail.
struct TileWG {
  int t, tx, ty, tz, i0, j0, k0;
  int RA, RA3, RAx, RAy, RAz, H;
  long long *sacc;
  int *s_nbr;             // the 27 neighbour tiles
  int *s_nslot, *s_ncap;  // slot offset and capacity of each neighbour tile
};

// Drift + wrap, deposit, and re-binning of one particle per thread.  Every lane of a
// wave calls this in uniform control flow (re-binning uses wave-wide ballots).
template <int MODE, int TX, int TY, int TZ, int HH>
__device__ inline void particle_tail(const TileArgs &a, TileWG &w, bool valid, long p, double px, double py,
                                     double pz, double vx, double vy, double vz, int pid) {
  const TileGeom &g = a.g;
  const double dx = a.pc.dx, dy = a.pc.dy, dz = a.pc.dz;
  bool lost = false, late = false;
  if ((MODE & M_DRIFT) && valid) {
    // UpdatePositions drift (src/positions.c:455-459) and ApplyBC (src/boundaries.c:19-28).
    px += vx * a.dc.dt;
    py += vy * a.dc.dt;
    pz += vz * a.dc.dt;
    if (px < a.dc.r0[0]) px += a.dc.L[0];
    if (px >= a.dc.r1[0]) px -= a.dc.L[0];
    if (py < a.dc.r0[1]) py += a.dc.L[1];
    if (py >= a.dc.r1[1]) py -= a.dc.L[1];
    if (pz < a.dc.r0[2]) pz += a.dc.L[2];
    if (pz >= a.dc.r1[2]) pz -= a.dc.L[2];
    lost = !(px >= a.dc.r0[0] && px <= a.dc.r1[0] && py >= a.dc.r0[1] && py <= a.dc.r1[1] && pz >= a.dc.r0[2] &&
             pz <= a.dc.r1[2]);
  }
  // Normalised coordinates, x = pos/dx (src/moments.c:154-156), shared by the tile
  // lookup and the deposit.
  const double x = cell_coord(px, dx, a.pc.inv[0], a.pc.pow2);
  const double y = cell_coord(py, dy, a.pc.inv[1], a.pc.pow2);
  const double z = cell_coord(pz, dz, a.pc.inv[2], a.pc.pow2);
  // Cell and tile of the new position (a coordinate equal to N stays in the last cell).
  const int ci = min((int)x, g.Nx - 1), cj = min((int)y, g.Ny - 1), ck = min((int)z, g.Nz - 1);
  if ((MODE & M_DEPOSIT) && valid && !lost) {
    // Collect3DFluidMoments body, src/moments.c:154-205.
    int ixl = (int)x, iyl = (int)y, izl = (int)z;
    double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    double wzh = z - (double)izl, wzl = 1.0 - wzh;
    double w8[8];
    w8[0] = wzh * wyh * wxh;
    w8[1] = wzl * wyh * wxh;
    w8[2] = wzh * wyl * wxh;
    w8[3] = wzl * wyl * wxh;
    w8[4] = wzh * wyh * wxl;
    w8[5] = wzl * wyh * wxl;
    w8[6] = wzh * wyl * wxl;
    w8[7] = wzl * wyl * wxl;
    constexpr int RAX = TX + 1 + 2 * HH, RAY = TY + 1 + 2 * HH, RAZ = TZ + 1 + 2 * HH, RA3 = RAX * RAY * RAZ;
    int la = ixl - (w.i0 - HH), lb = iyl - (w.j0 - HH), lc = izl - (w.k0 - HH);
    if (la < 0) la += g.Nx; else if (la > w.RAx - 2) la -= g.Nx;
    if (lb < 0) lb += g.Ny; else if (lb > w.RAy - 2) lb -= g.Ny;
    if (lc < 0) lc += g.Nz; else if (lc > w.RAz - 2) lc -= g.Nz;
    const bool inreg = la >= 0 && la <= w.RAx - 2 && lb >= 0 && lb <= w.RAy - 2 && lc >= 0 && lc <= w.RAz - 2;
    // Weights and fluxes go into LDS as exact 64-bit fixed point: weight w as
    // rint(w * fx_w), flux v*w as rint(v/vmax * w * fx_w) (vmax and fx_w are powers of
    // two, so the scalings are exact).  Ions with a component at or above vmax keep FP64
    // global atomics.
    const double ux = vx * a.inv_vmax, uy = vy * a.inv_vmax, uz = vz * a.inv_vmax;
    const bool slow = fabs(ux) < 1.0 && fabs(uy) < 1.0 && fabs(uz) < 1.0;
    if (inreg && slow) {
      const int base = (lc * RAY + lb) * RAX + la;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int ax = (c < 4) ? 1 : 0, ay = (c & 2) ? 0 : 1, az = (c & 1) ? 0 : 1;
        const int idx = base + (az * RAY + ay) * RAX + ax;
        const double wt = w8[c] * a.fx_w;
        lds_addfx(&w.sacc[idx], wt);
        lds_addfx(&w.sacc[RA3 + idx], ux * wt);
        lds_addfx(&w.sacc[2 * RA3 + idx], uy * wt);
        lds_addfx(&w.sacc[3 * RA3 + idx], uz * wt);
      }
    } else if ((MODE & M_REBIN) && slow) {
      late = true;
    } else {
      deposit_overflow(a.ovf, g, ixl, iyl, izl, w8, vx, vy, vz);
    }
  }
  if (MODE & M_REBIN) {
    // Destination code: neighbour index 0..26, 27 = further than one tile, 28 = none.
    int code = 28;
    if (valid && !lost) {
      const int ox = tile_step(ci / TX - w.tx, g.ntx), oy = tile_step(cj / TY - w.ty, g.nty);
      const int oz = tile_step(ck / TZ - w.tz, g.ntz);
      code = (ox < 2 && oy < 2 && oz < 2) ? ((oz + 1) * 3 + (oy + 1)) * 3 + (ox + 1) : 27;
    }
    const int slot = wave_reserve(code, w.s_nbr, a.next);
    if (code < 27) {
      const int nb = w.s_nslot[code], nc = w.s_ncap[code];
      place(a, nb, nc, slot, px, py, pz, vx, vy, vz, pid);
      if (late) late_append(a.late, a.late_cnt, a.ovf, g, a.pc, w.s_nbr[code], nb, nc, px, py, pz, vx, vy, vz);
    } else if (code == 27) {
      // further than one tile: k_rebin places it, and lists its deposit when owed
      const unsigned long long e = atomicAdd(&a.ctr[0], 1ull);
      store_particle(a.emi, (long)e, px, py, pz, vx, vy, vz, late ? (pid | EMI_DEPOSIT) : pid);
    }
    if (valid && lost) atomicAdd(&a.ctr[2], 1ull);
  } else if ((MODE & M_PUSH) && valid) {
    a.src.vx(p) = vx;
    a.src.vy(p) = vy;
    a.src.vz(p) = vz;
  }
}

// The fused pass of workgroup wi of one run (arguments at ap): phase 1 pushes, decides
// collisions and finishes the ions without attempts; phases 2 and 3 collide the deferred
// ones; the deposit region goes to the run's fold buffer.
template <int MODE, int TX, int TY, int TZ, int HH>
__device__ __attribute__((always_inline)) inline void tile_f64(ConstPtr<TileArgs> ap, int wi) {
  static_assert(!(MODE & M_COLLIDE) || (MODE & M_REBIN), "collisions run only in the re-binning pass");
  const TileArgs &a = fresh(ap);
  extern __shared__ double smem[];
  // Re-binning bookkeeping: particles per destination among the 27 neighbour tiles,
  // the slot range reserved in each; number of ions deferred for collisions.
  __shared__ int s_nbr[27], s_nslot[27], s_ncap[27], s_ndef, s_n3;
  const TileGeom g = a.g;
  constexpr int NEX = TX + 1, NEY = TY + 1, NE3 = NEX * NEY * (TZ + 1);
  constexpr int RAX = TX + 1 + 2 * HH, RAY = TY + 1 + 2 * HH, RAZ = TZ + 1 + 2 * HH;
  TileWG w;
  w.t = wi / g.split;
  const int s = wi % g.split;
  w.tx = w.t % g.ntx;
  w.ty = (w.t / g.ntx) % g.nty;
  w.tz = w.t / (g.ntx * g.nty);
  w.i0 = w.tx * TX;
  w.j0 = w.ty * TY;
  w.k0 = w.tz * TZ;
  w.H = HH;
  w.RA = RAX;
  w.RA3 = RAX * RAY * RAZ;
  w.RAx = min(TX, g.Nx - w.i0) + 1 + 2 * HH;
  w.RAy = min(TY, g.Ny - w.j0) + 1 + 2 * HH;
  w.RAz = min(TZ, g.Nz - w.k0) + 1 + 2 * HH;
  w.s_nbr = s_nbr;
  w.s_nslot = s_nslot;
  w.s_ncap = s_ncap;
  // LDS: the gather's difference numerators at the tile's (T+1)^3 corner nodes, three
  // components, then the deposit region.
  double *sE = smem;
  w.sacc = reinterpret_cast<long long *>(smem + ((MODE & M_PUSH) ? 3 * NE3 : 0));
  const int tid = threadIdx.x;
  const PhiGlobal gphi{a.phi, g.Nx, g.Ny, g.Nz};

  if (MODE & M_PUSH) {
    for (int q = tid; q < NE3; q += blockDim.x) {
      const int li = q % NEX, lj = (q / NEX) % NEY, lk = q / (NEX * NEY);
      double d[3];
      node_diff<double>(gphi, g.Nx, g.Ny, g.Nz, w.i0 + li, w.j0 + lj, w.k0 + lk, d);
      sE[q] = d[0];
      sE[NE3 + q] = d[1];
      sE[2 * NE3 + q] = d[2];
    }
  }
  if (MODE & M_DEPOSIT) {
    for (int q = tid; q < 4 * w.RA3; q += blockDim.x) w.sacc[q] = 0;
  }
  if ((MODE & M_REBIN) && tid < 27) {
    const int ox = tid % 3 - 1, oy = (tid / 3) % 3 - 1, oz = tid / 9 - 1;
    const int nb = (wrap_index(w.tz + oz, g.ntz) * g.nty + wrap_index(w.ty + oy, g.nty)) * g.ntx +
                   wrap_index(w.tx + ox, g.ntx);
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

  // Phase 1.  Software pipeline: the next batch's particles load while this one computes.
  double nxp = 0, nyp = 0, nzp = 0, nvx = 0, nvy = 0, nvz = 0;
  int nid = 0;
  if (beg + tid < end) {
    const long p = beg + tid;
    nxp = a.src.x(p);
    nyp = a.src.y(p);
    nzp = a.src.z(p);
    nvx = a.src.vx(p);
    nvy = a.src.vy(p);
    nvz = a.src.vz(p);
    nid = a.src.id(p);
  }
  for (long p0 = beg; p0 < end; p0 += blockDim.x) {
    const TileArgs &A = fresh(ap);
    const long p = p0 + tid;
    const bool valid = p < end;
    double px = nxp, py = nyp, pz = nzp, vx = nvx, vy = nvy, vz = nvz;
    const int pid = nid;
    {
      const long pn = p + blockDim.x;
      if (pn < end) {
        nxp = A.src.x(pn);
        nyp = A.src.y(pn);
        nzp = A.src.z(pn);
        nvx = A.src.vx(pn);
        nvy = A.src.vy(pn);
        nvz = A.src.vz(pn);
        nid = A.src.id(pn);
      }
    }
    if ((MODE & M_PUSH) && valid) {
      // BorisMoverBz3D body, src/velocities.c:342-356.
      // DifferenceVector3D at x = pos/d: il = (int)x, w = x - il (vector-math.c:607-617).
      double E[3];
      const double x0 = cell_coord(px, A.pc.dx, A.pc.inv[0], A.pc.pow2);
      const double y0 = cell_coord(py, A.pc.dy, A.pc.inv[1], A.pc.pow2);
      const double z0 = cell_coord(pz, A.pc.dz, A.pc.inv[2], A.pc.pow2);
      const int ixl = (int)x0, iyl = (int)y0, izl = (int)z0;
      const int la = ixl - w.i0, lb = iyl - w.j0, lc = izl - w.k0;
      if ((unsigned)la < TX && (unsigned)lb < TY && (unsigned)lc < TZ && ixl < g.Nx && iyl < g.Ny && izl < g.Nz)
        gather_nodes<double, NEX, NEY, NE3>(sE, (lc * NEY + lb) * NEX + la, x0 - (double)ixl, y0 - (double)iyl,
                                            z0 - (double)izl, E);
      else  // a coordinate that rounded to N: the node differences differ by corner role
        diff_vec3d_core<double>(PhiGlobal{A.phi, g.Nx, g.Ny, g.Nz}, g.Nx, g.Ny, g.Nz, ixl, iyl, izl,
                                x0 - (double)ixl, y0 - (double)iyl, z0 - (double)izl, E);
      double vxm = vx + A.pc.Escale[0] * E[0];
      double vym = vy + A.pc.Escale[1] * E[1];
      double vzm = vz + A.pc.Escale[2] * E[2];
      double vxt = vxm + vym * A.pc.t;
      double vyp = vym - vxt * A.pc.s;
      double vxp = vxt + vyp * A.pc.t;
      vx = vxp;
      vy = vyp;
      vz = vzm;
    }
    bool defer = false;
    if ((MODE & M_COLLIDE) && valid && poisson_any(A.pt, A.key0, A.key1, A.step, pid)) {
      // About one ion in seven has a collision attempt this step.  It is set aside
      // (post-push state, in this workgroup's own slot range) and finished in
      // phase 2 with every lane busy.
      defer = true;
      const int q = atomicAdd(&s_ndef, 1);
      store_particle(A.def, beg + q, px, py, pz, vx, vy, vz, pid);
    }
    particle_tail<MODE, TX, TY, TZ, HH>(A, w, valid && !defer, p, px, py, pz, vx, vy, vz, pid);
  }
  if (MODE & M_COLLIDE) {
    // Phase 2: the deferred ions, first attempt each (every lane busy), then drifted,
    // deposited and re-binned.  Ions with more attempts are parked after the deferred
    // ones in this workgroup's slots of D and finish in phase 3.
    __syncthreads();
    const int ndef = s_ndef;
    const long room = (end - beg) - ndef;
    for (int q0 = 0; q0 < ndef; q0 += blockDim.x) {
      const TileArgs &A = fresh(ap);
      const int q = q0 + tid;
      const bool valid = q < ndef;
      bool done = valid;
      double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
      int pid = 0;
      if (valid) {
        const long d = beg + q;
        px = A.def.x(d);
        py = A.def.y(d);
        pz = A.def.z(d);
        vx = A.def.vx(d);
        vy = A.def.vy(d);
        vz = A.def.vz(d);
        pid = A.def.id(d);
        const int k = poisson_attempts(A.pt, A.key0, A.key1, A.step, pid);
        collide_attempts<double>(A.cc, A.key0, A.key1, A.step, pid, 0, 1, vx, vy, vz, natt, nsucc, nfail);
        if (k > 1) {
          const int r = atomicAdd(&s_n3, 1);
          if (r < room) {
            store_particle(A.def, beg + ndef + r, px, py, pz, vx, vy, vz, pid);
            done = false;
          } else {
            collide_attempts<double>(A.cc, A.key0, A.key1, A.step, pid, 1, k, vx, vy, vz, natt, nsucc, nfail);
          }
        }
      }
      particle_tail<MODE, TX, TY, TZ, HH>(A, w, done, 0, px, py, pz, vx, vy, vz, pid);
    }
    // Phase 3: the remaining attempts of the parked ions.
    __syncthreads();
    const int n3 = (int)min((long)s_n3, room);
    for (int q0 = 0; q0 < n3; q0 += blockDim.x) {
      const TileArgs &A = fresh(ap);
      const int q = q0 + tid;
      const bool valid = q < n3;
      double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
      int pid = 0;
      if (valid) {
        const long d = beg + ndef + q;
        px = A.def.x(d);
        py = A.def.y(d);
        pz = A.def.z(d);
        vx = A.def.vx(d);
        vy = A.def.vy(d);
        vz = A.def.vz(d);
        pid = A.def.id(d);
        const int k = poisson_attempts(A.pt, A.key0, A.key1, A.step, pid);
        collide_attempts<double>(A.cc, A.key0, A.key1, A.step, pid, 1, k, vx, vy, vz, natt, nsucc, nfail);
      }
      particle_tail<MODE, TX, TY, TZ, HH>(A, w, valid, 0, px, py, pz, vx, vy, vz, pid);
    }
    __shared__ unsigned long long s_red[16][3];
    block_add3(s_red, natt, nsucc, nfail, a.ctr + 3);
  }
  if (MODE & M_DEPOSIT) {
    __syncthreads();
    long long *out = a.fold + (size_t)wi * 4 * w.RA3;
    for (int q = tid; q < 4 * w.RA3; q += blockDim.x) out[q] = w.sacc[q];
  }
}

template <int MODE, int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(1024) k_tile(TileArgs) {
  tile_f64<MODE, TX, TY, TZ, HH>(kernarg<TileArgs>(), xcd_order(blockIdx.x, gridDim.x));
}

// One launch over the tiles of several runs of the same tile shape and block size.
template <int MODE, int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(1024) k_tile_runs(RunTable) {
  const ConstPtr<RunTable> t = kernarg<RunTable>();
  const int gi = xcd_order(blockIdx.x, gridDim.x);
  const int r = run_of(t, gi);
  tile_f64<MODE, TX, TY, TZ, HH>((ConstPtr<TileArgs>)t->args + r, gi - t->start[r]);
}

// Deposit of the ions listed for tile t: those that arrived from beyond their source
// tile's deposit region.  They now lie in tile t, so fixed-point LDS accumulators over
// its (T+1)^3 nodes take them (the tile kernel's units), and the non-zero nodes are
// added into the tile's own slot of the fold buffer (its deposit region contains those
// nodes; no other workgroup touches the slot now), so k_fold sums them with everything
// else.  Resets the tile's list.
template <int TX, int TY, int TZ, int HH>
__global__ void __launch_bounds__(256) k_late(TileGeom g, const double *__restrict__ late,
                                             int *__restrict__ late_cnt, const int *__restrict__ base,
                                             const int *__restrict__ cap, PushConsts pc, double *__restrict__ ovf,
                                             long long *__restrict__ fold, double fx_w, double inv_vmax) {
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
  const double2 *list = reinterpret_cast<const double2 *>(late + 6 * (long)(base[t] >> 2));
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const double2 e0 = list[3 * q], e1 = list[3 * q + 1], e2 = list[3 * q + 2];
    const double px = e0.x, py = e0.y, pz = e1.x, vx = e1.y, vy = e2.x, vz = e2.y;
    const double x = cell_coord(px, pc.dx, pc.inv[0], pc.pow2);
    const double y = cell_coord(py, pc.dy, pc.inv[1], pc.pow2);
    const double z = cell_coord(pz, pc.dz, pc.inv[2], pc.pow2);
    const int ixl = (int)x, iyl = (int)y, izl = (int)z;
    const double wxh = x - (double)ixl, wxl = 1.0 - wxh;
    const double wyh = y - (double)iyl, wyl = 1.0 - wyh;
    const double wzh = z - (double)izl, wzl = 1.0 - wzh;
    const double w8[8] = {wzh * wyh * wxh, wzl * wyh * wxh, wzh * wyl * wxh, wzl * wyl * wxh,
                          wzh * wyh * wxl, wzl * wyh * wxl, wzh * wyl * wxl, wzl * wyl * wxl};
    const int la = ixl - i0, lb = iyl - j0, lc = izl - k0;
    if (la < 0 || la >= TX || lb < 0 || lb >= TY || lc < 0 || lc >= TZ) {
      deposit_overflow(ovf, g, ixl, iyl, izl, w8, vx, vy, vz);  // a coordinate on the upper face
      continue;
    }
    const double ux = vx * inv_vmax, uy = vy * inv_vmax, uz = vz * inv_vmax;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ax = (c < 4) ? 1 : 0, ay = (c & 2) ? 0 : 1, az = (c & 1) ? 0 : 1;
      const int idx = ((lc + az) * NY + (lb + ay)) * NX + (la + ax);
      const double wt = w8[c] * fx_w;
      lds_addfx(&acc[idx], wt);
      lds_addfx(&acc[N3 + idx], ux * wt);
      lds_addfx(&acc[2 * N3 + idx], uy * wt);
      lds_addfx(&acc[3 * N3 + idx], uz * wt);
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

// Emigrants -> destination tiles (slots after the stayers), spill when a tile is full;
// an emigrant whose deposit is owed joins its destination tile's late list.
__global__ void k_rebin(TileGeom g, ParticlesF64 emi, ParticlesF64 dst, ParticlesF64 spill, long scap,
                        const int *base, const int *cap, int *next, unsigned long long *ctr, PushConsts pc,
                        double *late, int *late_cnt, double *ovf) {
  const unsigned long long n = ctr[0];
  for (unsigned long long e = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; e < n;
       e += (unsigned long long)gridDim.x * blockDim.x) {
    const double x = emi.x(e), y = emi.y(e), z = emi.z(e), vx = emi.vx(e), vy = emi.vy(e), vz = emi.vz(e);
    const int tagged = emi.id(e), id = tagged & ~EMI_DEPOSIT;
    const int t = tile_of(g, x / pc.dx, y / pc.dy, z / pc.dz);
    const int slot = atomicAdd(&next[tile_ctr(t)], 1);
    if (slot < cap[t]) {
      store_particle(dst, (long)base[t] + slot, x, y, z, vx, vy, vz, id);
    } else {
      unsigned long long sp = atomicAdd(&ctr[1], 1ull);
      if ((long)sp < scap) store_particle(spill, (long)sp, x, y, z, vx, vy, vz, id);
      else atomicAdd(&ctr[2], 1ull);
    }
    if (tagged & EMI_DEPOSIT) late_append(late, late_cnt, ovf, g, pc, t, base[t], cap[t], x, y, z, vx, vy, vz);
  }
}

__global__ void k_finalize_counts(int ntiles, int *count, int *next, const int *cap) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles) {
    count[t] = min(next[tile_ctr(t)], cap[t]);
    next[tile_ctr(t)] = 0;
  }
}

// Sum of every workgroup region covering a node, in a fixed order, plus the
// overflow deposits: the moments in FluidNode layout.
// R = double for the FP64 path.  R = long long for the mixed path's fixed-point
// regions: the integer sum is exact, then scaled to weight (sn) and flux (sg) units.
template <typename R>
__global__ void k_fold(TileGeom g, const R *__restrict__ fold, const double *__restrict__ ovf,
                       const int *__restrict__ tab, int offx, int offy, int offz, double *__restrict__ m, double sn,
                       double sg) {
  const long nnode = (long)g.Nx * g.Ny * g.Nz;
  const int RAX = g.RAx, RAY = g.RAy, RA3 = g.RAx * g.RAy * g.RAz;
  for (long node = blockIdx.x * (long)blockDim.x + threadIdx.x; node < nnode; node += (long)gridDim.x * blockDim.x) {
    const int i = node % g.Nx, j = (node / g.Nx) % g.Ny, k = node / ((long)g.Nx * g.Ny);
    const int *tx = tab + offx, *ty = tab + offy, *tz = tab + offz;
    const int cx = tx[i], cy = ty[j], cz = tz[k];
    const int *px = tx + g.Nx + i * FOLD_MAXP * 2;
    const int *py = ty + g.Ny + j * FOLD_MAXP * 2;
    const int *pz = tz + g.Nz + k * FOLD_MAXP * 2;
    // Exact integer sum of the regions in two words (high parts signed, low 32-bit parts
    // unsigned), rounded to FP64 once at the end: the moments do not depend on how ions
    // were split among workgroups or ordered within them.
    long long hi[4] = {0, 0, 0, 0};
    unsigned long long lo[4] = {0, 0, 0, 0};
    for (int a = 0; a < cz; ++a)
      for (int b = 0; b < cy; ++b)
        for (int c = 0; c < cx; ++c) {
          const int t = (pz[2 * a] * g.nty + py[2 * b]) * g.ntx + px[2 * c];
          const int loc = (pz[2 * a + 1] * RAY + py[2 * b + 1]) * RAX + px[2 * c + 1];
          for (int s = 0; s < g.split; ++s) {
            const R *r = fold + (size_t)(t * g.split + s) * 4 * RA3 + loc;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
              const long long v = r[d * RA3];
              hi[d] += v >> 32;
              lo[d] += (unsigned long long)(v & 0xffffffffLL);
            }
          }
        }
    double acc[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const long long h = hi[d] + (long long)(lo[d] >> 32);
      acc[d] = (double)h * 4294967296.0 + (double)(lo[d] & 0xffffffffULL);
    }
    m[node * 4 + 0] = acc[0] * sn + ovf[node * 4 + 0];
    for (int d = 1; d < 4; ++d) m[node * 4 + d] = acc[d] * sg + ovf[node * 4 + d];
  }
}

__global__ void k_fill_f64(double *p, long n, double v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int *p, long n, int v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// Binning from an AoS staging buffer (pos[np][3] then vel[np][3]); staging index = id.
__global__ void k_bin_count(TileGeom g, const double *stage, long np, double dx, double dy, double dz, DriftConsts dc,
                            int *key, int *cnt, unsigned long long *ctr) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    double x = stage[p * 3], y = stage[p * 3 + 1], z = stage[p * 3 + 2];
    bool in = x >= dc.r0[0] && x <= dc.r1[0] && y >= dc.r0[1] && y <= dc.r1[1] && z >= dc.r0[2] && z <= dc.r1[2];
    if (in) {
      int t = tile_of(g, x / dx, y / dy, z / dz);
      key[2 * p] = t;
      key[2 * p + 1] = atomicAdd(&cnt[t], 1);
    } else {
      key[2 * p] = -1;
      atomicAdd(&ctr[2], 1ull);
    }
  }
}

__global__ void k_bin_scatter(const double *stage, long np, const int *key, const int *base, ParticlesF64 A) {
  const double *vel = stage + 3 * np;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    int t = key[2 * p];
    if (t < 0) continue;
    long slot = (long)base[t] + key[2 * p + 1];
    store_particle(A, slot, stage[p * 3], stage[p * 3 + 1], stage[p * 3 + 2], vel[p * 3], vel[p * 3 + 1],
                   vel[p * 3 + 2], (int)p);
  }
}

__global__ void k_unbin(ParticlesF64 A, const int *base, const int *count, double *stage, long np) {
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  double *vel = stage + 3 * np;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    long s = b + q;
    long id = A.id(s);
    stage[id * 3] = A.x(s);
    stage[id * 3 + 1] = A.y(s);
    stage[id * 3 + 2] = A.z(s);
    vel[id * 3] = A.vx(s);
    vel[id * 3 + 1] = A.vy(s);
    vel[id * 3 + 2] = A.vz(s);
  }
}

__global__ void k_unbin_list(ParticlesF64 S, long n, double *stage, long np) {
  double *vel = stage + 3 * np;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    long id = S.id(q);
    stage[id * 3] = S.x(q);
    stage[id * 3 + 1] = S.y(q);
    stage[id * 3 + 2] = S.z(q);
    vel[id * 3] = S.vx(q);
    vel[id * 3 + 1] = S.vy(q);
    vel[id * 3 + 2] = S.vz(q);
  }
}

// Initial conditions on the device: uniform y, z; x from the inverse CDF of
// 1 + amp*sin(2*pi*k*x/L) (Newton iterations from the uniform guess); Maxwellian
// velocities from Box-Muller pairs.  Counter = (particle index, 0x1C1C, 0, 0).
__global__ void k_init(double *stage, long np, double Lx, double Ly, double Lz, double x0, double y0, double z0,
                       double amp, int kk, double vtx, double vty, double vtz, double v0x, double v0y, double v0z,
                       uint32_t key0, uint32_t key1) {
  double *vel = stage + 3 * np;
  const double twopi = 6.283185307179586;
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    u32x4 c0 = {{(uint32_t)p, (uint32_t)(p >> 32), 0x1C1Cu, 0u}};
    u32x4 c1 = {{(uint32_t)p, (uint32_t)(p >> 32), 0x1C1Cu, 1u}};
    u32x4 r0 = philox4x32_10(c0, key0, key1);
    u32x4 r1 = philox4x32_10(c1, key0, key1);
    // 53-bit uniforms in [0, 1) for positions.
    double ux = ((double)(r0.v[0] >> 5) * 67108864.0 + (double)(r0.v[1] >> 6)) * (1.0 / 9007199254740992.0);
    double uy = ((double)(r0.v[2] >> 5) * 67108864.0 + (double)(r0.v[3] >> 6)) * (1.0 / 9007199254740992.0);
    double uz = ((double)(r1.v[0] >> 5) * 67108864.0 + (double)(r1.v[1] >> 6)) * (1.0 / 9007199254740992.0);
    double xs = ux;
    if (amp != 0.0 && kk != 0) {
      const double kw = twopi * kk;
      for (int it = 0; it < 8; ++it) {
        double F = xs - amp / kw * (cos(kw * xs) - 1.0) - ux;
        double dF = 1.0 + amp * sin(kw * xs);
        xs -= F / dF;
      }
      xs = xs - floor(xs);
    }
    double x = x0 + xs * Lx, y = y0 + uy * Ly, z = z0 + uz * Lz;
    if (x >= x0 + Lx) x = x0;
    if (y >= y0 + Ly) y = y0;
    if (z >= z0 + Lz) z = z0;
    stage[p * 3] = x;
    stage[p * 3 + 1] = y;
    stage[p * 3 + 2] = z;
    double u1 = ((double)r1.v[2] + 0.5) * (1.0 / 4294967296.0);
    double u2 = ((double)r1.v[3] + 0.5) * (1.0 / 4294967296.0);
    u32x4 c2 = {{(uint32_t)p, (uint32_t)(p >> 32), 0x1C1Cu, 2u}};
    u32x4 r2 = philox4x32_10(c2, key0, key1);
    double u3 = ((double)r2.v[0] + 0.5) * (1.0 / 4294967296.0);
    double u4 = ((double)r2.v[1] + 0.5) * (1.0 / 4294967296.0);
    double ra = sqrt(-2.0 * log(u1)), rb = sqrt(-2.0 * log(u3));
    vel[p * 3] = vtx * ra * cos(twopi * u2) + v0x;
    vel[p * 3 + 1] = vty * ra * sin(twopi * u2) + v0y;
    vel[p * 3 + 2] = vtz * rb * cos(twopi * u4) + v0z;
  }
}

__global__ void k_cell_counts(TileGeom g, ParticlesF64 A, const int *base, const int *count, double dx, double dy,
                              double dz, int *counts) {
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    long s = b + q;
    int i = wrap_index((int)(A.x(s) / dx), g.Nx);
    int j = wrap_index((int)(A.y(s) / dy), g.Ny);
    int k = wrap_index((int)(A.z(s) / dz), g.Nz);
    atomicAdd(&counts[((long)k * g.Ny + j) * g.Nx + i], 1);
  }
}

__global__ void k_apply_bc_aos(double *pos, long np, DriftConsts dc) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
    for (int d = 0; d < 3; ++d) {
      double r = pos[p * 3 + d];
      if (r < dc.r0[d]) r += dc.L[d];
      if (r >= dc.r1[d]) r -= dc.L[d];
      pos[p * 3 + d] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void fill_f64(ecsprt_hip_ctx *c, double *p, long n, double v) {
  k_fill_f64<<<grid_for(n, 256), 256, 0, c->stream>>>(p, n, v);
  ECS_HIP(hipGetLastError());
}
void fill_i32(ecsprt_hip_ctx *c, int *p, long n, int v) {
  k_fill_i32<<<grid_for(n, 256), 256, 0, c->stream>>>(p, n, v);
  ECS_HIP(hipGetLastError());
}

// The overflow moments (FP64 atomics of ions deposited outside the tile regions) start
// each deposit at zero.
static void clear_overflow(ecsprt_hip_ctx *c) {
  fill_f64(c, c->d_ovf, 4 * c->nnode, 0.0);
}
void zero_ctr(ecsprt_hip_ctx *c) {
  k_fill_u64<<<1, 64, 0, c->stream>>>(c->d_ctr, CTR_N);
  ECS_HIP(hipGetLastError());
}

static void free_particles(ParticlesF64 &P) {
  (void)hipFree(P.p);
  P = ParticlesF64{};
}

static void alloc_particles(ParticlesF64 &P, long n) {
  free_particles(P);
  n = std::max(n, 1L);
  hipError_t e = hipMalloc(&P.p, ParticlesF64::bytes(n));
  if (e != hipSuccess) {
    free_particles(P);
    throw Error{ECSPRT_ERR_NOMEM, "particle arrays: " + std::string(hipGetErrorString(e))};
  }
}

void ensure_stage(ecsprt_hip_ctx *c, long np) {
  if (c->stage_cap >= np && c->d_stage) return;
  (void)hipFree(c->d_stage);
  c->d_stage = nullptr;
  c->stage_cap = 0;
  if (hipMalloc(&c->d_stage, sizeof(double) * 6 * std::max(np, 1L)) != hipSuccess)
    throw Error{ECSPRT_ERR_NOMEM, "staging buffer"};
  c->stage_cap = np;
}

static void ensure_keys(ecsprt_hip_ctx *c, long np) {
  if (c->key_cap >= np && c->d_key) return;
  (void)hipFree(c->d_key);
  c->d_key = nullptr;
  if (hipMalloc(&c->d_key, sizeof(int) * 2 * std::max(np, 1L)) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "keys"};
  c->key_cap = np;
}

template <int MODE, int TX, int TY, int TZ, int HH>
static void set_lds_attr() {
  // Once per kernel instance and process, also when contexts on several host threads
  // launch their first step at the same time.
  static std::once_flag once;
  std::call_once(once, [] {
    // 160 KB of LDS per workgroup on gfx950, less what the kernel declares statically.
    hipFuncAttributes fa;
    ECS_HIP(hipFuncGetAttributes(&fa, (const void *)k_tile<MODE, TX, TY, TZ, HH>));
    const int dyn = 160 * 1024 - (int)fa.sharedSizeBytes;
    ECS_HIP(hipFuncSetAttribute((const void *)k_tile<MODE, TX, TY, TZ, HH>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    if constexpr ((MODE & (M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN)) == (M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN)) {
      ECS_HIP(hipFuncGetAttributes(&fa, (const void *)k_tile_runs<MODE, TX, TY, TZ, HH>));
      ECS_HIP(hipFuncSetAttribute((const void *)k_tile_runs<MODE, TX, TY, TZ, HH>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - (int)fa.sharedSizeBytes));
    }
  });
}

template <int MODE>
static void set_lds_attrs() {
  set_lds_attr<MODE, 4, 4, 4, 2>();
  set_lds_attr<MODE, 8, 8, 4, 2>();
  set_lds_attr<MODE, 8, 8, 8, 2>();
}

static size_t lds_for(const TileGeom &g, int mode, int) {
  size_t b = 0;
  if (mode & M_PUSH) b += sizeof(double) * 3 * (g.Tx + 1) * (g.Ty + 1) * (g.Tz + 1);
  if (mode & M_DEPOSIT) b += sizeof(double) * 4 * g.RAx * g.RAy * g.RAz;
  return b;
}

void tiles_init(ecsprt_hip_ctx *c) {
  const ecsprt_params &p = c->p;
  TileGeom &g = c->g;
  g.Nx = p.Nx;
  g.Ny = p.Ny;
  g.Nz = p.Nz;
  int T = c->o.tile;
  if (T <= 0) {
    long n8 = (long)((p.Nx + 7) / 8) * ((p.Ny + 7) / 8) * ((p.Nz + 7) / 8);
    T = n8 >= 512 ? 8 : 4;
  }
  // Measured with ecsprt_amd.autotune: at 128^3 x 100 ions per cell FP64 tiles run
  // fastest as 8x8x4 cells with 512 threads (two workgroups per CU fit in LDS), mixed
  // tiles as 8x8x8 with 1024 threads; 4x4x4 tiles (small grids) with 256 threads.
  const bool auto_shape = c->o.tile <= 0 && c->o.tile_z <= 0 && c->o.block <= 0;
  const int Tz = c->o.tile_z > 0 ? c->o.tile_z : ((auto_shape && T == 8 && !c->mixed) ? 4 : T);
  ECS_CHECK((T == 4 && Tz == 4) || (T == 8 && (Tz == 8 || Tz == 4)), ECSPRT_ERR_ARG,
            "tile shapes: 4x4x4, 8x8x8 or 8x8x4 cells");
  ECS_CHECK(c->o.halo <= 0 || c->o.halo == 2, ECSPRT_ERR_ARG, "the deposit halo is 2 cells");
  g.Tx = T;
  g.Ty = T;
  g.Tz = Tz;
  g.H = 2;
  g.ntx = (p.Nx + g.Tx - 1) / g.Tx;
  g.nty = (p.Ny + g.Ty - 1) / g.Ty;
  g.ntz = (p.Nz + g.Tz - 1) / g.Tz;
  g.RAx = g.Tx + 1 + 2 * g.H;
  g.RAy = g.Ty + 1 + 2 * g.H;
  g.RAz = g.Tz + 1 + 2 * g.H;
  c->ntiles = g.ntx * g.nty * g.ntz;
  // Workgroups per tile: one once the grid has a workgroup for each of the 512 slots of
  // the chip (two per CU); below that enough to reach about 1024.  A split tile pays its
  // staging, region clearing and region flush once per workgroup: at 64^3 and 80^3 (512
  // and 1000 tiles of 8^3) one workgroup per tile measured 5% faster than two.
  int split = c->o.split;
  if (split <= 0) {
    split = c->ntiles >= 512 ? 1 : (1024 + c->ntiles - 1) / c->ntiles;
    if (split > 32) split = 32;
    if (split < 1) split = 1;
  }
  g.split = split;
  int block = c->o.block > 0 ? c->o.block : (T == 4 ? 256 : (auto_shape && !c->mixed) ? 512 : 1024);
  c->block = block;
  ECS_CHECK(block % 64 == 0 && block >= 128 && block <= 1024, ECSPRT_ERR_ARG,
            "block must be a multiple of 64 in [128, 1024]");
  size_t lds = lds_for(g, M_PUSH | M_COLLIDE | M_DEPOSIT, block);
  ECS_CHECK(lds <= 159 * 1024, ECSPRT_ERR_ARG, "tile/halo too large for 160 KB of LDS");
  c->lds_bytes = lds;

  // Fold tables.
  std::vector<int> tab;
  const int Ns[3] = {p.Nx, p.Ny, p.Nz};
  const int nts[3] = {g.ntx, g.nty, g.ntz};
  const int Ts[3] = {g.Tx, g.Ty, g.Tz};
  for (int ax = 0; ax < 3; ++ax) {
    const int N = Ns[ax];
    c->foldtab_off[ax] = (int)tab.size();
    std::vector<int> cnt(N, 0), pairs((size_t)N * FOLD_MAXP * 2, 0);
    for (int ti = 0; ti < nts[ax]; ++ti) {
      int i0 = ti * Ts[ax], ext = std::min(Ts[ax], N - i0), R = ext + 1 + 2 * g.H;
      for (int a = 0; a < R; ++a) {
        int node = wrap_index(i0 - g.H + a, N);
        ECS_CHECK(cnt[node] < FOLD_MAXP, ECSPRT_ERR_ARG, "grid too small for the tile halo");
        pairs[(size_t)node * FOLD_MAXP * 2 + 2 * cnt[node]] = ti;
        pairs[(size_t)node * FOLD_MAXP * 2 + 2 * cnt[node] + 1] = a;
        cnt[node]++;
      }
    }
    tab.insert(tab.end(), cnt.begin(), cnt.end());
    tab.insert(tab.end(), pairs.begin(), pairs.end());
  }
  ECS_HIP(hipMalloc(&c->d_foldtab, sizeof(int) * tab.size()));
  ECS_HIP(hipMemcpy(c->d_foldtab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));

  const long nreg = (long)c->ntiles * g.split;
  const size_t fold_bytes = sizeof(double) * 4 * (size_t)g.RAx * g.RAy * g.RAz * nreg;
  if (hipMalloc(&c->d_fold, fold_bytes) != hipSuccess) throw Error{ECSPRT_ERR_NOMEM, "fold buffer"};
  ECS_HIP(hipMalloc(&c->d_ovf, sizeof(double) * 4 * c->nnode));
  ECS_HIP(hipMalloc(&c->d_ctr, sizeof(unsigned long long) * CTR_N));
  ECS_HIP(hipHostMalloc((void **)&c->h_ctr, sizeof(unsigned long long) * CTR_N, hipHostMallocDefault));
  ECS_HIP(hipMalloc(&c->d_base, sizeof(int) * c->ntiles));
  ECS_HIP(hipMalloc(&c->d_cap, sizeof(int) * c->ntiles));
  ECS_HIP(hipMalloc(&c->d_count, sizeof(int) * c->ntiles));
  ECS_HIP(hipMalloc(&c->d_next, sizeof(int) * tile_ctr(c->ntiles)));
  ECS_HIP(hipMalloc(&c->d_late_cnt, sizeof(int) * tile_ctr(c->ntiles)));
  fill_i32(c, c->d_late_cnt, tile_ctr(c->ntiles), 0);
  fill_i32(c, c->d_count, c->ntiles, 0);
  fill_i32(c, c->d_next, tile_ctr(c->ntiles), 0);
  clear_overflow(c);
  zero_ctr(c);
  set_lds_attrs<M_PUSH>();
  set_lds_attrs<M_DEPOSIT>();
  set_lds_attrs<M_DRIFT | M_REBIN>();
  set_lds_attrs<M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN>();
  set_lds_attrs<M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN>();
  if (c->mixed) mx_set_lds_attrs();
}

void tiles_free(ecsprt_hip_ctx *c) {
  free_particles(c->A);
  free_particles(c->B);
  free_particles(c->E);
  free_particles(c->S);
  free_particles(c->D);
  for (ParticlesMix *P : {&c->Am, &c->Bm, &c->Em, &c->Sm, &c->Dm}) mx_free(*P);
  (void)hipFree(c->d_fold);
  (void)hipFree(c->d_ovf);
  (void)hipFree(c->d_ctr);
  (void)hipFree(c->d_pilot);
  c->d_pilot = nullptr;
  c->pilot_cap = 0;
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  (void)hipFree(c->d_base);
  (void)hipFree(c->d_cap);
  (void)hipFree(c->d_count);
  (void)hipFree(c->d_next);
  (void)hipFree(c->d_late_cnt);
  (void)hipFree(c->d_late);
  (void)hipFree(c->d_foldtab);
  (void)hipFree(c->d_stage);
  (void)hipFree(c->d_key);
}

// Bytes per entry of the late-deposit lists: x, y, z, vx, vy, vz (FP64), or the packed
// cell, offsets and velocities padded to eight words (mixed).
static size_t late_entry_bytes(const ecsprt_hip_ctx *c) { return c->mixed ? 32 : 48; }

void particles_bin_from_stage(ecsprt_hip_ctx *c, long np, bool) {
  ECS_CHECK(np > 0 && np < 2147483647L, ECSPRT_ERR_ARG, "np must be in [1, 2^31-1)");
  ensure_keys(c, np);
  const TileGeom &g = c->g;
  const ecsprt_params &p = c->p;
  zero_ctr(c);
  fill_i32(c, c->d_count, c->ntiles, 0);
  DriftConsts dc = drift_consts(p, 0.0);
  k_bin_count<<<grid_for(np, 256), 256, 0, c->stream>>>(g, c->d_stage, np, p.dx, p.dy, p.dz, dc, c->d_key, c->d_count,
                                                         c->d_ctr);
  ECS_HIP(hipGetLastError());
  std::vector<int> cnt(c->ntiles);
  ECS_HIP(hipMemcpyAsync(cnt.data(), c->d_count, sizeof(int) * c->ntiles, hipMemcpyDeviceToHost, c->stream));
  read_ctr(c);
  const long lost = (long)c->h_ctr[2];
  // Capacities with slack for immigrants; bases in tile order.
  const double slack = c->o.slack > 0 ? c->o.slack : 0.25;
  c->h_base.assign(c->ntiles, 0);
  c->h_cap.assign(c->ntiles, 0);
  long total = 0;
  for (int t = 0; t < c->ntiles; ++t) {
    long cap = cnt[t] + (long)(slack * cnt[t]) + 64;
    cap = (cap + 63) / 64 * 64;  // tile bases on 64-slot (one wavefront) boundaries
    c->h_base[t] = (int)total;
    c->h_cap[t] = (int)cap;
    total += cap;
  }
  ECS_CHECK(total < 2147483647L, ECSPRT_ERR_ARG, "particle capacity exceeds int32 slots");
  {
    const int maxcap = *std::max_element(c->h_cap.begin(), c->h_cap.end());
    c->fx_bits = std::min(50, 62 - (int)std::ceil(std::log2(1.25 * maxcap + 1.0)));
  }
  const bool philox = c->o.collide_mode == ECSPRT_COLLIDE_PHILOX;
  if (c->mixed) {
    if (total > c->pcap || !c->Am.p) {
      mx_alloc(c->Am, total);
      mx_alloc(c->Bm, total);
      if (philox) mx_alloc(c->Dm, total);
      c->pcap = total;
      (void)hipFree(c->d_late);
      c->d_late = nullptr;
      ECS_HIP(hipMalloc(&c->d_late, late_entry_bytes(c) * (total / 4)));
    }
    if (!c->Em.p || c->np != np) {
      mx_alloc(c->Em, np);
      c->scap = std::max(np / 64, 65536L);
      mx_alloc(c->Sm, c->scap);
    }
  } else {
    if (total > c->pcap || !c->A.p) {
      alloc_particles(c->A, total);
      alloc_particles(c->B, total);
      if (philox) alloc_particles(c->D, total);
      c->pcap = total;
      (void)hipFree(c->d_late);
      c->d_late = nullptr;
      ECS_HIP(hipMalloc(&c->d_late, late_entry_bytes(c) * (total / 4)));
    }
    if (!c->E.p || c->np != np) {
      alloc_particles(c->E, np);
      c->scap = std::max(np / 64, 65536L);
      alloc_particles(c->S, c->scap);
    }
  }
  ECS_HIP(hipMemcpyAsync(c->d_base, c->h_base.data(), sizeof(int) * c->ntiles, hipMemcpyHostToDevice, c->stream));
  ECS_HIP(hipMemcpyAsync(c->d_cap, c->h_cap.data(), sizeof(int) * c->ntiles, hipMemcpyHostToDevice, c->stream));
  fill_i32(c, c->d_next, tile_ctr(c->ntiles), 0);
  if (c->mixed) {
    mx_bin_scatter(c, np);
  } else {
    k_bin_scatter<<<grid_for(np, 256), 256, 0, c->stream>>>(c->d_stage, np, c->d_key, c->d_base, c->A);
    ECS_HIP(hipGetLastError());
  }
  ECS_HIP(hipStreamSynchronize(c->stream));
  // Losing a particle is the reference's Abort (src/boundaries.c:101-109): the
  // context drops its particle state, since ids no longer cover [0, np).
  c->np = lost == 0 ? np : 0;
  ECS_CHECK(lost == 0, ECSPRT_ERR_COUNT_CHANGED,
            "particles outside the periodic domain were dropped (" + std::to_string(lost) + ")");
}

void particles_to_stage(ecsprt_hip_ctx *c) {
  ensure_stage(c, c->np);
  if (c->mixed) {
    mx_unbin(c);
    return;
  }
  k_unbin<<<c->ntiles, 256, 0, c->stream>>>(c->A, c->d_base, c->d_count, c->d_stage, c->np);
  ECS_HIP(hipGetLastError());
}

void particles_init_device(ecsprt_hip_ctx *c, long np, uint64_t seed, double amp, int k) {
  ensure_stage(c, np);
  const ecsprt_params &p = c->p;
  k_init<<<grid_for(np, 256), 256, 0, c->stream>>>(c->d_stage, np, p.Lx, p.Ly, p.Lz, p.x0, p.y0, p.z0, amp, k,
                                                    p.viTx, p.viTy, p.viTz, p.vi0x, p.vi0y, p.vi0z,
                                                    (uint32_t)seed, (uint32_t)(seed >> 32));
  ECS_HIP(hipGetLastError());
  particles_bin_from_stage(c, np, true);
}

static TileArgs make_args(ecsprt_hip_ctx *c) {
  TileArgs a{};
  a.g = c->g;
  a.src = c->A;
  a.dst = c->B;
  a.emi = c->E;
  a.spill = c->S;
  a.scap = c->scap;
  a.def = c->D;
  a.base = c->d_base;
  a.count = c->d_count;
  a.cap = c->d_cap;
  a.next = c->d_next;
  a.ctr = c->d_ctr;
  a.phi = c->d_phi;
  a.fold = reinterpret_cast<long long *>(c->d_fold);
  a.ovf = c->d_ovf;
  a.fx_w = std::ldexp(1.0, fx_bits(c));
  a.inv_vmax = 1.0 / mx_vmax(c->p);
  a.late = reinterpret_cast<double *>(c->d_late);
  a.late_cnt = c->d_late_cnt;
  a.pc = push_consts(c->p, 0.0);
  return a;
}

template <int MODE>
static void launch_tile(ecsprt_hip_ctx *c, const TileArgs &a) {
  const int block = c->block;
  const size_t lds = lds_for(c->g, MODE, block);
  // Tile edge and halo are compile-time so LDS offsets become immediates.
  const int tx = c->g.Tx, tz = c->g.Tz;
  if (tx == 8 && tz == 8) k_tile<MODE, 8, 8, 8, 2><<<c->ntiles * c->g.split, block, lds, c->stream>>>(a);
  else if (tx == 8 && tz == 4) k_tile<MODE, 8, 8, 4, 2><<<c->ntiles * c->g.split, block, lds, c->stream>>>(a);
  else k_tile<MODE, 4, 4, 4, 2><<<c->ntiles * c->g.split, block, lds, c->stream>>>(a);
  ECS_HIP(hipGetLastError());
}

static void rebuild_from_spill(ecsprt_hip_ctx *c, long nspill) {
  // A tile overflowed its capacity: gather everything by id and re-bin with new
  // capacities.
  particles_to_stage(c);
  if (nspill > 0 && c->mixed) {
    mx_unbin_list(c, nspill);
  } else if (nspill > 0) {
    k_unbin_list<<<grid_for(nspill, 256), 256, 0, c->stream>>>(c->S, nspill, c->d_stage, c->np);
    ECS_HIP(hipGetLastError());
  }
  particles_bin_from_stage(c, c->np, true);
}

// After a re-binning kernel: immigrants, counts, swap, conservation check.
static void finish_rebin(ecsprt_hip_ctx *c) {
  const TileGeom &g = c->g;
  const ecsprt_params &p = c->p;
  if (c->mixed) {
    mx_rebin(c);
  } else {
    k_rebin<<<2048, 256, 0, c->stream>>>(g, c->E, c->B, c->S, c->scap, c->d_base, c->d_cap, c->d_next, c->d_ctr,
                                         push_consts(p, 0.0), reinterpret_cast<double *>(c->d_late), c->d_late_cnt,
                                         c->d_ovf);
    ECS_HIP(hipGetLastError());
  }
  k_finalize_counts<<<(c->ntiles + 255) / 256, 256, 0, c->stream>>>(c->ntiles, c->d_count, c->d_next, c->d_cap);
  ECS_HIP(hipGetLastError());
  std::swap(c->A, c->B);
  std::swap(c->Am, c->Bm);
}

// The conservation check and spill recovery of a re-binning pass, on counters already
// read into h_ctr.
static void check_counters(ecsprt_hip_ctx *c) {
  const long lost = (long)c->h_ctr[2];
  const long nspill = (long)c->h_ctr[1];
  if (lost > 0) {
    c->np = 0;
    throw Error{ECSPRT_ERR_COUNT_CHANGED,
                "total number of ions changed (" + std::to_string(-lost) + ") under periodic boundaries"};
  }
  if (nspill > 0) rebuild_from_spill(c, nspill);
}

static void check_after_rebin(ecsprt_hip_ctx *c) {
  read_ctr(c);
  check_counters(c);
}

// The attempt rate of this step's Philox collisions, from the ions' state at the start
// of the step: the full pilot on first use, else the sampled prediction.  The fused pass
// and the push slot both set it before the push, so the staged push + collide draws
// exactly what the fused step draws.
static void prepare_lambda(ecsprt_hip_ctx *c, double dt, int Nc) {
  if (c->coll_lambda <= 0.0) run_collide_philox_pilot(c, dt);
  else collide_predict(c, Nc);
  c->lambda_step = c->step_index;
}

void run_push(ecsprt_hip_ctx *c, double dt) {
  if (c->o.collide_mode == ECSPRT_COLLIDE_PHILOX && c->p.mn > 0.0) {
    const int Nc = collision_count(&c->p, c->np, dt);
    if (Nc > 0) prepare_lambda(c, dt, Nc);
  }
  if (c->mixed) {
    mx_launch(c, M_PUSH, dt, nullptr, 0, 0, 0);
    return;
  }
  TileArgs a = make_args(c);
  a.pc = push_consts(c->p, dt);
  launch_tile<M_PUSH>(c, a);
}

void run_drift_rebin(ecsprt_hip_ctx *c, double dt) {
  zero_ctr(c);
  if (c->mixed) {
    mx_launch(c, M_DRIFT | M_REBIN, dt, nullptr, 0, 0, 0);
  } else {
    TileArgs a = make_args(c);
    a.dc = drift_consts(c->p, dt);
    launch_tile<M_DRIFT | M_REBIN>(c, a);
  }
  finish_rebin(c);
  check_after_rebin(c);
}

// Deposits of the ions listed by the fused pass (the re-binned arrays are now c->A).
static void launch_late(ecsprt_hip_ctx *c) {
  if (c->mixed) {
    mx_late(c);
    return;
  }
  const PushConsts pc = push_consts(c->p, 0.0);
  const double *late = reinterpret_cast<const double *>(c->d_late);
  long long *fold = reinterpret_cast<long long *>(c->d_fold);
  const double fw = std::ldexp(1.0, fx_bits(c)), iv = 1.0 / mx_vmax(c->p);
  const int tx = c->g.Tx, tz = c->g.Tz;
  if (tx == 8 && tz == 8)
    k_late<8, 8, 8, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base, c->d_cap,
                                                         pc, c->d_ovf, fold, fw, iv);
  else if (tx == 8 && tz == 4)
    k_late<8, 8, 4, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base, c->d_cap,
                                                         pc, c->d_ovf, fold, fw, iv);
  else
    k_late<4, 4, 4, 2><<<c->ntiles, 256, 0, c->stream>>>(c->g, late, c->d_late_cnt, c->d_base, c->d_cap,
                                                         pc, c->d_ovf, fold, fw, iv);
  ECS_HIP(hipGetLastError());
}

// Fixed-point bits of the deposit regions: 2^30 on the mixed path; on the FP64 path the
// finest scale at which no region can overflow: a workgroup's region sums at most
// 1.25 cap ions (its tile's ions and the tile's late list), each below 2^bits, so
// bits = 62 - ceil(log2(1.25 max cap)), at most 50 (the rounding add needs |v| < 2^51).
int fx_bits(const ecsprt_hip_ctx *c) { return c->mixed ? 30 : c->fx_bits; }

static void launch_fold(ecsprt_hip_ctx *c) {
  const double sn = std::ldexp(1.0, -fx_bits(c)), sg = mx_vmax(c->p) * sn;
  k_fold<long long><<<grid_for(c->nnode, 256), 256, 0, c->stream>>>(
      c->g, reinterpret_cast<const long long *>(c->d_fold), c->d_ovf, c->d_foldtab, c->foldtab_off[0],
      c->foldtab_off[1], c->foldtab_off[2], c->d_moments, sn, sg);
  ECS_HIP(hipGetLastError());
}

void run_deposit(ecsprt_hip_ctx *c) {
  clear_overflow(c);
  if (c->mixed) {
    mx_launch(c, M_DEPOSIT, 0.0, nullptr, 0, 0, 0);
  } else {
    TileArgs a = make_args(c);
    launch_tile<M_DEPOSIT>(c, a);
  }
  launch_fold(c);
}

static CollideConsts collide_consts(const ecsprt_params &p) {
  CollideConsts k;
  k.vn0[0] = p.vn0x;
  k.vn0[1] = p.vn0y;
  k.vn0[2] = p.vn0z;
  k.vnT = p.vnT;
  k.mi = p.mi;
  k.mn = p.mn;
  k.M = p.mn + p.mi;
  double tmp = (p.vi0x - p.vn0x) * (p.vi0x - p.vn0x) + (p.vi0y - p.vn0y) * (p.vi0y - p.vn0y) +
               (p.vi0z - p.vn0z) * (p.vi0z - p.vn0z);
  k.vrm = 4.0 * p.viT + sqrt(tmp);
  k.viT = p.viT;
  k.invM_f = (float)(1.0 / k.M);
  k.inv_vrm2_f = (float)(1.0 / (k.vrm * k.vrm));
  k.vf_max_f = (float)(10.0 * p.viT);
  return k;
}

CollideConsts collide_consts_of(const ecsprt_params &p) { return collide_consts(p); }

// The step's collision counters (read into h_ctr) into the context; ten refusals are
// the reference's Abort (src/collisions.c:141-148).
static void collision_totals(ecsprt_hip_ctx *c) {
  const long natt = (long)c->h_ctr[3], ns = (long)c->h_ctr[4], nf = (long)c->h_ctr[5];
  c->last_attempts = natt;
  c->last_Ns = ns;
  c->last_Nf = nf;
  ECS_CHECK(nf < 10, ECSPRT_ERR_COLLISION, "failed to collide " + std::to_string(nf) + " ion-neutral pairs");
}

void run_fused_particles(ecsprt_hip_ctx *c, double dt) {
  zero_ctr(c);
  clear_overflow(c);
  TileArgs a = make_args(c);
  a.pc = push_consts(c->p, dt);
  a.dc = drift_consts(c->p, dt);
  const bool coll = c->o.collide_mode == ECSPRT_COLLIDE_PHILOX && c->p.mn > 0.0;
  int Nc = 0;
  if (coll) {
    Nc = collision_count(&c->p, c->np, dt);
    c->last_Nc = Nc;
    prepare_lambda(c, dt, Nc);
    a.cc = collide_consts(c->p);
    a.key0 = (uint32_t)c->o.seed;
    a.key1 = (uint32_t)(c->o.seed >> 32);
    a.step = (uint32_t)c->step_index;
    a.pt = poisson_params(Nc > 0 ? c->coll_lambda : 0.0);
    if (c->o.timing) ECS_HIP(hipEventRecord(c->ev[6], c->stream));
    if (c->mixed) mx_launch(c, M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN, dt, &a.pt, a.key0, a.key1, a.step);
    else launch_tile<M_PUSH | M_COLLIDE | M_DRIFT | M_DEPOSIT | M_REBIN>(c, a);
  } else {
    if (c->o.timing) ECS_HIP(hipEventRecord(c->ev[6], c->stream));
    if (c->mixed) mx_launch(c, M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN, dt, nullptr, 0, 0, 0);
    else launch_tile<M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN>(c, a);
  }
  if (c->o.timing) ECS_HIP(hipEventRecord(c->ev[1], c->stream));
  finish_rebin(c);
  launch_late(c);
  launch_fold(c);
  if (c->o.timing) ECS_HIP(hipEventRecord(c->ev[2], c->stream));
  check_after_rebin(c);
  if (coll) collision_totals(c);
  c->step_index++;
}

template <int MODE, int TX, int TY, int TZ>
static void launch_tile_runs(hipStream_t st, const RunTable &t, int block, size_t lds) {
  k_tile_runs<MODE, TX, TY, TZ, 2><<<t.start[t.nrun], block, lds, st>>>(t);
  ECS_HIP(hipGetLastError());
}

// The fused FP64 pass over the runs cs (same tile shape and block size; with collisions
// when coll, then every run has them) in one launch on stream st; arguments staged at
// h_args (pinned) and copied to d_args first.
static void tile_runs_f64(hipStream_t st, const std::vector<ecsprt_hip_ctx *> &cs, double dt, bool coll,
                          const std::vector<PoissonParams> &pt, void *h_args, void *d_args) {
  const int n = (int)cs.size();
  TileArgs *h = static_cast<TileArgs *>(h_args);
  RunTable t{};
  t.args = d_args;
  t.nrun = n;
  long total = 0;
  for (int r = 0; r < n; ++r) {
    ecsprt_hip_ctx *c = cs[r];
    TileArgs a = make_args(c);
    a.pc = push_consts(c->p, dt);
    a.dc = drift_consts(c->p, dt);
    if (coll) {
      a.cc = collide_consts(c->p);
      a.key0 = (uint32_t)c->o.seed;
      a.key1 = (uint32_t)(c->o.seed >> 32);
      a.step = (uint32_t)c->step_index;
      a.pt = pt[r];
    }
    h[r] = a;
    t.start[r] = (int)total;
    total += (long)c->ntiles * c->g.split;
  }
  ECS_CHECK(total < (1L << 31), ECSPRT_ERR_ARG, "multi-run tile pass: too many workgroups");
  t.start[n] = (int)total;
  ECS_HIP(hipMemcpyAsync(d_args, h, sizeof(TileArgs) * n, hipMemcpyHostToDevice, st));
  constexpr int F = M_PUSH | M_DRIFT | M_DEPOSIT | M_REBIN, FC = F | M_COLLIDE;
  const size_t lds = lds_for(cs[0]->g, coll ? FC : F, cs[0]->block);
  const int block = cs[0]->block, tx = cs[0]->g.Tx, tz = cs[0]->g.Tz;
  if (tx == 8 && tz == 8) {
    if (coll) launch_tile_runs<FC, 8, 8, 8>(st, t, block, lds);
    else launch_tile_runs<F, 8, 8, 8>(st, t, block, lds);
  } else if (tx == 8 && tz == 4) {
    if (coll) launch_tile_runs<FC, 8, 8, 4>(st, t, block, lds);
    else launch_tile_runs<F, 8, 8, 4>(st, t, block, lds);
  } else {
    if (coll) launch_tile_runs<FC, 4, 4, 4>(st, t, block, lds);
    else launch_tile_runs<F, 4, 4, 4>(st, t, block, lds);
  }
}

size_t tile_args_bytes() { return std::max(sizeof(TileArgs), mx_args_bytes()); }

// One fused particle step of every run of cs (3-D contexts whose stream is st for the
// call): the stages of run_fused_particles, each queued for all runs before the next.
// The rate predictions of all runs share each host synchronisation; the tile passes of
// runs with the same precision, tile shape, block size and collision setting run as one
// launch (at most RUNS_MAX runs each); the counters of all runs come back with one
// synchronisation.  With rs, each run's pilot passes and its re-binning, late deposits,
// fold and counter copy go on the run's own stream (rs->st[r]) between a fork and a join
// event on st, so that the runs' small launches overlap; the tile passes stay on st.
// Every run's arithmetic and draws are those of run_fused_particles,
// so its ions and moments are the same bit for bit.  A run whose checks fail gets its
// error code in rc and its message in last_error and takes no further part; a HIP error
// is thrown.  Events ev_pass0/ev_pass1 (may be null) bracket the tile passes.
RunFork::RunFork(const std::vector<ecsprt_hip_ctx *> &c, hipStream_t s, const RunStreams *r)
    : cs(c), st(s), rs(r), out(c.size(), 0) {
  if (rs) ECS_HIP(hipEventRecord(rs->fork, st));
}
void RunFork::leave(int r) {
  if (!rs) return;
  ECS_HIP(hipStreamWaitEvent(rs->st[r], rs->fork, 0));
  cs[r]->stream = rs->st[r];
  out[r] = 1;
}
void RunFork::back(int r) {
  if (!rs || !out[r]) return;
  ECS_HIP(hipEventRecord(rs->done[r], rs->st[r]));
  cs[r]->stream = st;
}
void RunFork::join() {
  if (!rs) return;
  for (size_t r = 0; r < out.size(); ++r)
    if (out[r]) ECS_HIP(hipStreamWaitEvent(st, rs->done[r], 0));
}

void run_fused_particles_runs(const std::vector<ecsprt_hip_ctx *> &cs, hipStream_t st, double dt, void *h_args,
                              void *d_args, std::vector<int> &rc, hipEvent_t ev_pass0, hipEvent_t ev_pass1,
                              const RunStreams *rs) {
  const int n = (int)cs.size();
  std::vector<char> coll(n, 0);
  std::vector<PoissonParams> pt(n);
  std::vector<PredictState> ps(n);
  for (int r = 0; r < n; ++r) {
    ecsprt_hip_ctx *c = cs[r];
    zero_ctr(c);
    clear_overflow(c);
    coll[r] = c->o.collide_mode == ECSPRT_COLLIDE_PHILOX && c->p.mn > 0.0;
    if (!coll[r]) continue;
    const int Nc = collision_count(&c->p, c->np, dt);
    c->last_Nc = Nc;
    if (c->coll_lambda <= 0.0) run_collide_philox_pilot(c, dt);
    else predict_begin(c, Nc, ps[r]);
    c->lambda_step = c->step_index;
  }
  for (;;) {
    bool any = false;
    for (int r = 0; r < n && !any; ++r) any = ps[r].active;
    if (!any) break;
    RunFork f(cs, st, rs);
    for (int r = 0; r < n; ++r)
      if (ps[r].active) {
        f.leave(r);
        predict_enqueue(cs[r], ps[r]);
        f.back(r);
      }
    f.join();
    ECS_HIP(hipStreamSynchronize(st));
    for (int r = 0; r < n; ++r)
      if (ps[r].active) predict_collect(cs[r], ps[r]);
  }
  for (int r = 0; r < n; ++r) {
    predict_end(cs[r], ps[r]);
    if (coll[r]) pt[r] = poisson_params(cs[r]->last_Nc > 0 ? cs[r]->coll_lambda : 0.0);
  }
  // groups of runs that share one launch
  if (ev_pass0) ECS_HIP(hipEventRecord(ev_pass0, st));
  std::vector<char> done(n, 0);
  size_t off = 0;
  const size_t abytes = (tile_args_bytes() + 255) & ~(size_t)255;
  for (int r0 = 0; r0 < n; ++r0) {
    if (done[r0]) continue;
    const ecsprt_hip_ctx *c0 = cs[r0];
    std::vector<int> g;
    for (int r = r0; r < n && (int)g.size() < RUNS_MAX; ++r) {
      const ecsprt_hip_ctx *c = cs[r];
      if (!done[r] && c->mixed == c0->mixed && c->g.Tx == c0->g.Tx && c->g.Tz == c0->g.Tz && c->block == c0->block &&
          coll[r] == coll[r0]) {
        g.push_back(r);
        done[r] = 1;
      }
    }
    std::vector<ecsprt_hip_ctx *> gc;
    std::vector<PoissonParams> gpt;
    for (int r : g) {
      gc.push_back(cs[r]);
      gpt.push_back(pt[r]);
    }
    void *h = static_cast<char *>(h_args) + off, *d = static_cast<char *>(d_args) + off;
    off += abytes * g.size();
    if (c0->mixed) {
      std::vector<const PoissonParams *> pp;
      std::vector<uint32_t> k0, k1, sp;
      for (size_t i = 0; i < g.size(); ++i) {
        const ecsprt_hip_ctx *c = gc[i];
        pp.push_back(coll[r0] ? &gpt[i] : nullptr);
        k0.push_back(coll[r0] ? (uint32_t)c->o.seed : 0u);
        k1.push_back(coll[r0] ? (uint32_t)(c->o.seed >> 32) : 0u);
        sp.push_back(coll[r0] ? (uint32_t)c->step_index : 0u);
      }
      mx_launch_runs(st, gc, dt, pp, k0, k1, sp, h, d);
    } else {
      tile_runs_f64(st, gc, dt, coll[r0], gpt, h, d);
    }
  }
  if (ev_pass1) ECS_HIP(hipEventRecord(ev_pass1, st));
  {
    RunFork f(cs, st, rs);
    for (int r = 0; r < n; ++r) {
      ecsprt_hip_ctx *c = cs[r];
      f.leave(r);
      finish_rebin(c);
      launch_late(c);
      launch_fold(c);
      ECS_HIP(hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(unsigned long long) * CTR_N, hipMemcpyDeviceToHost,
                             c->stream));
      f.back(r);
    }
    f.join();
  }
  ECS_HIP(hipStreamSynchronize(st));
  for (int r = 0; r < n; ++r) {
    ecsprt_hip_ctx *c = cs[r];
    try {
      check_counters(c);
      if (coll[r]) collision_totals(c);
      c->step_index++;
      rc[r] = ECSPRT_OK;
    } catch (const Error &e) {
      if (e.code == ECSPRT_ERR_HIP) throw;
      c->last_error = e.msg;
      rc[r] = e.code;
    }
  }
}

// ScatterElastic (src/collisions.c:9-167) in the Philox statistical mode outside the
// fused pass, the CollideFunc slot: every binned ion draws its Poisson number of
// attempts for this step and runs them in order on its current velocity.  The draws
// (keyed by id, step and attempt) and the arithmetic are those of phases 2 and 3 of the
// fused pass, so push followed by this kernel gives the fused step's velocities bit for
// bit.  One workgroup per tile.
__global__ void __launch_bounds__(256) k_collide_philox(ParticlesF64 A, const int *__restrict__ base,
                                                        const int *__restrict__ count, CollideConsts cc,
                                                        PoissonParams pt, uint32_t key0, uint32_t key1, uint32_t step,
                                                        unsigned long long *ctr) {
  __syncthreads();
  const int t = blockIdx.x;
  const long b = base[t], n = count[t];
  unsigned natt = 0, nsucc = 0, nfail = 0;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const long s = b + q;
    const int pid = A.id(s);
    const int k = poisson_attempts(pt, key0, key1, step, pid);
    if (k == 0) continue;
    double vx = A.vx(s), vy = A.vy(s), vz = A.vz(s);
    collide_attempts<double>(cc, key0, key1, step, pid, 0, k, vx, vy, vz, natt, nsucc, nfail);
    A.vx(s) = vx;
    A.vy(s) = vy;
    A.vz(s) = vz;
  }
  const unsigned long long s0 = wave_sum(natt), s1 = wave_sum(nsucc), s2 = wave_sum(nfail);
  if (__lane_id() == 0) {
    if (s0) atomicAdd(&ctr[3], s0);
    if (s1) atomicAdd(&ctr[4], s1);
    if (s2) atomicAdd(&ctr[5], s2);
  }
}

// The CollideFunc slot in the Philox mode (3-D, FP64 or mixed ions): Nc from the
// reference formula, the attempt rate from the pilot on first use, one pass over the
// ions, the per-step counters and the rate steering, then the step counter advances (the
// draws of the next push + collide are those of the next fused step).  Ten refusals are
// the reference's Abort (src/collisions.c:141-148).
void run_collide_philox(ecsprt_hip_ctx *c, double dt) {
  const int Nc = collision_count(&c->p, c->np, dt);
  c->last_Nc = Nc;
  c->last_Ns = c->last_Nf = c->last_attempts = 0;
  if (Nc > 0 && c->p.mn > 0.0) {
    if (c->lambda_step != c->step_index) prepare_lambda(c, dt, Nc);
    const PoissonParams pt = poisson_params(c->coll_lambda);
    const CollideConsts cc = collide_consts(c->p);
    const uint32_t k0 = (uint32_t)c->o.seed, k1 = (uint32_t)(c->o.seed >> 32), step = (uint32_t)c->step_index;
    zero_ctr(c);
    if (c->mixed) {
      mx_collide_philox(c, pt, cc, k0, k1, step);
    } else {
      k_collide_philox<<<c->ntiles, 256, 0, c->stream>>>(c->A, c->d_base, c->d_count, cc, pt, k0, k1, step,
                                                          c->d_ctr);
      ECS_HIP(hipGetLastError());
    }
    read_ctr(c);
    const long natt = (long)c->h_ctr[3], ns = (long)c->h_ctr[4], nf = (long)c->h_ctr[5];
    c->last_attempts = natt;
    c->last_Ns = ns;
    c->last_Nf = nf;
    ECS_CHECK(nf < 10, ECSPRT_ERR_COLLISION, "failed to collide " + std::to_string(nf) + " ion-neutral pairs");
  }
  c->step_index++;
}

void run_cell_counts(ecsprt_hip_ctx *c, int *d_counts) {
  fill_i32(c, d_counts, c->nnode, 0);
  if (c->mixed) {
    mx_cell_counts(c, d_counts);
    return;
  }
  k_cell_counts<<<c->ntiles, 256, 0, c->stream>>>(c->g, c->A, c->d_base, c->d_count, c->p.dx, c->p.dy, c->p.dz,
                                                   d_counts);
  ECS_HIP(hipGetLastError());
}

void run_apply_bc_aos(ecsprt_hip_ctx *c, long np, double *pos) {
  DriftConsts dc = drift_consts(c->p, 0.0);
  k_apply_bc_aos<<<grid_for(np, 256), 256, 0, c->stream>>>(pos, np, dc);
  ECS_HIP(hipGetLastError());
}

}  // namespace ecsprt
