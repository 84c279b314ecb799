// Per-particle arithmetic of the hot path, shared by the tile kernels and the
// collision kernels.  Operation order follows the reference line by line so that
// FP64 results are bit-identical to a non-contracting CPU build (the library is
// compiled with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "common.hpp"
#include "philox.hpp"

namespace ecsprt {

__host__ __device__ inline int wrap_index(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// Field gather, DifferenceVector3D (src/vector-math.c:596-722).  F(k, j, i) returns the
// potential at global node (i, j, k) (unwrapped indices; the loader wraps them).  The
// particle sits in cell (ixl, iyl, izl) at in-cell weights (wxh, wyh, wzh).  Branch
// choices use the global indices exactly as the reference does; the interpolation order
// keeps the reference's y->wyh, z->wxh, x->wzh weight assignment (vector-math.c:645-651).
// T is the arithmetic type.  The tile kernels use this form only for an ion whose
// cell index reached N by rounding; all others go through gather_nodes below.
template <typename T, typename L>
__device__ inline void diff_vec3d_core(const L &F, int Nx, int Ny, int Nz, int ixl, int iyl, int izl, T wxh, T wyh,
                                       T wzh, T f[3]) {
  const int ixh = ixl + 1, iyh = iyl + 1, izh = izl + 1;
  const T c3 = T(3), c4 = T(4), c1 = T(1);
  T hhh, lhh, hlh, llh, hhl, lhl, hll, lll;
  T whh, whl, wlh, wll, Ewh, Ewl;
  // ixl >= 0 for every ion in the domain, so the reference's forward-difference branch
  // (vector-math.c:625-631) cannot occur.
  hhl = F(izh, iyh, ixl + 1) - F(izh, iyh, ixl - 1);
  lhl = F(izl, iyh, ixl + 1) - F(izl, iyh, ixl - 1);
  hll = F(izh, iyl, ixl + 1) - F(izh, iyl, ixl - 1);
  lll = F(izl, iyl, ixl + 1) - F(izl, iyl, ixl - 1);
  if (ixh < Nx) {
    hhh = F(izh, iyh, ixh + 1) - F(izh, iyh, ixh - 1);
    lhh = F(izl, iyh, ixh + 1) - F(izl, iyh, ixh - 1);
    hlh = F(izh, iyl, ixh + 1) - F(izh, iyl, ixh - 1);
    llh = F(izl, iyl, ixh + 1) - F(izl, iyl, ixh - 1);
  } else {
    hhh = +c3 * F(izh, iyh, ixh) - c4 * F(izh, iyh, ixh - 1) + c1 * F(izh, iyh, ixh - 2);
    lhh = +c3 * F(izl, iyh, ixh) - c4 * F(izl, iyh, ixh - 1) + c1 * F(izl, iyh, ixh - 2);
    hlh = +c3 * F(izh, iyl, ixh) - c4 * F(izh, iyl, ixh - 1) + c1 * F(izh, iyl, ixh - 2);
    llh = +c3 * F(izl, iyl, ixh) - c4 * F(izl, iyl, ixh - 1) + c1 * F(izl, iyl, ixh - 2);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[0] = Ewl + wzh * (Ewh - Ewl);

  hlh = F(izh, iyl + 1, ixh) - F(izh, iyl - 1, ixh);
  llh = F(izl, iyl + 1, ixh) - F(izl, iyl - 1, ixh);
  hll = F(izh, iyl + 1, ixl) - F(izh, iyl - 1, ixl);
  lll = F(izl, iyl + 1, ixl) - F(izl, iyl - 1, ixl);
  if (iyh < Ny) {
    hhh = F(izh, iyh + 1, ixh) - F(izh, iyh - 1, ixh);
    lhh = F(izl, iyh + 1, ixh) - F(izl, iyh - 1, ixh);
    hhl = F(izh, iyh + 1, ixl) - F(izh, iyh - 1, ixl);
    lhl = F(izl, iyh + 1, ixl) - F(izl, iyh - 1, ixl);
  } else {
    hhh = +c3 * F(izh, iyh, ixh) - c4 * F(izh, iyh - 1, ixh) + c1 * F(izh, iyh - 2, ixh);
    lhh = +c3 * F(izl, iyh, ixh) - c4 * F(izl, iyh - 1, ixh) + c1 * F(izl, iyh - 2, ixh);
    hhl = +c3 * F(izh, iyh, ixl) - c4 * F(izh, iyh - 1, ixl) + c1 * F(izh, iyh - 2, ixl);
    lhl = +c3 * F(izl, iyh, ixl) - c4 * F(izl, iyh - 1, ixl) + c1 * F(izl, iyh - 2, ixl);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[1] = Ewl + wzh * (Ewh - Ewl);

  lhh = F(izl + 1, iyh, ixh) - F(izl - 1, iyh, ixh);
  llh = F(izl + 1, iyl, ixh) - F(izl - 1, iyl, ixh);
  lhl = F(izl + 1, iyh, ixl) - F(izl - 1, iyh, ixl);
  lll = F(izl + 1, iyl, ixl) - F(izl - 1, iyl, ixl);
  if (izh < Nz) {
    hhh = F(izh + 1, iyh, ixh) - F(izh - 1, iyh, ixh);
    hlh = F(izh + 1, iyl, ixh) - F(izh - 1, iyl, ixh);
    hhl = F(izh + 1, iyh, ixl) - F(izh - 1, iyh, ixl);
    hll = F(izh + 1, iyl, ixl) - F(izh - 1, iyl, ixl);
  } else {
    hhh = +c3 * F(izh, iyh, ixh) - c4 * F(izh - 1, iyh, ixh) + c1 * F(izh - 2, iyh, ixh);
    hlh = +c3 * F(izh, iyl, ixh) - c4 * F(izh - 1, iyl, ixh) + c1 * F(izh - 2, iyl, ixh);
    hhl = +c3 * F(izh, iyh, ixl) - c4 * F(izh - 1, iyh, ixl) + c1 * F(izh - 2, iyh, ixl);
    hll = +c3 * F(izh, iyl, ixl) - c4 * F(izh - 1, iyl, ixl) + c1 * F(izh - 2, iyl, ixl);
  }
  whh = hlh + wyh * (hhh - hlh);
  whl = hll + wyh * (hhl - hll);
  wlh = llh + wyh * (lhh - llh);
  wll = lll + wyh * (lhl - lll);
  Ewh = wlh + wxh * (whh - wlh);
  Ewl = wll + wxh * (whl - wll);
  f[2] = Ewl + wzh * (Ewh - Ewl);
}

// Potential loader over the global periodic array (any unwrapped index).
struct PhiGlobal {
  const double *phi;
  int Nx, Ny, Nz;
  __device__ double operator()(int k, int j, int i) const {
    return phi[((long)wrap_index(k, Nz) * Ny + wrap_index(j, Ny)) * Nx + wrap_index(i, Nx)];
  }
};

// The difference numerators DifferenceVector3D takes at one grid node, per component:
// F[i+1]-F[i-1] for a node below N (the reference's central branches at both the low
// and the high corner, vector-math.c:619-637), the backward 3F[i]-4F[i-1]+F[i-2] for a
// node index equal to N (its high-corner branch, :638-644).  These depend on the node
// alone, so a tile computes them once for its (T+1)^3 corner nodes and every ion reads
// eight of them.  Same expressions, same order: FP64 results are bit-identical.
template <typename T, typename L>
__device__ inline void node_diff(const L &F, int Nx, int Ny, int Nz, int i, int j, int k, T d[3]) {
  const T c3 = T(3), c4 = T(4), c1 = T(1);
  d[0] = (i < Nx) ? T(F(k, j, i + 1) - F(k, j, i - 1)) : T(+c3 * F(k, j, i) - c4 * F(k, j, i - 1) + c1 * F(k, j, i - 2));
  d[1] = (j < Ny) ? T(F(k, j + 1, i) - F(k, j - 1, i)) : T(+c3 * F(k, j, i) - c4 * F(k, j - 1, i) + c1 * F(k, j - 2, i));
  d[2] = (k < Nz) ? T(F(k + 1, j, i) - F(k - 1, j, i)) : T(+c3 * F(k, j, i) - c4 * F(k - 1, j, i) + c1 * F(k - 2, j, i));
}

// The gather from per-node differences: E holds NX x NY x NZ nodes per component
// (component c at E + c*NE3), base is the index of the ion's low corner.  Corner names
// are (z, y, x) low/high as in the reference; interpolation as vector-math.c:645-651.
template <typename T, int NX, int NY, int NE3>
__device__ inline void gather_nodes(const T *__restrict__ E, int base, T wxh, T wyh, T wzh, T f[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T *e = E + c * NE3 + base;
    const T lll = e[0], llh = e[1], lhl = e[NX], lhh = e[NX + 1];
    const T hll = e[NX * NY], hlh = e[NX * NY + 1], hhl = e[NX * NY + NX], hhh = e[NX * NY + NX + 1];
    const T whh = hlh + wyh * (hhh - hlh);
    const T whl = hll + wyh * (hhl - hll);
    const T wlh = llh + wyh * (lhh - llh);
    const T wll = lll + wyh * (lhl - lll);
    const T Ewh = wlh + wxh * (whh - wlh);
    const T Ewl = wll + wxh * (whl - wll);
    f[c] = Ewl + wzh * (Ewh - Ewl);
  }
}

// Collision constants (src/collisions.c:18-52).
struct CollideConsts {
  double vn0[3];
  double vnT, mi, mn, M, vrm, viT;
  // FP32 path: 1/M, 1/vrm^2 and the refusal bound 10 viT, so its attempts multiply
  // where the FP64 path divides as the reference does.
  float invM_f, inv_vrm2_f, vf_max_f;
};

// Arithmetic of the scattering formula by precision.  FP64 is the reference's
// (sqrt, division, cos/sin of 2*pi*u); FP32 is the statistical mixed path's, on the
// hardware square root, reciprocal and sin/cos of revolutions (v_sqrt_f32, v_rcp_f32,
// v_sin_f32, v_cos_f32; about 1 ulp), which the FP32 draws cannot resolve anyway.
__device__ inline double c_sqrt(double x) { return sqrt(x); }
__device__ inline float c_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double c_div(double a, double b) { return a / b; }
__device__ inline float c_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// cos and sin of 2*pi*u: in FP64 through sincospi(2u), whose reduction by multiples of
// pi is exact (within an ulp or two of the reference's cos(2*PI*u), and one range
// reduction instead of two); in FP32 on the hardware's sin/cos of revolutions.
__device__ inline void c_sincos2pi(double u, double &s, double &c) { sincospi(2.0 * u, &s, &c); }
__device__ inline void c_sincos2pi(float u, float &s, float &c) {
  s = __builtin_amdgcn_sinf(u);
  c = __builtin_amdgcn_cosf(u);
}

// One attempt of the ScatterElastic loop body (src/collisions.c:77-154) on ion
// velocity vi.  Returns 0 rejected by the acceptance draw, 1 collided (vf written),
// 2 refused because |vf|/viT > 10 (collisions.c:141).  T = double follows the
// reference's arithmetic; T = float is the mixed-precision path's.
// ND = 2 is the two-dimensional run (ndim = 2): every loop covers x and y and vr[2]
// keeps its initial 0 (collisions.c:32), so cos(theta) = 0.
template <typename T, int ND = 3>
__device__ inline int collide_one(const CollideConsts &k, const float g[3], float u_acc1, float u_acc2, float u_z,
                                  float u_phi, const T vi[3], T vf[3]) {
  const T vnT = (T)k.vnT, mi = (T)k.mi, mn = (T)k.mn, M = (T)k.M, vrm = (T)k.vrm, viT = (T)k.viT;
  T vn[3] = {0, 0, 0}, vr[3] = {0, 0, 0}, vc[3] = {0, 0, 0}, u2[3];
  T tmp = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    vn[d] = g[d] * vnT + (T)k.vn0[d];
    vr[d] = vi[d] - vn[d];
    tmp += vr[d] * vr[d];
  }
  T vrr = c_sqrt(tmp);
  float acc = u_acc1 * u_acc2;
  constexpr bool f64 = sizeof(T) == sizeof(double);
  if constexpr (f64) {
    if (!(acc < vrr * vrr / (vrm * vrm))) return 0;
  } else {
    if (!(acc < vrr * vrr * k.inv_vrm2_f)) return 0;
  }
  tmp = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) {
    if constexpr (f64) vc[d] = (vi[d] * mi + vn[d] * mn) / M;
    else vc[d] = (vi[d] * mi + vn[d] * mn) * k.invM_f;
    tmp += (vi[d] - vc[d]) * (vi[d] - vc[d]);
  }
  T vcr = c_sqrt(tmp);
  T u0z = T(2) * (T)u_z - T(1);
  T uperp = c_sqrt(T(1) - u0z * u0z);
  T sphi, cphi;
  c_sincos2pi((T)u_phi, sphi, cphi);
  T u0x = uperp * cphi;
  T u0y = uperp * sphi;
  T costht, sintht, cosphi, sinphi;
  if constexpr (f64) {
    costht = vr[2] / vrr;
    sintht = sqrt(T(1) - costht * costht);
    cosphi = vr[0] / (vrr * sintht);
    sinphi = vr[1] / (vrr * sintht);
  } else {
    // In FP32 cos(theta) rounds to 1 once |vr_perp| < 3.5e-4 |vr| (about one attempt
    // in 3e7), where the reference's vr_x / (vrr sin(theta)) is 0/0.  The azimuth is
    // taken from vr_perp itself, with a fixed direction when it vanishes.
    const T vperp = c_sqrt(vr[0] * vr[0] + vr[1] * vr[1]);
    const T ivrr = c_div(T(1), vrr), ivp = c_div(T(1), vperp);
    costht = vr[2] * ivrr;
    sintht = vperp * ivrr;
    cosphi = vperp > T(0) ? vr[0] * ivp : T(1);
    sinphi = vperp > T(0) ? vr[1] * ivp : T(0);
  }
  T u1x = u0z * sintht + u0x * costht;
  T u1y = u0y;
  T u1z = u0z * costht - u0x * sintht;
  u2[0] = u1x * cosphi - u1y * sinphi;
  u2[1] = u1x * sinphi + u1y * cosphi;
  u2[2] = u1z;
  tmp = 0;
  T w[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < ND; d++) {
    w[d] = vc[d] + vcr * u2[d];
    tmp += w[d] * w[d];
  }
  T vfr = c_sqrt(tmp);
  if constexpr (f64) {
    if (vfr / viT > T(10)) return 2;
  } else {
    if (vfr > k.vf_max_f) return 2;  // |vf| / viT > 10
  }
#pragma unroll
  for (int d = 0; d < ND; d++) vf[d] = w[d];
  return 1;
}

// Box-Muller pair from two 32-bit draws (statistical collision mode only).
// Hardware log2, square root and sin/cos of revolutions (statistical mode only).
__device__ inline void box_muller(uint32_t a, uint32_t b, float &g0, float &g1) {
  const float u1 = u01(a), u2 = u01(b);
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));  // -2 ln u1
  g0 = r * __builtin_amdgcn_cosf(u2);
  g1 = r * __builtin_amdgcn_sinf(u2);
}

// Statistical collision mode: a Poisson(lambda) number of attempts on each ion, each
// attempt the reference's scattering formula on Philox draws keyed by (id, step,
// attempt).  With lambda = Nc / (np * acceptance) the expected number of collisions
// is Nc.  The count is drawn by inversion of one 24-bit uniform, searching the CDF from
// the mode m = floor(lambda) with the pmf recurrences p(k+1) = p(k) lambda/(k+1) and
// p(k-1) = p(k) k/lambda in FP64: about 0.8 sqrt(lambda) steps for an ion with
// attempts, one comparison (u <= P(K = 0)) for one without, and no bound on lambda:
// the cold-neutral runs, whose ions cool and whose acceptance falls step by step, need
// ever more attempts per collision, as the reference's loop does (src/collisions.c:70-153).
struct PoissonParams {
  double lam;     // mean attempts per ion
  double F0;      // P(K = 0)
  double F_mode;  // P(K <= mode)
  double p_mode;  // P(K = mode)
  int mode;       // floor(lam)
  int kmax;       // search bound, far beyond any representable tail
};

// K for the uniform u in (0, 1).
__host__ __device__ inline int poisson_invert(const PoissonParams &pp, double u) {
  if (u <= pp.F0) return 0;
  int k = pp.mode;
  double F = pp.F_mode, p = pp.p_mode;
  if (u <= F) {
    // smallest k with u <= F(k), walking down: F(k-1) = F(k) - p(k)
    while (k > 0) {
      const double Fd = F - p;
      if (u > Fd) break;
      p *= (double)k / pp.lam;
      F = Fd;
      --k;
    }
    return k;
  }
  while (k < pp.kmax) {
    p *= pp.lam / (double)(k + 1);
    F += p;
    ++k;
    if (u <= F) break;
  }
  return k;
}

// The attempt draw of ion pid at this step: one Philox2x32-10 draw, counter (pid, step),
// key mixed from both seed words.
__device__ inline float poisson_uniform(uint32_t key0, uint32_t key1, uint32_t step, int pid) {
  const u32x2 c0 = {{(uint32_t)pid, step}};
  const u32x2 r0 = philox2x32_10(c0, key0 ^ (key1 * 0x9E3779B9u));
  return u01(r0.v[0]);
}

// The number of attempts of ion pid at this step.
__device__ inline int poisson_attempts(const PoissonParams &pp, uint32_t key0, uint32_t key1, uint32_t step,
                                       int pid) {
  return poisson_invert(pp, (double)poisson_uniform(key0, key1, step, pid));
}

// Whether ion pid has at least one attempt at this step (the same draw, K > 0).
__device__ inline bool poisson_any(const PoissonParams &pp, uint32_t key0, uint32_t key1, uint32_t step, int pid) {
  return (double)poisson_uniform(key0, key1, step, pid) > pp.F0;
}

// Attempts first..k-1 of ion pid at this step, in order (ND = 2: vz is not touched).
template <typename T, int ND = 3>
__device__ inline void collide_attempts(const CollideConsts &cc, uint32_t key0, uint32_t key1, uint32_t step, int pid,
                                        int first, int k, T &vx, T &vy, T &vz, unsigned &natt, unsigned &nsucc,
                                        unsigned &nfail) {
  for (int at = first; at < k; ++at) {
    u32x4 c1 = {{(uint32_t)pid, step, (uint32_t)(at + 1), 0xA77E3B71u}};
    u32x4 c2 = {{(uint32_t)pid, step, (uint32_t)(at + 1), 0x5CA77E12u}};
    u32x4 r1 = philox4x32_10(c1, key0, key1);
    u32x4 r2 = philox4x32_10(c2, key0, key1);
    float g[3], spare;
    box_muller(r1.v[0], r1.v[1], g[0], g[1]);
    box_muller(r1.v[2], r1.v[3], g[2], spare);
    T vi[3] = {vx, vy, ND == 3 ? vz : T(0)}, vf[3];
    int res = collide_one<T, ND>(cc, g, u01(r2.v[0]), u01(r2.v[1]), u01(r2.v[2]), u01(r2.v[3]), vi, vf);
    natt += 1u;
    nsucc += res == 1 ? 1u : 0u;
    nfail += res == 2 ? 1u : 0u;
    if (res == 1) {
      vx = vf[0];
      vy = vf[1];
      if (ND == 3) vz = vf[2];
    }
  }
}

__device__ inline void collide_philox(const CollideConsts &cc, const PoissonParams &pp, uint32_t key0, uint32_t key1,
                                      uint32_t step, int pid, double &vx, double &vy, double &vz, unsigned &natt,
                                      unsigned &nsucc, unsigned &nfail) {
  const int k = poisson_attempts(pp, key0, key1, step, pid);
  collide_attempts<double>(cc, key0, key1, step, pid, 0, k, vx, vy, vz, natt, nsucc, nfail);
}

// Host: the search start for K ~ Poisson(lam), lam in [0, POISSON_LAMBDA_MAX]: the
// pmf at the mode from its logarithm, the CDF there summed downward from the mode
// (terms below 1e-300 of it dropped).
constexpr double POISSON_LAMBDA_MAX = 1.0e6;
inline PoissonParams poisson_params(double lam) {
  if (!(lam >= 0.0)) lam = 0.0;
  if (lam > POISSON_LAMBDA_MAX) throw Error{ECSPRT_ERR_COLLISION, "Poisson attempt rate above 1e6 per ion and step"};
  PoissonParams pp;
  pp.lam = lam;
  pp.F0 = exp(-lam);
  pp.mode = (int)floor(lam);
  pp.p_mode = lam > 0.0 ? exp(-lam + pp.mode * log(lam) - lgamma(pp.mode + 1.0)) : 1.0;
  double F = 0.0, p = pp.p_mode;
  for (int k = pp.mode; k >= 0 && p > 1e-300 * pp.p_mode; --k) {
    F += p;
    p *= (double)k / lam;
  }
  pp.F_mode = std::min(F, 1.0);
  pp.kmax = pp.mode + 64 + (int)(64.0 * sqrt(lam));
  return pp;
}

// One pilot pass's update of lambda: rescaled by Nc / (collisions of the dry run at
// min(lam, POISSON_LAMBDA_MAX)), by at most 16 per pass.  A result above the bound after
// the last pass is an error, not a cap.
inline double pilot_rescale(double lam, long Nc, double s) {
  return lam * (s > 0 ? std::min(16.0, std::max(0.25, (double)Nc / s)) : 16.0);
}
inline void pilot_check(double lam) {
  if (lam > POISSON_LAMBDA_MAX)
    throw Error{ECSPRT_ERR_COLLISION, "collision acceptance too low: the attempt rate needed (" + std::to_string(lam) +
                                          " per ion and step) exceeds 1e6"};
}

// Attempt-rate steering shared by the particle paths: lambda scaled by Nc / successes
// (within a factor two per step), kept within POISSON_LAMBDA_MAX.
inline double steer_lambda(double lam, long Nc, long ns) {
  if (ns > 0 && Nc > 0) lam *= std::min(2.0, std::max(0.5, (double)Nc / (double)ns));
  return std::min(lam, POISSON_LAMBDA_MAX);
}

// (PetscInt)(np*ran3(&seed)): the int*float product is a float (src/collisions.c:72);
// a product rounding up to np is clamped to np-1.
__host__ __device__ inline long pick_ip(long np, float u) {
  long ip = (long)(int)((float)np * u);
  if (ip >= np) ip = np - 1;
  if (ip < 0) ip = 0;
  return ip;
}


}  // namespace ecsprt
