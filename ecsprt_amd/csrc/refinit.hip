// Initial conditions drawn as the reference draws them on one rank, so a run can start
// from the reference's own particles: positions from ran3 (NormalDistribution,
// src/positions.c:42-137, which despite its name places every coordinate uniformly,
// pos = ran3 * L), velocities vT * gasdev + v0 per component (NormalVelocities,
// src/velocities.c:14-47), each stream seeded with getseed() = -(rank+1)*12345 = -12345
// (src/random.c:292-295).  The generators are Numerical Recipes (2nd edition) ran3 and
// gasdev over ran1, with the reference's float return types (src/random.c:24-204).  The
// streams are serial, so they run on the host; the arrays are then uploaded and binned
// like any other host particles.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "common.hpp"

namespace ecsprt {

void ensure_stage(ecsprt_hip_ctx *c, long np);
void particles_bin_from_stage(ecsprt_hip_ctx *c, long np, bool);

namespace {

// ran1: Park-Miller minimal standard with a 32-entry Bays-Durham shuffle.
struct Ran1 {
  static constexpr long IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
  static constexpr long NDIV = 1 + (IM - 1) / NTAB;
  static constexpr double AM = 1.0 / IM, RNMX = 1.0 - 1.2e-7;
  long iy = 0, iv[NTAB] = {};
  float next(long *idum) {
    long k;
    if (*idum <= 0 || !iy) {
      *idum = -(*idum) < 1 ? 1 : -(*idum);
      for (int j = NTAB + 7; j >= 0; --j) {
        k = *idum / IQ;
        *idum = IA * (*idum - k * IQ) - IR * k;
        if (*idum < 0) *idum += IM;
        if (j < NTAB) iv[j] = *idum;
      }
      iy = iv[0];
    }
    k = *idum / IQ;
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    const int j = (int)(iy / NDIV);
    iy = iv[j];
    iv[j] = *idum;
    const float temp = (float)(AM * iy);
    return temp > RNMX ? (float)RNMX : temp;
  }
};

// ran3: Knuth's subtractive generator, table of 55, lags 55 and 24.
struct Ran3 {
  static constexpr long MBIG = 1000000000, MSEED = 161803398, MZ = 0;
  static constexpr double FAC = 1.0 / MBIG;
  int inext = 0, inextp = 0, iff = 0;
  long ma[56] = {};
  float next(long *idum) {
    if (*idum < 0 || iff == 0) {
      iff = 1;
      long mj = std::labs(MSEED - std::labs(*idum)) % MBIG;
      ma[55] = mj;
      long mk = 1;
      for (int i = 1; i <= 54; ++i) {
        const int ii = (21 * i) % 55;
        ma[ii] = mk;
        mk = mj - mk;
        if (mk < MZ) mk += MBIG;
        mj = ma[ii];
      }
      for (int k = 1; k <= 4; ++k)
        for (int i = 1; i <= 55; ++i) {
          ma[i] -= ma[1 + (i + 30) % 55];
          if (ma[i] < MZ) ma[i] += MBIG;
        }
      inext = 0;
      inextp = 31;
      *idum = 1;
    }
    if (++inext == 56) inext = 1;
    if (++inextp == 56) inextp = 1;
    long mj = ma[inext] - ma[inextp];
    if (mj < MZ) mj += MBIG;
    ma[inext] = mj;
    return (float)(mj * FAC);
  }
};

// gasdev: Marsaglia's polar Box-Muller over ran1, returning the cached second deviate
// on every other call; a negative seed clears the cache.
struct Gasdev {
  Ran1 r1;
  int iset = 0;
  float gset = 0.0f;
  float next(long *idum) {
    if (*idum < 0) iset = 0;
    if (iset == 0) {
      float v1, v2, rsq;
      do {
        v1 = (float)(2.0 * r1.next(idum) - 1.0);
        v2 = (float)(2.0 * r1.next(idum) - 1.0);
        rsq = v1 * v1 + v2 * v2;
      } while (rsq >= 1.0 || rsq == 0.0);
      const float fac = (float)std::sqrt(-2.0 * std::log((double)rsq) / rsq);
      gset = v1 * fac;
      iset = 1;
      return v2 * fac;
    }
    iset = 0;
    return gset;
  }
};

}  // namespace

void particles_init_reference(ecsprt_hip_ctx *c, long np) {
  const ecsprt_params &p = c->p;
  std::vector<double> pos(3 * (size_t)np), vel(3 * (size_t)np);
  const double L[3] = {p.Lx, p.Ly, p.Lz};
  const double vT[3] = {p.viTx, p.viTy, p.viTz}, v0[3] = {p.vi0x, p.vi0y, p.vi0z};
  {
    Ran3 r3;
    long seed = -12345;
    for (long ip = 0; ip < np; ++ip)
      for (int d = 0; d < 3; ++d) pos[3 * ip + d] = (double)r3.next(&seed) * L[d];
  }
  {
    Gasdev g;
    long seed = -12345;
    for (long ip = 0; ip < np; ++ip)
      for (int d = 0; d < 3; ++d) vel[3 * ip + d] = vT[d] * (double)g.next(&seed) + v0[d];
  }
  ensure_stage(c, np);
  ECS_HIP(hipMemcpyAsync(c->d_stage, pos.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  ECS_HIP(hipMemcpyAsync(c->d_stage + 3 * np, vel.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice,
                         c->stream));
  particles_bin_from_stage(c, np, true);
}

}  // namespace ecsprt
