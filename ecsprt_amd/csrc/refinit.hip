// Initial conditions drawn as the reference draws them on one rank, so a run can start
// from the reference's own particles.  Positions (InitializePositions,
// src/positions.c:390-433):
//   normal      ran3 per coordinate (NormalDistribution, src/positions.c:42-137, which
//               despite its name places every coordinate uniformly, pos = ran3 * L),
//               seeded with getseed() = -(rank+1)*12345 = -12345 (src/random.c:292-295);
//   sobol       the reference's Sobseq, one 3-D point per ion, pos = s * L
//               (SobolDistribution, src/positions.c:139-234; src/random.c:233-289);
//   sinusoidal  rejection against 1 + 0.25 sin(2 pi x/Lx) over PETSc's default
//               PetscRandom, the rand48 generator seeded with 7 (Rejection,
//               src/positions.c:236-388).
// Each ion takes ndim coordinates and ndim velocity components (two in a
// two-dimensional run), in the streams' order.
// Velocities are vT * gasdev + v0 per component (NormalVelocities,
// src/velocities.c:14-47), seed -12345.  The generators are Numerical Recipes (2nd
// edition) ran3, gasdev over ran1 and sobseq, with the reference's return types
// (src/random.c:24-289), and the 48-bit LCG of srand48/drand48.  The streams are
// serial, so they run on the host; the arrays are then uploaded and binned like any
// other host particles.
#include <algorithm>
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
  // RNMX is 1 - PETSC_MACHINE_EPSILON (src/random.c:10-11), not the book's 1 - 1.2e-7:
  // the float temp is compared with it after promotion to double, so only a temp that
  // rounded to 1.0f is replaced, by (float)RNMX = 1.0f.
  static constexpr double AM = 1.0 / IM, RNMX = 1.0 - 2.220446049250313e-16;
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

// sobseq with MAXBIT 30 and MAXDIM 6: the direction numbers of the six primitive
// polynomials of degree up to 4 (degrees mdeg, interior coefficients ip, initial odd
// numerators of the listed table), the point sequence in Gray-code order.  For degrees
// above 2 the reference's recurrence XORs the previous direction number once per set
// coefficient bit (src/random.c:263-266), where the published algorithm XORs the one l
// places back; the reference's sequence is the one reproduced.
struct Sobol {
  static constexpr int MAXBIT = 30, MAXDIM = 6;
  unsigned long in = 0, ix[MAXDIM + 1] = {};
  unsigned long v[MAXBIT + 1][MAXDIM + 1] = {};  // v[j][k]: direction number j of dimension k
  Sobol() {
    static const int mdeg[MAXDIM + 1] = {0, 1, 2, 3, 3, 4, 4};
    static const unsigned ip[MAXDIM + 1] = {0, 0, 1, 1, 2, 1, 4};
    static const unsigned long m0[4][MAXDIM + 1] = {
        {0, 1, 1, 1, 1, 1, 1}, {0, 3, 1, 3, 3, 1, 1}, {0, 5, 7, 7, 3, 3, 5}, {0, 15, 11, 5, 15, 13, 9}};
    for (int k = 1; k <= MAXDIM; ++k) {
      for (int j = 1; j <= mdeg[k]; ++j) v[j][k] = m0[j - 1][k] << (MAXBIT - j);
      for (int j = mdeg[k] + 1; j <= MAXBIT; ++j) {
        unsigned long a = v[j - mdeg[k]][k];
        a ^= a >> mdeg[k];
        unsigned bits = ip[k];
        for (int l = mdeg[k] - 1; l >= 1; --l, bits >>= 1)
          if (bits & 1) a ^= v[j - 1][k];
        v[j][k] = a;
      }
    }
  }
  // Next point in x[0..2]; false once 2^30 - 1 points have been drawn.
  bool next3(double x[3]) {
    unsigned long c = in++;
    int j = 1;
    while (j <= MAXBIT && (c & 1)) {
      c >>= 1;
      ++j;
    }
    if (j > MAXBIT) return false;
    for (int k = 1; k <= 3; ++k) {
      ix[k] ^= v[j][k];
      x[k - 1] = (double)ix[k] * (1.0 / (double)(1L << MAXBIT));
    }
    return true;
  }
};

// srand48 / drand48: x <- (0x5DEECE66D x + 0xB) mod 2^48, seeded x = (seed << 16) | 0x330E,
// each value x / 2^48 (exact in FP64).
struct Rand48 {
  unsigned long long x;
  explicit Rand48(long seed) : x((((unsigned long long)(unsigned int)seed) << 16) | 0x330EULL) {}
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return std::ldexp((double)x, -48);
  }
};

double sinusoidal_density(const double r[3]) { return 1.0 + 0.25 * std::sin(2 * 3.14159265358979323846 * r[0]); }

}  // namespace

void particles_init_distribution(ecsprt_hip_ctx *c, long np, int pdist) {
  const ecsprt_params &p = c->p;
  ECS_CHECK(pdist == ECSPRT_PDIST_NORMAL || pdist == ECSPRT_PDIST_SOBOL || pdist == ECSPRT_PDIST_SINUSOIDAL,
            ECSPRT_ERR_ARG,
            "position distribution not available: reverse and gaussian are not implemented by the reference, "
            "uniform is PETSc's DMSwarmSetPointsUniformCoordinates lattice");
  ECS_CHECK(pdist != ECSPRT_PDIST_SOBOL || np < (1L << Sobol::MAXBIT), ECSPRT_ERR_ARG,
            "Sobol positions: at most 2^30 - 1 ions (MAXBIT)");
  const int nd = p.ndim == 2 ? 2 : 3;  // components per ion, as the reference's ndim
  std::vector<double> pos(nd * (size_t)np), vel(nd * (size_t)np);
  const double L[3] = {p.Lx, p.Ly, p.Lz};
  const double vT[3] = {p.viTx, p.viTy, p.viTz}, v0[3] = {p.vi0x, p.vi0y, p.vi0z};
  if (pdist == ECSPRT_PDIST_NORMAL) {
    Ran3 r3;
    long seed = -12345;
    for (long ip = 0; ip < np; ++ip)
      for (int d = 0; d < nd; ++d) pos[nd * ip + d] = (double)r3.next(&seed) * L[d];
  } else if (pdist == ECSPRT_PDIST_SOBOL) {
    // Sobseq(&ndim, ...) advances dimensions 1..ndim only; each dimension is independent,
    // so the first two of a three-dimensional draw are the two-dimensional point.
    Sobol sb;
    double s3[3];
    for (long ip = 0; ip < np; ++ip) {
      sb.next3(s3);
      for (int d = 0; d < nd; ++d) pos[nd * ip + d] = s3[d] * L[d];
    }
  } else {
    // The maximum over the grid nodes (i/Nx, j/Ny, k/Nz); the density depends on x alone.
    double maxVal = 0.0, s3[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < p.Nx; ++i) {
      s3[0] = (double)i / p.Nx;
      maxVal = std::max(maxVal, sinusoidal_density(s3));
    }
    const double normVal = 1.0 / maxVal;
    Rand48 r(7);
    for (long ip = 0; ip < np;) {
      for (int d = 0; d < nd; ++d) s3[d] = r.next();
      const double w = sinusoidal_density(s3);
      if (w * normVal > r.next()) {
        for (int d = 0; d < nd; ++d) pos[nd * ip + d] = s3[d] * L[d];
        ++ip;
      }
    }
  }
  {
    Gasdev g;
    long seed = -12345;
    for (long ip = 0; ip < np; ++ip)
      for (int d = 0; d < nd; ++d) vel[nd * ip + d] = vT[d] * (double)g.next(&seed) + v0[d];
  }
  if (nd == 2) {
    p2d_set(c, np, pos.data(), vel.data(), hipMemcpyHostToDevice);
    return;
  }
  ensure_stage(c, np);
  ECS_HIP(hipMemcpyAsync(c->d_stage, pos.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  ECS_HIP(hipMemcpyAsync(c->d_stage + 3 * np, vel.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice,
                         c->stream));
  particles_bin_from_stage(c, np, true);
}

}  // namespace ecsprt

extern "C" int ecsprt_hip_nr_stream(int generator, long seed, long n, float *out) {
  if (n < 0 || (n > 0 && !out) || generator < 0 || generator > 2) return ECSPRT_ERR_ARG;
  long idum = seed;
  if (generator == 0) {
    ecsprt::Ran1 r;
    for (long i = 0; i < n; ++i) out[i] = r.next(&idum);
  } else if (generator == 1) {
    ecsprt::Ran3 r;
    for (long i = 0; i < n; ++i) out[i] = r.next(&idum);
  } else {
    ecsprt::Gasdev g;
    for (long i = 0; i < n; ++i) out[i] = g.next(&idum);
  }
  return ECSPRT_OK;
}
