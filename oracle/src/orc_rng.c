/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * Counter-based synthetic inputs.  The reference draws its test data with the Fortran
 * intrinsic random_number (infrastructure/source/utilities/random_number_generator_mod.f90:
 * 40-78), whose stream is compiler-specific, so the synthetic operators of the benchmark
 * configurations are defined here by a counter-based generator that a CPU and a GPU
 * evaluate to the same bits:
 *
 *   bits(seed, stream, ctr) = mix(mix(seed + G*(stream+1)) ^ (ctr*K1 + K2))
 *     mix = SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)
 *   uniform(0,1)  = ((bits >> 11) + 0.5) * 2^-53
 *   normal(0,1)   = Marsaglia polar method on attempt pairs 2a, 2a+1 (a < 32) of the
 *                   sub-counter ctr*64
 *   log, exp      = written out with +,-,*,/ and fma only (no libm), so both sides round
 *                   identically: log by atanh series on m in [sqrt(1/2), sqrt(2)),
 *                   exp by Cody-Waite reduction and a degree-13 Taylor polynomial.
 *
 * Element operators (configs of the LMA benchmark):
 *   B_e(i,j) = normal/8 [* 10^u_j, u_j ~ U(-2,2) per element column when scaled]
 *   A_e(i,j) = sum_m fma(B(i,m), B(j,m), .) + 0.1*delta_ij, mirrored so A_e is exactly symmetric
 * stored in the reference LMA layout local_stencil(ncell_3d, ndf, ndf)
 * (infrastructure/source/operator/operator_mod.t90:113,119).
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_rng_bits(uint64_t seed, uint64_t stream, uint64_t ctr) {
  uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL * (stream + 1));
  return mix64(k ^ (ctr * 0xD1B54A32D192ED03ULL + 0x8CB92BA72F3D8DD7ULL));
}

double orc_rng_uniform(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return ((double)(orc_rng_bits(seed, stream, ctr) >> 11) + 0.5) * 0x1p-53;
}

static double bits2d(uint64_t b) {
  double d;
  memcpy(&d, &b, 8);
  return d;
}
static uint64_t d2bits(double d) {
  uint64_t b;
  memcpy(&b, &d, 8);
  return b;
}

#define LN2_HI 0x1.62e42fefa3800p-1
#define LN2_LO 0x1.ef35793c76730p-45

/* natural log of a positive normal double */
double orc_log(double x) {
  uint64_t b = d2bits(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
  if (m > 0x1.6a09e667f3bcdp+0) { /* sqrt(2) */
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  /* 2*atanh(s) = 2s * sum_{j>=0} z^j/(2j+1), j <= 12 */
  double p = 1.0 / 25.0;
  static const double inv[12] = {1.0 / 23, 1.0 / 21, 1.0 / 19, 1.0 / 17, 1.0 / 15, 1.0 / 13,
                                 1.0 / 11, 1.0 / 9,  1.0 / 7,  1.0 / 5,  1.0 / 3,  1.0};
  for (int j = 0; j < 12; ++j) p = fma(p, z, inv[j]);
  double lm = (2.0 * s) * p;
  double de = (double)e;
  return fma(de, LN2_HI, fma(de, LN2_LO, lm));
}

/* exp for |x| < 700 */
double orc_exp(double x) {
  double n = nearbyint(x * 0x1.71547652b82fep+0); /* 1/ln2 */
  double r = fma(-n, LN2_HI, x);
  r = fma(-n, LN2_LO, r);
  /* sum_{j=0}^{13} r^j/j! */
  static const double c[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
                               1.0 / 3628800.0,    1.0 / 362880.0,    1.0 / 40320.0,
                               1.0 / 5040.0,       1.0 / 720.0,       1.0 / 120.0,
                               1.0 / 24.0,         1.0 / 6.0,         0.5,
                               1.0,                1.0};
  double p = c[0];
  for (int j = 1; j < 14; ++j) p = fma(p, r, c[j]);
  int ni = (int)n;
  return p * bits2d((uint64_t)(ni + 1023) << 52);
}

double orc_rng_normal(uint64_t seed, uint64_t stream, uint64_t ctr) {
  for (uint64_t a = 0; a < 32; ++a) {
    double v1 = 2.0 * orc_rng_uniform(seed, stream, ctr * 64 + 2 * a) - 1.0;
    double v2 = 2.0 * orc_rng_uniform(seed, stream, ctr * 64 + 2 * a + 1) - 1.0;
    double s = fma(v1, v1, v2 * v2);
    if (s < 1.0 && s > 0.0) return v1 * sqrt(-2.0 * orc_log(s) / s);
  }
  return 0.0;
}

/* streams of the synthetic inputs */
enum { STREAM_B = 1, STREAM_SCALE = 2 };

void orc_gen_spd(int ndf, int64_t nc3, uint64_t seed, int scaled, double *A) {
  orc_gen_spd_shift(ndf, nc3, seed, scaled, 0.1, A);
}

void orc_gen_spd_shift(int ndf, int64_t nc3, uint64_t seed, int scaled, double shift, double *A) {
  double B[27 * 27];
  double sc[27];
  const double ln10 = 0x1.26bb1bbb55516p+1;
  for (int64_t ic = 0; ic < nc3; ++ic) {
    for (int j = 0; j < ndf; ++j)
      sc[j] = scaled ? orc_exp((4.0 * orc_rng_uniform(seed, STREAM_SCALE, (uint64_t)ic * ndf + j) -
                                2.0) * ln10)
                     : 1.0;
    for (int i = 0; i < ndf; ++i)
      for (int j = 0; j < ndf; ++j)
        B[i * ndf + j] =
            orc_rng_normal(seed, STREAM_B, (uint64_t)ic * ndf * ndf + (uint64_t)i * ndf + j) *
            0.125 * sc[j];
    for (int i = 0; i < ndf; ++i)
      for (int j = i; j < ndf; ++j) {
        double acc = 0.0;
        for (int m = 0; m < ndf; ++m) acc = fma(B[i * ndf + m], B[j * ndf + m], acc);
        if (i == j) acc += shift;
        A[ic + nc3 * (i + (int64_t)ndf * j)] = acc;
        A[ic + nc3 * (j + (int64_t)ndf * i)] = acc;
      }
  }
}

void orc_gen_uniform(int64_t n, uint64_t seed, uint64_t stream, double *v) {
  for (int64_t i = 0; i < n; ++i) v[i] = 2.0 * orc_rng_uniform(seed, stream, (uint64_t)i) - 1.0;
}

/* general element entries sign(j) * (lo + (hi-lo) u), counter ic*n1*n2 + i*n2 + j */
void orc_gen_op_uniform(int n1, int n2, int64_t nc3, uint64_t seed, uint64_t stream, double lo, double hi,
                        const double *sign, double *A) {
  for (int64_t ic = 0; ic < nc3; ++ic)
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n2; ++j) {
        double u = orc_rng_uniform(seed, stream, (uint64_t)ic * n1 * n2 + (uint64_t)i * n2 + j);
        double v = lo + (hi - lo) * u;
        if (sign) v = v * sign[j];
        A[ic + nc3 * (i + (int64_t)n1 * j)] = v;
      }
}
