/*
 * orc_rng.c -- the project random stream, restated for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference draws from boost::mt19937 through uniform_01, normal and
 * gamma distributions (src/rand.hpp:45-51,174-186).  Boost is absent and its
 * distribution algorithms are version dependent, so the project defines its
 * own stream, implemented identically here and in the HIP engine:
 *   - block generator: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
 *     random numbers: as easy as 1, 2, 3", SC'11), key = {seed, chain},
 *     counter = {lo, hi, c2, c3}; each draw consumes one block and then
 *     increments the 64-bit (lo, hi);
 *   - u01 = (w0 * 2^21 + (w1 >> 11)) * 2^-53, in [0, 1);
 *   - normal = sqrt(-2 log(1 - u_a)) * cos(2 pi u_b) (Box-Muller, one
 *     output per block; u_a from words 0-1, u_b from words 2-3);
 *   - gamma(a): Marsaglia & Tsang (2000), a >= 1 directly, a < 1 through
 *     gamma(a + 1) * u^(1/a);
 *   - sinvchi2(nu, s2) = nu s2 / (2 gamma(nu/2)) redrawn while not finite
 *     and positive, as src/rand.hpp:61-97 does.
 * Pinned by: Philox known-answer vectors of the Random123 distribution,
 * moment tests (the reference's own tests are moment tests,
 * src/tests/rand_tests.hpp:30-177) and CPU==GPU stream equality.
 */
#include "oracle.h"
#include <math.h>

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static void philox_round(uint32_t c[4], const uint32_t k[2])
{
  uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
  uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
  uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t n0 = hi1 ^ c[1] ^ k[0];
  uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4])
{
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  philox_round(c, k);
  for (int r = 1; r < 10; ++r) {
    k[0] += PHILOX_W0; k[1] += PHILOX_W1;
    philox_round(c, k);
  }
  out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

void orc_rng_init(orc_rng* r, uint32_t seed, uint32_t chain, uint32_t c2,
                  uint32_t c3)
{
  r->key[0] = seed; r->key[1] = chain;
  r->ctr[0] = 0; r->ctr[1] = 0; r->ctr[2] = c2; r->ctr[3] = c3;
}

static void next_block(orc_rng* r, uint32_t out[4])
{
  orc_philox4x32_10(r->ctr, r->key, out);
  if (++r->ctr[0] == 0) ++r->ctr[1];
}

static double words_to_u01(uint32_t w0, uint32_t w1)
{
  uint64_t m = ((uint64_t)w0 << 21) | (uint64_t)(w1 >> 11);
  return (double)m * 0x1.0p-53;
}

double orc_rng_u01(orc_rng* r)
{
  uint32_t o[4];
  next_block(r, o);
  return words_to_u01(o[0], o[1]);
}

double orc_rng_normal(orc_rng* r)
{
  uint32_t o[4];
  next_block(r, o);
  double ua = words_to_u01(o[0], o[1]);
  double ub = words_to_u01(o[2], o[3]);
  return sqrt(-2.0 * log(1.0 - ua)) * cos(6.283185307179586 * ub);
}

/* Marsaglia & Tsang, "A simple method for generating gamma variables",
 * ACM TOMS 26(3), 2000. */
double orc_rng_gamma(orc_rng* r, double a)
{
  if (a < 1.0) {
    double g = orc_rng_gamma(r, a + 1.0);
    double u = orc_rng_u01(r);
    return g * pow(u, 1.0 / a);
  }
  double d = a - 1.0 / 3.0;
  double c = 1.0 / sqrt(9.0 * d);
  for (;;) {
    double z = orc_rng_normal(r);
    double t = 1.0 + c * z;
    if (t <= 0.0) continue;
    double v = t * t * t;
    double u = orc_rng_u01(r);
    double z2 = z * z;
    if (u < 1.0 - 0.0331 * z2 * z2) return d * v;
    if (log(u) < 0.5 * z2 + d * (1.0 - v + log(v))) return d * v;
  }
}

/* src/rand.hpp:85-97 */
double orc_rng_sinvchi2(orc_rng* r, double nu, double s2)
{
  double val = -1.0;
  while (val <= 0 || !isfinite(val)) {
    val = nu * s2 / (2.0 * orc_rng_gamma(r, 0.5 * nu));
  }
  return val;
}

void orc_rng_fill(uint32_t seed, uint32_t chain, uint32_t c2, uint32_t c3,
                  int kind, double a, double b, int64_t cnt, double* out)
{
  orc_rng r;
  orc_rng_init(&r, seed, chain, c2, c3);
  for (int64_t i = 0; i < cnt; ++i) {
    switch (kind) {
      case 0: out[i] = orc_rng_u01(&r); break;
      case 1: out[i] = orc_rng_normal(&r); break;
      case 2: out[i] = orc_rng_gamma(&r, a); break;
      default: out[i] = orc_rng_sinvchi2(&r, a, b); break;
    }
  }
}
