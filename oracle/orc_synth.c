/*
 * orc_synth.c -- CPU restatement of the engine's synthetic genotype
 * generator, used to check the HIP generator byte for byte.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Model (no reference counterpart; the reference ships only a PLINK
 * --simulate-qt dataset, testdata/plinksimu.sh): individual i has two
 * haplotypes h; along the SNPs a latent Gaussian follows
 *   z_j = rho_j z_{j-1} + sqrt(1 - rho_j^2) e_j,  rho_j = 0 at block starts,
 * e_j = Box-Muller normal of the Philox block with key {seed, 0x5EED5EED}
 * and counter {j_lo, j_hi, 2 i + h, 0x40000000}.  The allele is 1[z_j <
 * thresh_j]; the genotype is the allele count, packed with the reference's
 * code mapping 0 -> 00, 1 -> 10, 2 -> 11 (src/data.cpp:289-306), individual i
 * at bits 2 (i % 4) of byte i / 4.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static double synth_normal(uint32_t seed, uint64_t j, uint32_t hap)
{
  uint32_t ctr[4] = {(uint32_t)j, (uint32_t)(j >> 32), hap, 0x40000000u};
  uint32_t key[2] = {seed, 0x5EED5EEDu};
  uint32_t o[4];
  orc_philox4x32_10(ctr, key, o);
  double ua = (double)((((uint64_t)o[0]) << 21) | (o[1] >> 11)) * 0x1.0p-53;
  double ub = (double)((((uint64_t)o[2]) << 21) | (o[3] >> 11)) * 0x1.0p-53;
  return sqrt(-2.0 * log(1.0 - ua)) * cos(6.283185307179586 * ub);
}

void orc_synth_genotypes(uint32_t seed, uint32_t n, uint64_t p,
                         const double* thresh, const double* rho,
                         uint8_t* body)
{
  static const uint8_t CODE[3] = {0x0, 0x2, 0x3};
  size_t stride = (n + 3) / 4;
  memset(body, 0, stride * p);
  for (uint32_t i = 0; i < n; ++i) {
    double z0 = 0.0, z1 = 0.0;
    for (uint64_t j = 0; j < p; ++j) {
      double e0 = synth_normal(seed, j, 2 * i);
      double e1 = synth_normal(seed, j, 2 * i + 1);
      double r = rho[j];
      if (r == 0.0) { z0 = e0; z1 = e1; }
      else {
        double s = sqrt(1.0 - r * r);
        z0 = r * z0 + s * e0;
        z1 = r * z1 + s * e1;
      }
      int g = (z0 < thresh[j]) + (z1 < thresh[j]);
      body[stride * j + i / 4] |= (uint8_t)(CODE[g] << (2 * (i % 4)));
    }
  }
}
