/*
 * orc_learn.c -- CPU oracle of the learned proposal's pair features.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  No reference counterpart: the
 * features of a candidate swap (j -> k = j + o) that the network scores.
 */
#include "oracle.h"
#include <math.h>

static float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

void orc_learn_pair_features(const double* pre, int off, int ot, uint32_t n,
                             const int32_t* band, int H, const float* z, const float* pbar,
                             uint64_t j, int o, float* f)
{
  uint64_t k = (uint64_t)((int64_t)j + o);
  double nn = (double)n;
  double sj = pre[j * off + ot], cj = pre[j * off + ot + 1];
  double sk = pre[k * off + ot], ck = pre[k * off + ot + 1];
  uint64_t lo = j < k ? j : k;
  int d = o < 0 ? -o : o;
  double cov = (double)band[lo * (uint64_t)(H + 1) + d] - sj * sk / nn;
  double r = (cj > 0.0 && ck > 0.0) ? cov / sqrt(cj * ck) : 0.0;
  r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
  float rf = (float)r;
  float zj = z[j], zk = z[k];
  float pj = clampf(pbar[j], 1e-6f, 1.0f - 1e-6f), pk = clampf(pbar[k], 1e-6f, 1.0f - 1e-6f);
  float lj = clampf(logf(pj / (1.0f - pj)), -10.0f, 10.0f) / 5.0f;
  float lk = clampf(logf(pk / (1.0f - pk)), -10.0f, 10.0f) / 5.0f;
  float one_r2 = 1.0f - rf * rf;
  float zc = (zk - rf * zj) / sqrtf(one_r2 > 1e-3f ? one_r2 : 1e-3f);
  float zcs = clampf(zc / 10.0f, -5.0f, 5.0f);
  float zj2 = zj * zj / 100.0f, zk2 = zk * zk / 100.0f;
  f[0] = rf;
  f[1] = rf * rf;
  f[2] = zj / 10.0f;
  f[3] = zk / 10.0f;
  f[4] = zj2 < 25.0f ? zj2 : 25.0f;
  f[5] = zk2 < 25.0f ? zk2 : 25.0f;
  f[6] = clampf(rf * zj * zk / 100.0f, -25.0f, 25.0f);
  f[7] = (float)d / (float)H;
  f[8] = (float)(sj / nn) * 0.5f;
  f[9] = (float)(sk / nn) * 0.5f;
  f[10] = lj;
  f[11] = lk;
  f[12] = zcs;
  f[13] = zcs * zcs;
  f[14] = o > 0 ? 1.0f : -1.0f;
  f[15] = 1.0f;
}
