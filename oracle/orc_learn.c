/*
 * orc_learn.c -- CPU oracle of the learned proposal's pair features.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  No reference counterpart: the
 * features of a candidate swap (j -> k = j + o) that the swap network
 * scores, and the features of a SNP in a chain's state that the per-SNP
 * proposal network scores.
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

static float clipped_logit(double q)
{
  double c = q < 1e-6 ? 1e-6 : (q > 1.0 - 1e-6 ? 1.0 - 1e-6 : q);
  double v = log(c / (1.0 - c));
  return (float)(v < -12.0 ? -12.0 : (v > 12.0 ? 12.0 : v));
}

/* SNP j of a chain whose model is given by model_inds (>= 0: in the model)
 * and whose running RB mean is p_proposal; the window is min(H, 20) */
void orc_learn_snp_features(const double* pre, int off, int ot, uint32_t n, uint64_t p,
                            const int32_t* band, int H, const float* z, const float* pbar,
                            const int32_t* model_inds, const double* p_proposal, uint64_t j, float* f)
{
  double nn = (double)n;
  double sj = pre[j * off + ot], cj = pre[j * off + ot + 1];
  float zj = z[j];
  int hw = H < 20 ? H : 20;
  float r2max = 0.0f, r2sum = 0.0f, zmax = 0.0f, rbest = 0.0f, zbest = 0.0f, pbest = 0.0f;
  for (int o = -hw; o <= hw; ++o) {
    int64_t k = (int64_t)j + o;
    if (o == 0 || k < 0 || k >= (int64_t)p) continue;
    float azk = fabsf(z[k]);
    if (azk > zmax) zmax = azk;
    if (model_inds[k] < 0) continue;
    double sk = pre[(uint64_t)k * off + ot], ck = pre[(uint64_t)k * off + ot + 1];
    uint64_t lo = (uint64_t)k < j ? (uint64_t)k : j;
    double cov = (double)band[lo * (uint64_t)(H + 1) + (o < 0 ? -o : o)] - sj * sk / nn;
    double r = (cj > 0.0 && ck > 0.0) ? cov / sqrt(cj * ck) : 0.0;
    r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
    float r2 = (float)(r * r);
    r2sum += r2;
    if (r2 > r2max) { r2max = r2; rbest = (float)r; zbest = z[k]; pbest = (float)p_proposal[k]; }
  }
  float one_r2 = 1.0f - r2max;
  float zc = r2max > 0.0f ? (zj - rbest * zbest) / sqrtf(one_r2 > 1e-3f ? one_r2 : 1e-3f) : zj;
  float zj2 = zj * zj / 100.0f, zc2 = zc * zc / 100.0f;
  f[0] = zj / 10.0f;
  f[1] = zj2 < 25.0f ? zj2 : 25.0f;
  f[2] = clipped_logit(p_proposal[j]) / 6.0f;
  f[3] = clipped_logit((double)pbar[j]) / 6.0f;
  f[4] = model_inds[j] >= 0 ? 1.0f : 0.0f;
  f[5] = r2max;
  f[6] = r2sum < 5.0f ? r2sum : 5.0f;
  f[7] = zmax / 10.0f < 5.0f ? zmax / 10.0f : 5.0f;
  f[8] = (float)(sj / nn) * 0.5f;
  f[9] = r2max > 0.0f ? clipped_logit((double)pbest) / 6.0f : -2.0f;
  f[10] = clampf(zc / 10.0f, -5.0f, 5.0f);
  f[11] = zc2 < 25.0f ? zc2 : 25.0f;
  f[12] = 0.0f;
  f[13] = 0.0f;
  f[14] = 0.0f;
  f[15] = 1.0f;
}
