// bmg_learn.hip -- online training of the learned swap proposal.
//
// Per dataset: the band X^T X over the window (int8 MFMA, exact), the
// marginal score of every SNP, a ring of swap labels filled from the burn-in
// RB scans, the network (bmg_mlp.hip) and the partner-weight table the chain
// kernel draws from.  Everything here runs between chain segments, on the
// batch's stream, and only during burn-in; the table is frozen before the
// first kept sample.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <random>
#include <vector>

#include "bmg_device.h"
#include "bmg_learn.h"
#include "bmg_mlp.h"
#include "bmg_rb.h"

namespace bmg {

#define LEARN_RING (1u << 19)
#define LEARN_MAX_ROWS (16 * 256)   // label rows of one RB group: chains x model indices
#define LEARN_CHUNK (1 << 20)
#define LEARN_SRING (1u << 19)      // per-SNP proposal samples kept
#define LEARN_SRAND 1024            // uniformly drawn SNPs per chain per RB scan
#define LEARN_SWIN 20               // per-SNP features: LD window (SNPs either side)

struct LearnView {
  const double* pre;   // p x offset self moments
  int off, ot;         // offset, offset of the sampler's type
  uint64_t p;
  uint32_t n;
  int H;
  const int32_t* band; // p x (H + 1): sum_i x_j x_{j+d}
  const float* z;      // marginal scores
  const float* pbar;   // mean p_proposal of the dataset's chains
};

struct LearnDS {
  LearnCfg cfg{};
  LearnView v{};
  int type = 0;
  int32_t* band = nullptr;
  float* z = nullptr;
  float* pbar = nullptr;
  uint32_t* rj = nullptr;
  int32_t* ro = nullptr;
  float* rl = nullptr;
  unsigned long long* counter = nullptr;
  uint32_t* lcnt = nullptr;               // labels per (chain of the group, model index)
  unsigned long long* lbase = nullptr;    // their first ring positions
  MlpArgs mlp{};
  uint16_t* Xb = nullptr;
  float* lb = nullptr;
  uint16_t* Xc = nullptr;
  float* oc = nullptr;
  float* table = nullptr;
  int* d_chains = nullptr;
  int n_chains = 0;
  int adam_t = 0;
  uint64_t rounds = 0;
  int frozen = 0;
  // learned per-SNP proposal: its network, a ring of (feature row, label)
  // samples from the burn-in RB scans and one chain's table at a time
  MlpArgs snet{};
  uint16_t* sx = nullptr;                 // LEARN_SRING x 16 bf16 feature rows
  float* sl = nullptr;                    // their labels
  unsigned long long* scounter = nullptr;
  uint32_t* scnt = nullptr;
  unsigned long long* sbase = nullptr;    // window rows' ring positions, then each chain's random rows
  double* ptab = nullptr;                 // p: one chain's table
  int sadam_t = 0;
  uint32_t sgroups = 0;                   // label groups written (seeds the random rows)
  int applied = 0;                        // tables set as the chains' proposals
};

__device__ __forceinline__ int learn_off(int H, int c) { return c < H ? c - H : c - H + 1; }

// Features of the pair (j, k = j + o), k inside the genome.
__device__ void pair_features(const LearnView& L, uint64_t j, int o, float* f) {
  const uint64_t k = (uint64_t)((int64_t)j + o);
  const double n = (double)L.n;
  const double sj = L.pre[j * L.off + L.ot], cj = L.pre[j * L.off + L.ot + 1];
  const double sk = L.pre[k * L.off + L.ot], ck = L.pre[k * L.off + L.ot + 1];
  const uint64_t lo = j < k ? j : k;
  const int d = o < 0 ? -o : o;
  const double cov = (double)L.band[lo * (L.H + 1) + d] - sj * sk / n;
  double r = (cj > 0.0 && ck > 0.0) ? cov / sqrt(cj * ck) : 0.0;
  r = fmin(1.0, fmax(-1.0, r));
  const float rf = (float)r;
  const float zj = L.z[j], zk = L.z[k];
  const float pj = fminf(fmaxf(L.pbar[j], 1e-6f), 1.0f - 1e-6f);
  const float pk = fminf(fmaxf(L.pbar[k], 1e-6f), 1.0f - 1e-6f);
  const float lj = fminf(fmaxf(logf(pj / (1.0f - pj)), -10.0f), 10.0f) / 5.0f;
  const float lk = fminf(fmaxf(logf(pk / (1.0f - pk)), -10.0f), 10.0f) / 5.0f;
  const float zc = (zk - rf * zj) / sqrtf(fmaxf(1.0f - rf * rf, 1e-3f));
  const float zcs = fminf(fmaxf(zc / 10.0f, -5.0f), 5.0f);
  f[0] = rf;
  f[1] = rf * rf;
  f[2] = zj / 10.0f;
  f[3] = zk / 10.0f;
  f[4] = fminf(zj * zj / 100.0f, 25.0f);
  f[5] = fminf(zk * zk / 100.0f, 25.0f);
  f[6] = fminf(fmaxf(rf * zj * zk / 100.0f, -25.0f), 25.0f);
  f[7] = (float)d / (float)L.H;
  f[8] = (float)(sj / n) * 0.5f;
  f[9] = (float)(sk / n) * 0.5f;
  f[10] = lj;
  f[11] = lk;
  f[12] = zcs;
  f[13] = zcs * zcs;
  f[14] = o > 0 ? 1.0f : -1.0f;
  f[15] = 1.0f;
}

__device__ __forceinline__ uint16_t f2bf_l(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return 0x7FC0u;
  u = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void store_row(uint16_t* X, int64_t row, const float* f) {
  uint16_t h[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) h[i] = f2bf_l(f[i]);
  uint4 a, b;
  __builtin_memcpy(&a, h, 16);
  __builtin_memcpy(&b, h + 8, 16);
  *(uint4*)(X + row * 16) = a;
  *(uint4*)(X + row * 16 + 8) = b;
}

// z_j = corr(x_j, y) sqrt(n), clipped to +-50; rx = x_j . (y - mean y)
__global__ void k_learn_static(LearnView L, const double* __restrict__ rx, double syy, float* __restrict__ z) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= L.p) return;
  double css = L.pre[j * L.off + L.ot + 1];
  double c = (css > 0.0 && syy > 0.0) ? rx[j] / sqrt(css * syy) : 0.0;
  double v = c * sqrt((double)L.n);
  z[j] = (float)fmin(50.0, fmax(-50.0, v));
}

__device__ __forceinline__ double lbf(double rx, double css, double d, double s2) {
  double q = css + d;
  return rx * rx / (2.0 * s2 * q) - 0.5 * log(q / d);
}

// label of lane `lane` (an offset) for model index m of chain list[v]; false
// when the pair gives no sample (partner outside the genome or in the model,
// degenerate moments, non-finite label)
__device__ __forceinline__ bool label_at(const LearnView& L, int type, const ChainScal* __restrict__ scal,
                                         const ChainArrays* __restrict__ arrays, const int* __restrict__ list,
                                         int kmax, const double* __restrict__ xa,
                                         const double* __restrict__ rsum, int v, int m, int lane,
                                         uint64_t* j_out, int* o_out, double* lab_out) {
  const int c = list[v];
  const ChainScal& CS = scal[c];
  const ChainArrays& A = arrays[c];
  if (m >= CS.nloci || lane >= 2 * L.H) return false;
  const uint64_t j = (uint64_t)A.idx[m];
  const int o = learn_off(L.H, lane);
  const int64_t k = (int64_t)j + o;
  if (k < 0 || k >= (int64_t)L.p) return false;
  if (A.model_inds[k] >= 0) return false;
  const double n = (double)L.n;
  const double mr = rsum[v] / n;
  const double sj = L.pre[j * L.off + L.ot], cj = L.pre[j * L.off + L.ot + 1];
  const double sk = L.pre[(uint64_t)k * L.off + L.ot], ck = L.pre[(uint64_t)k * L.off + L.ot + 1];
  const int x1 = A.idx[kmax + m];
  const double bj = A.vec[3 * kmax + x1];
  const uint64_t lo = (uint64_t)k < j ? (uint64_t)k : j;
  const double cov = (double)L.band[lo * (L.H + 1) + (o < 0 ? -o : o)] - sj * sk / n;
  const double rxj = xa[(size_t)v * L.p + j] - sj * mr + bj * cj;
  const double rxk = xa[(size_t)v * L.p + (uint64_t)k] - sk * mr + bj * cov;
  const double d = CS.inv_tau2_alpha2[type];
  const double s2 = CS.sigma2;
  if (!(cj > 0.0) || !(ck > 0.0) || !(d > 0.0) || !(s2 > 0.0)) return false;
  double lab = lbf(rxk, ck, d, s2) - lbf(rxj, cj, d, s2);
  if (!isfinite(lab)) return false;
  *j_out = j;
  *o_out = o;
  *lab_out = fmin(30.0, fmax(-30.0, lab));
  return true;
}

// Labels go to the ring in a fixed order -- chain of the group, model index,
// offset -- so the ring, and the network trained from it, do not depend on
// the order in which workgroups run: a count pass, one ordered prefix sum,
// then a write pass.  One workgroup (one wave) per (model index, chain of
// the group); one lane per offset.
__global__ void k_learn_count(LearnView L, int type, const ChainScal* __restrict__ scal,
                              const ChainArrays* __restrict__ arrays, const int* __restrict__ list, int kmax,
                              const double* __restrict__ xa, const double* __restrict__ rsum,
                              uint32_t* __restrict__ cnt) {
  const int v = blockIdx.y, m = blockIdx.x;
  uint64_t j; int o; double lab;
  const bool ok = label_at(L, type, scal, arrays, list, kmax, xa, rsum, v, m, threadIdx.x, &j, &o, &lab);
  const unsigned long long mask = __ballot(ok);
  if (threadIdx.x == 0) cnt[v * kmax + m] = (uint32_t)__popcll(mask);
}

// base[i] = ring position of row i's first label; the counter advances by
// the total
__global__ void k_learn_base(const uint32_t* __restrict__ cnt, int rows, unsigned long long* __restrict__ base,
                             unsigned long long* __restrict__ counter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long at = *counter;
  for (int i = 0; i < rows; ++i) { base[i] = at; at += cnt[i]; }
  *counter = at;
}

__global__ void k_learn_write(LearnView L, int type, const ChainScal* __restrict__ scal,
                              const ChainArrays* __restrict__ arrays, const int* __restrict__ list, int kmax,
                              const double* __restrict__ xa, const double* __restrict__ rsum,
                              const unsigned long long* __restrict__ base, uint32_t* __restrict__ rj,
                              int32_t* __restrict__ ro, float* __restrict__ rl) {
  const int v = blockIdx.y, m = blockIdx.x;
  uint64_t j; int o; double lab;
  const bool ok = label_at(L, type, scal, arrays, list, kmax, xa, rsum, v, m, threadIdx.x, &j, &o, &lab);
  const unsigned long long mask = __ballot(ok);
  if (!ok) return;
  const unsigned long long below = threadIdx.x ? (mask & ((1ull << threadIdx.x) - 1ull)) : 0ull;
  const unsigned long long slot = (base[v * kmax + m] + (unsigned long long)__popcll(below)) % LEARN_RING;
  rj[slot] = (uint32_t)j;
  ro[slot] = o;
  rl[slot] = (float)lab;
}

__global__ void k_learn_pbar(const ChainArrays* __restrict__ arrays, const int* __restrict__ chains, int nc,
                             uint64_t p, float* __restrict__ pbar) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= p) return;
  double s = 0.0;
  for (int i = 0; i < nc; ++i) s += arrays[chains[i]].p_proposal[j];
  pbar[j] = (float)(s / (double)nc);
}

// minibatch: rows drawn uniformly from the filled part of the ring
__global__ void k_learn_batch(LearnView L, const uint32_t* __restrict__ rj, const int32_t* __restrict__ ro,
                              const float* __restrict__ rl, uint64_t count, uint32_t seed, uint32_t round,
                              uint32_t step, int rows, uint16_t* __restrict__ X, float* __restrict__ lab) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= rows) return;
  Rng r = rng_make(seed, 0x1EA2Du, round, step);
  r.c0 = (uint32_t)b;
  uint32_t o4[4];
  rng_block(r, o4);
  uint64_t u = ((uint64_t)o4[0] << 32) | o4[1];
  uint64_t s = u % count;
  float f[16];
  pair_features(L, rj[s], ro[s], f);
  store_row(X, b, f);
  lab[b] = rl[s];
}

// features of table rows [r0, r0 + rows): row = j * 2H + c
__global__ void k_learn_table_feat(LearnView L, uint64_t r0, int64_t rows, uint16_t* __restrict__ X) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  uint64_t row = r0 + (uint64_t)i;
  uint64_t j = row / (uint64_t)(2 * L.H);
  int o = learn_off(L.H, (int)(row % (uint64_t)(2 * L.H)));
  int64_t k = (int64_t)j + o;
  float f[16];
  if (k >= 0 && k < (int64_t)L.p) {
    pair_features(L, j, o, f);
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) f[q] = 0.0f;
  }
  store_row(X, i, f);
}

__global__ void k_learn_table_store(LearnView L, uint64_t r0, int64_t rows, const float* __restrict__ out,
                                    float* __restrict__ table) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  uint64_t row = r0 + (uint64_t)i;
  uint64_t j = row / (uint64_t)(2 * L.H);
  int o = learn_off(L.H, (int)(row % (uint64_t)(2 * L.H)));
  int64_t k = (int64_t)j + o;
  float w = 0.0f;
  if (k >= 0 && k < (int64_t)L.p) {
    float v = out[i];
    v = isfinite(v) ? fminf(30.0f, fmaxf(-30.0f, v)) : 0.0f;
    w = expf(0.5f * v);
  }
  table[row] = w;
}

__global__ void k_learn_features_rows(LearnView L, const uint32_t* __restrict__ js, const int32_t* __restrict__ os,
                                      int64_t rows, float* __restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float f[16];
  pair_features(L, js[i], os[i], f);
  for (int q = 0; q < 16; ++q) out[i * 16 + q] = f[q];
}

// ------------------------------------------------ per-SNP proposal ------
// Features of SNP j for a chain in its present state: the marginal score,
// the chain's p_proposal (the running mean of its RB probabilities so far),
// the mean p_proposal over the dataset's chains, whether j is in the model,
// its LD to the model (largest and summed r^2 over model SNPs within the
// window), the largest marginal score around it, and its score conditional
// on the model SNP it is most correlated with.
__device__ __forceinline__ float clipped_logit(double q) {
  const double c = fmin(fmax(q, 1e-6), 1.0 - 1e-6);
  return (float)fmin(12.0, fmax(-12.0, log(c / (1.0 - c))));
}

__device__ void snp_features(const LearnView& L, const ChainArrays& A, uint64_t j, float* f) {
  const double n = (double)L.n;
  const double sj = L.pre[j * L.off + L.ot], cj = L.pre[j * L.off + L.ot + 1];
  const float zj = L.z[j];
  const int HW = L.H < LEARN_SWIN ? L.H : LEARN_SWIN;
  float r2max = 0.0f, r2sum = 0.0f, zmax = 0.0f, rbest = 0.0f, zbest = 0.0f, pbest = 0.0f;
  for (int o = -HW; o <= HW; ++o) {
    const int64_t k = (int64_t)j + o;
    if (o == 0 || k < 0 || k >= (int64_t)L.p) continue;
    zmax = fmaxf(zmax, fabsf(L.z[k]));
    if (A.model_inds[k] < 0) continue;
    const double sk = L.pre[(uint64_t)k * L.off + L.ot], ck = L.pre[(uint64_t)k * L.off + L.ot + 1];
    const uint64_t lo = (uint64_t)k < j ? (uint64_t)k : j;
    const double cov = (double)L.band[lo * (L.H + 1) + (o < 0 ? -o : o)] - sj * sk / n;
    double r = (cj > 0.0 && ck > 0.0) ? cov / sqrt(cj * ck) : 0.0;
    r = fmin(1.0, fmax(-1.0, r));
    const float r2 = (float)(r * r);
    r2sum += r2;
    if (r2 > r2max) { r2max = r2; rbest = (float)r; zbest = L.z[k]; pbest = (float)A.p_proposal[k]; }
  }
  const float zc = r2max > 0.0f ? (zj - rbest * zbest) / sqrtf(fmaxf(1.0f - r2max, 1e-3f)) : zj;
  f[0] = zj / 10.0f;
  f[1] = fminf(zj * zj / 100.0f, 25.0f);
  f[2] = clipped_logit(A.p_proposal[j]) / 6.0f;
  f[3] = clipped_logit((double)L.pbar[j]) / 6.0f;
  f[4] = A.model_inds[j] >= 0 ? 1.0f : 0.0f;
  f[5] = r2max;
  f[6] = fminf(r2sum, 5.0f);
  f[7] = fminf(zmax / 10.0f, 5.0f);
  f[8] = (float)(sj / n) * 0.5f;
  f[9] = r2max > 0.0f ? clipped_logit((double)pbest) / 6.0f : -2.0f;
  f[10] = fminf(fmaxf(zc / 10.0f, -5.0f), 5.0f);
  f[11] = fminf(zc * zc / 100.0f, 25.0f);
  f[12] = 0.0f;
  f[13] = 0.0f;
  f[14] = 0.0f;
  f[15] = 1.0f;
}

// Samples of one RB group: for every chain of the group, the SNPs within
// the window around each model SNP (one wave per (model index, chain), one
// lane per offset) and LEARN_SRAND SNPs drawn uniformly.  Features use the
// chain's state before this scan's update of p_proposal; the label is the
// scan's RB probability as a clipped logit.  Written in a fixed order
// (count, ordered prefix sum, write), like the swap labels.
__device__ __forceinline__ bool snp_window_at(const LearnView& L, const ChainScal& CS, const ChainArrays& A,
                                              int m, int lane, uint64_t* j_out) {
  const int HW = L.H < LEARN_SWIN ? L.H : LEARN_SWIN;
  if (m >= CS.nloci || lane > 2 * HW) return false;
  const int64_t k = (int64_t)A.idx[m] + (lane - HW);
  if (k < 0 || k >= (int64_t)L.p) return false;
  *j_out = (uint64_t)k;
  return true;
}

__global__ void k_snp_count(LearnView L, const ChainScal* __restrict__ scal, const ChainArrays* __restrict__ arrays,
                            const int* __restrict__ list, int kmax, uint32_t* __restrict__ cnt) {
  const int v = blockIdx.y, m = blockIdx.x, c = list[v];
  uint64_t j;
  const bool ok = snp_window_at(L, scal[c], arrays[c], m, threadIdx.x, &j);
  const unsigned long long mask = __ballot(ok);
  if (threadIdx.x == 0) cnt[v * kmax + m] = (uint32_t)__popcll(mask);
}

// sbase[v * kmax + m]: ring position of window row (v, m); sbase[nv * kmax
// + v]: chain v's first random row
__global__ void k_snp_base(const uint32_t* __restrict__ cnt, int nv, int kmax,
                           unsigned long long* __restrict__ base, unsigned long long* __restrict__ counter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long at = *counter;
  for (int v = 0; v < nv; ++v) {
    base[(size_t)nv * kmax + v] = at;
    at += LEARN_SRAND;
    for (int m = 0; m < kmax; ++m) { base[(size_t)v * kmax + m] = at; at += cnt[(size_t)v * kmax + m]; }
  }
  *counter = at;
}

__device__ __forceinline__ void snp_put(const LearnView& L, const ChainArrays& A, uint64_t j,
                                        unsigned long long slot, uint16_t* __restrict__ X,
                                        float* __restrict__ lab) {
  float f[16];
  snp_features(L, A, j, f);
  store_row(X, (int64_t)(slot % LEARN_SRING), f);
  lab[slot % LEARN_SRING] = clipped_logit(A.p_r[j]);
}

__global__ void k_snp_write(LearnView L, const ChainScal* __restrict__ scal, const ChainArrays* __restrict__ arrays,
                            const int* __restrict__ list, int kmax, const unsigned long long* __restrict__ base,
                            uint16_t* __restrict__ X, float* __restrict__ lab) {
  const int v = blockIdx.y, m = blockIdx.x, c = list[v];
  uint64_t j;
  const bool ok = snp_window_at(L, scal[c], arrays[c], m, threadIdx.x, &j);
  const unsigned long long mask = __ballot(ok);
  if (!ok) return;
  const unsigned long long below = threadIdx.x ? (mask & ((1ull << threadIdx.x) - 1ull)) : 0ull;
  snp_put(L, arrays[c], j, base[(size_t)v * kmax + m] + (unsigned long long)__popcll(below), X, lab);
}

__global__ void k_snp_write_rand(LearnView L, const ChainArrays* __restrict__ arrays, const int* __restrict__ list,
                                 int nv, int kmax, const unsigned long long* __restrict__ base, uint32_t seed,
                                 uint32_t group, uint16_t* __restrict__ X, float* __restrict__ lab) {
  const int v = blockIdx.y, c = list[v];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= LEARN_SRAND) return;
  Rng r = rng_make(seed, 0x5A9Du, group, (uint32_t)v);
  r.c0 = (uint32_t)i;
  uint32_t o4[4];
  rng_block(r, o4);
  const uint64_t j = (((uint64_t)o4[0] << 32) | o4[1]) % L.p;
  snp_put(L, arrays[c], j, base[(size_t)nv * kmax + v] + (unsigned long long)i, X, lab);
}

// minibatch rows drawn uniformly from the filled part of the ring
__global__ void k_snp_batch(const uint16_t* __restrict__ sx, const float* __restrict__ sl, uint64_t count,
                            uint32_t seed, uint32_t round, uint32_t step, int rows, uint16_t* __restrict__ X,
                            float* __restrict__ lab) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= rows) return;
  Rng r = rng_make(seed, 0x5B47u, round, step);
  r.c0 = (uint32_t)b;
  uint32_t o4[4];
  rng_block(r, o4);
  const uint64_t s = ((((uint64_t)o4[0] << 32) | o4[1])) % count;
  *(uint4*)(X + (size_t)b * 16) = *(const uint4*)(sx + s * 16);
  *(uint4*)(X + (size_t)b * 16 + 8) = *(const uint4*)(sx + s * 16 + 8);
  lab[b] = sl[s];
}

// one chain's features of SNPs [r0, r0 + rows)
__global__ void k_snp_table_feat(LearnView L, const ChainArrays* __restrict__ arrays, int c, uint64_t r0,
                                 int64_t rows, uint16_t* __restrict__ X) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float f[16];
  snp_features(L, arrays[c], r0 + (uint64_t)i, f);
  store_row(X, i, f);
}

// features of SNPs js[0..rows) in chain c's state, for checking
__global__ void k_snp_features_rows(LearnView L, const ChainArrays* __restrict__ arrays, int c,
                                    const uint32_t* __restrict__ js, int64_t rows, float* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float f[16];
  snp_features(L, arrays[c], js[i], f);
  for (int q = 0; q < 16; ++q) out[i * 16 + q] = f[q];
}

// the table: the network's logit through the logistic function
__global__ void k_snp_table_store(uint64_t r0, int64_t rows, const float* __restrict__ out,
                                  double* __restrict__ ptab) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float v = out[i];
  v = isfinite(v) ? fminf(12.0f, fmaxf(-12.0f, v)) : -12.0f;
  ptab[r0 + (uint64_t)i] = 1.0 / (1.0 + exp(-(double)v));
}

static unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int learn_create(LearnDS** out, bmg_ctx* ctx, const DevData& D, const LearnCfg& cfg, int n_chains,
                 hipStream_t s) {
  *out = nullptr;
  if (cfg.H < 1 || cfg.H > BMG_LH_MAX) { set_error("learned proposal: window must be 1..32"); return BMG_E_ARG; }
  if (D.type > 3) { set_error("learned proposal: one coding (A, H, D or R) only"); return BMG_E_UNSUPPORTED; }
  if (cfg.batch < 64 || cfg.steps < 0 || !(cfg.lr > 0.0f)) { set_error("learned proposal: bad training settings"); return BMG_E_ARG; }
  LearnDS* L = new LearnDS();
  L->cfg = cfg;
  L->type = D.type;
  L->n_chains = n_chains;
  const uint64_t p = D.p;
  auto fail = [&](int rc) { learn_destroy(L); return rc; };
  if (hipMalloc(&L->band, sizeof(int32_t) * p * (cfg.H + 1)) != hipSuccess ||
      hipMalloc(&L->z, sizeof(float) * p) != hipSuccess || hipMalloc(&L->pbar, sizeof(float) * p) != hipSuccess ||
      hipMalloc(&L->rj, sizeof(uint32_t) * LEARN_RING) != hipSuccess ||
      hipMalloc(&L->ro, sizeof(int32_t) * LEARN_RING) != hipSuccess ||
      hipMalloc(&L->rl, sizeof(float) * LEARN_RING) != hipSuccess ||
      hipMalloc(&L->counter, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&L->lcnt, sizeof(uint32_t) * LEARN_MAX_ROWS) != hipSuccess ||
      hipMalloc(&L->lbase, sizeof(unsigned long long) * LEARN_MAX_ROWS) != hipSuccess ||
      hipMalloc(&L->Xb, sizeof(uint16_t) * 16 * cfg.batch) != hipSuccess ||
      hipMalloc(&L->lb, sizeof(float) * cfg.batch) != hipSuccess ||
      hipMalloc(&L->Xc, sizeof(uint16_t) * 16 * LEARN_CHUNK) != hipSuccess ||
      hipMalloc(&L->oc, sizeof(float) * LEARN_CHUNK) != hipSuccess ||
      hipMalloc(&L->table, sizeof(float) * p * 2 * cfg.H) != hipSuccess ||
      hipMalloc(&L->d_chains, sizeof(int) * std::max(n_chains, 1)) != hipSuccess) {
    set_error("learned proposal: device allocation failed");
    return fail(BMG_E_HIP);
  }
  if (cfg.snp &&
      (hipMalloc(&L->sx, sizeof(uint16_t) * 16 * (size_t)LEARN_SRING) != hipSuccess ||
       hipMalloc(&L->sl, sizeof(float) * LEARN_SRING) != hipSuccess ||
       hipMalloc(&L->scounter, sizeof(unsigned long long)) != hipSuccess ||
       hipMalloc(&L->scnt, sizeof(uint32_t) * LEARN_MAX_ROWS) != hipSuccess ||
       hipMalloc(&L->sbase, sizeof(unsigned long long) * (LEARN_MAX_ROWS + 64)) != hipSuccess ||
       hipMalloc(&L->ptab, sizeof(double) * p) != hipSuccess)) {
    set_error("learned proposal: device allocation failed");
    return fail(BMG_E_HIP);
  }
  int rc = mlp_alloc(&L->mlp);
  if (rc) return fail(rc);
  if (cfg.snp) {
    rc = mlp_alloc(&L->snet);
    if (rc) return fail(rc);
  }
  L->v = LearnView{D.pre, D.offset, D.ot[D.type], p, D.n, cfg.H, L->band, L->z, L->pbar};
  // band X^T X over the window, exact int32 (int8 MFMA)
  rc = launch_band(ctx, (uint32_t)(cfg.H + 1), D.type, D.type, L->band, s);
  if (rc) return fail(rc);
  // marginal scores from one scan of y - mean(y)
  std::vector<double> r(ctx->n_pad, 0.0);
  double my = 0.0, syy = 0.0;
  for (uint32_t i = 0; i < ctx->n; ++i) my += ctx->y_host[i];
  my /= (double)ctx->n;
  for (uint32_t i = 0; i < ctx->n; ++i) { r[i] = ctx->y_host[i] - my; syy += r[i] * r[i]; }
  double *rd = nullptr, *rx = nullptr;
  if (hipMalloc(&rd, sizeof(double) * ctx->n_pad) != hipSuccess || hipMalloc(&rx, sizeof(double) * p) != hipSuccess) {
    hipFree(rd); set_error("learned proposal: device allocation failed"); return fail(BMG_E_HIP);
  }
  hipMemcpyAsync(rd, r.data(), sizeof(double) * ctx->n_pad, hipMemcpyHostToDevice, s);
  rc = launch_scan(ctx, rd, 1, D.type, 5, rx, s);
  if (!rc) {
    hipLaunchKernelGGL(k_learn_static, dim3(blocks((int64_t)p)), dim3(256), 0, s, L->v, rx, syy, L->z);
    if (hipGetLastError() != hipSuccess) { set_error("learned proposal: static kernel"); rc = BMG_E_HIP; }
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) { set_error("learned proposal: set-up failed"); rc = BMG_E_HIP; }
  hipFree(rd); hipFree(rx);
  if (rc) return fail(rc);
  hipMemsetAsync(L->counter, 0, sizeof(unsigned long long), s);
  std::vector<float> flat(p, 0.0f);
  hipMemcpyAsync(L->pbar, flat.data(), sizeof(float) * p, hipMemcpyHostToDevice, s);
  // He-initialised weights (fixed seed), small last layer
  std::vector<float> P(MLP_NPAR, 0.0f);
  std::mt19937_64 gen(0x5EEDull * 1000003ull + cfg.seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int i = 0; i < MLP_IN * MLP_HID; ++i) P[MLP_OFF_W1 + i] = nd(gen) * sqrtf(2.0f / MLP_IN);
  for (int i = 0; i < MLP_HID * MLP_HID; ++i) P[MLP_OFF_W2 + i] = nd(gen) * sqrtf(2.0f / MLP_HID);
  for (int i = 0; i < MLP_HID; ++i) P[MLP_OFF_W3 + i] = nd(gen) * 0.05f;
  hipMemcpyAsync(L->mlp.P, P.data(), sizeof(float) * MLP_NPAR, hipMemcpyHostToDevice, s);
  rc = mlp_launch_copies(L->mlp, s);
  if (!rc && cfg.snp) {
    // the per-SNP network: its own He-initialised weights
    std::vector<float> Q(MLP_NPAR, 0.0f);
    std::mt19937_64 g2(0x5EEDull * 1000003ull + cfg.seed + 0x9E37ull);
    for (int i = 0; i < MLP_IN * MLP_HID; ++i) Q[MLP_OFF_W1 + i] = nd(g2) * sqrtf(2.0f / MLP_IN);
    for (int i = 0; i < MLP_HID * MLP_HID; ++i) Q[MLP_OFF_W2 + i] = nd(g2) * sqrtf(2.0f / MLP_HID);
    for (int i = 0; i < MLP_HID; ++i) Q[MLP_OFF_W3 + i] = nd(g2) * 0.05f;
    hipMemcpyAsync(L->snet.P, Q.data(), sizeof(float) * MLP_NPAR, hipMemcpyHostToDevice, s);
    hipMemsetAsync(L->scounter, 0, sizeof(unsigned long long), s);
    rc = mlp_launch_copies(L->snet, s);
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) { set_error("learned proposal: set-up failed"); rc = BMG_E_HIP; }
  if (rc) return fail(rc);
  *out = L;
  return BMG_OK;
}

void learn_destroy(LearnDS* L) {
  if (!L) return;
  hipFree(L->band); hipFree(L->z); hipFree(L->pbar); hipFree(L->rj); hipFree(L->ro); hipFree(L->rl);
  hipFree(L->counter); hipFree(L->lcnt); hipFree(L->lbase); hipFree(L->Xb); hipFree(L->lb); hipFree(L->Xc); hipFree(L->oc); hipFree(L->table);
  hipFree(L->d_chains);
  hipFree(L->sx); hipFree(L->sl); hipFree(L->scounter); hipFree(L->scnt); hipFree(L->sbase); hipFree(L->ptab);
  mlp_free(&L->mlp);
  mlp_free(&L->snet);
  delete L;
}

float* learn_table(LearnDS* L) { return L->table; }
int learn_H(const LearnDS* L) { return L->cfg.H; }
int learn_frozen(const LearnDS* L) { return L->frozen; }
void learn_freeze(LearnDS* L) { L->frozen = 1; }
uint64_t learn_rounds(const LearnDS* L) { return L->rounds; }

int learn_labels(LearnDS* L, const DevData* d_data, const ChainConst* consts, const ChainScal* scal,
                 const ChainArrays* arrays, const int* d_list, int nv, int kmax, const double* xa,
                 const double* rsum, hipStream_t s) {
  (void)d_data; (void)consts;
  if (L->frozen || nv <= 0 || !L->cfg.swap) return BMG_OK;
  if (nv * kmax > LEARN_MAX_ROWS) { set_error("learned proposal: label rows exceed the buffer"); return BMG_E_ARG; }
  hipLaunchKernelGGL(k_learn_count, dim3(kmax, nv), dim3(64), 0, s, L->v, L->type, scal, arrays, d_list, kmax, xa,
                     rsum, L->lcnt);
  hipLaunchKernelGGL(k_learn_base, dim3(1), dim3(64), 0, s, L->lcnt, nv * kmax, L->lbase, L->counter);
  hipLaunchKernelGGL(k_learn_write, dim3(kmax, nv), dim3(64), 0, s, L->v, L->type, scal, arrays, d_list, kmax, xa,
                     rsum, L->lbase, L->rj, L->ro, L->rl);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int learn_round(LearnDS* L, const ChainArrays* d_arrays, const int* chains_host, int n, hipStream_t s) {
  if (L->frozen) return BMG_OK;
  const uint64_t p = L->v.p;
  BMG_HIP(hipMemcpyAsync(L->d_chains, chains_host, sizeof(int) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_learn_pbar, dim3(blocks((int64_t)p)), dim3(256), 0, s, d_arrays, L->d_chains, n, p, L->pbar);
  BMG_HIP(hipGetLastError());
  unsigned long long cnt = 0;
  BMG_HIP(hipMemcpyAsync(&cnt, L->counter, sizeof(cnt), hipMemcpyDeviceToHost, s));
  BMG_HIP(hipStreamSynchronize(s));
  uint64_t count = L->cfg.swap ? std::min<unsigned long long>(cnt, LEARN_RING) : 0;
  if (L->cfg.snp) {
    unsigned long long scount = 0;
    BMG_HIP(hipMemcpyAsync(&scount, L->scounter, sizeof(scount), hipMemcpyDeviceToHost, s));
    BMG_HIP(hipStreamSynchronize(s));
    const uint64_t sc = std::min<unsigned long long>(scount, LEARN_SRING);
    if (sc >= 64) {
      for (int st = 0; st < L->cfg.steps; ++st) {
        hipLaunchKernelGGL(k_snp_batch, dim3(blocks(L->cfg.batch)), dim3(256), 0, s, L->sx, L->sl, sc, L->cfg.seed,
                           (uint32_t)L->rounds, (uint32_t)st, L->cfg.batch, L->Xb, L->lb);
        BMG_HIP(hipGetLastError());
        int rc = mlp_launch_train(L->snet, L->Xb, L->lb, L->cfg.batch, nullptr, s);
        if (rc) return rc;
        rc = mlp_launch_adam(L->snet, L->cfg.lr, ++L->sadam_t, s);
        if (rc) return rc;
      }
    }
  }
  if (!L->cfg.swap) { L->rounds++; return BMG_OK; }
  if (count >= 64) {
    for (int st = 0; st < L->cfg.steps; ++st) {
      hipLaunchKernelGGL(k_learn_batch, dim3(blocks(L->cfg.batch)), dim3(256), 0, s, L->v, L->rj, L->ro, L->rl,
                         count, L->cfg.seed, (uint32_t)L->rounds, (uint32_t)st, L->cfg.batch, L->Xb, L->lb);
      BMG_HIP(hipGetLastError());
      int rc = mlp_launch_train(L->mlp, L->Xb, L->lb, L->cfg.batch, nullptr, s);
      if (rc) return rc;
      rc = mlp_launch_adam(L->mlp, L->cfg.lr, ++L->adam_t, s);
      if (rc) return rc;
    }
  }
  // the table over every (j, offset) pair
  const uint64_t total = p * (uint64_t)(2 * L->cfg.H);
  for (uint64_t r0 = 0; r0 < total; r0 += LEARN_CHUNK) {
    int64_t rows = (int64_t)std::min<uint64_t>(LEARN_CHUNK, total - r0);
    hipLaunchKernelGGL(k_learn_table_feat, dim3(blocks(rows)), dim3(256), 0, s, L->v, r0, rows, L->Xc);
    BMG_HIP(hipGetLastError());
    int rc = mlp_launch_forward(L->mlp, L->Xc, rows, L->oc, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_learn_table_store, dim3(blocks(rows)), dim3(256), 0, s, L->v, r0, rows, L->oc, L->table);
    BMG_HIP(hipGetLastError());
  }
  L->rounds++;
  return BMG_OK;
}

int learn_snp_labels(LearnDS* L, const ChainScal* scal, const ChainArrays* arrays, const int* d_list, int nv,
                     int kmax, hipStream_t s) {
  if (L->frozen || nv <= 0 || !L->cfg.snp) return BMG_OK;
  if (nv * kmax > LEARN_MAX_ROWS || nv > 64) { set_error("learned proposal: label rows exceed the buffer"); return BMG_E_ARG; }
  hipLaunchKernelGGL(k_snp_count, dim3(kmax, nv), dim3(64), 0, s, L->v, scal, arrays, d_list, kmax, L->scnt);
  hipLaunchKernelGGL(k_snp_base, dim3(1), dim3(64), 0, s, L->scnt, nv, kmax, L->sbase, L->scounter);
  hipLaunchKernelGGL(k_snp_write, dim3(kmax, nv), dim3(64), 0, s, L->v, scal, arrays, d_list, kmax, L->sbase, L->sx,
                     L->sl);
  hipLaunchKernelGGL(k_snp_write_rand, dim3((LEARN_SRAND + 255) / 256, nv), dim3(256), 0, s, L->v, arrays, d_list, nv,
                     kmax, L->sbase, L->cfg.seed, L->sgroups, L->sx, L->sl);
  BMG_HIP(hipGetLastError());
  L->sgroups++;
  return BMG_OK;
}

int learn_snp_apply(LearnDS* L, const DevData* d_data, const ChainConst* d_consts, ChainScal* d_scal,
                    const ChainArrays* d_arrays, const int* chains_host, int n, const uint64_t* lev_len, int top,
                    hipStream_t s) {
  if (!L->cfg.snp || L->sadam_t == 0) return BMG_OK;     // never trained: the plain proposal stays
  const uint64_t p = L->v.p;
  for (int i = 0; i < n; ++i) {
    const int c = chains_host[i];
    for (uint64_t r0 = 0; r0 < p; r0 += LEARN_CHUNK) {
      const int64_t rows = (int64_t)std::min<uint64_t>(LEARN_CHUNK, p - r0);
      hipLaunchKernelGGL(k_snp_table_feat, dim3(blocks(rows)), dim3(256), 0, s, L->v, d_arrays, c, r0, rows, L->Xc);
      BMG_HIP(hipGetLastError());
      int rc = mlp_launch_forward(L->snet, L->Xc, rows, L->oc, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_snp_table_store, dim3(blocks(rows)), dim3(256), 0, s, r0, rows, L->oc, L->ptab);
      BMG_HIP(hipGetLastError());
    }
    BMG_HIP(hipMemcpyAsync(L->d_chains, &chains_host[i], sizeof(int), hipMemcpyHostToDevice, s));
    if (launch_set_proposal(d_data, d_consts, d_arrays, c, L->ptab, p, s) ||
        launch_tree_build(d_data, d_consts, d_scal, d_arrays, L->d_chains, 1, lev_len, top, 0, 0, s)) {
      set_error("learned proposal: table update failed"); return BMG_E_HIP;
    }
    BMG_HIP(hipStreamSynchronize(s));
  }
  L->applied = 1;
  return BMG_OK;
}

int learn_snp_applied(const LearnDS* L) { return L->applied; }

int learn_snp_features(LearnDS* L, const ChainArrays* d_arrays, int chain, const uint32_t* j_host, int64_t rows,
                       float* out, hipStream_t s) {
  uint32_t* dj = nullptr;
  float* dout = nullptr;
  int rc = BMG_OK;
  if (hipMalloc(&dj, sizeof(uint32_t) * rows) != hipSuccess || hipMalloc(&dout, sizeof(float) * 16 * rows) != hipSuccess) {
    set_error("allocation failed"); rc = BMG_E_HIP;
  }
  if (!rc) {
    hipMemcpyAsync(dj, j_host, sizeof(uint32_t) * rows, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_snp_features_rows, dim3(blocks(rows)), dim3(256), 0, s, L->v, d_arrays, chain, dj, rows, dout);
    if (hipGetLastError() != hipSuccess) { set_error("features kernel"); rc = BMG_E_HIP; }
  }
  if (!rc) {
    hipMemcpyAsync(out, dout, sizeof(float) * 16 * rows, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) { set_error("features failed"); rc = BMG_E_HIP; }
  }
  hipFree(dj); hipFree(dout);
  return rc;
}

int learn_features(LearnDS* L, const uint32_t* j_host, const int32_t* off_host, int64_t rows, float* out,
                   hipStream_t s) {
  uint32_t* dj = nullptr;
  int32_t* dof = nullptr;
  float* dout = nullptr;
  int rc = BMG_OK;
  if (hipMalloc(&dj, sizeof(uint32_t) * rows) != hipSuccess || hipMalloc(&dof, sizeof(int32_t) * rows) != hipSuccess ||
      hipMalloc(&dout, sizeof(float) * 16 * rows) != hipSuccess) {
    set_error("allocation failed"); rc = BMG_E_HIP;
  }
  if (!rc) {
    hipMemcpyAsync(dj, j_host, sizeof(uint32_t) * rows, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(dof, off_host, sizeof(int32_t) * rows, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_learn_features_rows, dim3(blocks(rows)), dim3(256), 0, s, L->v, dj, dof, rows, dout);
    if (hipGetLastError() != hipSuccess) { set_error("features kernel"); rc = BMG_E_HIP; }
  }
  if (!rc) {
    hipMemcpyAsync(out, dout, sizeof(float) * 16 * rows, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) { set_error("features failed"); rc = BMG_E_HIP; }
  }
  hipFree(dj); hipFree(dof); hipFree(dout);
  return rc;
}

int learn_samples(LearnDS* L, uint32_t* j, int32_t* off, float* label, int64_t cap, int64_t* count,
                  hipStream_t s) {
  unsigned long long cnt = 0;
  BMG_HIP(hipMemcpyAsync(&cnt, L->counter, sizeof(cnt), hipMemcpyDeviceToHost, s));
  BMG_HIP(hipStreamSynchronize(s));
  int64_t k = (int64_t)std::min<unsigned long long>(cnt, LEARN_RING);
  *count = (int64_t)cnt;
  int64_t m = std::min(k, cap);
  if (m > 0) {
    if (j) BMG_HIP(hipMemcpy(j, L->rj, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
    if (off) BMG_HIP(hipMemcpy(off, L->ro, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
    if (label) BMG_HIP(hipMemcpy(label, L->rl, sizeof(float) * m, hipMemcpyDeviceToHost));
  }
  return BMG_OK;
}

int learn_params(LearnDS* L, float* params, hipStream_t s) {
  BMG_HIP(hipStreamSynchronize(s));
  BMG_HIP(hipMemcpy(params, L->mlp.P, sizeof(float) * MLP_NPAR, hipMemcpyDeviceToHost));
  return BMG_OK;
}

int learn_static(LearnDS* L, float* z, float* pbar, int32_t* band, hipStream_t s) {
  BMG_HIP(hipStreamSynchronize(s));
  const uint64_t p = L->v.p;
  if (z) BMG_HIP(hipMemcpy(z, L->z, sizeof(float) * p, hipMemcpyDeviceToHost));
  if (pbar) BMG_HIP(hipMemcpy(pbar, L->pbar, sizeof(float) * p, hipMemcpyDeviceToHost));
  if (band) BMG_HIP(hipMemcpy(band, L->band, sizeof(int32_t) * p * (L->cfg.H + 1), hipMemcpyDeviceToHost));
  return BMG_OK;
}

}  // namespace bmg
