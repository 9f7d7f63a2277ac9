// bmg_rb.hip -- the Rao-Blackwellised all-SNP scan of a batch of chains
// (src/sampler.cpp:32-261) and the proposal updates that follow it
// (src/sampler.cpp:739-810), plus the sum-tree (re)build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>

#include "bmg_chain.h"
#include "bmg_device.h"
#include "bmg_rb.h"

using namespace bmg;

// r_v = y - y_hat of chain list[v] (zero past n); sums of r per vector
// r = y - y_hat for chain list[v], and rsum[v] = sum of r: one workgroup per
// vector, so the sum is reduced in a fixed order (the RB PIPs do not depend
// on the order in which workgroups finish)
__global__ void k_residuals(const double* __restrict__ y, uint32_t n, uint32_t n_pad,
                            const ChainArrays* __restrict__ arrays, const int* __restrict__ list,
                            double* __restrict__ r, double* __restrict__ rsum) {
  __shared__ double red[8];
  int v = blockIdx.y;
  const double* yh = arrays[list[v]].y_hat;
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < n_pad; i += blockDim.x) {
    double ri = i < n ? y[i] - yh[i] : 0.0;
    r[(size_t)v * n_pad + i] = ri;
    s += ri;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) rsum[v] = s;
}

__device__ inline double rb_prior_add(const ChainConst& K, double Ns_t, double n_old) {
  return (log(K.g_a + n_old) - log(K.g_b + K.m_g - n_old - 1)) +
         (log(K.tp_type + Ns_t) - log(K.tp_sum + n_old));
}

// p_r of every SNP for chain list[v] from x_j . r (xa: coding of the type or
// A for AH; xh: H for AH)
__global__ void k_rb_finalize(const DevData* __restrict__ data, const ChainConst* __restrict__ consts,
                              const ChainScal* __restrict__ scal,
                              const ChainArrays* __restrict__ arrays, const int* __restrict__ list,
                              int kmax, const double* __restrict__ xa, const double* __restrict__ xh,
                              const double* __restrict__ rsum) {
  int v = blockIdx.y;
  int c = list[v];
  const ChainConst& K = consts[c];
  const ChainScal& CS = scal[c];
  const ChainArrays& A = arrays[c];
  const DevData& D = data[K.ds];
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= D.p) return;
  const int type = D.type;
  const double n = (double)D.n;
  const double mr = rsum[v] / n;
  const double s2x2 = 2.0 * CS.sigma2;
  const double nl = (double)CS.nloci;
  // tabulated prior changes (src/sampler.cpp:51-76), single type
  const double lmp_add = CS.nloci >= K.max_size ? -INFINITY : rb_prior_add(K, nl, nl);
  const double lmp_rem = rb_prior_add(K, nl - 1, nl - 1);
  double dv[4] = {CS.inv_tau2_alpha2[0], CS.inv_tau2_alpha2[1], CS.inv_tau2_alpha2[2],
                  CS.inv_tau2_alpha2[3]};
  if (K.use_ind) {
    Rng rr = rng_make(K.seed, K.chain_id, (uint32_t)j, 0x80000000u | CS.rb_call_idx);
    for (int t = 0; t < 4; ++t)
      if (K.allow_terms[t]) dv[t] = 1.0 / (CS.alpha2 * rng_sinvchi2(rr, K.nu_tau2[t], K.s2_tau2[t]));
  }
  const double* pre = D.pre + j * D.offset;
  int mind = A.model_inds[j];
  double lmp = mind < 0 ? lmp_add : lmp_rem;
  double p;
  if (type == 4) {
    double sa = pre[0], css_a = pre[1], sh = pre[2], css_h = pre[3], vah = pre[D.offset - 1];
    double rxa = xa[(size_t)v * D.p + j] - sa * mr;
    double rxh = xh[(size_t)v * D.p + j] - sh * mr;
    if (mind >= 0) {
      int c1 = A.idx[kmax + mind], c2 = A.idx[2 * kmax + mind];
      double ba = A.vec[3 * kmax + c1], bh = A.vec[3 * kmax + c2];
      if (K.use_ind) { dv[0] = A.vec[2 * kmax + c1]; dv[1] = A.vec[2 * kmax + c2]; }
      rxa += ba * css_a + bh * vah;
      rxh += bh * css_h + ba * vah;
    }
    double va = css_a + dv[0], vh = css_h + dv[1];
    double slq = -log(dv[0]) - log(dv[1]);
    double det = va * vh - vah * vah;
    double et = (rxa * rxa * vh - 2.0 * rxa * rxh * vah + rxh * rxh * va) / det;
    p = et / s2x2 - 0.5 * (log(det) + slq) + lmp;
  } else {
    int off = D.ot[type];
    double sx = pre[off], css = pre[off + 1];
    double rx = xa[(size_t)v * D.p + j] - sx * mr;
    if (mind >= 0) {
      int c1 = A.idx[kmax + mind];
      rx += A.vec[3 * kmax + c1] * css;
      if (K.use_ind) dv[type] = A.vec[2 * kmax + c1];
    }
    double det = css + dv[type];
    double slq = -log(dv[type]);
    double et = rx * rx / det;
    p = et / s2x2 - 0.5 * (log(det) + slq) + lmp;
  }
  p = exp(p);
  A.p_r[j] = isfinite(p) ? p / (1 + p) : 1.0;
}

// src/sampler.cpp:739-810 element-wise part: p_rao and p_proposal running
// means, q_add / q_rem and the new leaf values (zeroed leaves stay 0)
__global__ void k_rb_update(const DevData* __restrict__ data, const ChainConst* __restrict__ consts,
                            const ChainScal* __restrict__ scal, const ChainArrays* __restrict__ arrays,
                            const int* __restrict__ list) {
  int c = list[blockIdx.y];
  const ChainConst& K = consts[c];
  const ChainScal& CS = scal[c];
  const ChainArrays& A = arrays[c];
  const DevData& D = data[K.ds];
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= D.p) return;
  double pr = A.p_r[j];
  if (CS.rb_accum_rao) A.p_rao[j] = CS.rao_z2 * A.p_rao[j] + pr / CS.rao_z1;
  if (CS.rb_update_prop) A.p_proposal[j] = CS.prop_z2 * A.p_proposal[j] + pr / CS.prop_z1;
  if (CS.rb_rebuild) {
    double pp = A.p_proposal[j];
    double qa = fmax(pp, K.q_add_min), qr = fmax(1 - pp, K.q_rem_min);
    A.q_add[j] = qa; A.q_rem[j] = qr;
    double* la = A.tree_add + D.lev_off[0];
    double* lr = A.tree_rem + D.lev_off[0];
    if (la[j] != 0.0) la[j] = qa;
    if (lr[j] != 0.0) lr[j] = qr;
  }
}

// a proposal table from outside for chain c: p_proposal <- p, q_add and
// q_rem from it (src/sampler.cpp:517-524, 797-806) and the leaves of the
// sum trees (a zeroed leaf stays 0: every weight is at least q_add_min or
// q_rem_min, both positive); the upper levels are rebuilt after it
__global__ void k_set_proposal(const DevData* __restrict__ data, const ChainConst* __restrict__ consts,
                               const ChainArrays* __restrict__ arrays, int c, const double* __restrict__ p) {
  const ChainConst& K = consts[c];
  const DevData& D = data[K.ds];
  const ChainArrays& A = arrays[c];
  double* la = A.tree_add + D.lev_off[0];
  double* lr = A.tree_rem + D.lev_off[0];
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < D.p;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const double pp = p[j];
    const double qa = fmax(pp, K.q_add_min), qr = fmax(1 - pp, K.q_rem_min);
    A.p_proposal[j] = pp; A.q_add[j] = qa; A.q_rem[j] = qr;
    if (la[j] != 0.0) la[j] = qa;
    if (lr[j] != 0.0) lr[j] = qr;
  }
}

// one level of the sum trees: entry g of level h = group sum of group g of
// level h - 1 (inclusive wave scan, last lane); one wave per group
__global__ void k_tree_level(const DevData* __restrict__ data, const ChainConst* __restrict__ consts,
                             ChainScal* __restrict__ scal, const ChainArrays* __restrict__ arrays,
                             const int* __restrict__ list, int h, int only_rebuild) {
  int c = list[blockIdx.y];
  const ChainConst& K = consts[c];
  ChainScal& CS = scal[c];
  if (only_rebuild && !CS.rb_rebuild) return;
  const DevData& D = data[K.ds];
  const ChainArrays& A = arrays[c];
  int lane = threadIdx.x & 63;
  uint64_t g = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  int tr = blockIdx.z;
  double* t = tr == 0 ? A.tree_add : A.tree_rem;
  if (h > D.top) return;
  uint64_t n_groups = D.lev_len[h - 1] / 64;
  uint64_t len_h = D.lev_len[h];
  if (g >= len_h) return;
  double s = 0.0;
  if (g < n_groups) {
    double x = t[D.lev_off[h - 1] + 64 * g + lane];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      double y = __shfl_up(x, d, 64);
      if (lane >= d) x = x + y;
    }
    s = __shfl(x, 63, 64);
  }
  if (lane == 0) t[D.lev_off[h] + g] = s;
}

// totals: group sum of the top level; clears the pending RB flag
__global__ void k_tree_total(const DevData* __restrict__ data, const ChainConst* __restrict__ consts,
                             ChainScal* __restrict__ scal, const ChainArrays* __restrict__ arrays,
                             const int* __restrict__ list, int clear_pending) {
  int c = list[blockIdx.x];
  const ChainConst& K = consts[c];
  ChainScal& CS = scal[c];
  const DevData& D = data[K.ds];
  const ChainArrays& A = arrays[c];
  int lane = threadIdx.x & 63;
  int tr = threadIdx.x / 64;   // two waves: add, rem
  double* t = tr == 0 ? A.tree_add : A.tree_rem;
  double x = t[D.lev_off[D.top] + lane];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double y = __shfl_up(x, d, 64);
    if (lane >= d) x = x + y;
  }
  double s = __shfl(x, 63, 64);
  if (lane == 0) { if (tr == 0) CS.tot_add = s; else CS.tot_rem = s; }
  __syncthreads();
  if (threadIdx.x == 0 && clear_pending) {
    CS.rb_pending = 0; CS.rb_rebuild = 0; CS.rb_accum_rao = 0; CS.rb_update_prop = 0;
  }
}

namespace bmg {

int launch_residuals(const double* y, uint32_t n, uint32_t n_pad, const ChainArrays* arrays,
                     const int* list, int nv, double* r, double* rsum, hipStream_t s) {
  hipLaunchKernelGGL(k_residuals, dim3(1, nv), dim3(256), 0, s, y, n, n_pad, arrays, list, r, rsum);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_rb_finalize(const DevData* data, const ChainConst* consts, const ChainScal* scal,
                       const ChainArrays* arrays, const int* list, int nv, int kmax, uint64_t p,
                       const double* xa, const double* xh, const double* rsum, hipStream_t s) {
  hipLaunchKernelGGL(k_rb_finalize, dim3((unsigned)((p + 255) / 256), nv), dim3(256), 0, s, data, consts,
                     scal, arrays, list, kmax, xa, xh, rsum);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_rb_update(const DevData* data, const ChainConst* consts, const ChainScal* scal,
                     const ChainArrays* arrays, const int* list, int nv, uint64_t p, hipStream_t s) {
  hipLaunchKernelGGL(k_rb_update, dim3((unsigned)((p + 255) / 256), nv), dim3(256), 0, s, data, consts,
                     scal, arrays, list);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_set_proposal(const DevData* data, const ChainConst* consts, const ChainArrays* arrays, int chain,
                        const double* p, uint64_t np, hipStream_t s) {
  hipLaunchKernelGGL(k_set_proposal, dim3((unsigned)std::min<uint64_t>((np + 255) / 256, 4096)), dim3(256), 0, s,
                     data, consts, arrays, chain, p);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_tree_build(const DevData* data, const ChainConst* consts, ChainScal* scal,
                      const ChainArrays* arrays, const int* list, int nv, const uint64_t* lev_len,
                      int top, int only_rebuild, int clear_pending, hipStream_t s) {
  for (int h = 1; h <= top; ++h) {
    uint64_t groups = lev_len[h];
    unsigned gx = (unsigned)((groups + 3) / 4);
    hipLaunchKernelGGL(k_tree_level, dim3(gx, nv, 2), dim3(256), 0, s, data, consts, scal, arrays, list,
                       h, only_rebuild);
  }
  hipLaunchKernelGGL(k_tree_total, dim3(nv), dim3(128), 0, s, data, consts, scal, arrays, list,
                     clear_pending);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bmg
