// bmg_device.h -- device-side building blocks of the gfx950 engine:
// the project random stream, wavefront/workgroup reductions and the bit-plane
// view of 2-bit packed genotypes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BMG_WAVE 64
#define BMG_BLOCK 256

namespace bmg {

// ------------------------------------------------------------------ RNG --
// Philox4x32-10 (Salmon et al., SC'11).  Stream: key {seed, chain},
// counter {lo, hi, c2, c3}; one block per draw, then (lo, hi) += 1.
// u01 = (w0 * 2^21 + (w1 >> 11)) * 2^-53; normal = Box-Muller on the two
// halves of the block; gamma = Marsaglia-Tsang (2000).
struct Rng {
  uint32_t k0, k1;
  uint32_t c0, c1, c2, c3;
};

__host__ __device__ inline void philox10(uint32_t& x0, uint32_t& x1, uint32_t& x2,
                                         uint32_t& x3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
    x1 = (uint32_t)p1; x3 = (uint32_t)p0; x0 = n0; x2 = n2;
  }
}

__host__ __device__ inline double words_u01(uint32_t a, uint32_t b) {
  return (double)((((uint64_t)a) << 21) | (uint64_t)(b >> 11)) * 0x1.0p-53;
}

__host__ __device__ inline void rng_block(Rng& r, uint32_t o[4]) {
  o[0] = r.c0; o[1] = r.c1; o[2] = r.c2; o[3] = r.c3;
  philox10(o[0], o[1], o[2], o[3], r.k0, r.k1);
  if (++r.c0 == 0) ++r.c1;
}

__device__ __forceinline__ double rng_u01(Rng& r) {
  uint32_t o[4]; rng_block(r, o);
  return words_u01(o[0], o[1]);
}

__device__ __forceinline__ double rng_normal(Rng& r) {
  uint32_t o[4]; rng_block(r, o);
  double ua = words_u01(o[0], o[1]), ub = words_u01(o[2], o[3]);
  return sqrt(-2.0 * log(1.0 - ua)) * cos(6.283185307179586 * ub);
}

// Trials are capped so that a non-finite shape cannot stall a wavefront:
// each trial of these samplers succeeds with probability > 0.9 for any
// finite positive shape, so 1000 failures in a row never happen in a valid
// state, while a non-finite state fails fast and is flagged by the caller.
#define BMG_MAX_TRIALS 1000
__device__ __forceinline__ double rng_gamma_ge1(Rng& r, double a) {
  double d = a - 1.0 / 3.0;
  double c = 1.0 / sqrt(9.0 * d);
  for (int trial = 0; trial < BMG_MAX_TRIALS; ++trial) {
    double z = rng_normal(r);
    double t = 1.0 + c * z;
    if (t <= 0.0) continue;
    double v = t * t * t;
    double u = rng_u01(r);
    double z2 = z * z;
    if (u < 1.0 - 0.0331 * z2 * z2) return d * v;
    if (log(u) < 0.5 * z2 + d * (1.0 - v + log(v))) return d * v;
  }
  return NAN;
}

__device__ __forceinline__ double rng_gamma_impl(Rng& r, double a) {
  if (a < 1.0) {
    double g = rng_gamma_ge1(r, a + 1.0);
    double u = rng_u01(r);
    return g * pow(u, 1.0 / a);
  }
  return rng_gamma_ge1(r, a);
}

// The gamma draw is one out-of-line function that takes the stream by value
// and returns the advanced counter, so that a caller's stream (and the state
// it sits in) stays in registers.
struct GammaDraw { double v; uint32_t c0, c1; };
static __device__ __noinline__ GammaDraw rng_gamma_by_value(Rng r, double a) {
  double v = rng_gamma_impl(r, a);
  return GammaDraw{v, r.c0, r.c1};
}
__device__ __forceinline__ double rng_gamma(Rng& r, double a) {
  GammaDraw g = rng_gamma_by_value(r, a);
  r.c0 = g.c0; r.c1 = g.c1;
  return g.v;
}

// scaled inverse chi^2, redrawn until finite and positive (src/rand.hpp:85-97)
__device__ __forceinline__ double rng_sinvchi2(Rng& r, double nu, double s2) {
  double val = -1.0;
  for (int trial = 0; trial < BMG_MAX_TRIALS && (val <= 0 || !isfinite(val)); ++trial)
    val = nu * s2 / (2.0 * rng_gamma(r, 0.5 * nu));
  return val;
}

__device__ inline Rng rng_make(uint32_t seed, uint32_t chain, uint32_t c2, uint32_t c3) {
  Rng r; r.k0 = seed; r.k1 = chain; r.c0 = 0; r.c1 = 0; r.c2 = c2; r.c3 = c3;
  return r;
}

// ------------------------------------------------------------ reductions --
// Wave sums with DPP: row_shr 1, 2, 4, 8 leave each row's total in its
// lane 15, row_bcast 15 and 31 carry them up to lane 63, whose value every
// lane then reads.  Six dependent VALU steps instead of six LDS permutes.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  int2 h = *reinterpret_cast<int2*>(&v);
  h.x = __builtin_amdgcn_update_dpp(0, h.x, CTRL, ROWS, 0xf, false);
  h.y = __builtin_amdgcn_update_dpp(0, h.y, CTRL, ROWS, 0xf, false);
  return *reinterpret_cast<double*>(&h);
}
__device__ __forceinline__ int wave_sum_i32(int v) {
  v += dpp_i32<0x111, 0xf>(v);
  v += dpp_i32<0x112, 0xf>(v);
  v += dpp_i32<0x114, 0xf>(v);
  v += dpp_i32<0x118, 0xf>(v);
  v += dpp_i32<0x142, 0xa>(v);
  v += dpp_i32<0x143, 0xc>(v);
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  int2 h = *reinterpret_cast<int2*>(&v);
  h.x = __builtin_amdgcn_readlane(h.x, 63);
  h.y = __builtin_amdgcn_readlane(h.y, 63);
  return *reinterpret_cast<double*>(&h);
}
__device__ inline long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Workgroup-wide sum; every thread gets the total.  `red` is LDS scratch of
// at least BMG_BLOCK / BMG_WAVE entries.  Contains barriers: call uniformly.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  int w = threadIdx.x / BMG_WAVE;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < (int)(BMG_BLOCK / BMG_WAVE); ++i) t += red[i];
  return t;
}

__device__ inline long long block_sum_i64(long long v, long long* red) {
  v = wave_sum_i64(v);
  int w = threadIdx.x / BMG_WAVE;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  long long t = 0;
#pragma unroll
  for (int i = 0; i < (int)(BMG_BLOCK / BMG_WAVE); ++i) t += red[i];
  return t;
}

// --------------------------------------------------------- genotype bits --
// A 32-bit word holds 16 individuals, individual i at bits 2 (i % 16).
// Codes (src/data.cpp:36): 00 -> 0, 01 -> missing, 10 -> 1, 11 -> 2.
// Per coding the value is p1 + p2 with two bit planes (even bit positions):
//   A: p1 = hi, p2 = hi & lo      H: p1 = hi & ~lo      D: p1 = hi
//   R: p1 = hi & lo               (missing counts as 0, as the reference's
// precompute sees it before missing values are sampled).
#define BMG_EVEN 0x55555555u

__device__ __host__ inline void planes(uint32_t w, int coding, uint32_t& p1, uint32_t& p2) {
  uint32_t lo = w & BMG_EVEN, hi = (w >> 1) & BMG_EVEN;
  switch (coding) {
    case 0: p1 = hi; p2 = hi & lo; break;
    case 1: p1 = hi & ~lo; p2 = 0; break;
    case 2: p1 = hi; p2 = 0; break;
    default: p1 = hi & lo; p2 = 0; break;
  }
}

// exact integer dot of two SNP columns (codings ca, cb) over `nw` words
__device__ inline long long dot_planes(uint32_t a, uint32_t b, int ca, int cb) {
  uint32_t a1, a2, b1, b2;
  planes(a, ca, a1, a2);
  planes(b, cb, b1, b2);
  return (long long)(__popc(a1 & b1) + __popc(a1 & b2) + __popc(a2 & b1) + __popc(a2 & b2));
}

// 16 genotype values of a word as four words of int8 (value 0, 1 or 2)
__device__ inline void word_to_i8x16(uint32_t w, int coding, uint32_t out[4]) {
  uint32_t p1, p2;
  planes(w, coding, p1, p2);
  uint32_t v = p1 + p2;  // each 2-bit field holds the value (<= 2, no carry)
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    uint32_t x = (v >> (8 * b)) & 0xFFu;            // fields f3 f2 f1 f0
    uint32_t t = (x | (x << 12)) & 0x000F000Fu;     // f1f0 | f3f2 << 16
    out[b] = (t | (t << 6)) & 0x03030303u;          // f0, f1, f2, f3 bytes
  }
}

__device__ inline double code_value(uint32_t c, int coding) {
  // reference tables (src/genotype_tables.hpp), -1 for missing
  const double A[4] = {0.0, -1.0, 1.0, 2.0}, H[4] = {0.0, -1.0, 1.0, 0.0},
               D[4] = {0.0, -1.0, 1.0, 1.0}, R[4] = {0.0, -1.0, 0.0, 1.0};
  switch (coding) {
    case 0: return A[c];
    case 1: return H[c];
    case 2: return D[c];
    default: return R[c];
  }
}

// genotype value with missing -> 0 (the value the models see)
__device__ inline int code_ival(uint32_t c, int coding) {
  uint32_t hi = c >> 1, lo = c & 1u;
  switch (coding) {
    case 0: return (int)(hi + (hi & lo));
    case 1: return (int)(hi & (lo ^ 1u));
    case 2: return (int)hi;
    default: return (int)(hi & lo);
  }
}

}  // namespace bmg
