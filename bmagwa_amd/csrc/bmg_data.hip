// bmg_data.hip -- dataset kernels: layout, decode, recode, integer per-SNP
// moments, synthetic genotypes, the X^T r scan and band X^T X (int8 MFMA).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "bmg_device.h"
#include "bmg_internal.h"

using namespace bmg;

typedef int v4i __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- layout ----
// The reference stores each SNP in ceil(n/4) bytes (src/data.cpp:245-273).
// In HBM every SNP starts on a 16-byte boundary (W32 words, multiple of 4),
// trailing fields are code 00 (value 0), so word loads never straddle SNPs.
__global__ void k_restride(const uint8_t* __restrict__ raw, uint64_t p, uint32_t stride_raw,
                           uint32_t W32, uint8_t* __restrict__ out) {
  uint64_t j = blockIdx.y + (uint64_t)blockIdx.z * gridDim.y;
  if (j >= p) return;
  uint32_t bytes = W32 * 4;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < bytes; b += gridDim.x * blockDim.x)
    out[j * bytes + b] = b < stride_raw ? raw[j * stride_raw + b] : 0;
}

// src/data.cpp:324-339: flip 0 <-> 2 where the allele frequency exceeds 0.5.
// Per word: fields with hi == lo (00 or 11) are inverted, others kept; the
// fields past n are left at 00.
__global__ void k_recode(uint32_t* g, uint64_t p, uint32_t W32, uint32_t n) {
  __shared__ long long red[8];
  uint64_t j = blockIdx.x;
  if (j >= p) return;
  uint32_t* w = g + j * W32;
  long long s = 0, ng = 0;
  for (uint32_t k = threadIdx.x; k < W32; k += blockDim.x) {
    uint32_t x = w[k];
    uint32_t lo = x & BMG_EVEN, hi = (x >> 1) & BMG_EVEN;
    s += __popc(hi) + __popc(hi & lo);
    uint32_t valid = k * 16 + 16 <= n ? 0xFFFFFFFFu
                     : (k * 16 >= n ? 0u : ((1u << (2 * (n - k * 16))) - 1u));
    ng += __popc(~(lo & ~hi) & BMG_EVEN & valid);   // not missing
  }
  s = block_sum_i64(s, red);
  ng = block_sum_i64(ng, red);
  if ((double)s / (2.0 * (double)ng) > 0.5) {
    for (uint32_t k = threadIdx.x; k < W32; k += blockDim.x) {
      uint32_t x = w[k];
      uint32_t lo = x & BMG_EVEN, hi = (x >> 1) & BMG_EVEN;
      uint32_t valid = k * 16 + 16 <= n ? 0xFFFFFFFFu
                       : (k * 16 >= n ? 0u : ((1u << (2 * (n - k * 16))) - 1u));
      uint32_t eq = ~(hi ^ lo) & BMG_EVEN & valid;
      w[k] = x ^ (eq | (eq << 1));
    }
  }
}

// Per SNP integer counts: c1 = #(code 10), c2 = #(code 11), nm = #not missing.
// All moments of every coding are integer functions of these.
__global__ void k_counts(const uint32_t* __restrict__ g, uint64_t p, uint32_t W32, uint32_t n,
                         long long* __restrict__ out) {
  __shared__ long long red[8];
  uint64_t j = blockIdx.x;
  if (j >= p) return;
  const uint32_t* w = g + j * W32;
  long long c1 = 0, c2 = 0, miss = 0;
  for (uint32_t k = threadIdx.x; k < W32; k += blockDim.x) {
    uint32_t x = w[k];
    uint32_t lo = x & BMG_EVEN, hi = (x >> 1) & BMG_EVEN;
    c1 += __popc(hi & ~lo);
    c2 += __popc(hi & lo);
    miss += __popc(lo & ~hi);
  }
  c1 = block_sum_i64(c1, red);
  c2 = block_sum_i64(c2, red);
  miss = block_sum_i64(miss, red);
  if (threadIdx.x == 0) {
    out[3 * j] = c1; out[3 * j + 1] = c2; out[3 * j + 2] = (long long)n - miss;
  }
}

// src/precomputed_snp_covariances.hpp:95-131 from the integer counts; the
// sums are exact, so the doubles equal the reference's loop bit for bit.
__global__ void k_self_moments(const long long* __restrict__ cnt, uint64_t p, uint32_t n,
                               int at0, int at1, int at2, int at3, int ah, int offset,
                               double* __restrict__ pre) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (j >= p) return;
  long long c1 = cnt[3 * j], c2 = cnt[3 * j + 1];
  const int allow[4] = {at0, at1, at2, at3};
  // sum and sum of squares of A, H, D, R
  const long long sum[4] = {c1 + 2 * c2, c1, c1 + c2, c2};
  const long long sq[4] = {c1 + 4 * c2, c1, c1 + c2, c2};
  double* o = pre + j * offset;
  int k = 0;
  for (int t = 0; t < 4; ++t) {
    if (!allow[t]) continue;
    double s = (double)sum[t];
    o[k++] = s;
    o[k++] = (double)sq[t] - s * s / (double)n;
  }
  if (ah) {
    double as = (double)sum[0], hs = (double)sum[1];
    o[k++] = (double)c1 - as * hs / (double)n;  // sum a*h = #(code 10)
  }
}

// DataModel::get_genotypes[coding] for a list of SNPs (values incl. -1)
__global__ void k_decode(const uint32_t* __restrict__ g, uint32_t W32, uint32_t n,
                         const uint64_t* __restrict__ snps, uint64_t count, int coding,
                         double* __restrict__ out) {
  uint64_t s = blockIdx.y + (uint64_t)blockIdx.z * gridDim.y;
  if (s >= count) return;
  const uint32_t* w = g + snps[s] * W32;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t c = (w[i >> 4] >> (2 * (i & 15))) & 3u;
    out[s * n + i] = code_value(c, coding);
  }
}

// ---------------------------------------------------------- synthetic ----
// Latent AR(1) haplotypes (see the header for the model).  One thread owns
// four individuals (one byte) within one LD block and walks the block's SNPs.
__device__ inline double synth_normal(uint32_t seed, uint64_t j, uint32_t hap) {
  uint32_t x0 = (uint32_t)j, x1 = (uint32_t)(j >> 32), x2 = hap, x3 = 0x40000000u;
  philox10(x0, x1, x2, x3, seed, 0x5EED5EEDu);
  double ua = words_u01(x0, x1), ub = words_u01(x2, x3);
  return sqrt(-2.0 * log(1.0 - ua)) * cos(6.283185307179586 * ub);
}

__global__ void k_synth(uint32_t seed, uint32_t n, uint64_t p, uint32_t W32,
                        const double* __restrict__ thresh, const double* __restrict__ rho,
                        const uint64_t* __restrict__ blk_start, uint64_t n_blocks,
                        uint8_t* __restrict__ out) {
  uint32_t nq = (n + 3) / 4;
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t b = blockIdx.y + (uint64_t)blockIdx.z * gridDim.y;
  if (q >= nq || b >= n_blocks) return;
  uint64_t j0 = blk_start[b], j1 = (b + 1 < n_blocks) ? blk_start[b + 1] : p;
  double z[8];
  uint32_t bytes = W32 * 4;
  for (uint64_t j = j0; j < j1; ++j) {
    double r = rho[j], t = thresh[j];
    double sq = (r == 0.0) ? 0.0 : sqrt(1.0 - r * r);
    uint8_t byte = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t i = 4 * q + k;
      if (i >= n) break;
      double e0 = synth_normal(seed, j, 2 * i), e1 = synth_normal(seed, j, 2 * i + 1);
      if (r == 0.0) { z[2 * k] = e0; z[2 * k + 1] = e1; }
      else {
        z[2 * k] = r * z[2 * k] + sq * e0;
        z[2 * k + 1] = r * z[2 * k + 1] + sq * e1;
      }
      int gv = (z[2 * k] < t) + (z[2 * k + 1] < t);
      uint8_t code = gv == 0 ? 0 : (gv == 1 ? 2 : 3);
      byte |= (uint8_t)(code << (2 * k));
    }
    out[j * bytes + q] = byte;
  }
}

// --------------------------------------------------------- X^T r scan ----
// The residual r (FP64) is rounded to fixed point R = round(r / 2^e) with
// |R| < 2^(8D - 2) and split into D balanced base-256 digits in [-128, 127].
// x^T R is then an exact integer product done with int8 MFMA
// (v_mfma_i32_16x16x64_i8): A = 16 SNPs x 64 individuals (2-bit codes
// expanded to int8 in registers), B = 64 individuals x 16 digit columns
// (vector v, digit d at column v*D + d).  x^T r = 2^e sum_d 256^d C_d.
__global__ void k_absmax(const double* __restrict__ r, uint32_t n_pad, int nvec,
                         double* __restrict__ scale, int digits) {
  __shared__ double red[8];
  int v = blockIdx.x;
  double m = 0.0;
  for (uint32_t i = threadIdx.x; i < n_pad; i += blockDim.x) m = fmax(m, fabs(r[v * (size_t)n_pad + i]));
  m = wave_max(m);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < (int)(blockDim.x / 64); ++i) t = fmax(t, red[i]);
    int e = 0;
    if (t > 0) { frexp(t, &e); }          // t < 2^e
    scale[v] = ldexp(1.0, e - (8 * digits - 2));
  }
}

__global__ void k_digits(const double* __restrict__ r, uint32_t n_pad, int nvec, int digits,
                         const double* __restrict__ scale, int8_t* __restrict__ dig) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  int v = blockIdx.y;
  if (i >= n_pad) return;
  long long R = llrint(r[v * (size_t)n_pad + i] / scale[v]);
  // within each group of 16 individuals, individual 4j + s sits at byte
  // 4s + j: the order in which the scan kernel splits a genotype word
  const uint32_t l = i & 15u, pos = (i & ~15u) | (4u * (l & 3u) + (l >> 2));
  for (int d = 0; d < digits; ++d) {
    int dg = (int)(((R + 128) & 0xFF)) - 128;
    dig[(size_t)(v * digits + d) * n_pad + pos] = (int8_t)dg;
    R = (R - dg) >> 8;
  }
}

// Genotype values of a word (16 individuals) as 16 int8 in the split order:
// out[s] byte j = value of individual 4j + s (field s of byte j), so each
// output word is one shift and one mask of the value word.
template <int CODING>
__device__ __forceinline__ void word_split_i8x16(uint32_t w, uint32_t out[4]) {
  const uint32_t lo = w & BMG_EVEN, hi = (w >> 1) & BMG_EVEN;
  uint32_t v;
  if (CODING == 0) v = hi + (hi & lo);     // 0, 1, 2 per field (no carry)
  else if (CODING == 1) v = hi & ~lo;
  else if (CODING == 2) v = hi;
  else v = hi & lo;
#pragma unroll
  for (int b = 0; b < 4; ++b) out[b] = (v >> (2 * b)) & 0x03030303u;
}

// grid.x: workgroups, each looping over blocks of WAVES waves x NT tiles x
// 16 SNPs (so the digit tile it stages in LDS serves many blocks); grid.y:
// individual tiles.  A step covers 32 words (512 individuals) of 16 SNPs:
// each SNP row's 128-byte line is read whole within the step, the first
// half by one load of the lane groups and the second by the next.  Each LDS
// column holds the tile's digits rounded up to whole steps with zeros, plus
// 16 bytes of padding so the 16 columns a lane group reads fall in distinct
// banks.
//
// A wave walks its steps (block, slice) in order with two register buffers
// of genotype words: while one step's MFMAs run, the loads of the step
// after next are in flight in the other buffer.  Every load address is
// clamped into the SNP's row and to the last SNP; what a clamped load
// brings in meets a zero digit or lands in a row that is never stored.
template <int NT, int CODING, int WAVES, bool ATOMIC>
__global__ __launch_bounds__(64 * WAVES) void k_scan_mfma(const uint32_t* __restrict__ g, uint64_t p,
                                                          uint32_t W32, const int8_t* __restrict__ dig,
                                                          int NC, uint32_t n_pad, uint32_t tile_words,
                                                          int32_t* __restrict__ acc) {
  extern __shared__ int8_t sdig[];  // NC x (tile_words * 16 + 16), tile_words a multiple of 32
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t w0 = blockIdx.y * tile_words;
  const uint32_t w1 = min(W32, w0 + tile_words);
  const uint32_t nind = (w1 - w0) * 16;
  const uint32_t steps = (w1 - w0 + 31) / 32;
  const uint32_t span = steps * 512;  // digit bytes the steps cover
  const uint32_t ld = span + 16;
  // stage the digit columns of this individual tile, zero past its end
  for (uint32_t e = threadIdx.x * 16; e < (uint32_t)NC * span; e += 64 * WAVES * 16) {
    uint32_t col = e / span, off = e % span;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (off < nind) v = *(const uint4*)(dig + (size_t)col * n_pad + w0 * 16 + off);
    *(uint4*)(sdig + (size_t)col * ld + off) = v;
  }
  __syncthreads();
  const int row = lane & 15, grp = lane >> 4;
  const int col = row;
  const int8_t* bcol = sdig + (size_t)min(col, NC - 1) * ld + grp * 64;
  const uint64_t per_blk = (uint64_t)WAVES * NT * 16;
  const uint64_t n_blk = (p + per_blk - 1) / per_blk;
  if ((uint64_t)blockIdx.x >= n_blk) return;
  const uint64_t total = (n_blk - blockIdx.x + gridDim.x - 1) / gridDim.x * steps;

  // load cursor (block lb, step lw) and compute cursor (cb, cw)
  uint64_t lb = blockIdx.x, cb = blockIdx.x;
  uint32_t lw = 0, cw = 0;
  auto load = [&](uint4 (&buf)[NT][2]) {
    const uint32_t wa = min(w0 + lw * 32 + 4 * grp, W32 - 4);
    const uint32_t wb = min(w0 + lw * 32 + 16 + 4 * grp, W32 - 4);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      uint64_t snp = min((lb * WAVES + wave) * (NT * 16) + t * 16 + row, p - 1);
      const uint32_t* r = g + snp * W32;
      buf[t][0] = *(const uint4*)(r + wa);
      buf[t][1] = *(const uint4*)(r + wb);
    }
    if (++lw == steps) { lw = 0; lb += gridDim.x; }
  };
  v4i c[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) c[t] = v4i{0, 0, 0, 0};
  auto compute = [&](const uint4 (&a4)[NT][2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v4i b[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        // columns past NC read column NC - 1; their results are never stored
        uint4 bb = *(const uint4*)(bcol + (cw * 32 + h * 16 + s) * 16);
        b[s] = v4i{(int)bb.x, (int)bb.y, (int)bb.z, (int)bb.w};
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const uint4& q = a4[t][h];
          uint32_t w = s == 0 ? q.x : (s == 1 ? q.y : (s == 2 ? q.z : q.w));
          uint32_t o[4];
          word_split_i8x16<CODING>(w, o);
          v4i a = {(int)o[0], (int)o[1], (int)o[2], (int)o[3]};
          c[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b[s], c[t], 0, 0, 0);
        }
      }
    }
    if (++cw == steps) {
      // lane holds C[4 * grp + r][col] of this block; rows past p land in
      // the accumulator's padding, columns past NC in slots the combine
      // never reads: the stores need no guard
      const uint64_t base = (cb * WAVES + wave) * (NT * 16);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int32_t* dst = acc + (base + t * 16 + 4 * grp + r) * 16 + col;
          if (ATOMIC) atomicAdd(dst, c[t][r]);
          else *dst = c[t][r];
        }
        c[t] = v4i{0, 0, 0, 0};
      }
      cw = 0;
      cb += gridDim.x;
    }
  };
  uint4 bufA[NT][2], bufB[NT][2];
  load(bufA);
  load(bufB);
  // every trip of the loop issues both buffers' loads, so the wait for a
  // buffer never has to cover the other's (the scheduling barriers keep
  // each buffer's new loads after the last read of its old words, so the
  // buffers keep their registers round the loop instead of being copied)
  uint64_t k = 0;
  for (; k + 1 < total; k += 2) {
    compute(bufA);
    __builtin_amdgcn_sched_barrier(0);
    load(bufA);
    compute(bufB);
    __builtin_amdgcn_sched_barrier(0);
    load(bufB);
  }
  if (k < total) compute(bufA);
}

__global__ void k_fill_i32(int32_t* __restrict__ p, uint64_t n, int32_t v) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_scan_combine(const int32_t* __restrict__ acc, uint64_t p, int nvec, int digits,
                               const double* __restrict__ scale, double* __restrict__ out) {
  uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  int v = blockIdx.y;
  if (j >= p) return;
  long long s = 0;
  for (int d = digits - 1; d >= 0; --d) s = s * 256 + (long long)acc[j * 16 + v * digits + d];
  out[(size_t)v * p + j] = ldexp((double)s, 0) * scale[v];
}

// ------------------------------------------------------ band X^T X ------
// out[j * W + d] = sum_i x_j(i) x_{j+d}(i).  A workgroup of 8 waves owns
// 128 row SNPs [r0, r0 + 128) (eight 16-row tiles) and the column SNPs
// [r0, r0 + 128 + W).  Wave w takes the four row tiles 4 (w & 1) ..
// 4 (w & 1) + 3 and a quarter of the column offsets, d in [g NDQ, (g + 1)
// NDQ) with g = w >> 1 (a 16-row tile needs offsets 0 .. ND - 1, ND =
// (W + 31) / 16 <= 4 NDQ).  Column tile c is offset c - rt of row tile rt,
// so each column tile the wave reads serves up to four MFMAs.
//
// LDS holds the packed 2-bit words (a chunk of 16 words, 256 individuals,
// per SNP; slots of 20 words so that the 64 lanes of a read hit distinct
// banks), not expanded bytes: each lane reads the one word its MFMA operand
// needs and expands it to sixteen int8 values in registers (the split byte
// order of word_split_i8x16, the same for both operands, so they agree on
// which individual each k is), which cuts the LDS traffic per MFMA by four
// and lets the row and column codings differ without a second copy.  Each
// 16x16x64 int8 MFMA takes one word per lane of each operand.
//
// Chunks are double-buffered in LDS: chunk c + 1's loads are issued after
// the barrier that ends chunk c - 1 and written to LDS after chunk c's
// MFMAs, so a chunk of MFMAs covers their latency; one barrier per chunk.
#define BAND_ROWS 128
#define BAND_WAVES 8
#define BAND_MAXP 36   // column tiles per row tile: (W + 31) / 16 <= 36, W <= 560
#define BAND_CW 16     // packed words per SNP per chunk
#define BAND_SW 20     // LDS words per SNP slot
template <int CA, int CB, int NDQ>
__global__ __launch_bounds__(64 * BAND_WAVES) void k_band(const uint32_t* __restrict__ g, uint64_t p,
                                                          uint32_t W32, uint32_t W,
                                                          int32_t* __restrict__ out, uint64_t n_row_blocks) {
  extern __shared__ uint32_t sx[];  // 2 buffers x nslot slots x BAND_SW words
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // XCD-aware ordering: workgroups are dealt round-robin over the 8 XCDs, so
  // row block rb goes to XCD rb / per; neighbouring row blocks (whose column
  // windows overlap) then share that XCD's L2.
  const uint64_t nb = n_row_blocks;
  const uint64_t per = (nb + 7) / 8;
  const uint64_t bid = blockIdx.x;
  const uint64_t rb = (bid % 8) * per + bid / 8;
  if (rb >= nb) return;
  const uint64_t r0 = rb * BAND_ROWS;
  const uint32_t ncol = ((BAND_ROWS + W + 15) / 16) * 16;
  const uint32_t nslot = ncol + (CA == CB ? 0 : BAND_ROWS);   // per buffer
  const int ND = (int)((W + 31) / 16);
  const uint32_t nch = W32 / BAND_CW;
  // HBM -> registers -> LDS: uint4 e of a chunk is word 4 (e % 4) of slot e / 4
  constexpr int UPS = BAND_CW / 4;
  constexpr int LPT = ((BAND_ROWS + 560 + 15) / 16 * 16 * UPS + 64 * BAND_WAVES - 1) / (64 * BAND_WAVES);
  uint4 raw[LPT];
  auto load = [&](uint32_t ch) {
    const uint32_t wc = (ch < nch ? ch : nch - 1) * BAND_CW;
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const uint32_t e = threadIdx.x + i * 64 * BAND_WAVES;
      const uint32_t sn = min(e / UPS, ncol - 1);
      const uint64_t snp = min(r0 + sn, p - 1);
      raw[i] = *(const uint4*)(g + snp * W32 + wc + (e % UPS) * 4);
    }
  };
  // each word is stored as its value word: a 2-bit field per individual
  // holding the coded genotype (0..2), so that an operand's expansion is a
  // shift and a mask per output register; the row codings' copy (when they
  // differ from the column coding) follows the ncol column slots
  auto value_word = [](uint32_t w, int cod) -> uint32_t {
    const uint32_t lo = w & BMG_EVEN, hi = (w >> 1) & BMG_EVEN;
    return cod == 0 ? hi + (hi & lo) : (cod == 1 ? hi & ~lo : (cod == 2 ? hi : hi & lo));
  };
  auto store = [&](uint32_t* buf) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const uint32_t e = threadIdx.x + i * 64 * BAND_WAVES;
      const uint32_t sn = e / UPS;
      if (sn < ncol) {
        const uint4 w = r0 + sn < p ? raw[i] : make_uint4(0, 0, 0, 0);
        *(uint4*)(buf + sn * BAND_SW + (e % UPS) * 4) =
            make_uint4(value_word(w.x, CB), value_word(w.y, CB), value_word(w.z, CB), value_word(w.w, CB));
        if (CA != CB && sn < BAND_ROWS)
          *(uint4*)(buf + (ncol + sn) * BAND_SW + (e % UPS) * 4) =
              make_uint4(value_word(w.x, CA), value_word(w.y, CA), value_word(w.z, CA), value_word(w.w, CA));
      }
    }
  };
  auto expand = [](uint32_t v) -> v4i {
    return v4i{(int)(v & 0x03030303u), (int)((v >> 2) & 0x03030303u), (int)((v >> 4) & 0x03030303u),
               (int)((v >> 6) & 0x03030303u)};
  };
  v4i c[4][NDQ];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int e = 0; e < NDQ; ++e) c[i][e] = v4i{0, 0, 0, 0};
  const int row = lane & 15, grp = lane >> 4;
  const int rt0 = 4 * (wave & 1);          // row tiles rt0 .. rt0 + 3
  const int dlo = (wave >> 1) * NDQ;       // offsets [dlo, dlo + NDQ)
  const uint32_t last_tile = ncol / 16 - 1;
  constexpr int NB = NDQ + 3;              // column tiles rt0 + dlo + t, t < NB
  load(0);
  store(sx);
  __syncthreads();
  load(1);
  for (uint32_t ch = 0; ch < nch; ++ch) {
    const uint32_t* buf = sx + (size_t)(ch & 1) * nslot * BAND_SW;
#pragma unroll 1
    for (int st = 0; st < BAND_CW / 4; ++st) {
      const int wo = st * 4 + grp;         // this lane's word of the 64 individuals
      uint32_t aw[4], bw[NB];
      const uint32_t* abuf = buf + (CA == CB ? 0 : (size_t)ncol * BAND_SW);
#pragma unroll
      for (int i = 0; i < 4; ++i) aw[i] = abuf[((rt0 + i) * 16 + row) * BAND_SW + wo];
#pragma unroll
      for (int t = 0; t < NB; ++t)
        bw[t] = buf[(min((uint32_t)(rt0 + dlo + t), last_tile) * 16 + row) * BAND_SW + wo];
      v4i a[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = expand(aw[i]);
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const v4i b = expand(bw[t]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int e = t - i;
          if (e >= 0 && e < NDQ) c[i][e] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b, c[i][e], 0, 0, 0);
        }
      }
      // issue order: the LDS reads, the row operands' expansion, then per
      // column tile its expansion (7 VALU) and its MFMAs, so that only one
      // expanded column operand is live at a time
      __builtin_amdgcn_sched_group_barrier(0x0100, 4 + NB, 1);
      __builtin_amdgcn_sched_group_barrier(0x0002, 28, 1);
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int lo = t - 3 > 0 ? t - 3 : 0, hi = t < NDQ - 1 ? t : NDQ - 1;
        __builtin_amdgcn_sched_group_barrier(0x0002, 7, 1);
        switch (hi - lo + 1) {
          case 1: __builtin_amdgcn_sched_group_barrier(0x0008, 1, 1); break;
          case 2: __builtin_amdgcn_sched_group_barrier(0x0008, 2, 1); break;
          case 3: __builtin_amdgcn_sched_group_barrier(0x0008, 3, 1); break;
          default: __builtin_amdgcn_sched_group_barrier(0x0008, 4, 1); break;
        }
      }
    }
    if (ch + 1 < nch) store(sx + (size_t)((ch + 1) & 1) * nslot * BAND_SW);
    __syncthreads();
    if (ch + 2 < nch) load(ch + 2);
  }
  const int col = lane & 15;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rt = rt0 + i;
#pragma unroll
    for (int e = 0; e < NDQ; ++e) {
      const int d = dlo + e;
      if (d < ND) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint64_t j = r0 + rt * 16 + 4 * grp + r;    // row SNP
          const uint64_t jj = r0 + (rt + d) * 16 + col;     // column SNP
          if (j < p && jj >= j && jj - j < W && jj < p) out[j * W + (jj - j)] = c[i][e][r];
        }
      }
    }
  }
}

// -------------------------------------------------------------- host ----
namespace bmg {

int ensure(void** ptr, size_t* have, size_t need) {
  if (*have >= need) return BMG_OK;
  if (*ptr) BMG_HIP(hipFree(*ptr));
  *ptr = nullptr; *have = 0;
  BMG_HIP(hipMalloc(ptr, need));
  *have = need;
  return BMG_OK;
}

// out SNP j, word w, field f <- pool SNP j, individual sub[16 w + f] (code
// 00 past n_sub); one thread per output word, the pool row's words read
// through L2 (a gather of 2-bit fields within one SNP's row)
__global__ void k_subset(const uint32_t* __restrict__ pool, uint32_t W32p, const uint32_t* __restrict__ sub,
                         uint32_t n_sub, uint64_t p, uint32_t W32, uint32_t* __restrict__ out) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W32) return;
  for (uint64_t j = blockIdx.y; j < p; j += gridDim.y) {
    const uint32_t* row = pool + j * W32p;
    uint32_t v = 0;
#pragma unroll
    for (int f = 0; f < 16; ++f) {
      const uint32_t i = w * 16 + f;
      if (i < n_sub) {
        const uint32_t k = sub[i];
        v |= ((row[k >> 4] >> (2 * (k & 15))) & 3u) << (2 * f);
      }
    }
    out[j * W32 + w] = v;
  }
}

int launch_subset(const bmg_ctx* pool, const uint32_t* sub_dev, bmg_ctx* c, hipStream_t s) {
  const unsigned gy = (unsigned)std::min<uint64_t>(c->p, 65535);
  hipLaunchKernelGGL(k_subset, dim3((c->W32 + 255) / 256, gy), dim3(256), 0, s, pool->g, pool->W32, sub_dev,
                     c->n, c->p, c->W32, c->g);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int ctx_prepare(bmg_ctx* c, int recode) {
  hipStream_t s = c->stream;
  if (recode) {
    hipLaunchKernelGGL(k_recode, dim3((unsigned)c->p), dim3(256), 0, s, c->g, c->p, c->W32, c->n);
    BMG_HIP(hipGetLastError());
  }
  return ctx_self_moments(c);
}

int ctx_self_moments(bmg_ctx* c) {
  hipStream_t s = c->stream;
  long long* cnt = nullptr;
  BMG_HIP(hipMalloc(&cnt, sizeof(long long) * 3 * c->p));
  hipLaunchKernelGGL(k_counts, dim3((unsigned)c->p), dim3(256), 0, s, c->g, c->p, c->W32, c->n, cnt);
  BMG_HIP(hipGetLastError());
  if (c->pre) BMG_HIP(hipFree(c->pre));
  BMG_HIP(hipMalloc(&c->pre, sizeof(double) * c->offset * c->p));
  hipLaunchKernelGGL(k_self_moments, dim3((unsigned)((c->p + 255) / 256)), dim3(256), 0, s, cnt,
                     c->p, c->n, c->allow_terms[0], c->allow_terms[1], c->allow_terms[2],
                     c->allow_terms[3], c->allow_types[BMG_AH], c->offset, c->pre);
  BMG_HIP(hipGetLastError());
  // dataset statistics in SNP order (src/data.cpp:403-434), exact integer sums
  std::vector<long long> h(3 * c->p);
  BMG_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(long long) * 3 * c->p, hipMemcpyDeviceToHost, s));
  BMG_HIP(hipStreamSynchronize(s));
  BMG_HIP(hipFree(cnt));
  double tv = 0, tm = 0; uint64_t nsm = 0, nsv = 0, miss = 0;
  for (uint64_t j = 0; j < c->p; ++j) {
    long long c1 = h[3 * j], c2 = h[3 * j + 1], ng = h[3 * j + 2];
    miss += (uint64_t)(c->n - ng);
    double sq = (double)(c1 + 4 * c2), sm = (double)(c1 + 2 * c2);
    if (ng > 1) {
      tm += sm / ng; tv += (sq - sm * sm / ng) / (ng - 1); ++nsm; ++nsv;
    } else if (ng == 1) {
      tm += sm / ng; ++nsm;
    }
  }
  c->var_x = tv / nsv;
  c->mean_x = tm / nsm;
  c->n_missing = miss;
  return BMG_OK;
}

int launch_decode(bmg_ctx* c, const uint64_t* snps_dev, uint64_t count, int coding,
                  double* out_dev, hipStream_t s) {
  unsigned gy = (unsigned)std::min<uint64_t>(count, 65535);
  unsigned gz = (unsigned)((count + gy - 1) / gy);
  unsigned gx = (unsigned)std::min<uint64_t>((c->n + 255) / 256, 64);
  hipLaunchKernelGGL(k_decode, dim3(gx, gy, gz), dim3(256), 0, s, c->g, c->W32, c->n, snps_dev,
                     count, coding, out_dev);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int launch_synth(bmg_ctx* c, uint32_t seed, const double* thresh_dev, const double* rho_dev,
                 const uint64_t* blk_start_dev, uint64_t n_blocks, hipStream_t s) {
  uint32_t nq = (c->n + 3) / 4;
  unsigned gx = (nq + 63) / 64;
  unsigned gy = (unsigned)std::min<uint64_t>(n_blocks, 65535);
  unsigned gz = (unsigned)((n_blocks + gy - 1) / gy);
  BMG_HIP(hipMemsetAsync(c->g, 0, (size_t)c->p * c->W32 * 4, s));
  hipLaunchKernelGGL(k_synth, dim3(gx, gy, gz), dim3(64), 0, s, seed, c->n, c->p, c->W32,
                     thresh_dev, rho_dev, blk_start_dev, n_blocks, (uint8_t*)c->g);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

// the scan's scratch: digit planes (NC x n_pad int8), the int32
// accumulator ((p + 1024) x 16) and the per-vector scales (16 doubles)
struct ScanBufs {
  int8_t* dig;
  int32_t* acc;
  double* scale;
};

static int scan_check(int nvec, int digits) {
  if (nvec < 1 || digits < 1 || digits > 6 || nvec * digits > 16) {
    set_error("scan: need 1 <= nvec*digits <= 16, digits <= 6");
    return BMG_E_ARG;
  }
  return BMG_OK;
}

static int scan_digits(bmg_ctx* c, const double* r_dev, int nvec, int digits, const ScanBufs& B,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_absmax, dim3(nvec), dim3(256), 0, s, r_dev, c->n_pad, nvec, B.scale, digits);
  hipLaunchKernelGGL(k_digits, dim3((c->n_pad + 255) / 256, nvec), dim3(256), 0, s, r_dev,
                     c->n_pad, nvec, digits, B.scale, B.dig);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

static int scan_prepare(bmg_ctx* c, const double* r_dev, int nvec, int digits, hipStream_t s, ScanBufs* B) {
  int rc = scan_check(nvec, digits);
  if (rc) return rc;
  int NC = nvec * digits;
  if (ensure((void**)&c->dig, &c->dig_bytes, (size_t)NC * c->n_pad)) return BMG_E_HIP;
  // p rows plus one block of the largest scan shape (16 waves x 4 tiles x 16
  // SNPs) for the unguarded stores of the last block
  if (ensure((void**)&c->acc, &c->acc_bytes, (size_t)(c->p + 1024) * 16 * sizeof(int32_t))) return BMG_E_HIP;
  if (!c->scale) BMG_HIP(hipMalloc(&c->scale, sizeof(double) * 16));
  *B = ScanBufs{c->dig, c->acc, c->scale};
  return scan_digits(c, r_dev, nvec, digits, *B, s);
}

struct ScanShape {
  int nt, waves;
  uint32_t tile_words, n_tiles;
  unsigned gx;
  int atomic;
  size_t lds;
};

// a shape: NT SNP tiles per wave, individual tiles of tile_words words,
// gx workgroups looping over the SNP blocks
static ScanShape make_shape(const bmg_ctx* c, int NC, int nt, int waves, uint32_t tile_words, unsigned gx) {
  ScanShape sh;
  sh.nt = nt;
  sh.waves = waves;
  sh.tile_words = std::min<uint32_t>(tile_words, (c->W32 + 31) / 32 * 32);
  sh.n_tiles = (c->W32 + sh.tile_words - 1) / sh.tile_words;
  const uint64_t per_blk = (uint64_t)waves * nt * 16;
  const uint64_t n_blk = (c->p + per_blk - 1) / per_blk;
  sh.gx = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(n_blk, gx));
  sh.atomic = sh.n_tiles > 1;
  sh.lds = (size_t)NC * (sh.tile_words * 16 + 16);
  return sh;
}

// widest digit tile (a multiple of 32 words) whose padded columns fit `bytes`
static uint32_t tile_for(int NC, size_t bytes) {
  return (uint32_t)std::max<size_t>(32, (bytes / NC - 16) / 16 / 32 * 32);
}

// zero the accumulator when several individual tiles add into it (a fill
// kernel, so the sequence stays valid inside captured graphs)
static void scan_clear(bmg_ctx* c, const ScanShape& sh, const ScanBufs& B, hipStream_t s) {
  if (sh.atomic)
    hipLaunchKernelGGL(k_fill_i32, dim3(2048), dim3(256), 0, s, B.acc, (uint64_t)c->p * 16, 0);
}

template <int NT, int CODING, bool ATOMIC>
static void scan_mfma_w(bmg_ctx* c, int NC, const ScanShape& sh, const ScanBufs& B, hipStream_t s) {
  dim3 grid(sh.gx, sh.n_tiles);
  switch (sh.waves) {
    case 16:
      hipLaunchKernelGGL((k_scan_mfma<NT, CODING, 16, ATOMIC>), grid, dim3(1024), sh.lds, s, c->g, c->p, c->W32,
                         B.dig, NC, c->n_pad, sh.tile_words, B.acc);
      break;
    case 8:
      hipLaunchKernelGGL((k_scan_mfma<NT, CODING, 8, ATOMIC>), grid, dim3(512), sh.lds, s, c->g, c->p, c->W32,
                         B.dig, NC, c->n_pad, sh.tile_words, B.acc);
      break;
    default:
      hipLaunchKernelGGL((k_scan_mfma<NT, CODING, 4, ATOMIC>), grid, dim3(256), sh.lds, s, c->g, c->p, c->W32,
                         B.dig, NC, c->n_pad, sh.tile_words, B.acc);
  }
}

template <int CODING>
static void scan_mfma_c(bmg_ctx* c, int NC, const ScanShape& sh, const ScanBufs& B, hipStream_t s) {
  if (sh.atomic) {
    if (sh.nt == 4) scan_mfma_w<4, CODING, true>(c, NC, sh, B, s);
    else scan_mfma_w<2, CODING, true>(c, NC, sh, B, s);
  } else {
    if (sh.nt == 4) scan_mfma_w<4, CODING, false>(c, NC, sh, B, s);
    else scan_mfma_w<2, CODING, false>(c, NC, sh, B, s);
  }
}

static void scan_mfma(bmg_ctx* c, int NC, int coding, const ScanShape& sh, const ScanBufs& B, hipStream_t s) {
  switch (coding) {
    case 0: scan_mfma_c<0>(c, NC, sh, B, s); break;
    case 1: scan_mfma_c<1>(c, NC, sh, B, s); break;
    case 2: scan_mfma_c<2>(c, NC, sh, B, s); break;
    default: scan_mfma_c<3>(c, NC, sh, B, s); break;
  }
}

static int scan_kernel(bmg_ctx* c, int nvec, int coding, int digits, const ScanShape& sh, const ScanBufs& B,
                       hipStream_t s) {
  scan_clear(c, sh, B, s);
  scan_mfma(c, nvec * digits, coding, sh, B, s);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

// Autotuner: the first scan of a context with NC digit columns times every
// candidate shape (SNP tiles per wave, LDS tile width, grid size) on the
// prepared digits -- HIP events on the scan's stream, best of two launches,
// the accumulator clear included -- and keeps the fastest for the context's
// lifetime.  A candidate the runtime refuses to launch is skipped.
static int scan_tuned_shape(bmg_ctx* c, int nvec, int coding, int digits, const ScanBufs& B, hipStream_t s,
                            ScanShape* out) {
  const int NC = nvec * digits;
  if (c->scan_tw[NC]) {
    *out = make_shape(c, NC, c->scan_nt[NC], c->scan_wv[NC], c->scan_tw[NC], c->scan_gx[NC]);
    return BMG_OK;
  }
  if (c->no_tune) {  // a stream capture cannot time: a fixed shape, not remembered
    *out = make_shape(c, NC, 2, 8, tile_for(NC, 163840), 256);
    return BMG_OK;
  }
  // 256 CUs: grids of 1 to 8 workgroups per CU, over the individual tiles
  std::vector<ScanShape> cand;
  const uint32_t tws[2] = {tile_for(NC, 65536), tile_for(NC, 163840)};
  for (int waves : {4, 8, 16})
    for (int nt : {2, 4})
      for (int ti = 0; ti < 2; ++ti) {
        if (nt == 4 && waves == 16) continue;  // spills at the 128-VGPR budget of 1,024 lanes
        ScanShape probe = make_shape(c, NC, nt, waves, tws[ti], 1);
        if (ti == 1 && probe.tile_words == make_shape(c, NC, nt, waves, tws[0], 1).tile_words) continue;
        for (unsigned per_cu : {1u, 2u, 4u, 8u})
          cand.push_back(make_shape(c, NC, nt, waves, tws[ti], std::max(1u, 256u * per_cu / probe.n_tiles)));
      }
  // BMG_SCAN_TUNE_LOG=1 prints every candidate's time to stderr
  const char* log_env = getenv("BMG_SCAN_TUNE_LOG");
  const bool log = log_env && log_env[0] == '1';
  hipEvent_t e0, e1;
  BMG_HIP(hipEventCreate(&e0));
  BMG_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int picked = -1;
  for (size_t i = 0; i < cand.size(); ++i) {
    float t_min = 1e30f;
    bool ok = true;
    for (int r = 0; r < 2 && ok; ++r) {
      (void)hipEventRecord(e0, s);
      scan_clear(c, cand[i], B, s);
      scan_mfma(c, NC, coding, cand[i], B, s);
      if (hipGetLastError() != hipSuccess) { ok = false; break; }
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) { ok = false; break; }
      float t = 0;
      (void)hipEventElapsedTime(&t, e0, e1);
      t_min = std::min(t_min, t);
    }
    if (ok && t_min < best) { best = t_min; picked = (int)i; }
    if (log)
      fprintf(stderr, "scan tune NC=%d waves=%d nt=%d tile_words=%u tiles=%u gx=%u: %s %.4f ms\n", NC,
              cand[i].waves, cand[i].nt, cand[i].tile_words, cand[i].n_tiles, cand[i].gx, ok ? "ok" : "failed",
              ok ? t_min : 0.0f);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (picked < 0) { set_error("scan: no candidate shape could be launched"); return BMG_E_HIP; }
  const ScanShape& pick = cand[picked];
  c->scan_nt[NC] = pick.nt; c->scan_wv[NC] = pick.waves;
  c->scan_tw[NC] = pick.tile_words; c->scan_gx[NC] = pick.gx;
  *out = pick;
  return BMG_OK;
}

int launch_scan(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                double* out_dev, hipStream_t s) {
  ScanBufs B;
  int rc = scan_prepare(c, r_dev, nvec, digits, s, &B);
  if (rc) return rc;
  ScanShape sh;
  rc = scan_tuned_shape(c, nvec, coding, digits, B, s, &sh);
  if (rc) return rc;
  rc = scan_kernel(c, nvec, coding, digits, sh, B, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_combine, dim3((unsigned)((c->p + 255) / 256), nvec), dim3(256), 0, s,
                     B.acc, c->p, nvec, digits, B.scale, out_dev);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int launch_scan_timed(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                      double* out_dev, int reps, double* ms, hipStream_t s) {
  ScanBufs B;
  int rc = scan_prepare(c, r_dev, nvec, digits, s, &B);
  if (rc) return rc;
  ScanShape sh;
  rc = scan_tuned_shape(c, nvec, coding, digits, B, s, &sh);
  if (rc) return rc;
  hipEvent_t e0, e1;
  BMG_HIP(hipEventCreate(&e0));
  BMG_HIP(hipEventCreate(&e1));
  double total = 0.0;
  // the clear stays outside the timed region: the events bracket the MFMA
  // kernel alone, as a kernel trace reports it
  for (int r = 0; r < reps && !rc; ++r) {
    scan_clear(c, sh, B, s);
    BMG_HIP(hipEventRecord(e0, s));
    scan_mfma(c, nvec * digits, coding, sh, B, s);
    BMG_HIP(hipGetLastError());
    BMG_HIP(hipEventRecord(e1, s));
    BMG_HIP(hipEventSynchronize(e1));
    float t = 0;
    BMG_HIP(hipEventElapsedTime(&t, e0, e1));
    total += t;
  }
  hipEventDestroy(e0); hipEventDestroy(e1);
  *ms = total / (reps > 0 ? reps : 1);
  hipLaunchKernelGGL(k_scan_combine, dim3((unsigned)((c->p + 255) / 256), nvec), dim3(256), 0, s,
                     B.acc, c->p, nvec, digits, B.scale, out_dev);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

// ------------------------------------------------ allocation-free scan --
// The scan's scratch for callers that cannot allocate (graph capture, tight
// serving loops): one file-scope device array per GPU, held until the
// process ends.  Host code never passes `scan_scratch` itself to a kernel
// (on the host the name is a stand-in, not a device address): the device
// address is looked up once per device with hipGetSymbolAddress and kept.
// Layout: [accumulator (p + 1024) x 16 int32 | digits 16 x n_pad int8 |
// scales 16 doubles], at most SCAN_SCRATCH_BYTES.
constexpr size_t SCAN_SCRATCH_ACC_ROWS = size_t(1) << 20;    // p + 1024 <= 2^20 SNP rows
constexpr size_t SCAN_SCRATCH_DIG = size_t(16) << 16;        // 16 columns x n_pad <= 65,536
constexpr size_t SCAN_SCRATCH_BYTES = SCAN_SCRATCH_ACC_ROWS * 16 * 4 + SCAN_SCRATCH_DIG + 16 * 8;
static __device__ __align__(256) uint8_t scan_scratch[SCAN_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {nullptr};

static int scratch_ptr(int device, void** out) {
  if (device < 0 || device >= 64) { set_error("scan_noalloc: device index out of range"); return BMG_E_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    BMG_HIP(hipGetDevice(&cur));
    if (cur != device) BMG_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scan_scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr) return hip_fail(e, "hipGetSymbolAddress(scan_scratch)");
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return BMG_OK;
}

int launch_scan_noalloc(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                        double* out_dev, hipStream_t s) {
  int rc = scan_check(nvec, digits);
  if (rc) return rc;
  if (c->p + 1024 > SCAN_SCRATCH_ACC_ROWS || (size_t)16 * c->n_pad > SCAN_SCRATCH_DIG) {
    set_error("scan_noalloc: the allocation-free scan holds p <= 1,047,552 SNPs and n <= 65,536 "
              "individuals; use bmg_scan_xtr_dev");
    return BMG_E_CAPACITY;
  }
  void* base = nullptr;
  rc = scratch_ptr(c->device, &base);
  if (rc) return rc;
  uint8_t* b8 = (uint8_t*)base;
  ScanBufs B{(int8_t*)(b8 + SCAN_SCRATCH_ACC_ROWS * 16 * 4), (int32_t*)b8,
             (double*)(b8 + SCAN_SCRATCH_ACC_ROWS * 16 * 4 + SCAN_SCRATCH_DIG)};
  rc = scan_digits(c, r_dev, nvec, digits, B, s);
  if (rc) return rc;
  // the shape the tuner picked for this width if it ran, else the fixed one
  // (no timing here: tuning creates events)
  const int NC = nvec * digits;
  ScanShape sh = c->scan_tw[NC] ? make_shape(c, NC, c->scan_nt[NC], c->scan_wv[NC], c->scan_tw[NC], c->scan_gx[NC])
                                : make_shape(c, NC, 2, 8, tile_for(NC, 163840), 256);
  rc = scan_kernel(c, nvec, coding, digits, sh, B, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_combine, dim3((unsigned)((c->p + 255) / 256), nvec), dim3(256), 0, s,
                     B.acc, c->p, nvec, digits, B.scale, out_dev);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

// plain 16-byte copy: the reference kernel of known traffic against which
// the FETCH_SIZE / WRITE_SIZE counters of the other kernels are scaled
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

int launch_copy_timed(const void* src, void* dst, uint64_t bytes, int reps, double* ms, hipStream_t s) {
  if (bytes % 16 || ((uintptr_t)src | (uintptr_t)dst) % 16) {
    set_error("copy: 16-byte aligned sizes and pointers only"); return BMG_E_ARG;
  }
  uint64_t n16 = bytes / 16;
  unsigned grid = (unsigned)std::min<uint64_t>((n16 + 255) / 256, 256 * 64);
  hipEvent_t e0, e1;
  BMG_HIP(hipEventCreate(&e0));
  BMG_HIP(hipEventCreate(&e1));
  double total = 0.0;
  for (int r = 0; r < reps; ++r) {
    BMG_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
    BMG_HIP(hipEventRecord(e1, s));
    BMG_HIP(hipEventSynchronize(e1));
    float t = 0;
    BMG_HIP(hipEventElapsedTime(&t, e0, e1));
    total += t;
  }
  hipEventDestroy(e0); hipEventDestroy(e1);
  *ms = total / (reps > 0 ? reps : 1);
  return BMG_OK;
}

// Streaming read of n16 16-byte words, four loads in flight per thread; the
// XOR of what a thread read is stored only if it equals a value the host
// never plants, so the loads cannot be dropped and nothing is written.
__global__ void k_read16(const uint4* __restrict__ src, uint64_t n16, uint32_t* __restrict__ sink) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t x = 0;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    x ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
  }
  for (; i < n16; i += stride) {
    uint4 a = src[i];
    x ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (x == 0x9E3779B9u) sink[0] = x;
}

// Best mean time over a few grid sizes: the streaming-read rate the device
// reaches, the ceiling for the read-bound scan.
int launch_read_timed(const void* src, uint64_t bytes, void* sink, int reps, double* ms, hipStream_t s) {
  if (bytes % 16 || (uintptr_t)src % 16) { set_error("read: 16-byte aligned size and pointer only"); return BMG_E_ARG; }
  const uint64_t n16 = bytes / 16;
  hipEvent_t e0, e1;
  BMG_HIP(hipEventCreate(&e0));
  BMG_HIP(hipEventCreate(&e1));
  double best = 1e30;
  for (unsigned grid : {1024u, 2048u, 4096u, 8192u}) {
    double total = 0.0;
    for (int r = 0; r <= reps; ++r) {  // the first launch is a warm-up
      BMG_HIP(hipEventRecord(e0, s));
      hipLaunchKernelGGL(k_read16, dim3(grid), dim3(256), 0, s, (const uint4*)src, n16, (uint32_t*)sink);
      BMG_HIP(hipGetLastError());
      BMG_HIP(hipEventRecord(e1, s));
      BMG_HIP(hipEventSynchronize(e1));
      float t = 0;
      BMG_HIP(hipEventElapsedTime(&t, e0, e1));
      if (r > 0) total += t;
    }
    best = std::min(best, total / (reps > 0 ? reps : 1));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = best;
  return BMG_OK;
}

template <int CA, int CB>
static void band_p(bmg_ctx* c, uint32_t W, int32_t* out_dev, uint64_t nrb, unsigned grid, hipStream_t s) {
  const uint32_t ncol = ((BAND_ROWS + W + 15) / 16) * 16;
  const size_t lds = (size_t)2 * (ncol + (CA == CB ? 0 : BAND_ROWS)) * BAND_SW * sizeof(uint32_t);
  const dim3 blk(64 * BAND_WAVES);
  const int nd = (int)((W + 31) / 16);
  // offsets per quarter: the smallest bucket with 4 NDQ >= nd
  if (nd <= 4)
    hipLaunchKernelGGL((k_band<CA, CB, 1>), dim3(grid), blk, lds, s, c->g, c->p, c->W32, W, out_dev, nrb);
  else if (nd <= 12)
    hipLaunchKernelGGL((k_band<CA, CB, 3>), dim3(grid), blk, lds, s, c->g, c->p, c->W32, W, out_dev, nrb);
  else
    hipLaunchKernelGGL((k_band<CA, CB, BAND_MAXP / 4>), dim3(grid), blk, lds, s, c->g, c->p, c->W32, W, out_dev,
                       nrb);
}

template <int CA>
static void band_cb(bmg_ctx* c, uint32_t W, int cb, int32_t* out_dev, uint64_t nrb, unsigned grid, hipStream_t s) {
  switch (cb) {
    case 0: band_p<CA, 0>(c, W, out_dev, nrb, grid, s); break;
    case 1: band_p<CA, 1>(c, W, out_dev, nrb, grid, s); break;
    case 2: band_p<CA, 2>(c, W, out_dev, nrb, grid, s); break;
    default: band_p<CA, 3>(c, W, out_dev, nrb, grid, s); break;
  }
}

int launch_band(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, hipStream_t s) {
  if (W < 1 || W > 560) { set_error("band: 1 <= W <= 560"); return BMG_E_ARG; }
  if ((W + 31) / 16 > BAND_MAXP) { set_error("band: W too large"); return BMG_E_ARG; }
  uint64_t nrb = (c->p + BAND_ROWS - 1) / BAND_ROWS;
  unsigned grid = (unsigned)(((nrb + 7) / 8) * 8);
  switch (ca) {
    case 0: band_cb<0>(c, W, cb, out_dev, nrb, grid, s); break;
    case 1: band_cb<1>(c, W, cb, out_dev, nrb, grid, s); break;
    case 2: band_cb<2>(c, W, cb, out_dev, nrb, grid, s); break;
    default: band_cb<3>(c, W, cb, out_dev, nrb, grid, s); break;
  }
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}
int launch_band_timed(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, int reps, double* ms,
                      hipStream_t s) {
  hipEvent_t e0, e1;
  BMG_HIP(hipEventCreate(&e0));
  BMG_HIP(hipEventCreate(&e1));
  double total = 0.0;
  int rc = BMG_OK;
  for (int r = 0; r < reps && !rc; ++r) {
    BMG_HIP(hipEventRecord(e0, s));
    rc = launch_band(c, W, ca, cb, out_dev, s);
    BMG_HIP(hipEventRecord(e1, s));
    BMG_HIP(hipEventSynchronize(e1));
    float t = 0;
    BMG_HIP(hipEventElapsedTime(&t, e0, e1));
    total += t;
  }
  hipEventDestroy(e0); hipEventDestroy(e1);
  *ms = total / (reps > 0 ? reps : 1);
  return rc;
}

}  // namespace bmg
