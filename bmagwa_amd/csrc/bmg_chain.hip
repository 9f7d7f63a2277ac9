// bmg_chain.hip -- the device-resident PMV sampler: FP64 model state in LDS,
// Cholesky add / remove / full refactorisation, sum-tree proposals, the
// moves of src/sampler.hpp:899-1669 and the schedule of src/sampler.cpp:
// 626-834, one workgroup per chain.
//
// All threads of a workgroup run the scalar control flow redundantly on
// identical register copies (the random stream included), so every decision
// is uniform; the O(n k) parts use all 256 threads and the O(k^2) triangular
// and Givens sweeps use wave 0.  Triangular solves and the Cholesky
// factorisation are right-looking: each element receives its subtractions in
// the same order as the netlib loops of the CPU oracle, so with contraction
// off the factors agree bit for bit; only the inner products differ in
// summation order.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include <type_traits>
#include "bmg_chain.h"
#include "bmg_device.h"
#include "../../include/bmagwa_hip.h"

using namespace bmg;

#define LOG_HALF (-0.69314718055994528622676398299518041312694549560546875)
#define TLEV_LDS 512

__device__ __forceinline__ int P(int c) { return c * (c + 1) / 2; }

// A log-likelihood that is not finite (+inf when the residual term reaches
// zero, NaN) can only come from rounding in a near-singular design; it is
// taken as -inf, so such a proposal is rejected (as in the oracle).
__device__ __forceinline__ double finite_ll(double v) { return isfinite(v) ? v : -INFINITY; }

// element i (runtime) of a small register array without dynamic indexing,
// which would place the whole per-lane state in scratch memory.  Each
// element passes through an empty register constraint, so the compiler
// cannot fold the selects back into one load at a computed address.
template <typename T, int N>
__device__ __forceinline__ T sel(const T (&a)[N], int i) {
  T r = a[0];
  asm volatile("" : "+v"(r));
#pragma unroll
  for (int q = 1; q < N; ++q) {
    T v = a[q];
    asm volatile("" : "+v"(v));
    r = (i == q) ? v : r;
  }
  return r;
}
template <typename T, int N>
__device__ __forceinline__ void put(T (&a)[N], int i, T v) {
#pragma unroll
  for (int q = 0; q < N; ++q) a[q] = (i == q) ? v : a[q];
}

// ------------------------------------------------------------- state -----
// The proposed model's factor (packed upper): an LDS array up to 160
// columns; a larger capacity keeps it in a per-chain global array (L2/HBM),
// whose address thread 0 stores in the LDS member at kernel entry.  The
// member is a plain array in the LDS case (as a base-class member the
// compiler's code for the triangular loops was a fifth slower).
constexpr int KMAX_LDS = 160;
template <int KMAX>
using FactorT = std::conditional_t<(KMAX <= KMAX_LDS), double[KMAX * (KMAX + 1) / 2], double*>;

template <int KMAX>
struct Smem {
  FactorT<KMAX> L;                       // proposed model's factor (packed upper)
  double xy[KMAX], v[KMAX], d[KMAX];     // proposed model vectors
  double beta[KMAX], mu[KMAX];           // the chain's beta / mu_beta
  double work[KMAX + 64];
  double c_xy[KMAX], c_v[KMAX], c_d[KMAX];
  int32_t slot[KMAX], snp[KMAX];         // per column: slot in xxs, SNP
  int32_t c_slot[KMAX], c_snp[KMAX];
  uint32_t loci[KMAX], c_loci[KMAX];     // per locus
  int32_t xi1[KMAX], xi2[KMAX], c_xi1[KMAX], c_xi2[KMAX];
  int8_t xtype[KMAX], c_xtype[KMAX], tind[KMAX], c_tind[KMAX];
  uint64_t mi[BMG_MAXMS], mi_add[BMG_MAXMS], mi_rem[BMG_MAXMS];
  int64_t mi_map[BMG_MAXMS];
  uint8_t mi_isadd[BMG_MAXMS];
  uint64_t rm_list[2 * BMG_MAXMS];
  uint64_t marks[BMG_MAXMS];
  int32_t mark_val[BMG_MAXMS];
  uint64_t sc_perm[BMG_MAXMS], sc_from[BMG_MAXMS], sc_to[BMG_MAXMS];
  double tlev[2][TLEV_LDS];              // sum-tree levels >= 2 (add, rem)
  int32_t win[512];                      // window of the current model_inds
  uint64_t slots_cur[(KMAX + 63) / 64], slots_prop[(KMAX + 63) / 64];
  int32_t iacc[KMAX + 1];                // integer column dots of new_column_dots
  int32_t iacc2[KMAX + 1];               // the next addition's dots, formed ahead
  // delayed rejection: factor block, v, d and log d of the toggled loci's
  // columns, their column offsets and types, and the moved SNPs' proposal
  // weights
  double ex_L[BMG_DR_COLS * (BMG_DR_COLS + 1) / 2];
  double ex_v[BMG_DR_COLS], ex_d[BMG_DR_COLS], ex_ld[BMG_DR_COLS];
  double ex_pa[BMG_DR_MAX + 1], ex_pr[BMG_DR_MAX + 1];   // prior change by model size
  double dq_add[BMG_DR_MAX], dq_rem[BMG_DR_MAX];
  int32_t ex_x1[BMG_DR_MAX], ex_x2[BMG_DR_MAX];
  int8_t ex_t[BMG_DR_MAX];
  uint8_t ex_b2n[BMG_DR_MAX];
  // undo journal of an add/remove move's sum-tree writes: the global
  // entries it overwrote (leaf and level 1, in order) and a copy of the LDS
  // levels taken when the move began
  double tj_old[2 * BMG_MAXMS][2];
  uint64_t tj_ind[2 * BMG_MAXMS];        // leaf index, tree in bit 63
  double tj_tlev[2][TLEV_LDS];
  double bq[BMG_MAXMS];                  // proposal weights of the reverse move
  double red[8];
  long long redi[8];
  double bcd[4];
  long long bci[4];
};

struct Reg {
  Rng rng;
  int cols, nloci, Ns[5];
  double syx, ldq, ldl, ll;
  int c_cols, c_nloci, c_Ns[5];
  double c_syx, c_ldq, c_ldl, c_ll;
  double sigma2, alpha, alpha2, ita2[4];
  int beta_len, mu_computed;
  int lmin;
  int n_rm, n_marks;
  int movesize;
  double tot_add, tot_rem;
  int err;
  int learn_on;
  int tj_on, tj_n;                       // journal active, entries
  long long pre_snp;                     // SNP whose dots are in iacc2 (or -1)
  int pre_slot;                          // and the slot its value words went to
  double tj_tot[2];                      // tree totals when the move began
};

template <int KMAX>
struct Ctx {
  Smem<KMAX>& S;
  Reg& R;
  const DevData& D;
  const ChainConst& K;
  const ChainArrays& A;
  uint64_t* prof;   // the chain's phase counters (thread 0 writes), or null
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// broadcast a value computed on wave 0 lane 0 to every thread
template <int KMAX>
__device__ __forceinline__ double bcast_d(Smem<KMAX>& S, double v) {
  __syncthreads();
  if (threadIdx.x == 0) S.bcd[0] = v;
  __syncthreads();
  double r = S.bcd[0];
  __syncthreads();
  return r;
}
template <int KMAX>
__device__ __forceinline__ long long bcast_i(Smem<KMAX>& S, long long v) {
  __syncthreads();
  if (threadIdx.x == 0) S.bci[0] = v;
  __syncthreads();
  long long r = S.bci[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------ genotypes ---
__device__ __forceinline__ int coding_of_type(int type, int second) {
  // AH splits into A then H (src/model.hpp:254-258)
  return type == 4 ? (second ? 1 : 0) : type;
}

// x_j^T y (FP64) for one SNP/coding
// x . y of one SNP's packed row: lane-strided words, 16 fields each in
// order, then the workgroup sum (the order both the per-add dot and the
// set-up table use)
__device__ __forceinline__ double dot_y_block(const uint32_t* __restrict__ w, uint32_t W,
                                              const double* __restrict__ y, int coding, double* red) {
  double s = 0.0;
  for (uint32_t k = threadIdx.x; k < W; k += BMG_BLOCK) {
    uint32_t x = w[k];
    const double* yy = y + 16 * (size_t)k;
#pragma unroll
    for (int f = 0; f < 16; ++f) {
      int iv = code_ival((x >> (2 * f)) & 3u, coding);
      if (iv) s += (double)iv * yy[f];
    }
  }
  return block_sum(s, red);
}

// The Makefile compiles this file once per model capacity (BMG_KMAX_ONLY =
// 64, 128, 160, 256) so the four instantiations build in parallel; the
// non-template code lives in the 64-column unit.
#ifndef BMG_KMAX_ONLY
#define BMG_KMAX_ONLY 0
#endif
#define BMG_INST(K) (BMG_KMAX_ONLY == 0 || BMG_KMAX_ONLY == (K))

#if BMG_INST(64)
__global__ __launch_bounds__(BMG_BLOCK) void k_xy_all(DevData D, int coding, double* __restrict__ out) {
  __shared__ double red[8];
  for (uint64_t j = blockIdx.x; j < D.p; j += gridDim.x) {
    const double v = dot_y_block(D.g + j * D.W32, D.W32, D.y, coding, red);
    if (threadIdx.x == 0) out[j] = v;
  }
}
#endif

template <int KMAX>
__device__ __forceinline__ double snp_dot_y(Ctx<KMAX>& C, uint64_t snp, int coding) {
  const double* pre = C.D.xy[coding];
  if (pre) return pre[snp];   // precomputed at set-up in this same order
  return dot_y_block(C.D.g + snp * C.D.W32, C.D.W32, C.D.y, coding, C.S.red);
}

// value planes of a word under a coding known at compile time (CODING >= 0)
// or given at run time (CODING < 0)
template <int CODING>
__device__ __forceinline__ void planes_c(uint32_t w, int coding, uint32_t& p1, uint32_t& p2) {
  if (CODING < 0) {
    planes(w, coding, p1, p2);
  } else {
    const uint32_t lo = w & BMG_EVEN, hi = (w >> 1) & BMG_EVEN;
    if (CODING == 0) { p1 = hi; p2 = hi & lo; }
    else if (CODING == 1) { p1 = hi & ~lo; p2 = 0u; }
    else if (CODING == 2) { p1 = hi; p2 = 0u; }
    else { p1 = hi & lo; p2 = 0u; }
  }
}

// S.work[c] = x_c^T x_new for c in [0, cols], column `cols` being the new
// one itself; column 0 is the constant.  Exact integer dots: the slot
// copies of the model's columns hold value words (the coded genotype, 0..2,
// in each 2-bit field), whose even fields (masked in place) and odd fields
// (the rest, four times their value) meet the new column's in two 8-way
// unsigned 4-bit dot products (v_dot8_u32_u4); the odd sum is divided by
// four at the end.  The constant column and the new column's own dot come
// from the new column's nibbles alone.
//
// A lane owns one 16-word chunk of the individuals and a group of columns:
// it decodes the new column's chunk once, then streams its columns' chunks,
// four columns' words in flight at a time, each column's loads issued as
// soon as the previous column in its register set is summed; the chunk
// sums meet in LDS integer accumulators (exact, so the order is free).
// W / 16 lanes cover a column, so NTH / (W / 16) column groups run side by
// side; the first group also writes the new column's value words to its
// slot copy (xnew).  CODING: the new column's coding; threads [T0, T0 + NTH)
// take part (the whole workgroup, or waves 1-3 while wave 0 runs a
// triangular sweep); acc: the integer accumulators, zeroed by the caller.
template <int KMAX, int CODING, int T0, int NTH>
__device__ __forceinline__ void column_dots(Ctx<KMAX>& C, uint64_t snp_new, int cols, uint32_t* __restrict__ xnew,
                                            int32_t* acc) {
  const uint32_t W = C.D.W32;                 // a multiple of 32 words
  const uint32_t* __restrict__ wn = C.D.g + snp_new * W;
  const uint32_t* __restrict__ xcol = C.A.xcol;
  Smem<KMAX>& S = C.S;
  const int me = C.D.me;
  const uint32_t cpc = W / 16;                // chunks per column
  const int tid = (int)threadIdx.x - T0;
  if (tid < 0 || tid >= NTH) return;
  for (uint32_t ch0 = 0; ch0 < cpc; ch0 += NTH) {
    const uint32_t span = min(cpc - ch0, (uint32_t)NTH);
    const int groups = (int)(NTH / span);
    const int grp = (int)(tid / span);
    const uint32_t w0 = (ch0 + tid % span) * 16;
    if (grp < groups) {
      uint32_t ne[16], no[16];   // the new column's even / odd fields as nibbles
      uint32_t s_one = 0, s_self = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 v = *(const uint4*)(wn + w0 + 4 * q);
        const uint32_t xs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          uint32_t p1, p2;
          planes_c<CODING>(xs[u], CODING, p1, p2);
          const uint32_t vw = p1 + p2;
          const int k = 4 * q + u;
          if (grp == 0) xnew[w0 + k] = vw;
          ne[k] = vw & 0x33333333u;
          no[k] = (vw >> 2) & 0x33333333u;
          s_one = __builtin_amdgcn_udot8(0x11111111u, ne[k], s_one, false);
          s_one = __builtin_amdgcn_udot8(0x11111111u, no[k], s_one, false);
          s_self = __builtin_amdgcn_udot8(ne[k], ne[k], s_self, false);
          s_self = __builtin_amdgcn_udot8(no[k], no[k], s_self, false);
        }
      }
      // column of round r, register set u: grp + (4 r + u) groups; every
      // group runs the rounds of group 0, the one with the most columns
      const int rounds = (cols / groups + 1 + 3) / 4;
      auto src = [&](int cu) -> const uint32_t* {
        const int cc = (cu >= me && cu < cols) ? cu : 0;
        return (cu >= me && cu < cols) ? xcol + (size_t)S.slot[cc] * W : wn;
      };
      uint4 cur[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t* pc = src(grp + u * groups);
#pragma unroll
        for (int q = 0; q < 4; ++q) cur[u][q] = *(const uint4*)(pc + w0 + 4 * q);
      }
      for (int r = 0; r < rounds; ++r) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int cu = grp + (4 * r + u) * groups;
          uint32_t se = 0, so = 0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t xs[4] = {cur[u][q].x, cur[u][q].y, cur[u][q].z, cur[u][q].w};
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
              const int k = 4 * q + k4;
              const uint32_t e = xs[k4] & 0x33333333u;
              se = __builtin_amdgcn_udot8(e, ne[k], se, false);
              so = __builtin_amdgcn_udot8(xs[k4] ^ e, no[k], so, false);
            }
          }
          // next round's column of this register set
          const uint32_t* pn = src(cu + 4 * groups);
#pragma unroll
          for (int q = 0; q < 4; ++q) cur[u][q] = *(const uint4*)(pn + w0 + 4 * q);
          const uint32_t sum = cu == 0 ? s_one : (cu == cols ? s_self : se + (so >> 2));
          // covariate columns (0 < c < me): cov_dots
          if (cu <= cols && (cu == 0 || cu >= me)) atomicAdd(&acc[cu], (int)sum);
        }
      }
    }
  }
}

// S.work[c] = e_c^T x_new for the covariate columns c in [1, me): each
// thread sums over its individuals, then a fixed-order block reduction
template <int KMAX>
__device__ __forceinline__ void cov_dots(Ctx<KMAX>& C, uint64_t snp_new, int coding_new) {
  const int me = C.D.me;
  const uint32_t* __restrict__ wn = C.D.g + snp_new * C.D.W32;
  for (int c = 1; c < me; ++c) {
    const double* __restrict__ e = C.D.E + (size_t)(c - 1) * C.D.n_pad;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < C.D.n; i += BMG_BLOCK) {
      const int iv = code_ival((wn[i >> 4] >> (2 * (i & 15))) & 3u, coding_new);
      if (iv > 0) acc += (double)iv * e[i];
    }
    acc = block_sum(acc, C.S.red);
    if (threadIdx.x == 0) C.S.work[c] = acc;
  }
  __syncthreads();
}

// S.work[c] = x_c^T x_new for c in [0, cols] (covariates included), and
// the new column's value words into its slot copy xnew
template <int KMAX>
__device__ __forceinline__ void new_column_dots(Ctx<KMAX>& C, uint64_t snp_new, int coding_new, int cols,
                                                uint32_t* xnew) {
  Smem<KMAX>& S = C.S;
  for (int c = threadIdx.x; c <= cols; c += BMG_BLOCK) S.iacc[c] = 0;
  __syncthreads();
  switch (coding_new) {
    case 0: column_dots<KMAX, 0, 0, BMG_BLOCK>(C, snp_new, cols, xnew, S.iacc); break;
    case 1: column_dots<KMAX, 1, 0, BMG_BLOCK>(C, snp_new, cols, xnew, S.iacc); break;
    case 2: column_dots<KMAX, 2, 0, BMG_BLOCK>(C, snp_new, cols, xnew, S.iacc); break;
    default: column_dots<KMAX, 3, 0, BMG_BLOCK>(C, snp_new, cols, xnew, S.iacc); break;
  }
  __syncthreads();
  for (int c = threadIdx.x; c <= cols; c += BMG_BLOCK) S.work[c] = (double)S.iacc[c];
  __syncthreads();
  if (C.D.me > 1) cov_dots(C, snp_new, coding_new);
}

// ----------------------------------------------------- triangular (wave 0)
// value of lane `src` (uniform) of a double held by every lane
__device__ __forceinline__ double readlane_d(double v, int src) {
  int2 h = *reinterpret_cast<int2*>(&v);
  h.x = __builtin_amdgcn_readlane(h.x, src);
  h.y = __builtin_amdgcn_readlane(h.y, src);
  return *reinterpret_cast<double*>(&h);
}

// x := U^-T x for x[0..m) in LDS (dtrsv Upper/Trans order).  Wave 0 only.
// Lane l holds x[l + 64 t] in b[t] and the reciprocal of that row's
// diagonal entry in dg[t] (all computed at once before the sweep); the
// outer loop over t is unrolled so the register arrays are only ever
// indexed by constants.  Row i: the owner's scaled entry is read across
// lanes and every lane subtracts its coefficient times it, so the serial
// chain per row is one multiply, one cross-lane read and one fused
// multiply-subtract, with no division.  The sweep has no branches and no
// masks: every lane applies every row with the entry at its own column's
// row i (a lane at or past its row reads another column's entry into a
// value it no longer needs; its result was kept in xo when its row came),
// and the coefficients are read two rows ahead from clamped addresses, so
// no row waits on LDS.  Per element the subtractions come in row order, as
// in the sequential loop; the reciprocal and the fused update round
// differently from its divisions (by an ulp), which the log-likelihoods'
// 1e-10 agreement with the oracle absorbs.
template <int KMAX>
__device__ __forceinline__ void trsv_ut_w0(Smem<KMAX>& S, double* x, int m) {
  constexpr int NT = (KMAX + 63) / 64;
  constexpr int LIM = KMAX * (KMAX + 1) / 2 - 1;
  const int lane = lane_id();
  double b[NT], dg[NT], xo[NT];
  int base[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = lane + 64 * t;
    b[t] = j < m ? x[j] : 0.0;
    dg[t] = j < m ? 1.0 / S.L[P(j) + j] : 1.0;
    base[t] = j < m ? P(j) : 0;
    xo[t] = 0.0;
  }
  auto ld = [&](int t, int i) -> double {
    const int a = base[t] + i;
    return S.L[a < LIM ? a : LIM];
  };
#pragma unroll
  for (int ti = 0; ti < NT; ++ti) {
    const int i0 = 64 * ti;
    const int iend = m < 64 * (ti + 1) ? m : 64 * (ti + 1);
    if (i0 >= iend) break;
    double ca[NT], cb[NT];
#pragma unroll
    for (int t = ti; t < NT; ++t) { ca[t] = ld(t, i0); cb[t] = ld(t, i0 + 1); }
    auto row = [&](int i, double (&c)[NT]) {
      const int owner = i & 63;
      const double xi = readlane_d(b[ti] * dg[ti], owner);
      xo[ti] = lane == owner ? xi : xo[ti];
#pragma unroll
      for (int t = ti; t < NT; ++t)
        if (64 * t < m) b[t] = fma(-c[t], xi, b[t]);
#pragma unroll
      for (int t = ti; t < NT; ++t) c[t] = ld(t, i + 2);
    };
    int i = i0;
    for (; i + 1 < iend; i += 2) { row(i, ca); row(i + 1, cb); }
    if (i < iend) row(i, ca);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) { const int j = lane + 64 * t; if (j < m) x[j] = xo[t]; }
}

// x := U^-1 x (dtrsv Upper/NoTrans, back substitution).  Wave 0 only.  Lane
// l holds x[l + 64 t] and the reciprocal of that row's diagonal; step j
// reads the owner's scaled entry across lanes and every lane subtracts its
// entry of column j (one column: consecutive addresses) times it, so the
// serial chain per step is a multiply, a cross-lane read and a fused update
// (rounding within an ulp of the divide-and-subtract loop).  As in the
// forward sweep there are no branches or masks (a lane at or past its row
// updates a value it no longer needs; its result is in xo) and the columns
// are read two steps ahead.  A zero x_j leaves every entry as it was, as
// the reference's skip of zero entries does.
template <int KMAX>
__device__ __forceinline__ void trsv_un_w0(Smem<KMAX>& S, double* x, int m) {
  constexpr int NT = (KMAX + 63) / 64;
  constexpr int LIM = KMAX * (KMAX + 1) / 2 - 1;
  const int lane = lane_id();
  if (m <= 0) return;
  double b[NT], rd[NT], xo[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int j = lane + 64 * t;
    b[t] = j < m ? x[j] : 0.0;
    rd[t] = j < m ? 1.0 / S.L[P(j) + j] : 1.0;
    xo[t] = 0.0;
  }
  // entry (lane + 64 q) of column j, clamped into the factor
  auto ld = [&](int q, int j) -> double {
    const int a = (j > 0 ? P(j) : 0) + lane + 64 * q;
    return S.L[a < LIM ? a : LIM];
  };
#pragma unroll
  for (int tj = NT - 1; tj >= 0; --tj) {
    const int jend = m < 64 * (tj + 1) ? m : 64 * (tj + 1);
    if (64 * tj >= jend) continue;
    double ca[NT], cb[NT];
#pragma unroll
    for (int q = 0; q <= tj; ++q) { ca[q] = ld(q, jend - 1); cb[q] = ld(q, jend - 2); }
    auto step = [&](int j, double (&c)[NT]) {
      const int owner = j & 63;
      const double xj = readlane_d(b[tj] * rd[tj], owner);
      xo[tj] = lane == owner ? xj : xo[tj];
#pragma unroll
      for (int q = 0; q <= tj; ++q) b[q] = fma(-xj, c[q], b[q]);
#pragma unroll
      for (int q = 0; q <= tj; ++q) c[q] = ld(q, j - 2);
    };
    int j = jend - 1;
    for (; j - 1 >= 64 * tj; j -= 2) { step(j, ca); step(j - 1, cb); }
    if (j >= 64 * tj) step(j, ca);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) { const int j = lane + 64 * t; if (j < m) x[j] = xo[t]; }
}

// netlib drotg (as the oracle restates it)
__device__ __forceinline__ void drotg(double& a, double& b, double& c, double& s) {
  double sa = a, sb = b;
  double roe = (fabs(sa) > fabs(sb)) ? sa : sb;
  double scale = fabs(sa) + fabs(sb);
  double r, z;
  if (scale == 0.0) { c = 1.0; s = 0.0; r = 0.0; z = 0.0; }
  else {
    double as = sa / scale, bs = sb / scale;
    r = scale * sqrt(as * as + bs * bs);
    r = (roe >= 0.0 ? 1.0 : -1.0) * r;
    c = sa / r; s = sb / r;
    z = 1.0;
    if (fabs(sa) > fabs(sb)) z = s;
    if (fabs(sb) >= fabs(sa) && c != 0.0) z = 1.0 / c;
  }
  a = r; b = z;
}

// The rotation drotg generates (same sign convention: r takes the sign of
// the larger input), without its overflow scaling, which a Cholesky factor's
// entries never need.  1 / r comes from the hardware reciprocal square root
// of a^2 + b^2 refined by two Newton steps (each roughly doubles the correct
// bits, ending within an ulp or two of the correctly rounded value), so the
// downdate's serial chain per step is a handful of multiplies instead of a
// square root and a division.  c, s and r agree with drotg's to rounding.
__device__ __forceinline__ void rot_gen(double& a, double b, double& c, double& s) {
  const double roe = fabs(a) > fabs(b) ? a : b;
  const double q = fma(a, a, b * b);
  double y = __builtin_amdgcn_rsq(q);
  const double h = 0.5 * q;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  const double ri = q > 0.0 ? (roe >= 0.0 ? y : -y) : 0.0;
  c = q > 0.0 ? a * ri : 1.0;
  s = b * ri;
  a = q * ri;
}

// Cholesky downdate removing column cr of the m x m factor (LINPACK dchex
// job 2 + drop + sign fix, src/symmmatrix.cpp:166-195).  Logical column jl >=
// cr lives in physical column jl + 1 until it is written back; rotation j is
// generated from logical column j's rows (j, j + 1) and applied to every
// later column at once (per element the same sequence as the sequential
// loop).  The two rows a rotation touches stay in registers: lane jl % 64
// keeps rows (j, j + 1) of its columns, reads rows two steps ahead from LDS
// (rows below j + 1 are not rotated yet) and stores only the finished row,
// and the owner of column j hands the rotation's inputs to all lanes with a
// cross-lane read, so no step waits on an LDS round trip.
//
// Finished rows go straight to their place in the compacted factor
// (physical column jl, whose entries at and past row cr were read two or
// more steps earlier), with the sign fix folded in: a row whose diagonal
// comes out negative is stored negated, which is what the separate pass
// over the finished rows does.  Meanwhile wave 1 moves the rows above cr of
// the later columns down by one column, in ascending order of destination
// (every source lies above every earlier destination, so one wave's
// batches never overwrite what a later batch reads); the two touch disjoint
// entries.
//
// v (U^T v = X^T y) goes through the same rotations and sign fix as the
// factor's rows, so after the removal it solves the new factor's system
// without a second triangular solve (the reference re-solves,
// src/model.hpp:530-545; the two agree to rounding).
template <int KMAX, typename Extra>
__device__ __forceinline__ void chol_downdate_lds(Smem<KMAX>& S, int m, int cr, Extra extra) {
  constexpr int NT = (KMAX + 63) / 64;
  constexpr int LIM = KMAX * (KMAX + 1) / 2 - 1;
  const int lane = lane_id();
  if (wave_id() == 0) {
    double vt = S.v[cr], vb = S.v[cr + 1];   // rows (j, j + 1) of v, all lanes
    double top[NT], nA[NT], nB[NT];
    int pb[NT], lb[NT];   // physical and logical column offsets
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jl = lane + 64 * t;
      pb[t] = P(jl + 1);
      lb[t] = P(jl);
      top[t] = S.L[min(pb[t] + cr, LIM)];
      nA[t] = S.L[min(pb[t] + cr + 1, LIM)];
      nB[t] = S.L[min(pb[t] + cr + 2, LIM)];
    }
#pragma unroll
    for (int ti = 0; ti < NT; ++ti) {
      const int jbeg = cr > 64 * ti ? cr : 64 * ti;
      const int jend = (m - 1) < 64 * (ti + 1) ? (m - 1) : 64 * (ti + 1);
      // step j: n1 holds row j + 1 of each column and takes row j + 3
      auto step = [&](int j, double (&n1)[NT]) {
        const int owner = j & 63;
        // every lane generates the rotation from the same two values
        double a = readlane_d(top[ti], owner), b = readlane_d(n1[ti], owner);
        const double vbn = S.v[j + 2 < m ? j + 2 : m - 1];   // next row of v, read ahead
        double cc, ss;
        rot_gen(a, b, cc, ss);
        const double sg = a < 0.0 ? -1.0 : 1.0;   // sign fix of row j
        if (lane == owner) S.L[lb[ti] + j] = a * sg;
        {
          const double tv = cc * vt + ss * vb;
          vt = cc * vb - ss * vt;
          if (lane == 0) S.v[j] = tv * sg;
          vb = vbn;
        }
#pragma unroll
        for (int t = ti; t < NT; ++t) {
          if (64 * t <= m - 2) {
            const int jl = lane + 64 * t;
            const double x1 = top[t], x2 = n1[t];
            const double tt = cc * x1 + ss * x2;
            const bool on = jl > j && jl <= m - 2;
            top[t] = on ? cc * x2 - ss * x1 : top[t];
            if (on) S.L[lb[t] + j] = tt * sg;
            n1[t] = S.L[min(pb[t] + j + 3, LIM)];
          }
        }
      };
      int j = jbeg;
      for (; j + 1 < jend; j += 2) { step(j, nA); step(j + 1, nB); }
      if (j < jend) {
        step(j, nA);
#pragma unroll
        for (int t = 0; t < NT; ++t) { const double tmp = nA[t]; nA[t] = nB[t]; nB[t] = tmp; }
      }
    }
  } else {
    if (wave_id() == 1) {
      // rows [0, cr) of logical columns [cr, m - 2]: physical column jl + 1
      // -> jl, element k = (jl - cr) * cr + i in ascending order
      const int total = (m - 1 - cr) * cr;
      const float rcr = cr > 0 ? 1.0f / (float)cr : 0.0f;
      for (int k0 = 0; k0 < total; k0 += 64) {
        const int k = k0 + lane;
        if (k < total) {
          int q = (int)((float)k * rcr);
          if (q * cr > k) --q;
          if ((q + 1) * cr <= k) ++q;
          const int jl = cr + q, i = k - q * cr;
          S.L[P(jl) + i] = S.L[P(jl + 1) + i];
        }
      }
    }
    extra();   // waves 1-3: work that touches neither the factor nor v
  }
  __syncthreads();
}

// The same downdate for a factor in global memory (capacity above the LDS
// factor's): rows rotated in place, then the sign fix and the compaction as
// separate passes (LINPACK dchex job 2 + drop + sign fix,
// src/symmmatrix.cpp:166-195).  Logical column jl >=
// cr lives in physical column jl + 1 until the final compaction; rotation j
// is generated from logical column j's rows (j, j + 1) and applied to every
// later column at once (per element the same sequence as the sequential
// loop).  The two rows a rotation touches stay in registers: lane jl % 64
// keeps rows (j, j + 1) of its columns, reads the next row ahead from LDS
// (rows below j + 1 are not rotated yet) and stores only the finished row,
// and the owner of column j hands the rotation's inputs to all lanes with a
// cross-lane read, so no step waits on an LDS round trip.
//
// v (U^T v = X^T y) goes through the same rotations and sign fix as the
// factor's rows, so after the removal it solves the new factor's system
// without a second triangular solve (the reference re-solves,
// src/model.hpp:530-545; the two agree to rounding).
template <int KMAX, typename Extra>
__device__ __forceinline__ void chol_downdate_glb(Smem<KMAX>& S, int m, int cr, Extra extra) {
  constexpr int NT = (KMAX + 63) / 64;
  if (wave_id() == 0) {
    const int lane = lane_id();
    double vt = S.v[cr], vb = S.v[cr + 1];   // rows (j, j + 1) of v, all lanes
    double top[NT], nxt[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jl = lane + 64 * t;
      const bool on = jl >= cr && jl <= m - 2;
      top[t] = on ? S.L[P(jl + 1) + cr] : 0.0;
      nxt[t] = on ? S.L[P(jl + 1) + cr + 1] : 0.0;
    }
#pragma unroll
    for (int ti = 0; ti < NT; ++ti) {
      const int jbeg = cr > 64 * ti ? cr : 64 * ti;
      const int jend = (m - 1) < 64 * (ti + 1) ? (m - 1) : 64 * (ti + 1);
      for (int j = jbeg; j < jend; ++j) {
        const int owner = j & 63;
        // every lane generates the rotation from the same two values
        double a = readlane_d(top[ti], owner), b = readlane_d(nxt[ti], owner);
        const double vbn = j + 2 < m ? S.v[j + 2] : 0.0;   // next row of v, read ahead
        double cc, ss;
        rot_gen(a, b, cc, ss);
        if (lane == owner) S.L[P(j + 1) + j] = a;
        {
          const double tv = cc * vt + ss * vb;
          vt = cc * vb - ss * vt;
          if (lane == 0) S.v[j] = tv;
          vb = vbn;
        }
#pragma unroll
        for (int t = ti; t < NT; ++t) {
          const int jl = lane + 64 * t;
          if (jl > j && jl <= m - 2) {
            const double x1 = top[t], x2 = nxt[t];
            const double tt = cc * x1 + ss * x2;
            top[t] = cc * x2 - ss * x1;
            S.L[P(jl + 1) + j] = tt;
            nxt[t] = S.L[P(jl + 1) + j + 2];   // j + 2 <= jl + 1: inside the column
          }
        }
      }
    }
  } else {
    extra();   // waves 1-3: work that touches neither the factor nor v
  }
  __syncthreads();
  // sign fix over the kept rows i in [cr, m-2]: flags first, then negation
  for (int i = cr + threadIdx.x; i <= m - 2; i += BMG_BLOCK) S.work[i] = S.L[P(i + 1) + i] < 0 ? -1.0 : 1.0;
  __syncthreads();
  // eight rows at a time with unconditional stores: a row past jl is
  // clamped to jl, so its store rewrites col[jl] with the value the last
  // real row gives it (a conditional store per row waits out an LDS round
  // trip each)
  for (int jl = cr + threadIdx.x; jl <= m - 2; jl += BMG_BLOCK) {
    double* col = S.L + P(jl + 1);
    for (int i0 = cr; i0 <= jl; i0 += 8) {
      double c[8], f[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = i0 + q <= jl ? i0 + q : jl;
        c[q] = col[i];
        f[q] = S.work[i];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int i = i0 + q <= jl ? i0 + q : jl;
        col[i] = f[q] < 0 ? -c[q] : c[q];
      }
    }
    if (S.work[jl] < 0) S.v[jl] = -S.v[jl];
  }
  __syncthreads();
  // compaction: logical column jl rows 0..jl <- physical column jl+1
  const int e0 = P(cr), e1 = P(m - 1);
  for (int base = e0; base < e1; base += BMG_BLOCK * 8) {
    double tmp[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int e = base + q * BMG_BLOCK + threadIdx.x;
      if (e < e1) {
        int jl = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (P(jl + 1) <= e) ++jl;
        while (P(jl) > e) --jl;
        tmp[q] = S.L[P(jl + 1) + (e - P(jl))];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      int e = base + q * BMG_BLOCK + threadIdx.x;
      if (e < e1) S.L[e] = tmp[q];
    }
    __syncthreads();
  }
}

template <int KMAX, typename Extra>
__device__ __forceinline__ void chol_downdate(Smem<KMAX>& S, int m, int cr, Extra extra) {
  if constexpr (KMAX <= KMAX_LDS) {
    chol_downdate_lds(S, m, cr, extra);
  } else {
    chol_downdate_glb(S, m, cr, extra);
  }
}

// Full factorisation of S.L (m x m, in place) in panels of CF_PB rows;
// returns 0 if not positive definite.  Per panel [k0, k1): wave 0 factors
// the diagonal block in registers (lane l holds column k0 + l; the pivot
// and each row-k entry reach the other lanes by cross-lane reads, so the
// block needs no barrier); then a thread per later column computes that
// column's panel rows from the block (a forward substitution of CF_PB
// steps); then the trailing update gives a lane to each row i >= k1 and
// the four waves every fourth column j, so a wave's 64 lanes read and
// write 64 consecutive entries of one column and all read the same panel
// entries of column j.  Every element still receives the sequential loop's
// subtractions in k order, each a multiply then a subtract, and its
// division or square root after them, so the factor is that loop's bit for
// bit; the panels cost three barriers each instead of two per column.
constexpr int CF_PB = 16;
template <int KMAX>
__device__ __forceinline__ int chol_full(Smem<KMAX>& S, int m) {
  constexpr int PB = CF_PB;
  constexpr int NT = (KMAX + 63) / 64;
  __syncthreads();
  const int lane = lane_id(), wv = wave_id();
  for (int k0 = 0; k0 < m; k0 += PB) {
    const int k1 = k0 + PB < m ? k0 + PB : m;
    const int nb = k1 - k0;
    if (wv == 0) {
      const int j = k0 + lane;
      const bool own = lane < nb;
      double blk[PB];
#pragma unroll
      for (int r = 0; r < PB; ++r) blk[r] = (own && r <= lane) ? S.L[P(j) + k0 + r] : 0.0;
      int ok = 1;
#pragma unroll
      for (int kk = 0; kk < PB; ++kk) {
        if (kk < nb) {
          const double akk = readlane_d(blk[kk], kk);
          if (!(akk > 0.0)) { ok = 0; break; }
          const double dk = sqrt(akk);
          const double u = blk[kk] / dk;
          blk[kk] = lane == kk ? dk : (lane > kk ? u : blk[kk]);
#pragma unroll
          for (int r = kk + 1; r < PB; ++r) {
            if (r < nb) {
              const double ur = readlane_d(u, r);
              if (lane >= r) blk[r] = blk[r] - ur * u;
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < PB; ++r) if (own && r <= lane) S.L[P(j) + k0 + r] = blk[r];
      if (lane == 0) S.bci[0] = ok;
    }
    __syncthreads();
    if (!S.bci[0]) { __syncthreads(); return 0; }
    // the panel rows of the columns past the block
    {
      const int j = k1 + (int)threadIdx.x;
      if (j < m) {
        double c[PB];
        double* col = S.L + P(j) + k0;
#pragma unroll
        for (int r = 0; r < PB; ++r) c[r] = r < nb ? col[r] : 0.0;
#pragma unroll
        for (int i = 0; i < PB; ++i) {
          if (i < nb) {
            const double* bc = S.L + P(k0 + i) + k0;   // block column k0 + i
#pragma unroll
            for (int k = 0; k < i; ++k) c[i] = c[i] - bc[k] * c[k];
            c[i] = c[i] / bc[i];
          }
        }
#pragma unroll
        for (int r = 0; r < PB; ++r) if (r < nb) col[r] = c[r];
      }
    }
    __syncthreads();
    // trailing update: a_ij -= sum_k u_ki u_kj over the panel, i in [k1, j]
    for (int q = 0; q < NT; ++q) {
      const int i = k1 + lane + 64 * q;
      if (64 * q >= m - k1) break;
      double ui[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) ui[k] = (i < m && k < nb) ? S.L[P(i) + k0 + k] : 0.0;
      for (int j = k1 + 64 * q + wv; j < m; j += 4) {
        const double* cj = S.L + P(j);
        double uj[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) uj[k] = k < nb ? cj[k0 + k] : 0.0;
        const int ic = i <= j ? i : j;
        double a = cj[ic];
#pragma unroll
        for (int k = 0; k < PB; ++k) if (k < nb) a = a - ui[k] * uj[k];
        if (i <= j) S.L[P(j) + i] = a;
      }
    }
    __syncthreads();
  }
  return 1;
}

// ------------------------------------------------------------- sum trees --
// Levels 0 and 1 in global memory, levels >= 2 in LDS; see the oracle's
// orc_sumtree for the definition (64-ary, Hillis-Steele group sums).
template <int KMAX>
__device__ __forceinline__ double tree_val(Ctx<KMAX>& C, int tr, int h, uint64_t i) {
  if (h < 2) {
    const double* base = tr == 0 ? C.A.tree_add : C.A.tree_rem;
    return base[C.D.lev_off[h] + i];
  }
  return C.S.tlev[tr][C.D.lev_off[h] - C.D.lev_off[2] + i];
}

// a double moved across lanes by a DPP control (both halves; lanes whose
// source is out of range get 0)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  int2 h = *reinterpret_cast<int2*>(&v);
  h.x = __builtin_amdgcn_update_dpp(0, h.x, CTRL, 0xf, 0xf, false);
  h.y = __builtin_amdgcn_update_dpp(0, h.y, CTRL, 0xf, 0xf, false);
  return *reinterpret_cast<double*>(&h);
}

// the value of the lane 16 below (lanes 16..63): rows 1 and 3 from a row
// swap, row 2 from a half swap followed by a row swap
__device__ __forceinline__ double from_row_below(double v) {
  const int row = (threadIdx.x & 63) >> 4;
  int2 h = *reinterpret_cast<int2*>(&v), o;
  {
    const auto a = __builtin_amdgcn_permlane32_swap((unsigned)h.x, (unsigned)h.x, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap((unsigned)a[0], (unsigned)h.x, false, false);
    o.x = (int)(row == 2 ? b[1] : b[0]);
  }
  {
    const auto a = __builtin_amdgcn_permlane32_swap((unsigned)h.y, (unsigned)h.y, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap((unsigned)a[0], (unsigned)h.y, false, false);
    o.y = (int)(row == 2 ? b[1] : b[0]);
  }
  return *reinterpret_cast<double*>(&o);
}

// the value of the lane 32 below (lanes 32..63)
__device__ __forceinline__ double from_half_below(double v) {
  int2 h = *reinterpret_cast<int2*>(&v), o;
  o.x = (int)__builtin_amdgcn_permlane32_swap((unsigned)h.x, (unsigned)h.x, false, false)[0];
  o.y = (int)__builtin_amdgcn_permlane32_swap((unsigned)h.y, (unsigned)h.y, false, false)[0];
  return *reinterpret_cast<double*>(&o);
}

// inclusive Hillis-Steele scan over the wave (offsets 1, 2, ..., 32, each
// lane adding the value `offset` lanes below).  Offset 1 is a wave shift
// (DPP), 16 and 32 are lane swaps, 2-8 cross-lane reads; every lane adds
// the same operands in the same order at every offset whichever move
// brings them, so the sums are the plain shuffle scan's bit for bit.
__device__ __forceinline__ double wave_scan(double v) {
  const int lane = threadIdx.x & 63;
  { const double t = dpp_d<0x138>(v); if (lane >= 1) v = v + t; }   // wave_shr:1
#pragma unroll
  for (int d = 2; d < 16; d <<= 1) {
    const double t = __shfl_up(v, d, 64);
    if (lane >= d) v = v + t;
  }
  { const double t = from_row_below(v); if (lane >= 16) v = v + t; }
  { const double t = from_half_below(v); if (lane >= 32) v = v + t; }
  return v;
}

// the last element of wave_scan(v), in every lane: lane 63's scan value is
// a balanced pairwise sum ((v62 + v63) + (v60 + v61)) + ..., which a
// butterfly over pairs, quads, eights, rows and halves forms with the same
// operands (partners by DPP quad permutations and mirrors, then lane
// swaps), so it is bit-identical without the scan's serial reads
__device__ __forceinline__ double group_total(double v) {
  v = v + dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
  v = v + dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
  v = v + dpp_d<0x141>(v);   // row_half_mirror
  v = v + dpp_d<0x140>(v);   // row_mirror
  {
    const bool odd = ((threadIdx.x & 63) >> 4) & 1;
    int2 h = *reinterpret_cast<int2*>(&v), o;
    const auto a = __builtin_amdgcn_permlane16_swap((unsigned)h.x, (unsigned)h.x, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap((unsigned)h.y, (unsigned)h.y, false, false);
    o.x = (int)(odd ? a[0] : a[1]);
    o.y = (int)(odd ? b[0] : b[1]);
    v = v + *reinterpret_cast<double*>(&o);
  }
  return readlane_d(v, 0) + readlane_d(v, 32);
}

// draw an index with probability w_i / total (wave 0 computes, all return)
template <int KMAX>
__device__ __forceinline__ uint64_t tree_pick(Ctx<KMAX>& C, int tr, double u) {
  if (wave_id() == 0) {
    const int lane = lane_id();
    double total = tr == 0 ? C.R.tot_add : C.R.tot_rem;
    double r = u * total, carry = 0.0;
    uint64_t g = 0;
    long long res = 0;
    for (int h = C.D.top; h >= 0; --h) {
      double c = tree_val(C, tr, h, 64 * g + lane);
      double s = wave_scan(c);
      unsigned long long m1 = __ballot(c > 0.0 && r < carry + s);
      unsigned long long nz = __ballot(c > 0.0);
      int pick = m1 ? (__ffsll((long long)m1) - 1) : (nz ? 63 - __clzll((long long)nz) : -1);
      if (pick < 0) { res = -1; break; }
      double sp = readlane_d(s, pick > 0 ? pick - 1 : 0);   // uniform lane: a scalar read
      if (pick > 0) carry = carry + sp;
      g = 64 * g + (uint64_t)pick;
      res = (long long)g;
    }
    if (lane == 0) C.S.bci[1] = res;
  }
  __syncthreads();
  long long r = C.S.bci[1];
  __syncthreads();
  if (r < 0) { C.R.err |= ERR_TREE; r = 0; }
  return (uint64_t)r;
}

// set leaf i to its weight (unzero) or 0 (zero) and update the path
template <int KMAX>
__device__ __forceinline__ void tree_set(Ctx<KMAX>& C, int tr, uint64_t i, int zero) {
  double* g0 = tr == 0 ? C.A.tree_add : C.A.tree_rem;
  const double* w = tr == 0 ? C.A.q_add : C.A.q_rem;
  double total;
  if (wave_id() == 0) {
    const int lane = lane_id();
    double nv = zero ? 0.0 : w[i];
    // issue the level-0 and level-1 group loads together
    uint64_t g1 = i >> 6;
    double c0 = g0[C.D.lev_off[0] + 64 * g1 + lane];
    double c1 = C.D.top >= 1 ? g0[C.D.lev_off[1] + 64 * (g1 >> 6) + lane] : 0.0;
    if (C.R.tj_on) {
      const double o0 = readlane_d(c0, (int)(i & 63)), o1 = readlane_d(c1, (int)(g1 & 63));
      if (lane == 0) {
        C.S.tj_old[C.R.tj_n][0] = o0; C.S.tj_old[C.R.tj_n][1] = o1;
        C.S.tj_ind[C.R.tj_n] = i | ((uint64_t)tr << 63);
      }
    }
    if (lane == (int)(i & 63)) c0 = nv;
    double gs = group_total(c0);
    if (lane == 0) g0[C.D.lev_off[0] + i] = nv;
    if (C.D.top == 0) {
      total = gs;
    } else {
      if (lane == 0) g0[C.D.lev_off[1] + g1] = gs;
      if (lane == (int)(g1 & 63)) c1 = gs;
      gs = group_total(c1);
      uint64_t gi = g1 >> 6;   // entry index at level 2
      for (int h = 2; h <= C.D.top; ++h) {
        double* lv = C.S.tlev[tr] + (C.D.lev_off[h] - C.D.lev_off[2]);
        if (lane == 0) lv[gi] = gs;
        uint64_t grp = gi >> 6;
        double c = lv[64 * grp + lane];
        if (lane == (int)(gi & 63)) c = gs;
        gs = group_total(c);
        gi = grp;
      }
      total = gs;
    }
    if (lane == 0) C.S.bcd[1] = total;
  }
  __syncthreads();
  total = C.S.bcd[1];
  __syncthreads();
  if (tr == 0) C.R.tot_add = total; else C.R.tot_rem = total;
  if (C.R.tj_on) C.R.tj_n++;
}

// Pick a leaf with probability w_i / total and zero it (a tree_pick then a
// tree_set(zero)), reusing the groups the descent read: the path's sums are
// recomputed from them in tree_set's order, so the tree ends up exactly as
// after the two calls.  *w_out: the picked leaf's weight.
template <int KMAX>
__device__ __forceinline__ uint64_t tree_pick_zero(Ctx<KMAX>& C, int tr, double u, double* w_out) {
  double* g0 = tr == 0 ? C.A.tree_add : C.A.tree_rem;
  if (wave_id() == 0) {
    const int lane = lane_id();
    double total = tr == 0 ? C.R.tot_add : C.R.tot_rem;
    double r = u * total, carry = 0.0;
    uint64_t g = 0;
    long long res = 0;
    double cl[BMG_TREE_LEV];             // this lane's value in the group read at each level
#pragma unroll
    for (int h = 0; h < BMG_TREE_LEV; ++h) cl[h] = 0.0;
    for (int h = C.D.top; h >= 0; --h) {
      double c = tree_val(C, tr, h, 64 * g + lane);
#pragma unroll
      for (int q = 0; q < BMG_TREE_LEV; ++q) if (q == h) cl[q] = c;
      double s = wave_scan(c);
      unsigned long long m1 = __ballot(c > 0.0 && r < carry + s);
      unsigned long long nz = __ballot(c > 0.0);
      int pick = m1 ? (__ffsll((long long)m1) - 1) : (nz ? 63 - __clzll((long long)nz) : -1);
      if (pick < 0) { res = -1; break; }
      double sp = readlane_d(s, pick > 0 ? pick - 1 : 0);   // uniform lane: a scalar read
      if (pick > 0) carry = carry + sp;
      g = 64 * g + (uint64_t)pick;
      res = (long long)g;
    }
    if (res >= 0) {
      const uint64_t i = (uint64_t)res;
      const uint64_t g1 = i >> 6;
      const double wv = readlane_d(cl[0], (int)(i & 63));
      if (C.R.tj_on) {
        const double o1 = C.D.top >= 1 ? readlane_d(cl[1], (int)(g1 & 63)) : 0.0;
        if (lane == 0) {
          C.S.tj_old[C.R.tj_n][0] = wv; C.S.tj_old[C.R.tj_n][1] = o1;
          C.S.tj_ind[C.R.tj_n] = i | ((uint64_t)tr << 63);
        }
      }
      // tree_set(zero) on the groups already read
      double c0 = cl[0];
      if (lane == (int)(i & 63)) c0 = 0.0;
      double gs = group_total(c0);
      if (lane == 0) g0[C.D.lev_off[0] + i] = 0.0;
      if (C.D.top >= 1) {
        if (lane == 0) g0[C.D.lev_off[1] + g1] = gs;
        double c1 = cl[1];
        if (lane == (int)(g1 & 63)) c1 = gs;
        gs = group_total(c1);
        uint64_t gi = g1 >> 6;
        for (int h = 2; h <= C.D.top; ++h) {
          double* lv = C.S.tlev[tr] + (C.D.lev_off[h] - C.D.lev_off[2]);
          if (lane == 0) lv[gi] = gs;
          double c = 0.0;
#pragma unroll
          for (int q = 2; q < BMG_TREE_LEV; ++q) if (q == h) c = cl[q];
          if (lane == (int)(gi & 63)) c = gs;
          gs = group_total(c);
          gi = gi >> 6;
        }
      }
      if (lane == 0) { C.S.bcd[1] = gs; C.S.bcd[2] = wv; }
    }
    if (lane == 0) C.S.bci[1] = res;
  }
  __syncthreads();
  long long r = C.S.bci[1];
  const double total = C.S.bcd[1], wv = C.S.bcd[2];
  __syncthreads();
  if (r < 0) { C.R.err |= ERR_TREE; *w_out = 1.0; return 0; }
  if (tr == 0) C.R.tot_add = total; else C.R.tot_rem = total;
  if (C.R.tj_on) C.R.tj_n++;
  *w_out = wv;
  return (uint64_t)r;
}

// start journalling the sum-tree writes of a move (all threads)
template <int KMAX>
__device__ __forceinline__ void journal_begin(Ctx<KMAX>& C) {
  const int nl = C.D.top >= 2 ? (int)(C.D.lev_off[C.D.top] + C.D.lev_len[C.D.top] - C.D.lev_off[2]) : 0;
  for (int e = threadIdx.x; e < nl; e += BMG_BLOCK) {
    C.S.tj_tlev[0][e] = C.S.tlev[0][e];
    C.S.tj_tlev[1][e] = C.S.tlev[1][e];
  }
  __syncthreads();
  C.R.tj_on = 1; C.R.tj_n = 0;
  C.R.tj_tot[0] = C.R.tot_add; C.R.tj_tot[1] = C.R.tot_rem;
}

// put the trees back as they were when the journal began (all threads):
// the LDS levels from their copy, the global entries from the journal in
// reverse order, the totals
template <int KMAX>
__device__ __forceinline__ void journal_restore(Ctx<KMAX>& C) {
  const int nl = C.D.top >= 2 ? (int)(C.D.lev_off[C.D.top] + C.D.lev_len[C.D.top] - C.D.lev_off[2]) : 0;
  for (int e = threadIdx.x; e < nl; e += BMG_BLOCK) {
    C.S.tlev[0][e] = C.S.tj_tlev[0][e];
    C.S.tlev[1][e] = C.S.tj_tlev[1][e];
  }
  if (threadIdx.x == 0) {
    for (int k = C.R.tj_n - 1; k >= 0; --k) {
      const uint64_t v = C.S.tj_ind[k];
      const uint64_t i = v & ~(1ull << 63);
      double* g0 = (v >> 63) ? C.A.tree_rem : C.A.tree_add;
      g0[C.D.lev_off[0] + i] = C.S.tj_old[k][0];
      if (C.D.top >= 1) g0[C.D.lev_off[1] + (i >> 6)] = C.S.tj_old[k][1];
    }
  }
  __syncthreads();
  C.R.tot_add = C.R.tj_tot[0]; C.R.tot_rem = C.R.tj_tot[1];
  C.R.tj_on = 0; C.R.tj_n = 0;
}

// -------------------------------------------------------------- model -----
// Column bookkeeping helpers operate on LDS arrays with all threads.
template <int KMAX>
__device__ __forceinline__ int alloc_slot(Ctx<KMAX>& C) {
  // lowest slot free in both the current and proposed models
  int s = -1;
  for (int q = 0; q < (KMAX + 63) / 64 && s < 0; ++q) {
    uint64_t used = C.S.slots_cur[q] | C.S.slots_prop[q];
    if (~used) s = 64 * q + __ffsll((long long)~used) - 1;
  }
  return s < KMAX ? s : -1;
}

template <int KMAX>
__device__ __forceinline__ void set_slot_bit(Ctx<KMAX>& C, int s, int on) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (on) C.S.slots_prop[s >> 6] |= (1ull << (s & 63));
    else C.S.slots_prop[s >> 6] &= ~(1ull << (s & 63));
  }
  __syncthreads();
}

// The residual term S = nu s2 + y'y - v'v of the log-likelihood, formed the
// reference's way, cancels when the factor is nearly singular (strongly
// correlated columns: v'v comes close to y'y and rounding can leave S at
// zero or below).  Where it has lost its significant digits (below 1e-8 of
// nu s2 + y'y; a fit that close never occurs otherwise) it is formed
// directly instead, as nu s2 + |y - X mu|^2 + mu'D mu with mu = U^-1 v,
// a sum of non-negative terms equal to it in exact arithmetic.
#define BMG_S_TINY 1e-8
template <int KMAX>
__device__ __noinline__ double direct_syx(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  const int k = C.R.cols;
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += BMG_BLOCK) S.work[i] = S.v[i];
  __syncthreads();
  if (wave_id() == 0) trsv_un_w0(S, S.work, k);
  __syncthreads();
  double o[8];
  xcoef(C, S.work, k, o, 0);
  double md = 0.0;
  for (int i = threadIdx.x; i < k; i += BMG_BLOCK) md += S.d[i] * S.work[i] * S.work[i];
  md = block_sum(md, S.red);
  return C.K.nus2 + o[7] + md;
}

template <int KMAX>
__device__ __forceinline__ void guard_syx(Ctx<KMAX>& C) {
  if (!(C.R.syx >= BMG_S_TINY * C.K.nus2_plus_yy)) C.R.syx = direct_syx(C);
}

// src/model.hpp:432-506 for one column (type A, H, D or R)
template <int KMAX>
__device__ __forceinline__ void update_on_add(Ctx<KMAX>& C, uint32_t ind, int model_ind, int coding,
                              int is_x_ind2, double dval, long long next = -1) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int col = R.cols;
  int slot = alloc_slot(C);
  // dots formed ahead by the previous addition of this move (same SNP,
  // same slot, every column since unchanged)
  const bool pre = R.pre_snp == (long long)ind && R.pre_slot == slot;
  R.pre_snp = -1;
  if (slot < 0 || col >= KMAX) {
    // current and proposed models together need more slots than compiled:
    // the proposal is rejected (log-likelihood -inf) and flagged
    R.err |= ERR_CAPACITY;
    R.ll = -INFINITY; R.syx = INFINITY;
    return;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (is_x_ind2) S.xi2[model_ind] = col;
    else { S.xi1[model_ind] = col; S.xi2[model_ind] = 0; }
    S.snp[col] = (int32_t)ind;
    S.xtype[col] = (int8_t)coding;
    S.slot[col] = slot;
    S.d[col] = dval;
  }
  __syncthreads();
  set_slot_bit(C, slot, 1);
  R.cols = col + 1;
  if (R.lmin > col) R.lmin = col;
  const uint64_t ta = wall_clock64();
  double xy = snp_dot_y(C, ind, coding);
  const uint64_t tb = wall_clock64();
  // the slot keeps the column's value words (coded genotype in each 2-bit
  // field), which the dot products and X beta read without the coding;
  // the dot products write them as they decode the column
  if (pre) {
    for (int c = threadIdx.x; c <= col; c += BMG_BLOCK) S.work[c] = (double)S.iacc2[c];
    __syncthreads();
  } else {
    new_column_dots(C, ind, coding, col, C.A.xcol + (size_t)slot * C.D.W32);
  }
  const uint64_t tc = wall_clock64();
  // record the new X^T X row/column in the slot matrix
  for (int c = threadIdx.x; c <= col; c += BMG_BLOCK) {
    int sc = c == col ? slot : S.slot[c];
    C.A.xxs[(size_t)slot * BMG_KMAXG + sc] = S.work[c];
    C.A.xxs[(size_t)sc * BMG_KMAXG + slot] = S.work[c];
  }
  if (threadIdx.x == 0) S.xy[col] = xy;
  // the next addition of the move (next >= 0): its column's dots with every
  // column up to this one are formed by waves 1-3 while wave 0 runs this
  // column's sweep (one effect type, no covariates); its slot is the one it
  // will be given, the lowest free once this column's is taken
  int nslot = -1;
  if (next >= 0 && !C.D.mixed && C.D.type != BMG_AH && C.D.me <= 1 && col + 1 < KMAX) nslot = alloc_slot(C);
  if (nslot >= 0)
    for (int c = threadIdx.x; c <= col + 1; c += BMG_BLOCK) S.iacc2[c] = 0;
  // bordered update (src/symmmatrix.cpp:144-164)
  for (int i = threadIdx.x; i < col; i += BMG_BLOCK) S.L[P(col) + i] = S.work[i];
  __syncthreads();
  if (wave_id() == 0) {
    trsv_ut_w0(S, S.L + P(col), col);
  } else if (nslot >= 0) {
    uint32_t* xn = C.A.xcol + (size_t)nslot * C.D.W32;
    switch (C.D.type) {
      case 0: column_dots<KMAX, 0, 64, BMG_BLOCK - 64>(C, (uint64_t)next, col + 1, xn, S.iacc2); break;
      case 1: column_dots<KMAX, 1, 64, BMG_BLOCK - 64>(C, (uint64_t)next, col + 1, xn, S.iacc2); break;
      case 2: column_dots<KMAX, 2, 64, BMG_BLOCK - 64>(C, (uint64_t)next, col + 1, xn, S.iacc2); break;
      default: column_dots<KMAX, 3, 64, BMG_BLOCK - 64>(C, (uint64_t)next, col + 1, xn, S.iacc2); break;
    }
  }
  __syncthreads();
  if (nslot >= 0) { R.pre_snp = next; R.pre_slot = nslot; }
  const uint64_t td = wall_clock64();
  double ll2 = 0.0, lv = 0.0;
  for (int i = threadIdx.x; i < col; i += BMG_BLOCK) {
    double li = S.L[P(col) + i];
    ll2 += li * li;
    lv += li * S.v[i];
  }
  ll2 = block_sum(ll2, S.red);
  lv = block_sum(lv, S.red);
  double lmm2 = S.work[col] + dval - ll2;
  if (lmm2 <= 0) {
    R.ll = -INFINITY; R.syx = INFINITY;
    R.mu_computed = 0;
    __syncthreads();
    return;
  }
  double lcc = sqrt(lmm2);
  double vc = (xy - lv) / lcc;
  __syncthreads();
  if (threadIdx.x == 0) { S.L[P(col) + col] = lcc; S.v[col] = vc; }
  __syncthreads();
  R.syx = R.syx - vc * vc;
  guard_syx(C);
  R.ldq += 0.5 * log(dval);
  R.ldl += log(lcc);
  R.ll = finite_ll(R.ldq - R.ldl + C.K.minus_n_plus_nu_2 * log(R.syx));
  if (threadIdx.x == 0 && C.prof) {
    const uint64_t te = wall_clock64();
    C.prof[8] += tb - ta; C.prof[9] += tc - tb; C.prof[10] += td - tc; C.prof[11] += te - td;
    C.prof[12] += 1;
  }
}

// src/model.hpp:239-268
template <int KMAX>
__device__ __forceinline__ void add_term(Ctx<KMAX>& C, uint32_t ind, int t_ind, const double* d2,
                                         long long next = -1) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int type = C.D.types[t_ind];
  int is_ah = type == 4;
  if (R.cols + 1 + is_ah > KMAX) {
    R.err |= ERR_CAPACITY;
    R.ll = -INFINITY; R.syx = INFINITY;
    return;
  }
  int model_ind = R.nloci;
  __syncthreads();
  if (threadIdx.x == 0) { S.loci[model_ind] = ind; S.tind[model_ind] = (int8_t)t_ind; }
  __syncthreads();
  R.nloci++;
  put(R.Ns, type, sel(R.Ns, type) + 1);
  R.mu_computed = 0;
  if (is_ah) {
    update_on_add(C, ind, model_ind, 0, 0, d2[0]);
    if (isfinite(R.syx)) update_on_add(C, ind, model_ind, 1, 1, d2[1]);
    else if (R.cols < KMAX && alloc_slot(C) >= 0) {
      // keep the bookkeeping of the second column (not factorised)
      int col = R.cols, slot = alloc_slot(C);
      __syncthreads();
      if (threadIdx.x == 0) {
        S.xi2[model_ind] = col; S.snp[col] = (int32_t)ind; S.xtype[col] = 1;
        S.slot[col] = slot; S.d[col] = d2[1];
      }
      __syncthreads();
      set_slot_bit(C, slot, 1);
      R.cols = col + 1;
    }
  } else {
    update_on_add(C, ind, model_ind, type, 0, d2[0], next);
  }
}

// src/model.hpp:508-556
template <int KMAX>
__device__ __forceinline__ void update_on_remove_impl(Ctx<KMAX>& C, int x_ind, long long next);

template <int KMAX>
__device__ __forceinline__ void update_on_remove(Ctx<KMAX>& C, int x_ind, long long next = -1) {
  const uint64_t t0 = wall_clock64();
  update_on_remove_impl(C, x_ind, next);
  if (threadIdx.x == 0 && C.prof) { C.prof[13] += 1; C.prof[14] += wall_clock64() - t0; }
}

template <int KMAX>
__device__ __forceinline__ void update_on_remove_impl(Ctx<KMAX>& C, int x_ind, long long next) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int k = R.cols;
  double inv_val = S.d[x_ind];
  int slot_rm = S.slot[x_ind];
  // shift column arrays left (two-phase)
  double a0 = 0, a1 = 0; int i0 = 0, i1 = 0; int8_t t0 = 0;
  int t = threadIdx.x;
  bool mv = t >= x_ind && t < k - 1;
  if (mv) { a0 = S.xy[t + 1]; a1 = S.d[t + 1]; i0 = S.slot[t + 1]; i1 = S.snp[t + 1]; t0 = S.xtype[t + 1]; }
  __syncthreads();
  if (mv) { S.xy[t] = a0; S.d[t] = a1; S.slot[t] = i0; S.snp[t] = i1; S.xtype[t] = t0; }
  __syncthreads();
  set_slot_bit(C, slot_rm, 0);
  R.cols = k - 1;
  if (R.lmin > x_ind) R.lmin = x_ind;
  if (x_ind == R.cols) {
    double lxx = S.L[P(x_ind) + x_ind];
    double vx = S.v[x_ind];
    R.ldl -= log(lxx);
    R.syx += vx * vx;
  } else {
    // the move's first addition after this removal (next >= 0): its column
    // dots with the remaining columns are formed by waves 1-3 during the
    // downdate, into the slot it will be given (as during a forward sweep)
    int nslot = -1;
    if (next >= 0 && !C.D.mixed && C.D.type != BMG_AH && C.D.me <= 1 && R.cols < KMAX)
      nslot = alloc_slot(C);
    if (nslot >= 0) {
      for (int c = threadIdx.x; c <= R.cols; c += BMG_BLOCK) S.iacc2[c] = 0;
      __syncthreads();
    }
    auto ahead = [&]() {
      if (nslot < 0) return;
      uint32_t* xn = C.A.xcol + (size_t)nslot * C.D.W32;
      switch (C.D.type) {
        case 0: column_dots<KMAX, 0, 64, BMG_BLOCK - 64>(C, (uint64_t)next, R.cols, xn, S.iacc2); break;
        case 1: column_dots<KMAX, 1, 64, BMG_BLOCK - 64>(C, (uint64_t)next, R.cols, xn, S.iacc2); break;
        case 2: column_dots<KMAX, 2, 64, BMG_BLOCK - 64>(C, (uint64_t)next, R.cols, xn, S.iacc2); break;
        default: column_dots<KMAX, 3, 64, BMG_BLOCK - 64>(C, (uint64_t)next, R.cols, xn, S.iacc2); break;
      }
    };
    const uint64_t ta = wall_clock64();
    // v follows the downdate's rotations (no re-solve)
    chol_downdate(S, k, x_ind, ahead);
    if (nslot >= 0) { R.pre_snp = next; R.pre_slot = nslot; }
    const uint64_t tb = wall_clock64();
    if (threadIdx.x == 0 && C.prof) { C.prof[16] += tb - ta; C.prof[17] += wall_clock64() - tb; }
    double ld = 0.0, vv = 0.0;
    for (int i = threadIdx.x; i < R.cols; i += BMG_BLOCK) {
      ld += log(S.L[P(i) + i]);
      vv += S.v[i] * S.v[i];
    }
    R.ldl = block_sum(ld, S.red);
    vv = block_sum(vv, S.red);
    R.syx = C.K.nus2_plus_yy - vv;
    guard_syx(C);
  }
  R.ldq -= 0.5 * log(inv_val);
  R.ll = finite_ll(R.ldq - R.ldl + C.K.minus_n_plus_nu_2 * log(R.syx));
}

// src/model.hpp:270-312
template <int KMAX>
__device__ __forceinline__ void remove_term(Ctx<KMAX>& C, int model_ind, long long next = -1) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int type = C.D.types[S.tind[model_ind]];
  int xi1 = S.xi1[model_ind], xi2 = S.xi2[model_ind];
  uint32_t ind = S.loci[model_ind];
  put(R.Ns, type, sel(R.Ns, type) - 1);
  int nl = R.nloci;
  int xmove = type == 4 ? 2 : 1;
  int t = threadIdx.x;
  bool mv = t >= model_ind && t < nl - 1;
  uint32_t l0 = 0; int8_t ti = 0; int a1 = 0, a2 = 0;
  if (mv) { l0 = S.loci[t + 1]; ti = S.tind[t + 1]; a1 = S.xi1[t + 1] - xmove; a2 = S.xi2[t + 1] - xmove; }
  __syncthreads();
  if (mv) { S.loci[t] = l0; S.tind[t] = ti; S.xi1[t] = a1; S.xi2[t] = a2; }
  if (threadIdx.x == 0) S.rm_list[R.n_rm] = ind;
  __syncthreads();
  R.n_rm++;
  R.nloci = nl - 1;
  if (type == 4) {
    if (xi1 > xi2) { update_on_remove(C, xi1); update_on_remove(C, xi2); }
    else { update_on_remove(C, xi2); update_on_remove(C, xi1); }
  } else {
    update_on_remove(C, xi1, next);
  }
  R.mu_computed = 0;
}

// index of SNP `ind` in the proposed model, or -1 (all threads)
template <int KMAX>
__device__ __forceinline__ int find_locus(Ctx<KMAX>& C, uint64_t ind) {
  __syncthreads();
  if (threadIdx.x == 0) C.S.bci[2] = -1;
  __syncthreads();
  for (int t = threadIdx.x; t < C.R.nloci; t += BMG_BLOCK)
    if (C.S.loci[t] == (uint32_t)ind) C.S.bci[2] = t;
  __syncthreads();
  int r = (int)C.S.bci[2];
  __syncthreads();
  return r;
}

// src/model.hpp:199-237 on the proposed model
template <int KMAX>
__device__ __forceinline__ void compute_log_likelihood(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int k = R.cols;
  if (k < 1) { R.ll = NAN; return; }
  // X'X + D from the slot matrix into the packed factor: a thread per
  // column j, its rows i <= j read eight at a time (row slot(j) of the slot
  // matrix holds column j's entries), so the loads are in flight together
  if ((int)threadIdx.x < k) {
    const int j = (int)threadIdx.x;
    const double* __restrict__ row = C.A.xxs + (size_t)S.slot[j] * BMG_KMAXG;
    double* col = S.L + P(j);
    for (int i0 = 0; i0 <= j; i0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = row[S.slot[i0 + u <= j ? i0 + u : j]];   // past j: a valid slot
      const double dj = S.d[j];
      // unconditional stores: a row past j is clamped to j and rewrites the
      // diagonal with the value row j gives it
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u <= j ? i0 + u : j;
        col[i] = i == j ? v[u] + dj : v[u];
      }
    }
  }
  double ls = 0.0;
  for (int i = 1 + threadIdx.x; i < k; i += BMG_BLOCK) ls += log(S.d[i]);
  ls = block_sum(ls, S.red);
  R.ldq = 0.5 * ls;
  R.lmin = 0;
  if (!chol_full(S, k)) {
    R.ll = -INFINITY; R.syx = INFINITY; R.mu_computed = 0;
    return;
  }
  for (int i = threadIdx.x; i < k; i += BMG_BLOCK) S.v[i] = S.xy[i];
  __syncthreads();
  if (wave_id() == 0) trsv_ut_w0(S, S.v, k);
  __syncthreads();
  double ld = 0.0, vv = 0.0;
  for (int i = threadIdx.x; i < k; i += BMG_BLOCK) {
    ld += log(S.L[P(i) + i]);
    vv += S.v[i] * S.v[i];
  }
  R.ldl = block_sum(ld, S.red);
  vv = block_sum(vv, S.red);
  R.syx = C.K.nus2_plus_yy - vv;
  guard_syx(C);
  R.ll = finite_ll(R.ldq - R.ldl + C.K.minus_n_plus_nu_2 * log(R.syx));
  R.mu_computed = 0;
}

// --------------------------------------------------- commit / restore -----
// Accept (*current = *new) or reject (*new = *current) for the single
// working copy: only factor columns >= lmin differ from the stored current
// factor.

// dst[e] = src[e] for e in [e0, e1) (all threads): eight elements a thread
// are read before any is written, so the reads of one pass are in flight
// together instead of one round trip per element
__device__ __forceinline__ void copy_factor(double* __restrict__ dst, const double* __restrict__ src, int e0,
                                            int e1) {
  for (int base = e0; base < e1; base += BMG_BLOCK * 8) {
    double t[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = base + q * BMG_BLOCK + (int)threadIdx.x;
      t[q] = e < e1 ? src[e] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = base + q * BMG_BLOCK + (int)threadIdx.x;
      if (e < e1) dst[e] = t[q];
    }
  }
}

template <int KMAX>
__device__ __forceinline__ void commit(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  __syncthreads();
  copy_factor(C.A.Lcur, S.L, P(R.lmin < R.cols ? R.lmin : R.cols), P(R.cols));
  for (int i = threadIdx.x; i < R.cols; i += BMG_BLOCK) {
    S.c_xy[i] = S.xy[i]; S.c_v[i] = S.v[i]; S.c_d[i] = S.d[i];
    S.c_slot[i] = S.slot[i]; S.c_snp[i] = S.snp[i]; S.c_xtype[i] = S.xtype[i];
  }
  for (int i = threadIdx.x; i < R.nloci; i += BMG_BLOCK) {
    S.c_loci[i] = S.loci[i]; S.c_xi1[i] = S.xi1[i]; S.c_xi2[i] = S.xi2[i]; S.c_tind[i] = S.tind[i];
  }
  // current model_inds: removed loci first, then the positions of all loci
  for (int i = threadIdx.x; i < R.n_rm; i += BMG_BLOCK) C.A.model_inds[S.rm_list[i]] = -1;
  __syncthreads();
  for (int i = threadIdx.x; i < R.nloci; i += BMG_BLOCK) C.A.model_inds[S.loci[i]] = i;
  if (threadIdx.x < (KMAX + 63) / 64) S.slots_cur[threadIdx.x] = S.slots_prop[threadIdx.x];
  __syncthreads();
  R.c_cols = R.cols; R.c_nloci = R.nloci;
  for (int q = 0; q < 5; ++q) R.c_Ns[q] = R.Ns[q];
  R.c_syx = R.syx; R.c_ldq = R.ldq; R.c_ldl = R.ldl; R.c_ll = R.ll;
  R.lmin = KMAX; R.n_rm = 0;
  R.mu_computed = 0;
}

template <int KMAX>
__device__ __forceinline__ void restore(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  __syncthreads();
  int lm = R.lmin < R.c_cols ? R.lmin : R.c_cols;
  copy_factor(S.L, C.A.Lcur, P(lm), P(R.c_cols));
  for (int i = threadIdx.x; i < R.c_cols; i += BMG_BLOCK) {
    S.xy[i] = S.c_xy[i]; S.v[i] = S.c_v[i]; S.d[i] = S.c_d[i];
    S.slot[i] = S.c_slot[i]; S.snp[i] = S.c_snp[i]; S.xtype[i] = S.c_xtype[i];
  }
  for (int i = threadIdx.x; i < R.c_nloci; i += BMG_BLOCK) {
    S.loci[i] = S.c_loci[i]; S.xi1[i] = S.c_xi1[i]; S.xi2[i] = S.c_xi2[i]; S.tind[i] = S.c_tind[i];
  }
  if (threadIdx.x < (KMAX + 63) / 64) S.slots_prop[threadIdx.x] = S.slots_cur[threadIdx.x];
  __syncthreads();
  R.cols = R.c_cols; R.nloci = R.c_nloci;
  for (int q = 0; q < 5; ++q) R.Ns[q] = R.c_Ns[q];
  R.syx = R.c_syx; R.ldq = R.c_ldq; R.ldl = R.c_ldl; R.ll = R.c_ll;
  R.lmin = KMAX; R.n_rm = 0;
}

// --------------------------------------------------------------- prior ----
// src/prior.hpp:144-158 (single effect type, Ns[type] == n_loci)
__device__ __forceinline__ double prior_add_untruncated(const ChainConst& K, const int (&Ns)[5], int n_old,
                                                        int type) {
  return (log(K.g_a + n_old) - log(K.g_b + K.m_g - n_old - 1)) +
         (log(K.tp_type + sel(Ns, type)) - log(K.tp_sum + n_old));
}
// with a model-size limit the prior of larger models is 0
__device__ __forceinline__ double prior_add(const ChainConst& K, const int (&Ns)[5], int n_old, int type) {
  if (n_old >= K.max_size) return -INFINITY;
  return prior_add_untruncated(K, Ns, n_old, type);
}
__device__ __forceinline__ double prior_rem(const ChainConst& K, const int (&Ns)[5], int n_old, int type) {
  return (log(K.g_b + K.m_g - n_old) - log(K.g_a + n_old - 1)) +
         (log(K.tp_sum + n_old - 1) - log(K.tp_type + sel(Ns, type) - 1));
}

template <int KMAX>
__device__ __forceinline__ double draw_inv_tau2(Ctx<KMAX>& C, int term) {
  if (C.K.use_ind)
    return 1.0 / (C.R.alpha2 * rng_sinvchi2(C.R.rng, C.K.nu_tau2[term], C.K.s2_tau2[term]));
  return sel(C.R.ita2, term);
}

template <int KMAX>
__device__ __forceinline__ void prepare_add_new_term(Ctx<KMAX>& C, double* d2) {
  int type = C.D.type;
  if (type == 4) { d2[0] = draw_inv_tau2(C, 0); d2[1] = draw_inv_tau2(C, 1); }
  else d2[0] = draw_inv_tau2(C, type);
}

// ------------------------------------------------- beta / sigma2 / pve ----
template <int KMAX>
__device__ __forceinline__ void compute_mu_beta(Ctx<KMAX>& C) {
  if (!C.R.mu_computed) {
    __syncthreads();
    for (int i = threadIdx.x; i < C.R.cols; i += BMG_BLOCK) C.S.mu[i] = C.S.v[i];
    __syncthreads();
    if (wave_id() == 0) trsv_un_w0(C.S, C.S.mu, C.R.cols);
    __syncthreads();
    C.R.mu_computed = 1;
  }
}

// counter-based access to draw `k` of the main stream from its current
// position (the stream itself is sequential: draw k uses counter base + k)
__device__ __forceinline__ double normal_at(const Rng& r, uint32_t k) {
  Rng q = r;
  uint64_t c = ((uint64_t)q.c1 << 32 | q.c0) + k;
  q.c0 = (uint32_t)c; q.c1 = (uint32_t)(c >> 32);
  return rng_normal(q);
}
__device__ __forceinline__ void rng_skip(Rng& r, uint32_t k) {
  uint64_t c = ((uint64_t)r.c1 << 32 | r.c0) + k;
  r.c0 = (uint32_t)c; r.c1 = (uint32_t)(c >> 32);
}

// src/model.hpp:326-342 with src/rand.hpp:61-78,144-168
template <int KMAX>
__device__ __forceinline__ void sample_beta_sigma2(Ctx<KMAX>& C) {
  Reg& R = C.R;
  double nu = C.K.n_plus_nu;
  double s2 = R.syx / C.K.n_plus_nu;
  double val = -1.0;
  for (int trial = 0; trial < BMG_MAX_TRIALS && (val <= 0 || !isfinite(val)); ++trial)
    val = nu * s2 / (2.0 * rng_gamma(R.rng, 0.5 * nu));
  if (!(val > 0) || !isfinite(val)) R.err |= ERR_SIGMA2;
  R.sigma2 = val;
  compute_mu_beta(C);
  R.beta_len = R.cols;
  for (int i = threadIdx.x; i < R.cols; i += BMG_BLOCK) C.S.beta[i] = normal_at(R.rng, i);
  rng_skip(R.rng, R.cols);
  __syncthreads();
  if (wave_id() == 0) trsv_un_w0(C.S, C.S.beta, R.beta_len);
  __syncthreads();
  double sc = sqrt(R.sigma2);
  for (int i = threadIdx.x; i < R.beta_len; i += BMG_BLOCK)
    C.S.beta[i] = sc * C.S.beta[i] + C.S.mu[i];
  __syncthreads();
}

template <int KMAX>
__device__ __forceinline__ void xbeta_impl(Ctx<KMAX>& C, const double* coef, int L, double* out8,
                                           int write_yhat) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const double b0 = coef[0];
  const int me = C.D.me;
  const uint32_t n = C.D.n, W = C.D.W32;
  const uint32_t* __restrict__ xcol = C.A.xcol;
  const double* __restrict__ y = C.D.y;
  double* __restrict__ yhat = C.A.y_hat;
  double sxb = 0, sxb2 = 0, sy = 0, sy2 = 0, syx = 0, se = 0, se2 = 0, srr = 0;
  for (uint32_t w = threadIdx.x; w * 16 < n; w += BMG_BLOCK) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int c0 = me; c0 < L; c0 += 16) {
      uint32_t wd[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int c = c0 + u;
        wd[u] = (c < L && coef[c] != 0.0) ? xcol[(size_t)S.slot[c] * W + w] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int c = c0 + u;
        if (c < L) {
          const double t = coef[c];
          if (t != 0.0) {
            // t * value is exact (value 0, 1 or 2), so the fused form
            // rounds the sum as the multiply-then-add loop does
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = fma(t, (double)((wd[u] >> (2 * q)) & 3u), acc[q]);
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const uint32_t i = w * 16 + q;
      if (i < n) {
        const double a = acc[q];
        sxb += a; sxb2 += a * a;
        // E beta_E of this individual in column order (the constant first)
        double e = b0;
        for (int c = 1; c < me; ++c) e = e + C.D.E[(size_t)(c - 1) * C.D.n_pad + i] * coef[c];
        se += e; se2 += e * e;
        // with covariates and no loci the reference adds E beta_E to the
        // y_hat it holds (src/model.hpp:369-373 leaves the genetic part
        // unset), and so does this
        const double yh = R.nloci > 0 ? a + e : (me > 1 && write_yhat ? yhat[i] + e : e);
        sy += yh; sy2 += yh * yh;
        syx += (e - y[i]) * a;
        const double ri = y[i] - (R.nloci > 0 ? a + e : e);
        srr += ri * ri;
        if (write_yhat) yhat[i] = yh;
      }
    }
  }
  out8[0] = block_sum(sxb, S.red);
  out8[1] = block_sum(sxb2, S.red);
  out8[2] = block_sum(sy, S.red);
  out8[3] = block_sum(sy2, S.red);
  out8[4] = block_sum(syx, S.red);
  out8[5] = block_sum(se, S.red);
  out8[6] = block_sum(se2, S.red);
  out8[7] = block_sum(srr, S.red);
}

// y_hat = X_g beta_g + E beta_E into the chain's y_hat (write_yhat) and
// out7: sum xb, sum xb^2, sum y_hat, sum y_hat^2, sum (e - y) xb, sum e,
// sum e^2, with xb = X_g beta_g and e = E beta_E (the constant and the
// covariates).  One packed word (16 individuals) per lane: the columns'
// words are fetched sixteen at a time (coalesced across lanes, all in
// flight together) and each individual's sum is formed in model-column
// order, as the per-individual loop does.
template <int KMAX>
__device__ __forceinline__ void xcoef(Ctx<KMAX>& C, const double* coef, int L, double* out8, int write_yhat) {
  xbeta_impl<KMAX>(C, coef, L, out8, write_yhat);
}

template <int KMAX>
__device__ __forceinline__ void xbeta(Ctx<KMAX>& C, double* out8, int write_yhat) {
  xcoef(C, C.S.beta, C.R.beta_len, out8, write_yhat);
}

// src/model.hpp:345-392: shares of var(E beta_E) (covariates beyond the
// constant), var(X_g beta_g) and var(y_hat) in var(y_hat) + sigma2
template <int KMAX>
__device__ __forceinline__ void compute_pve(Ctx<KMAX>& C, double* pves) {
  double s[8];
  xbeta(C, s, 1);
  const double n = (double)C.D.n;
  const bool ce = C.D.me > 1, cg = C.R.nloci > 0;
  pves[2] = ce ? (s[6] - s[5] * s[5] / n) / (n - 1) : 0.0;
  pves[1] = cg ? (s[1] - s[0] * s[0] / n) / (n - 1) : 0.0;
  if (!ce && !cg) { pves[0] = pves[1] = pves[2] = 0.0; return; }
  pves[0] = !ce ? pves[1] : (!cg ? pves[2] : (s[3] - s[2] * s[2] / n) / (n - 1));
  const double z = pves[0] + C.R.sigma2;
  pves[0] /= z; pves[1] /= z; pves[2] /= z;
}

// src/prior.cpp:30-141 on the working model
template <int KMAX>
__device__ __forceinline__ void sample_alpha_and_tau2(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const ChainConst& K = C.K;
  // sample_tau2
  double a2s2 = R.alpha2 * R.sigma2;
  if (K.use_ind) {
    // one draw per term, in term order on the chain's stream, computed a
    // term per thread: a thread starts its term where the stream would be
    // if every earlier term of the round took one trial (two Philox
    // blocks); the first term after which that does not hold ends the
    // round, and the next round restarts from the stream position it left.
    // The draws are the sequential loop's, value for value.
    uint64_t base = (uint64_t)R.rng.c0 | ((uint64_t)R.rng.c1 << 32);
    int first = C.D.me;   // the covariates' precisions are fixed
    while (first < R.beta_len) {
      const int i = first + (int)threadIdx.x;
      long long stop = R.beta_len;     // first term whose stream use differs
      uint64_t next = 0;
      if (i < R.beta_len) {
        const uint64_t start = base + 2ull * (uint64_t)threadIdx.x;
        Rng r = R.rng;
        r.c0 = (uint32_t)start; r.c1 = (uint32_t)(start >> 32);
        const int t = S.xtype[i];
        const double nu = K.nu_tau2[t] + 1.0;
        const double s2 = (K.nus2_tau2[t] + S.beta[i] * S.beta[i] / a2s2) / nu;
        const double dv = 1.0 / (R.alpha2 * rng_sinvchi2(r, nu, s2));
        next = (uint64_t)r.c0 | ((uint64_t)r.c1 << 32);
        if (next - start != 2) stop = i;
        S.work[i] = dv;
        if (i == first + BMG_BLOCK - 1 && stop == R.beta_len) stop = i;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const long long w = __shfl_xor(stop, o, 64);
        stop = w < stop ? w : stop;
      }
      __syncthreads();
      if (lane_id() == 0) S.redi[wave_id()] = stop;
      __syncthreads();
      stop = S.redi[0];
      for (int w = 1; w < BMG_BLOCK / 64; ++w) stop = S.redi[w] < stop ? S.redi[w] : stop;
      // terms first..stop are final; the stream continues after term stop
      if (i <= stop && i < R.beta_len) S.d[i] = S.work[i];
      if (i == stop) S.bci[3] = (long long)next;
      __syncthreads();
      if (stop >= R.beta_len) {
        base += 2ull * (uint64_t)(R.beta_len - first);
        first = R.beta_len;
      } else {
        base = (uint64_t)S.bci[3];
        first = (int)stop + 1;
      }
      __syncthreads();
    }
    R.rng.c0 = (uint32_t)base; R.rng.c1 = (uint32_t)(base >> 32);
  } else {
    double b2[4] = {0, 0, 0, 0}; double mt[4] = {0, 0, 0, 0};
    for (int i = C.D.me; i < R.beta_len; ++i) {
      const int xt = S.xtype[i];
      put(b2, xt, sel(b2, xt) + S.beta[i] * S.beta[i]);
      put(mt, xt, sel(mt, xt) + 1.0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (!K.allow_terms[t]) continue;
      double nu = K.nu_tau2[t] + mt[t];
      double s2 = (K.nus2_tau2[t] + b2[t] / a2s2) / nu;
      R.ita2[t] = 1.0 / (R.alpha2 * rng_sinvchi2(R.rng, nu, s2));
    }
    __syncthreads();
    for (int i = C.D.me + threadIdx.x; i < R.beta_len; i += BMG_BLOCK) S.d[i] = sel(R.ita2, (int)S.xtype[i]);
  }
  R.mu_computed = 0;
  __syncthreads();
  // sample_alpha
  int nterms = R.beta_len - C.D.me;
  double a = R.alpha;
  double xbxb = 0, ebxb = 0;
  if (nterms > 0) {
    // x_b = X_g beta_g, e_b = E beta_E - y
    double s[8];
    xbeta(C, s, 0);
    ebxb = s[4];
    xbxb = s[1];
  }
  int trials = 0;
  do {
    if (++trials > BMG_MAX_TRIALS) { R.err |= ERR_ALPHA; a = R.alpha; break; }
    if (nterms > 0) {
      if ((a <= 0.0) || log(rng_u01(R.rng)) <= -2.0 * K.mu_alpha * a) a = -a;
      double as2 = a * R.sigma2;
      double var = 1.0 / (1.0 + xbxb / (a * as2));
      double mu = var * (K.mu_alpha - ebxb / as2);
      a = sqrt(var) * rng_normal(R.rng) + mu;
    } else {
      a = rng_normal(R.rng) + K.mu_alpha;
    }
  } while (a == 0 || a * a == 0);
  // set_alpha
  double old = R.alpha2;
  R.alpha = a; R.alpha2 = a * a;
  __syncthreads();
  if (K.use_ind) {
    for (int i = C.D.me + threadIdx.x; i < R.beta_len; i += BMG_BLOCK) S.d[i] = S.d[i] * old / R.alpha2;
  } else {
#pragma unroll
    for (int t = 0; t < 4; ++t) if (K.allow_terms[t]) R.ita2[t] = R.ita2[t] * old / R.alpha2;
    for (int i = C.D.me + threadIdx.x; i < R.beta_len; i += BMG_BLOCK) S.d[i] = sel(R.ita2, (int)S.xtype[i]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------- moves ---
__device__ __forceinline__ int sample_naive(const double* cumsum, int m, double u) {
  double r = u * cumsum[m - 1];
  for (int i = 0; i < m; ++i) if (r < cumsum[i]) return i;
  return m - 1;
}

template <int KMAX>
__device__ __forceinline__ void undo_dd(Ctx<KMAX>& C, int ms) {
  for (int i = 0; i < ms; ++i) {
    uint64_t ind = C.S.mi[i];
    if (C.S.mi_isadd[i]) { tree_set(C, 0, ind, 0); tree_set(C, 1, ind, 1); }
    else { tree_set(C, 0, ind, 1); tree_set(C, 1, ind, 0); }
  }
}

// ------------------------------------------------ delayed rejection -----
// ExhModel (src/model.hpp:585-844) on the proposed model once every moved SNP
// is in it: loci from position `cl` on are the ones the second stage toggles,
// and their columns c0..cols-1 form the trailing block.  The walk's adjacent
// swaps only permute the rows above c0 of those columns, and its
// log-likelihoods read only the block's diagonal, v and d, so the block (in
// indices relative to c0) is all that is copied; the working model is left
// as it is.  Wave 0 walks the 2^ms sub-models: the scalar state is
// replicated in its lanes, the swaps give each lane one row or column.
// With one effect type the prior terms depend on the model size only and
// are tabulated; log(d) is kept beside d and permuted with it.  Both give
// the values the reference's expressions give.
struct Exh {
  int ms;            // loci of the block in the walk's model (model_size - cl)
  int vs;            // columns of the block in it (v_size - c0)
  int nT;            // columns of the block
  double syx, ldq, ldl, lmp;
};

template <int KMAX>
__device__ __forceinline__ double& exl(Smem<KMAX>& S, int r, int c) { return S.ex_L[P(c) + r]; }

// orders wave 0's LDS accesses across its lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// src/model.hpp:673-701
template <int KMAX>
__device__ __forceinline__ double exh_add(Smem<KMAX>& S, const ChainConst& K, Exh& X) {
  const int ncol = S.ex_t[X.ms] == 4 ? 2 : 1;
  for (int q = 0; q < ncol; ++q) {
    const int s = X.vs;
    X.syx -= S.ex_v[s] * S.ex_v[s];
    X.ldq += 0.5 * S.ex_ld[s];
    X.ldl += log(exl(S, s, s));
    X.vs = s + 1;
  }
  const double ll = X.ldq - X.ldl + K.minus_n_plus_nu_2 * log(X.syx);
  X.lmp += S.ex_pa[X.ms];
  ++X.ms;
  return ll + X.lmp;
}

// src/symmmatrix.cpp:220-265 on the block (columns s, s+1 relative to c0):
// lane r swaps row r < s, lane i - s - 1 rotates column i > s
template <int KMAX>
__device__ __forceinline__ void exh_swapadj(Smem<KMAX>& S, int nT, int s) {
  const int lane = lane_id();
  if (lane < s) {
    const double t = exl(S, lane, s);
    exl(S, lane, s) = exl(S, lane, s + 1);
    exl(S, lane, s + 1) = t;
  }
  double a = exl(S, s, s + 1), b = exl(S, s + 1, s + 1);
  const double dcc = exl(S, s, s);
  const double v1 = S.ex_v[s], v2 = S.ex_v[s + 1];
  wave_sync();
  double c, sn;
  drotg(a, b, c, sn);
  if (a < 0) { a = -a; sn = -sn; c = -c; }
  // column s+1 holds (dcc, 0) in rows s, s+1 after the swap
  const int i = s + 1 + lane;
  if (i < nT) {
    const double x1 = i == s + 1 ? dcc : exl(S, s, i);
    const double x2 = i == s + 1 ? 0.0 : exl(S, s + 1, i);
    const double tmp = x1 * sn - x2 * c;
    exl(S, s, i) = x2 * sn + x1 * c;
    exl(S, s + 1, i) = tmp;
  }
  if (lane == 0) {
    exl(S, s, s) = a;
    const double tmp = v1 * sn - v2 * c;
    S.ex_v[s] = v2 * sn + v1 * c;
    S.ex_v[s + 1] = tmp;
  }
  wave_sync();
}

// d and log(d) of two columns trade places (every lane does the same)
template <int KMAX>
__device__ __forceinline__ void exh_swap_d(Smem<KMAX>& S, int a, int b) {
  const double da = S.ex_d[a], db = S.ex_d[b], la = S.ex_ld[a], lb = S.ex_ld[b];
  wave_sync();
  S.ex_d[a] = db; S.ex_d[b] = da; S.ex_ld[a] = lb; S.ex_ld[b] = la;
  wave_sync();
}

// src/model.hpp:703-832: the last locus of the walk's model moves one
// position left and the locus it passes leaves the model
template <int KMAX>
__device__ __forceinline__ double exh_moveleft(Smem<KMAX>& S, const ChainConst& K, Exh& X) {
  --X.ms;
  const int ik = X.ms, ir = X.ms - 1;
  const int tk = S.ex_t[ik], tr = S.ex_t[ir];
  const double* v = S.ex_v;
  const double* lgd = S.ex_ld;
  const int nT = X.nT;
  --X.vs;
  X.syx += v[X.vs] * v[X.vs];
  X.ldl -= log(exl(S, X.vs, X.vs));
  if (tk == 4) {
    --X.vs;
    X.syx += v[X.vs] * v[X.vs];
    X.ldl -= log(exl(S, X.vs, X.vs));
  }
  --X.vs;
  X.syx += v[X.vs] * v[X.vs];
  X.ldq -= 0.5 * lgd[X.vs];
  X.ldl -= log(exl(S, X.vs, X.vs));
  if (tr == 4) {
    --X.vs;
    X.syx += v[X.vs] * v[X.vs];
    X.ldq -= 0.5 * lgd[X.vs];
    X.ldl -= log(exl(S, X.vs, X.vs));
  }
  const int x1k = S.ex_x1[ik], x2k = S.ex_x2[ik], x1r = S.ex_x1[ir], x2r = S.ex_x2[ir];
  wave_sync();
  S.ex_t[ik] = (int8_t)tr;
  S.ex_t[ir] = (int8_t)tk;
  if (tr == 4) {
    if (tk == 4) {
      int si = x2r;
      exh_swapadj(S, nT, si); ++si;
      exh_swapadj(S, nT, si); si -= 2;
      exh_swapadj(S, nT, si); ++si;
      exh_swapadj(S, nT, si);
      exh_swap_d(S, x1k, x2r); exh_swap_d(S, x1k, x2k);
      exh_swap_d(S, x1r, x2r); exh_swap_d(S, x1k, x2r);
    } else {
      int si = x2r;
      exh_swapadj(S, nT, si); --si;
      exh_swapadj(S, nT, si);
      exh_swap_d(S, x2r, x1k); exh_swap_d(S, x1r, x2r);
      S.ex_x2[ik] = x1k;
      S.ex_x1[ik] = x1k - 1;
    }
  } else {
    if (tk == 4) {
      int si = x1r;
      exh_swapadj(S, nT, si); ++si;
      exh_swapadj(S, nT, si);
      exh_swap_d(S, x1r, x1k); exh_swap_d(S, x1k, x2k);
      S.ex_x2[ir] = x1k;
      S.ex_x1[ik] = x2k;
    } else {
      exh_swapadj(S, nT, x1r);
      exh_swap_d(S, x1k, x1r);
    }
  }
  wave_sync();
  X.syx -= v[X.vs] * v[X.vs];
  X.ldl += log(exl(S, X.vs, X.vs));
  ++X.vs;
  if (tk == 4) {
    X.syx -= v[X.vs] * v[X.vs];
    X.ldl += log(exl(S, X.vs, X.vs));
    ++X.vs;
  }
  const double ll = X.ldq - X.ldl + K.minus_n_plus_nu_2 * log(X.syx);
  X.lmp += S.ex_pr[X.ms + 1];
  return ll + X.lmp;
}

// The walk over all 2^n sub-models (src/sampler.cpp:882-980): additions,
// left moves and "two new + swap" steps in the reference's order, so that
// every log-probability is reached by the same sequence of updates.  Wave 0
// only (lane 0 stores); returns the largest log-probability.  Out of line:
// inlined, its registers crowd the kernel's hot loops.
template <int KMAX>
__device__ __noinline__ double exh_walk(Smem<KMAX>& S, const ChainConst& K, Exh X, int n,
                                        double ll0, double* lp) {
  const bool st = lane_id() == 0;
  uint32_t ci[BMG_DR_MAX + 2];   // walk order; indexed only through `sel`/`put`
  for (int i = 0; i < BMG_DR_MAX + 2; ++i) ci[i] = (uint32_t)i;
  uint64_t cmb = 0;
  uint32_t tmp;
  int msz = 0;
  double mx = ll0;
  if (st) lp[0] = ll0;
  for (int i = 0; i < n; ++i) {
    ++msz;
    if (i > 1) { ++msz; exh_add(S, K, X); }
    cmb = (1ull << msz) - 1;
    double v = exh_add(S, K, X);
    if (st) lp[cmb] = v;
    if (v > mx) mx = v;
    for (int j = 0; j < i; ++j) {
      --msz;
      tmp = sel(ci, msz - 1); put(ci, msz - 1, sel(ci, msz)); put(ci, msz, tmp);
      cmb &= ~(1ull << tmp);
      v = exh_moveleft(S, K, X);
      if (st) lp[cmb] = v;
      if (v > mx) mx = v;
    }
    const uint64_t nmodels = (1ull << i) - i - 1;
    uint64_t j = 0, nK = 0;
    int do_lefts = 0;
    while (j < nmodels) {
      if (do_lefts < 2) {
        ++msz;
        tmp = sel(ci, msz); put(ci, msz, sel(ci, msz - 1)); put(ci, msz - 1, tmp);
        cmb |= 1ull << tmp;
        exh_add(S, K, X);
        exh_add(S, K, X);
        v = exh_moveleft(S, K, X);
        if (st) lp[cmb] = v;
        if (v > mx) mx = v;
        ++j; ++do_lefts;
      } else {
        ++nK;
        const int kb = __builtin_ctzll(nK);
        for (int kk = 0; kk <= kb; ++kk) {
          --msz;
          tmp = sel(ci, msz - 1); put(ci, msz - 1, sel(ci, msz)); put(ci, msz, tmp);
          cmb &= ~(1ull << tmp);
          v = exh_moveleft(S, K, X);
          if (st) lp[cmb] = v;
          if (v > mx) mx = v;
          ++j;
        }
        do_lefts = 0;
      }
    }
  }
  return mx;
}

// src/model.hpp:602-671 (ExhModel::update_to_model) and the walk: fills
// lp[0 .. 2^n) with the log posterior (up to a constant) of every sub-model,
// bit i of the index standing for the proposed model's locus cl + i.  All
// threads; returns the largest entry.
template <int KMAX>
__device__ __forceinline__ double exhaustive_modelset(Ctx<KMAX>& C, int cl, int n, double* lp) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int c0 = C.D.me;   // the first column after the constant and the covariates
  if (cl > 0) c0 = (C.D.types[S.tind[cl - 1]] == 4 ? S.xi2[cl - 1] : S.xi1[cl - 1]) + 1;
  const int nT = R.cols - c0;
  for (int c = 0; c < nT; ++c)
    for (int r = threadIdx.x; r <= c; r += BMG_BLOCK) S.ex_L[P(c) + r] = S.L[P(c0 + c) + c0 + r];
  for (int i = threadIdx.x; i < nT; i += BMG_BLOCK) {
    S.ex_v[i] = S.v[c0 + i];
    S.ex_d[i] = S.d[c0 + i];
    S.ex_ld[i] = log(S.d[c0 + i]);
  }
  for (int q = threadIdx.x; q < R.nloci - cl; q += BMG_BLOCK) {
    S.ex_x1[q] = S.xi1[cl + q] - c0;
    S.ex_x2[q] = S.xi2[cl + q] - c0;
    S.ex_t[q] = (int8_t)C.D.types[S.tind[cl + q]];
  }
  // prior changes by model size: adding to n_old = cl + q loci, and removing
  // from n_old = cl + q (src/prior.hpp:144-158, one effect type)
  if (threadIdx.x <= n) {
    const int q = threadIdx.x, type = C.D.type;
    int Ns[5] = {0, 0, 0, 0, 0};
    put(Ns, type, cl + q);
    S.ex_pa[q] = prior_add_untruncated(C.K, Ns, cl + q, type);
    S.ex_pr[q] = q > 0 || cl > 0 ? prior_rem(C.K, Ns, cl + q, type) : 0.0;
  }
  double vv = 0.0, ldl = 0.0, ldq = 0.0;
  for (int i = threadIdx.x; i < c0; i += BMG_BLOCK) {
    vv += S.v[i] * S.v[i];
    ldl += log(S.L[P(i) + i]);
    if (i >= 1) ldq += log(S.d[i]);
  }
  vv = block_sum(vv, S.red);
  ldl = block_sum(ldl, S.red);
  ldq = block_sum(ldq, S.red);
  __syncthreads();
  if (wave_id() == 0) {
    Exh X;
    X.ms = 0; X.vs = 0; X.nT = nT;
    X.syx = C.K.nus2_plus_yy - vv;
    X.ldl = ldl;
    X.ldq = 0.5 * ldq;
    X.lmp = 0.0;
    const double ll0 = X.ldq - X.ldl + C.K.minus_n_plus_nu_2 * log(X.syx);
    const double mx = exh_walk(S, C.K, X, n, ll0, lp);
    if (lane_id() == 0) S.bcd[0] = mx;
  }
  __syncthreads();
  double mx = S.bcd[0];
  __syncthreads();
  if (cl + n > C.K.max_size) {
    // the walk scores with the untruncated prior (its steps pass through
    // larger models); sub-models above the limit get prior 0 afterwards
    double m = -INFINITY;
    for (uint64_t i = threadIdx.x; i < (1ull << n); i += BMG_BLOCK) {
      if (cl + __popcll(i) > C.K.max_size) lp[i] = -INFINITY;
      else m = fmax(m, lp[i]);
    }
    m = wave_max(m);
    if (lane_id() == 0) S.red[wave_id()] = m;
    __syncthreads();
    mx = -INFINITY;
    for (int w = 0; w < BMG_BLOCK / 64; ++w) mx = fmax(mx, S.red[w]);
    __syncthreads();
  }
  return mx;
}

// src/sampler.cpp:982-1049 with one effect type: adds the log proposal
// probability of reaching each sub-model by the multi-step add/remove move
// from the current model to lp[i].  Thread per sub-model; the terms of one
// sub-model are summed in the reference's order.
template <int KMAX>
__device__ __forceinline__ void dr_proposal_probs(Smem<KMAX>& S, int n, double z_add, double z_rem,
                                                  uint64_t cl, uint64_t m_g, double* lp) {
  const uint64_t nmodels = 1ull << n;
  for (uint64_t i = threadIdx.x; i < nmodels; i += BMG_BLOCK) {
    uint64_t max_adds = m_g - cl, max_rems = cl;
    double za = z_add, zr = z_rem;
    uint64_t isadd = 0;
    for (int j = 0; j < n; ++j) {
      const int nind = S.ex_b2n[j];
      if ((i >> j) & 1) {
        za -= S.dq_add[nind]; zr += S.dq_rem[nind];
        ++max_rems; --max_adds;
      } else {
        isadd |= 1ull << nind;
      }
    }
    double acc = lp[i];
    int lrp = n - 1;
    for (int j = 0; j < n; ++j) {
      if (max_adds > 0 && max_rems > 0) acc += LOG_HALF;
      if ((isadd >> j) & 1) {
        --max_adds;
        acc += log(S.dq_add[j]) - log(za);
        za -= S.dq_add[j];
      } else {
        --max_rems;
        while ((isadd >> lrp) & 1) --lrp;
        acc += log(S.dq_rem[lrp]) - log(zr);
        zr -= S.dq_rem[lrp];
        --lrp;
      }
    }
    lp[i] = acc;
  }
}

// Second-stage draw (src/sampler.hpp:1008-1028 and 1592-1612): pairs each
// sub-model with its complement, keeps the better of the two with the
// probability of the difference, and draws one pair from the cumulative
// weights (src/utils.cpp:73-100 with `ms/2 - 3` bisection steps).  The
// pairwise transform runs a thread per pair; wave 0 forms the cumulative
// sums in order, one chunk of 64 at a time passed along the lanes; the
// linear part of the search is a workgroup minimum over the first index
// whose sum exceeds the draw.
template <int KMAX>
__device__ __forceinline__ uint64_t dr_second_stage(Ctx<KMAX>& C, int n, double max_log, double* p) {
  const uint64_t nmodels = 1ull << n, mask = nmodels - 1, half = nmodels / 2;
  for (uint64_t i = threadIdx.x; i < half; i += BMG_BLOCK) {
    const uint64_t j = (~i) & mask;
    const double pi = p[i], pj = p[j];
    const double a = pj - pi;
    // a pair of two zero-prior sub-models (model-size limit) has weight 0
    if (pi == -INFINITY && pj == -INFINITY) { p[i] = 0.0; p[j] = 1.0; }
    else if (a > 0) { p[i] = exp(pj - max_log) * (1 - exp(-a)); p[j] = -1.0; }
    else { p[i] = exp(pi - max_log) * (1 - exp(a)); p[j] = 1.0; }
  }
  __syncthreads();
  if (wave_id() == 0) {
    const int lane = lane_id();
    double sum = 0.0;
    for (uint64_t base = 0; base < half; base += 64) {
      const uint64_t idx = base + lane;
      const double x = idx < half ? p[idx] : 0.0;
      const int cnt = half - base < 64 ? (int)(half - base) : 64;
      double mine = 0.0;
      for (int k = 0; k < cnt; ++k) {
        sum = readlane_d(x, k) + sum;
        if (lane == k) mine = sum;
      }
      if (idx < half) p[idx] = mine;
    }
  }
  __syncthreads();
  const double u = rng_u01(C.R.rng);
  const double r = u * p[half - 1];
  int level = n / 2 - 3;
  uint64_t a = 0, b = half - 1;
  while (level > 0) {
    const uint64_t c = (a + b) / 2;
    if (r < p[c]) b = c; else a = c + 1;
    --level;
  }
  long long first = (long long)half;
  for (uint64_t i = a + threadIdx.x; i <= b; i += BMG_BLOCK)
    if (r < p[i]) { first = (long long)i; break; }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const long long w = __shfl_xor(first, o, 64);
    first = w < first ? w : first;
  }
  __syncthreads();
  if (lane_id() == 0) C.S.redi[wave_id()] = first;
  __syncthreads();
  for (int w = 0; w < BMG_BLOCK / 64; ++w) first = C.S.redi[w] < first ? C.S.redi[w] : first;
  uint64_t sm = first < (long long)half ? (uint64_t)first : half - 1;
  if (p[(~sm) & mask] < 0) sm = (~sm) & mask;
  __syncthreads();
  return sm;
}

// the second stage re-adds the SNPs the first stage removed, each with the
// prior precision it has in the current model (src/sampler.hpp:931-960);
// false when the model capacity does not allow it
template <int KMAX>
__device__ __forceinline__ bool readd_removed(Ctx<KMAX>& C, uint64_t ind) {
  Smem<KMAX>& S = C.S;
  const int mi = C.A.model_inds[ind];
  const int t_ind = S.c_tind[mi];
  double d2[2];
  d2[0] = S.c_d[S.c_xi1[mi]];
  d2[1] = C.D.types[t_ind] == 4 ? S.c_d[S.c_xi2[mi]] : 0.0;
  add_term(C, (uint32_t)ind, t_ind, d2);
  return !(C.R.err & ERR_CAPACITY);
}

// removes the proposed model's loci cl + i whose bit of `sm` is clear
// (src/sampler.hpp:1081-1091)
template <int KMAX>
__device__ __forceinline__ void dr_keep(Ctx<KMAX>& C, int cl, int n, uint64_t sm) {
  uint64_t removals = sm & ((1ull << n) - 1);
  int c = cl;
  for (int i = 0; i < n; ++i) {
    if ((removals & 1) == 0) { remove_term(C, c + i); --c; }
    removals >>= 1;
  }
}

// Second stage of the multi-step add/remove move after a first-stage
// rejection (src/sampler.hpp:931-1099): the removed SNPs are re-added, every
// sub-model between the current and the proposed one is scored with its
// proposal probability, and one is drawn.  Returns the number of moved SNPs
// kept (0: the current model stays).
template <int KMAX>
__device__ __forceinline__ int dr_addrem(Ctx<KMAX>& C, int ms, double* adapt) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  int ms_rem = 0;
  for (int i = 0; i < ms; ++i) ms_rem += S.mi_isadd[i] ? 0 : 1;
  const int cl = R.c_nloci - ms_rem;
  const int ms_add = ms - ms_rem;
  double z_add = R.tot_add, z_rem = R.tot_rem;
  const uint64_t nmb = (1ull << ms_add) - 1, mask = (1ull << ms) - 1;
  bool ok = true;
  for (int i = 0; i < ms && ok; ++i) {
    const uint64_t ind = S.mi[i];
    if (S.mi_isadd[i]) { z_add += C.A.q_add[ind]; z_rem -= C.A.q_rem[ind]; }
    else ok = readd_removed(C, ind);
  }
  if (!ok) { restore(C); undo_dd(C, ms); return 0; }
  for (int i = 0; i < ms; ++i) {
    const uint64_t ind = (uint64_t)S.mi_map[i];
    const int pos = find_locus(C, ind);
    if (threadIdx.x == 0 && pos >= cl && pos - cl < ms) {
      S.ex_b2n[pos - cl] = (uint8_t)i;
      S.dq_add[i] = C.A.q_add[ind];
      S.dq_rem[i] = C.A.q_rem[ind];
    }
  }
  __syncthreads();
  double* lp = C.A.dr_lp;
  const uint64_t t0 = wall_clock64();
  const double max_log = exhaustive_modelset(C, cl, ms, lp);
  dr_proposal_probs(S, ms, z_add, z_rem, (uint64_t)cl, C.D.p, lp);
  __syncthreads();
  const uint64_t sm = dr_second_stage(C, ms, max_log, lp);
  if (threadIdx.x == 0 && C.prof) C.prof[19] += wall_clock64() - t0;
  if (sm == ((~nmb) & mask)) { restore(C); undo_dd(C, ms); return 0; }
  int moved = 0;
  for (int i = 0; i < ms; ++i) {
    const int bit = (int)((sm >> i) & 1), nb = (int)((nmb >> i) & 1);
    if (bit == nb) { ++moved; continue; }
    const uint64_t ind = (uint64_t)S.mi_map[S.ex_b2n[i]];
    if (bit) { tree_set(C, 0, ind, 1); tree_set(C, 1, ind, 0); }
    else { tree_set(C, 0, ind, 0); tree_set(C, 1, ind, 1); }
  }
  dr_keep(C, cl, ms, sm);
  if (threadIdx.x == 0) adapt[3 * 256 + ms - 1] += (double)moved / (double)ms;
  commit(C);
  return moved;
}

// Second stage of the state-change move (src/sampler.hpp:1547-1667): as
// above, scored by model size instead of proposal probabilities.
template <int KMAX>
__device__ __forceinline__ int dr_statechange(Ctx<KMAX>& C, int ms, int ms_add, int ms_rem) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const int cl = R.c_nloci - ms_rem;
  const uint64_t nmb = (1ull << ms_add) - 1, mask = (1ull << ms) - 1;
  for (int i = 0; i < ms_rem; ++i)
    if (!readd_removed(C, S.mi_rem[i])) { restore(C); return 0; }
  double* lp = C.A.dr_lp;
  const uint64_t t0 = wall_clock64();
  const double max_log = exhaustive_modelset(C, cl, ms, lp);
  for (uint64_t i = threadIdx.x; i <= mask; i += BMG_BLOCK) lp[i] -= log((double)(cl + __popcll(i)));
  __syncthreads();
  const uint64_t sm = dr_second_stage(C, ms, max_log, lp);
  if (threadIdx.x == 0 && C.prof) C.prof[19] += wall_clock64() - t0;
  if (sm == ((~nmb) & mask)) { restore(C); return 0; }
  int moved = 0;
  for (int i = 0; i < ms; ++i) {
    const int bit = (int)((sm >> i) & 1), nb = (int)((nmb >> i) & 1);
    if (bit != nb) continue;
    ++moved;
    if (bit) { const uint64_t ind = S.mi_add[i]; tree_set(C, 0, ind, 1); tree_set(C, 1, ind, 0); }
    else { const uint64_t ind = S.mi_rem[i - ms_add]; tree_set(C, 0, ind, 0); tree_set(C, 1, ind, 1); }
  }
  dr_keep(C, cl, ms, sm);
  commit(C);
  return moved;
}

// src/sampler.hpp:899-1123
template <int KMAX>
__device__ __forceinline__ int move_addrem(Ctx<KMAX>& C, double* adapt) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const ChainConst& K = C.K;
  const int type = C.D.type;
  int max_ms = (int)K.max_ms;
  int ms = sample_naive(adapt, max_ms, rng_u01(R.rng)) + 1;
  R.movesize = ms;
  double lqf = 0, lqb = 0, lmpc = 0;
  // the move's tree writes are journalled, so a rejection puts the trees
  // back without recomputing their sums
  const uint64_t tq0 = wall_clock64();
  journal_begin(C);
  const uint64_t tq1 = wall_clock64();
  // prepare_addrem (src/sampler.hpp:1126-1165)
  uint64_t max_adds = C.D.p - R.c_nloci, max_rems = R.c_nloci;
  for (int i = 0; i < ms; ++i) {
    int add;
    if (max_rems == 0) add = 1;
    else if (max_adds == 0) add = 0;
    else { lqf += LOG_HALF; add = rng_u01(R.rng) < 0.5; }
    double u = rng_u01(R.rng);
    uint64_t ind;
    double wv;
    if (add) {
      --max_adds;
      const double tot = R.tot_add;
      ind = tree_pick_zero(C, 0, u, &wv);   // wv = q_add[ind]
      lqf += log(wv) - log(tot);
    } else {
      --max_rems;
      const double tot = R.tot_rem;
      ind = tree_pick_zero(C, 1, u, &wv);   // wv = q_rem[ind]
      lqf += log(wv) - log(tot);
    }
    if (threadIdx.x == 0) {
      S.mi_isadd[i] = (uint8_t)add; S.mi[i] = ind; S.mi_map[i] = add ? (int64_t)ind : -1;
    }
    __syncthreads();
  }
  const uint64_t tq2 = wall_clock64();
  // do_addrem (src/sampler.hpp:1204-1268): removals first; the last removal's
  // downdate also forms the first addition's column dots
  int last_rem = -1;
  long long first_add = -1;
  for (int i = 0; i < ms; ++i) {
    if (!S.mi_isadd[i]) last_rem = i;
    else if (first_add < 0) first_add = (long long)S.mi[i];
  }
  R.pre_snp = -1;
  int last = ms - 1;
  for (int i = 0; i < ms; ++i) {
    if (S.mi_isadd[i]) continue;
    uint64_t ind = S.mi[i];
    tree_set(C, 0, ind, 0);
    while (S.mi_map[last] >= 0) --last;
    __syncthreads();
    if (threadIdx.x == 0) S.mi_map[last] = (int64_t)ind;
    __syncthreads();
    int mind = find_locus(C, ind);
    lmpc += prior_rem(K, R.Ns, R.nloci, type);
    remove_term(C, mind, i == last_rem ? first_add : -1);
  }
  double d2[2];
  for (int i = 0; i < ms; ++i) {
    if (!S.mi_isadd[i]) continue;
    uint64_t ind = S.mi[i];
    long long next = -1;   // the move's next addition, whose dots are formed ahead
    for (int j = i + 1; j < ms; ++j)
      if (S.mi_isadd[j]) { next = (long long)S.mi[j]; break; }
    tree_set(C, 1, ind, 0);
    lmpc += prior_add(K, R.Ns, R.nloci, type);
    prepare_add_new_term(C, d2);
    add_term(C, (uint32_t)ind, 0, d2, next);
  }
  R.pre_snp = -1;
  const uint64_t tq3 = wall_clock64();
  // backward_prepare_addrem (src/sampler.hpp:1168-1202); the reverse
  // move's weights are loaded together first (thread i loads step i's)
  {
    if ((int)threadIdx.x < ms) {
      const int i = (int)threadIdx.x;
      uint64_t ind;
      if (S.mi_isadd[i]) {
        int k = 0;   // step i is the k-th addition; it pairs with the k-th from last
        for (int q = 0; q < i; ++q) k += S.mi_isadd[q];
        int lr = ms - 1, seen = 0;
        for (;; --lr) if (S.mi_isadd[lr] && seen++ == k) break;
        ind = (uint64_t)S.mi_map[lr];
        S.bq[i] = C.A.q_rem[ind];
      } else {
        ind = (uint64_t)S.mi_map[i];
        S.bq[i] = C.A.q_add[ind];
      }
    }
    __syncthreads();
    uint64_t madd = C.D.p - R.nloci, mrem = R.nloci;
    double w_add = R.tot_add, w_rem = R.tot_rem;
    for (int i = 0; i < ms; ++i) {
      if (madd > 0 && mrem > 0) lqb += LOG_HALF;
      const double q = S.bq[i];
      if (S.mi_isadd[i]) {
        --mrem;
        lqb += log(q) - log(w_rem);
        w_rem -= q;
      } else {
        --madd;
        lqb += log(q) - log(w_add);
        w_add -= q;
      }
    }
    __syncthreads();
  }
  double log_r = lqb - lqf;
  log_r += lmpc + R.ll - R.c_ll;
  const uint64_t tq4 = wall_clock64();
  if (threadIdx.x == 0 && C.prof) {
    C.prof[20] += tq2 - tq1; C.prof[21] += tq3 - tq2; C.prof[22] += tq4 - tq3; C.prof[23] += tq1 - tq0;
  }
  double* r_ms_sum = adapt + 3 * 256;
  double* r_ms_n = adapt + 4 * 256;
  if (threadIdx.x == 0) {
    if (K.dr == 0) r_ms_sum[ms - 1] += (log_r >= 0 ? 1.0 : exp(log_r));
    r_ms_n[ms - 1] += 1.0;
  }
  if (log_r >= 0 || log(rng_u01(R.rng)) <= log_r) {
    R.tj_on = 0;
    commit(C);
    if (K.dr != 0 && threadIdx.x == 0) r_ms_sum[ms - 1] += 1.0;
    if (threadIdx.x == 0 && C.prof) C.prof[23] += wall_clock64() - tq4;
    return ms;
  }
  // a proposal past the compiled capacity has no second stage
  if (ms <= K.dr && ms > 1 && !(R.err & ERR_CAPACITY)) {
    R.tj_on = 0;
    const uint64_t t0 = wall_clock64();
    const int moved = dr_addrem(C, ms, adapt);
    if (threadIdx.x == 0 && C.prof) C.prof[18] += wall_clock64() - t0;
    return moved;
  }
  restore(C);
  journal_restore(C);
  if (threadIdx.x == 0 && C.prof) C.prof[23] += wall_clock64() - tq4;
  return 0;
}

// is SNP j in the current model (with the switch move's temporary marks)?
template <int KMAX>
__device__ __forceinline__ int cur_in_model(Ctx<KMAX>& C, int64_t j, int64_t wbase) {
  if (j < 0 || j >= (int64_t)C.D.p) return 0;   // outside the genome: stops a walk
  for (int q = 0; q < C.R.n_marks; ++q) if ((int64_t)C.S.marks[q] == j) return 1;
  if (j >= wbase && j < wbase + 512) return C.S.win[j - wbase] >= 0;
  return C.A.model_inds[j] >= 0;
}

template <int KMAX>
__device__ __forceinline__ void load_window(Ctx<KMAX>& C, int64_t center, int64_t* wbase) {
  int64_t b = center - 256;
  if (b < 0) b = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < 512; t += BMG_BLOCK) {
    int64_t j = b + t;
    C.S.win[t] = (j < (int64_t)C.D.p) ? C.A.model_inds[j] : -1;
  }
  __syncthreads();
  *wbase = b;
}

// src/sampler.hpp:1271-1418
template <int KMAX>
__device__ __forceinline__ int move_switch(Ctx<KMAX>& C, double* adapt) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const ChainConst& K = C.K;
  uint64_t full = C.D.p - R.c_nloci;
  uint64_t mm2 = (K.max_ms + 1) / 2;
  uint64_t mx = mm2 < (full + 1) / 2 ? mm2 : (full + 1) / 2;
  mx = mx < (uint64_t)R.c_nloci ? mx : (uint64_t)R.c_nloci;
  int ms = sample_naive(adapt + 256, (int)mx, rng_u01(R.rng)) + 1;
  R.movesize = ms;
  uint64_t nbh = (full - ms) / 2;
  if (nbh < 1) nbh = 1;
  if (nbh > K.max_nbh) nbh = K.max_nbh;
  int64_t m_g = (int64_t)C.D.p;
  R.n_marks = 0;
  for (int i = 0; i < ms; ++i) {
    int mir = (int)floor(rng_u01(R.rng) * R.nloci);
    int64_t ir = S.loci[mir];
    int64_t wb;
    load_window(C, ir, &wb);
    int64_t ia;
    int j = 0;
    int move = (int)floor(rng_u01(R.rng) * nbh) + 1;
    if (rng_u01(R.rng) < 0.5) {
      ia = ir; j = 0;
      while (j < move) {
        ++ia;
        if (ia < m_g) { if (!cur_in_model(C, ia, wb)) ++j; }
        else break;
      }
      if (ia == m_g) {
        ia = ir; move = move + move - j; j = 0;
        while (j < move) {
          --ia;
          while (cur_in_model(C, ia, wb)) --ia;
          ++j;
        }
      }
    } else {
      ia = ir; j = 0;
      if (ia > 0) {
        while (j < move) {
          --ia;
          if (ia > 0) { if (!cur_in_model(C, ia, wb)) ++j; }
          else break;
        }
      }
      if (j < move) {
        if (!cur_in_model(C, ia, wb)) ++j;
        if (j < move) {
          ia = ir; move = move + move - j; j = 0;
          while (j < move) {
            ++ia;
            while (cur_in_model(C, ia, wb)) ++ia;
            ++j;
          }
        }
      }
    }
    if (ia < 0 || ia >= m_g) { R.err |= ERR_WALK; ia = ia < 0 ? 0 : m_g - 1; }
    int t_ind = S.tind[mir];
    __syncthreads();
    if (threadIdx.x == 0) {
      S.mi_rem[i] = (uint64_t)ir; S.mi_add[i] = (uint64_t)ia;
      S.marks[R.n_marks] = (uint64_t)ia; S.mark_val[R.n_marks] = t_ind;
    }
    __syncthreads();
    R.n_marks++;
    remove_term(C, mir);
  }
  double d2[2];
  for (int i = 0; i < ms; ++i) {
    uint64_t ia = S.mi_add[i];
    int t_ind = S.mark_val[i];
    prepare_add_new_term(C, d2);
    add_term(C, (uint32_t)ia, t_ind, d2);
  }
  R.n_marks = 0;
  double log_r = R.ll - R.c_ll;
  if (log_r >= 0 || log(rng_u01(R.rng)) <= log_r) {
    commit(C);
    for (int i = 0; i < ms; ++i) {
      tree_set(C, 0, S.mi_rem[i], 0); tree_set(C, 1, S.mi_rem[i], 1);
      tree_set(C, 0, S.mi_add[i], 1); tree_set(C, 1, S.mi_add[i], 0);
    }
    return 2 * ms;
  }
  restore(C);
  return 0;
}

// Learned swap (no counterpart in the reference): the swap of
// src/sampler.hpp:1271-1418 with the partner of the removed SNP j drawn from
// the frozen weights W[j][.] over the SNPs within lH of j that are not in
// the model, and the exact Hastings ratio
//   log W[k][j] - log Z_r - log W[j][k] + log Z_f
// (Z_f, Z_r: forward / reverse normalisers; the uniform choice of the
// removed SNP and the prior cancel because the size is unchanged).  Column
// c of a weight row is the offset c < lH ? c - lH : c - lH + 1.  Lanes
// gather the masked weights into LDS; every lane then runs the same
// sequential sums, so the draw and the ratio are those of the oracle.
__device__ __forceinline__ int lw_off(int H, int c) { return c < H ? c - H : c - H + 1; }
__device__ __forceinline__ int lw_col(int H, int o) { return o < 0 ? o + H : o + H - 1; }

template <int KMAX>
__device__ __forceinline__ int move_learned_swap(Ctx<KMAX>& C) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const int H = C.D.lH;
  const int64_t m_g = (int64_t)C.D.p;
  R.movesize = 1;
  int mir = (int)floor(rng_u01(R.rng) * R.nloci);
  if (mir >= R.nloci) mir = R.nloci - 1;
  const int64_t j = S.loci[mir];
  double* wf = S.work;            // 2H forward weights (0: not a candidate)
  double* wr = S.work + 2 * H;    // 2H reverse weights
  __syncthreads();
  if (threadIdx.x < 2 * H) {
    int64_t k = j + lw_off(H, threadIdx.x);
    double w = 0.0;
    if (k >= 0 && k < m_g && C.A.model_inds[k] < 0) w = (double)C.D.lw[(size_t)j * 2 * H + threadIdx.x];
    wf[threadIdx.x] = w;
  }
  __syncthreads();
  double zf = 0.0;
  for (int c = 0; c < 2 * H; ++c) zf += wf[c];
  if (!(zf > 0.0)) return 0;
  double u = rng_u01(R.rng) * zf, acc = 0.0;
  int pick = -1;
  for (int c = 0; c < 2 * H; ++c) {
    if (wf[c] > 0.0) {
      acc += wf[c];
      pick = c;
      if (u < acc) break;
    }
  }
  if (pick < 0) return 0;
  const int64_t k = j + lw_off(H, pick);
  const double wjk = wf[pick];
  __syncthreads();
  if (threadIdx.x < 2 * H) {
    int64_t q = k + lw_off(H, threadIdx.x);
    double w = 0.0;
    if (q >= 0 && q < m_g && (q == j || C.A.model_inds[q] < 0))
      w = (double)C.D.lw[(size_t)k * 2 * H + threadIdx.x];
    wr[threadIdx.x] = w;
  }
  __syncthreads();
  double zr = 0.0;
  for (int c = 0; c < 2 * H; ++c) zr += wr[c];
  const double wkj = wr[lw_col(H, (int)(j - k))];
  if (!(wkj > 0.0)) return 0;
  const double lq = log(wkj) - log(zr) - log(wjk) + log(zf);
  const int t_ind = S.tind[mir];
  __syncthreads();
  if (threadIdx.x == 0) { S.mi_rem[0] = (uint64_t)j; S.mi_add[0] = (uint64_t)k; }
  remove_term(C, mir);
  double d2[2];
  prepare_add_new_term(C, d2);
  add_term(C, (uint32_t)k, t_ind, d2);
  double log_r = R.ll - R.c_ll + lq;
  if (log_r >= 0 || log(rng_u01(R.rng)) <= log_r) {
    commit(C);
    tree_set(C, 0, (uint64_t)j, 0); tree_set(C, 1, (uint64_t)j, 1);
    tree_set(C, 0, (uint64_t)k, 1); tree_set(C, 1, (uint64_t)k, 0);
    return 2;
  }
  restore(C);
  return 0;
}

// src/sampler.hpp:1421-1669
template <int KMAX>
__device__ __forceinline__ int move_statechange(Ctx<KMAX>& C, double* adapt) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const ChainConst& K = C.K;
  const int type = C.D.type;
  uint64_t m_g = C.D.p;
  uint64_t a = (uint64_t)K.max_nbh * 2, b = m_g - 1 < K.max_ms ? m_g - 1 : K.max_ms;
  uint64_t mx = a < b ? a : b;
  int ms = sample_naive(adapt + 2 * 256, (int)mx, rng_u01(R.rng)) + 1;
  R.movesize = ms;
  uint64_t nbh = K.max_nbh < m_g / 2 ? K.max_nbh : m_g / 2;
  int mic = (int)floor(rng_u01(R.rng) * R.nloci);
  uint64_t ic = S.loci[mic];
  // partial Fisher-Yates over the 2 nbh offsets (only the first ms entries
  // are used): positions that differ from the identity are kept as a sparse
  // map in LDS by thread 0; every thread consumes the same draws
  __syncthreads();
  int nsw = 0;
  for (int i = 0; i < ms; ++i) {
    uint64_t ind = (uint64_t)floor(rng_u01(R.rng) * (double)(nbh * 2 - i)) + i;
    if (threadIdx.x == 0) {
      uint64_t vi = i, vind = ind;
      for (int q = 0; q < nsw; ++q) {
        if (S.sc_from[q] == (uint64_t)i) vi = S.sc_to[q];
        if (S.sc_from[q] == ind) vind = S.sc_to[q];
      }
      S.sc_perm[i] = vind;
      int found = 0;
      for (int q = 0; q < nsw; ++q) if (S.sc_from[q] == ind) { S.sc_to[q] = vi; found = 1; }
      if (!found) { S.sc_from[nsw] = ind; S.sc_to[nsw] = vi; }
    }
    // every thread tracks the map size identically
    {
      int found = 0;
      __syncthreads();
      for (int q = 0; q < nsw; ++q) if (S.sc_from[q] == ind) found = 1;
      if (!found) ++nsw;
      __syncthreads();
    }
  }
  __syncthreads();
  uint64_t wrap = ic + nbh;
  if (ic < nbh) wrap += (nbh - ic);
  else if (wrap > m_g - 1) wrap = m_g - 1;
  int ms_rem = 0, ms_add = 0;
  for (int i = 0; i < ms; ++i) {
    uint64_t ind = S.sc_perm[i] + ic + 1;
    if (ind > wrap) ind = ic - (ind - wrap);
    if (ind >= m_g) { R.err |= ERR_WALK; ind = m_g - 1; }
    int fl = find_locus(C, ind);
    if (threadIdx.x == 0) {
      if (fl < 0) S.mi_add[ms_add] = ind; else S.mi_rem[ms_rem] = ind;
    }
    if (fl < 0) ++ms_add; else ++ms_rem;
  }
  __syncthreads();
  double lmpc = 0.0;
  R.pre_snp = -1;
  for (int i = 0; i < ms_rem; ++i) {
    int mir = find_locus(C, S.mi_rem[i]);
    lmpc += prior_rem(K, R.Ns, R.nloci, type);
    // the last removal's downdate also forms the first addition's dots
    remove_term(C, mir, i + 1 == ms_rem && ms_add > 0 ? (long long)S.mi_add[0] : -1);
  }
  double d2[2];
  int t_ind = 0;
  for (int i = 0; i < ms_add; ++i) {
    uint64_t ia = S.mi_add[i];
    lmpc += prior_add(K, R.Ns, R.nloci, type);
    prepare_add_new_term(C, d2);
    // the next addition's dots are formed during this one's sweep
    add_term(C, (uint32_t)ia, t_ind, d2, i + 1 < ms_add ? (long long)S.mi_add[i + 1] : -1);
  }
  R.pre_snp = -1;
  double log_r = lmpc + R.ll - R.c_ll;
  log_r += log((double)R.c_nloci) - log((double)R.nloci);
  if (log_r >= 0 || log(rng_u01(R.rng)) <= log_r) {
    commit(C);
    for (int i = 0; i < ms_add; ++i) { tree_set(C, 0, S.mi_add[i], 1); tree_set(C, 1, S.mi_add[i], 0); }
    for (int i = 0; i < ms_rem; ++i) { tree_set(C, 0, S.mi_rem[i], 0); tree_set(C, 1, S.mi_rem[i], 1); }
    return ms;
  }
  if (ms <= K.dr && ms > 1 && !(R.err & ERR_CAPACITY)) {
    const uint64_t t0 = wall_clock64();
    const int moved = dr_statechange(C, ms, ms_add, ms_rem);
    if (threadIdx.x == 0 && C.prof) C.prof[18] += wall_clock64() - t0;
    return moved;
  }
  restore(C);
  return 0;
}

// src/sampler.hpp:1728-1769
__device__ __forceinline__ void adapt_jd_mb(double* p_ms, double* adapt, int max_ms) {
  double* q_p0 = adapt + 5 * 256;
  const double* r_ms = adapt + 3 * 256;
  const double* r_n = adapt + 4 * 256;
  double q = 1 - *p_ms;
  double zp = 1 - pow(q, max_ms);
  double q_p = *p_ms / zp;
  for (int i = 0; i < max_ms; ++i) { q_p0[i] += q_p; q_p *= q; }
  double p_best = 0.01, p = 0.01, h_best = -INFINITY, h;
  double step = 0.98 / 49.0, ratio;
  for (int j = 0; j < 50; ++j) {
    h = 0.0; zp = 0.0; q = 1.0 - p; q_p = p;
    for (int i = 0; i < max_ms; ++i) {
      ratio = q_p / q_p0[i];
      h += (i + 1) * r_ms[i] * ratio;
      zp += r_n[i] * ratio;
      q_p *= q;
    }
    h /= zp;
    if (h > h_best) { h_best = h; p_best = p; }
    p += step;
  }
  *p_ms = p_best;
}

__device__ __forceinline__ void geometric_cdf(int maxsize, double p, double* values) {
  double q = 1 - p, qp = q;
  for (int i = 0; i < maxsize; ++i) { values[i] = 1 - q; q *= qp; }
}

// --------------------------------------------------------------- kernel ---

template <int KMAX>
__device__ __forceinline__ void load_chain(Ctx<KMAX>& C, const ChainScal& CS) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  R.rng = rng_make(C.K.seed, C.K.chain_id, 0, 0);
  R.rng.c0 = CS.rng_c0; R.rng.c1 = CS.rng_c1;
  R.cols = R.c_cols = CS.cols;
  R.nloci = R.c_nloci = CS.nloci;
  for (int q = 0; q < 5; ++q) R.Ns[q] = R.c_Ns[q] = CS.Ns[q];
  R.syx = R.c_syx = CS.syx; R.ldq = R.c_ldq = CS.ldq; R.ldl = R.c_ldl = CS.ldl;
  R.ll = R.c_ll = CS.ll;
  R.sigma2 = CS.sigma2; R.alpha = CS.alpha; R.alpha2 = CS.alpha2;
  for (int q = 0; q < 4; ++q) R.ita2[q] = CS.inv_tau2_alpha2[q];
  R.beta_len = CS.beta_len; R.mu_computed = CS.mu_computed;
  R.lmin = KMAX; R.n_rm = 0; R.n_marks = 0; R.movesize = CS.movesize;
  R.tot_add = CS.tot_add; R.tot_rem = CS.tot_rem;
  R.err = CS.err;
  R.learn_on = CS.learn_on && C.D.lw != nullptr && C.D.lH >= 1 && C.D.lH <= BMG_LH_MAX;
  R.tj_on = 0; R.tj_n = 0; R.tj_tot[0] = R.tj_tot[1] = 0.0;
  R.pre_snp = -1; R.pre_slot = -1;
  const double* vec = C.A.vec;
  const int32_t* idx = C.A.idx;
  for (int e = threadIdx.x; e < P(R.cols); e += BMG_BLOCK) S.L[e] = C.A.Lcur[e];
  for (int i = threadIdx.x; i < KMAX; i += BMG_BLOCK) {
    S.xy[i] = S.c_xy[i] = vec[i];
    S.v[i] = S.c_v[i] = vec[BMG_KMAXG + i];
    S.d[i] = S.c_d[i] = vec[2 * BMG_KMAXG + i];
    S.beta[i] = vec[3 * BMG_KMAXG + i];
    S.mu[i] = vec[4 * BMG_KMAXG + i];
    S.loci[i] = S.c_loci[i] = (uint32_t)idx[i];
    S.xi1[i] = S.c_xi1[i] = idx[BMG_KMAXG + i];
    S.xi2[i] = S.c_xi2[i] = idx[2 * BMG_KMAXG + i];
    S.slot[i] = S.c_slot[i] = idx[3 * BMG_KMAXG + i];
    S.snp[i] = S.c_snp[i] = idx[4 * BMG_KMAXG + i];
    S.xtype[i] = S.c_xtype[i] = (int8_t)idx[5 * BMG_KMAXG + i];
    S.tind[i] = S.c_tind[i] = (int8_t)idx[6 * BMG_KMAXG + i];
  }
  if (threadIdx.x < (KMAX + 63) / 64) {
    uint64_t m = 0;
    for (int i = 0; i < R.cols; ++i)
      if ((idx[3 * BMG_KMAXG + i] >> 6) == (int)threadIdx.x) m |= 1ull << (idx[3 * BMG_KMAXG + i] & 63);
    S.slots_cur[threadIdx.x] = S.slots_prop[threadIdx.x] = m;
  }
  // sum-tree levels >= 2
  if (C.D.top >= 2) {
    uint64_t o2 = C.D.lev_off[2];
    for (uint64_t e = threadIdx.x; e < C.D.lds_lev_size; e += BMG_BLOCK) {
      S.tlev[0][e] = C.A.tree_add[o2 + e];
      S.tlev[1][e] = C.A.tree_rem[o2 + e];
    }
  }
  __syncthreads();
}

template <int KMAX>
__device__ __forceinline__ void store_chain(Ctx<KMAX>& C, ChainScal& CS) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  __syncthreads();
  double* vec = C.A.vec;
  int32_t* idx = C.A.idx;
  for (int i = threadIdx.x; i < KMAX; i += BMG_BLOCK) {
    vec[i] = S.c_xy[i]; vec[BMG_KMAXG + i] = S.c_v[i]; vec[2 * BMG_KMAXG + i] = S.c_d[i];
    vec[3 * BMG_KMAXG + i] = S.beta[i]; vec[4 * BMG_KMAXG + i] = S.mu[i];
    idx[i] = (int32_t)S.c_loci[i]; idx[BMG_KMAXG + i] = S.c_xi1[i]; idx[2 * BMG_KMAXG + i] = S.c_xi2[i];
    idx[3 * BMG_KMAXG + i] = S.c_slot[i]; idx[4 * BMG_KMAXG + i] = S.c_snp[i];
    idx[5 * BMG_KMAXG + i] = S.c_xtype[i]; idx[6 * BMG_KMAXG + i] = S.c_tind[i];
  }
  if (C.D.top >= 2) {
    uint64_t o2 = C.D.lev_off[2];
    for (uint64_t e = threadIdx.x; e < C.D.lds_lev_size; e += BMG_BLOCK) {
      C.A.tree_add[o2 + e] = S.tlev[0][e];
      C.A.tree_rem[o2 + e] = S.tlev[1][e];
    }
  }
  if (threadIdx.x == 0) {
    CS.rng_c0 = R.rng.c0; CS.rng_c1 = R.rng.c1;
    CS.cols = R.c_cols; CS.nloci = R.c_nloci;
    for (int q = 0; q < 5; ++q) CS.Ns[q] = R.c_Ns[q];
    CS.syx = R.c_syx; CS.ldq = R.c_ldq; CS.ldl = R.c_ldl; CS.ll = R.c_ll;
    CS.sigma2 = R.sigma2; CS.alpha = R.alpha; CS.alpha2 = R.alpha2;
    for (int q = 0; q < 4; ++q) CS.inv_tau2_alpha2[q] = R.ita2[q];
    CS.beta_len = R.beta_len; CS.mu_computed = R.mu_computed;
    CS.movesize = R.movesize;
    CS.tot_add = R.tot_add; CS.tot_rem = R.tot_rem;
    CS.err = R.err;
  }
  __syncthreads();
}

// the tau2/alpha refresh at the top of an iteration (src/sampler.cpp:628-635)
// acts on the current model; here the working copy equals it, and the
// refreshed model is committed in full
template <int KMAX>
__device__ __forceinline__ void refresh_tau2(Ctx<KMAX>& C) {
  Reg& R = C.R;
  // a refresh whose factorisation fails is undone (as in the oracle): the
  // previous tau2 / alpha and the committed model are kept
  const double al = R.alpha, al2 = R.alpha2;
  double ita[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) ita[q] = R.ita2[q];
  const uint64_t t0 = wall_clock64();
  sample_alpha_and_tau2(C);
  if (threadIdx.x == 0 && C.prof) C.prof[15] += wall_clock64() - t0;
  compute_log_likelihood(C);
  R.lmin = 0;
  if (isfinite(R.ll)) {
    commit(C);
  } else {
    restore(C);
    R.alpha = al; R.alpha2 = al2;
#pragma unroll
    for (int q = 0; q < 4; ++q) R.ita2[q] = ita[q];
  }
}

template <int KMAX>
__device__ __forceinline__ void run_iterations(Ctx<KMAX>& C, ChainScal& CS, uint64_t n_iter) {
  Smem<KMAX>& S = C.S;
  Reg& R = C.R;
  const ChainConst& K = C.K;
  double* adapt = C.A.adapt;
  uint64_t it0 = CS.n_iter;
  uint64_t pr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint64_t iter = it0; iter < it0 + n_iter; ++iter) {
    // a chain whose current model turned non-finite stops here; the host
    // reports it (the oracle stops with an error at the same point)
    if (R.err & ERR_STOP) break;
    // a model within one iteration's growth of this kernel's capacity
    // continues in the 256-column kernel from this iteration, before any
    // draw of it, so the trajectory is the one a 256-column launch gives
    if (KMAX < BMG_KMAXG && R.c_cols + K.col_reserve > KMAX) {
      if (threadIdx.x == 0) {
        CS.big = 1;
#pragma unroll
        for (int q = 0; q < 8; ++q) CS.prof[q] += pr[q];
      }
      break;
    }
    const uint64_t t_it = wall_clock64();
    uint64_t t0 = t_it;
    if ((iter + 1) % K.n_st == 0) {
      refresh_tau2(C);
      if (!isfinite(R.c_ll)) R.err |= ERR_NONFINITE;
      uint64_t t1 = wall_clock64(); pr[0] += t1 - t0; t0 = t1;
    }
    // PMV move mix with one effect type (src/sampler.cpp:468-476); with a
    // learned table, a drawn swap is a learned swap with probability
    // p_learned (both swaps keep the posterior; the choice does not depend
    // on the state, and add/remove and state-change moves keep their share)
    int move;
    {
      const double pm[3] = {0.7, 0.7 + 0.15, 0.7 + 0.15 + 0.15};
      double u = rng_u01(R.rng);
      double r = u * pm[2];
      move = r < pm[0] ? 0 : (r < pm[1] ? 1 : 2);
    }
    if (move == 1 && R.learn_on && rng_u01(R.rng) < K.p_learned) move = 3;
    int jd = 0;
    if (move == 3) {
      if (R.c_nloci > 0) jd = move_learned_swap(C);
    } else if (move == 0) jd = move_addrem(C, adapt);
    else if (move == 1) {
      if (R.c_nloci > 0 && (uint64_t)R.c_nloci < C.D.p - 1) jd = move_switch(C, adapt);
    } else {
      if (R.c_nloci > 0 && C.D.p > 2) jd = move_statechange(C, adapt);
    }
    {
      const uint64_t t1 = wall_clock64(), dt = t1 - t0;
      if (move == 0) pr[1] += dt;
      else if (move == 1) pr[2] += dt;
      else if (move == 2) pr[3] += dt;
      else pr[4] += dt;
      t0 = t1;
    }
    if (threadIdx.x == 0) {
      CS.n_moves[move]++;
      CS.n_acpt[move] += jd > 0;
      CS.n_accepted += jd > 0;
    }
    if ((iter + 1) % K.thin == 0 || (iter + 2) % K.n_st == 0) {
      sample_beta_sigma2(C);
      uint64_t t1 = wall_clock64(); pr[5] += t1 - t0; t0 = t1;
    }
    if (K.trace_iter && threadIdx.x == 0) {
      uint64_t t = CS.tr_count;
      if (t < K.tr_cap) {
        C.A.tr_move[t] = (uint8_t)move; C.A.tr_jd[t] = (uint8_t)jd;
        C.A.tr_ms[t] = (uint8_t)R.movesize; C.A.tr_size[t] = (uint32_t)R.c_nloci;
        C.A.tr_ll[t] = R.c_ll;
      } else {
        R.err |= ERR_TRACE;
      }
      CS.tr_count = t + 1;
    }
    if ((iter + 1) % K.thin == 0) {
      double pves[3];
      compute_pve(C, pves);
      if (threadIdx.x == 0) {
        uint64_t t = CS.th_count;
        uint64_t ll0 = CS.th_loci_len;
        if (t < K.th_cap && ll0 + R.c_nloci <= K.th_loci_cap) {
          C.A.th_size[t] = (uint32_t)R.c_nloci;
          for (int q = 0; q < R.c_nloci; ++q) C.A.th_loci[ll0 + q] = S.c_loci[q];
          C.A.th_ll[t] = R.c_ll; C.A.th_s2[t] = R.sigma2; C.A.th_alpha[t] = R.alpha;
          C.A.th_pve[3 * t] = pves[0]; C.A.th_pve[3 * t + 1] = pves[1]; C.A.th_pve[3 * t + 2] = pves[2];
          CS.th_count = t + 1; CS.th_loci_len = ll0 + R.c_nloci;
          // save_beta (src/sampler.cpp:710-713): the beta drawn for this sample
          if (K.th_beta_cap) {
            const uint64_t b0 = CS.th_beta_len;
            if (b0 + R.beta_len <= K.th_beta_cap) {
              for (int q = 0; q < R.beta_len; ++q) C.A.th_beta[b0 + q] = S.beta[q];
              CS.th_beta_len = b0 + R.beta_len;
            } else {
              R.err |= ERR_TRACE;
            }
          }
        } else {
          R.err |= ERR_TRACE;
        }
        CS.pves[0] = pves[0]; CS.pves[1] = pves[1]; CS.pves[2] = pves[2];
      }
      uint64_t t1 = wall_clock64(); pr[6] += t1 - t0; t0 = t1;
    }
    pr[7] += wall_clock64() - t_it;
    if ((iter + 1) % K.n_rao == 0 || iter + 1 == it0 + n_iter) {
      if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) CS.prof[q] += pr[q];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) pr[q] = 0;
    }
    if ((iter + 1) % K.n_rao == 0) {
      // decide the RB block now; the scan and proposal update run as
      // separate kernels before the next iteration (src/sampler.cpp:731-811)
      if (threadIdx.x == 0) {
        CS.rb_pending = 1;
        CS.rb_do_scan = (!K.flat || CS.n_rao_burnin <= 0);
        if (CS.rb_do_scan) { CS.rb_call_idx = CS.rb_calls; CS.rb_calls++; }
        if (CS.n_rao_burnin <= 0) {
          CS.rb_accum_rao = 1;
          CS.rao_z1 = (double)(CS.p_rao_n + 1);
          CS.rao_z2 = (double)CS.p_rao_n / CS.rao_z1;
          CS.p_rao_n++;
        } else {
          CS.rb_accum_rao = 0;
          CS.n_rao_burnin--;
        }
        CS.rb_update_prop = 0; CS.rb_rebuild = 0;
        if (K.adaptation || CS.n_rao_burnin > 0) {
          if (!K.flat) {
            CS.rb_update_prop = 1;
            CS.prop_z1 = (double)(CS.p_proposal_n + 1);
            CS.prop_z2 = (double)CS.p_proposal_n / CS.prop_z1;
          }
          if (K.adapt_p_ms) {
            if (K.acpt_goal > 0) {
              double t = fmax(1.0, K.k_move_size * (double)CS.p_proposal_n);
              double ar = (double)CS.n_acpt[0] / (double)CS.n_moves[0];
              double pm2 = CS.p_ms + (K.acpt_goal - ar) / (2.0 + t);
              CS.p_ms = fmax(fmin(pm2, 0.99), 0.01);
              CS.n_acpt[0] = 1; CS.n_moves[0] = 2;
            } else {
              double pms = CS.p_ms;
              adapt_jd_mb(&pms, adapt, (int)K.max_ms);
              CS.p_ms = pms;
            }
            geometric_cdf((int)K.max_ms, CS.p_ms, adapt);
          }
          CS.p_proposal_n++;
          if (!K.flat) CS.rb_rebuild = 1;
        }
      }
      if (threadIdx.x == 0) CS.n_iter = iter + 1;
      __syncthreads();
      return;   // the RB block runs before the next iteration
    }
    if (threadIdx.x == 0) CS.n_iter = iter + 1;
  }
}

template <int KMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_chains(KernelArgs args) {
  __shared__ Smem<KMAX> S;
  int c = blockIdx.x;
  if (c >= args.n_chains) return;
  ChainScal& CS = args.scal[c];
  const ChainConst& K = args.consts[c];
  const DevData& D = args.data[K.ds];
  const ChainArrays& A = args.arrays[c];
  if constexpr (KMAX > KMAX_LDS) {
    if (threadIdx.x == 0) S.L = A.Lwork;
    __syncthreads();
  }
  Reg R;
  Ctx<KMAX> C{S, R, D, K, A, CS.prof};
  if (args.mode == 1) {
    // setup (src/sampler.cpp:591-620): state was initialised on the host
    // with an empty model; compute its factor, then sample
    load_chain(C, CS);
    compute_log_likelihood(C);
    commit(C);
    sample_beta_sigma2(C);
    refresh_tau2(C);
    store_chain(C, CS);
    return;
  }
  // role 1 fixes the segment's end for every chain; role 2 runs the chains
  // that left the role-1 kernel's capacity, up to that end
  if (args.role == 1 && threadIdx.x == 0) CS.seg_end = CS.n_iter + args.n_iter;
  if (CS.rb_pending) return;   // the RB block must run first
  uint64_t n_iter = args.n_iter;
  if (args.role == 1 && CS.big) return;
  if (args.role == 2) {
    if (!CS.big || CS.seg_end <= CS.n_iter) return;
    n_iter = CS.seg_end - CS.n_iter;
  }
  load_chain(C, CS);
  run_iterations(C, CS, n_iter);
  store_chain(C, CS);
  if (args.role == 2 && threadIdx.x == 0) {
    // back to the launch capacity once the model is well inside it and
    // every column's slot exists there
    int top = 0;
    for (int i = 0; i < R.c_cols; ++i) top = max(top, (int)S.c_slot[i]);
    if (!(R.err & ERR_STOP) && R.c_cols + K.col_reserve + 8 <= args.small_kmax && top < args.small_kmax)
      CS.big = 0;
  }
}

// explicit instantiations used by the host
#if BMG_INST(64)
template __global__ void k_chains<64>(KernelArgs);
#endif
#if BMG_INST(128)
template __global__ void k_chains<128>(KernelArgs);
#endif
#if BMG_INST(160)
template __global__ void k_chains<160>(KernelArgs);
#endif
#if BMG_INST(256)
template __global__ void k_chains<256>(KernelArgs);
#endif

// ------------------------------------------------------- model op kernel --

template <int KMAX>
__device__ __forceinline__ void swapadj_w0(Smem<KMAX>& S, int m, int col) {
  if (threadIdx.x == 0) {
    // src/symmmatrix.cpp:220-265 on the packed factor and v
    for (int i = 0; i < col; ++i) {
      double t = S.L[P(col) + i]; S.L[P(col) + i] = S.L[P(col + 1) + i]; S.L[P(col + 1) + i] = t;
    }
    double a = S.L[P(col + 1) + col];   // old (col, col+1)
    double b = S.L[P(col + 1) + col + 1];
    double dcc = S.L[P(col) + col];
    S.L[P(col + 1) + col] = dcc;         // column col+1 gets old column col
    S.L[P(col + 1) + col + 1] = 0.0;
    double c, s;
    drotg(a, b, c, s);
    if (a < 0) { a = -a; s = -s; c = -c; }
    S.L[P(col) + col] = a;
    for (int i = col + 1; i < m; ++i) {
      double x1 = S.L[P(i) + col], x2 = S.L[P(i) + col + 1];
      double tmp = x1 * s - x2 * c;
      S.L[P(i) + col] = x2 * s + x1 * c;
      S.L[P(i) + col + 1] = tmp;
    }
    double x1 = S.v[col], x2 = S.v[col + 1];
    double tmp = x1 * s - x2 * c;
    S.v[col] = x2 * s + x1 * c;
    S.v[col + 1] = tmp;
  }
  __syncthreads();
}

template <int KMAX>
__global__ __launch_bounds__(256) void k_model_op(ModelOpArgs a) {
  __shared__ Smem<KMAX> S;
  ChainScal& CS = a.scal[0];
  const ChainConst& K = a.consts[0];
  const DevData& D = a.data[K.ds];
  const ChainArrays& A = a.arrays[0];
  if constexpr (KMAX > KMAX_LDS) {
    if (threadIdx.x == 0) S.L = A.Lwork;
    __syncthreads();
  }
  Reg R;
  Ctx<KMAX> C{S, R, D, K, A, nullptr};
  load_chain(C, CS);
  R.err &= ~ERR_CAPACITY;   // a capacity error belongs to the call that hit it
  if (a.op == 0) {
    double d2[2] = {a.d0, a.d1};
    add_term(C, a.ind, a.t_ind, d2);
  } else if (a.op == 1) {
    remove_term(C, a.model_ind);
  } else if (a.op == 2) {
    compute_log_likelihood(C);
  } else if (a.op == 3) {
    swapadj_w0(S, R.cols, a.model_ind);
    R.lmin = a.model_ind;
  } else if (a.op == 4) {
    sample_beta_sigma2(C);   // from the model's own stream
  } else if (a.op == 5) {
    double pv[3];
    compute_pve(C, pv);      // leaves y_hat = X beta
    if (threadIdx.x == 0) { CS.pves[0] = pv[0]; CS.pves[1] = pv[1]; CS.pves[2] = pv[2]; }
  }
  R.lmin = 0;
  commit(C);
  store_chain(C, CS);
}

#if BMG_INST(64)
template __global__ void k_model_op<64>(ModelOpArgs);
#endif
#if BMG_INST(128)
template __global__ void k_model_op<128>(ModelOpArgs);
#endif
#if BMG_INST(160)
template __global__ void k_model_op<160>(ModelOpArgs);
#endif
#if BMG_INST(256)
template __global__ void k_model_op<256>(ModelOpArgs);
#endif

#if BMG_INST(64)
namespace bmg {

int launch_xy_all(const DevData& D, int coding, double* out, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<uint64_t>(D.p, 65536);
  hipLaunchKernelGGL(k_xy_all, dim3(grid), dim3(BMG_BLOCK), 0, s, D, coding, out);
  return hipGetLastError() == hipSuccess ? BMG_OK : BMG_E_HIP;
}

}  // namespace bmg
#endif
