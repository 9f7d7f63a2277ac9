// bmg_chain.h -- structs of the batched chain engine shared by the device
// code (bmg_chain.hip) and the host side (bmg_capi.hip).
#pragma once
#include <stdint.h>

#define BMG_MAXMS 64     // max_move_size supported on the device
#define BMG_KMAXG 256    // largest compiled model capacity (columns)
#define BMG_TREE_LEV 8
#define BMG_LH_MAX 32    // largest learned-swap window
#define BMG_DR_MAX 20    // largest delayed-rejection move size (2^20 sub-models)
#define BMG_DR_COLS (2 * BMG_DR_MAX)   // columns of the toggled loci (AH: two each)

namespace bmg {

// One dataset as the chain kernel sees it.
struct DevData {
  const uint32_t* g;     // p x W32 packed words
  const double* y;       // n_pad
  const double* pre;     // p x offset self moments
  uint64_t p;
  uint32_t n, W32, n_pad;
  int offset;
  int ot[4];             // offset_type
  int type;              // the sampler's single effect type (A, H, D, R or AH)
  int n_types;
  int types[5];          // DataModel::types (Model API terms may mix them)
  int top;               // sum-tree top level
  uint64_t lev_len[BMG_TREE_LEV], lev_off[BMG_TREE_LEV];
  uint64_t tree_size;    // doubles per tree
  uint64_t lds_lev_size; // doubles of levels >= 2
  const float* lw;       // learned-swap weights p x 2*lH (null: none)
  int lH;                // learned-swap window (SNPs each side, <= 32)
  const double* xy[4];   // x_j . y of every SNP per coding (null: computed per add)
  int mixed;             // columns may carry codings other than `type` (Model API)
  int me;                // fixed columns: the constant plus the covariates
  const double* E;       // covariate c >= 1 at E + (c - 1) * n_pad (zero past n)
};

// Per-chain constants (options + solved prior, src/sampler.hpp:107-203,
// src/prior.hpp:45-122).
struct ChainConst {
  uint32_t seed, chain_id, ds;
  uint32_t thin, n_rao, n_st, max_ms, max_nbh;
  int32_t adapt_p_ms, adaptation, flat, trace_iter;
  double acpt_goal, q_add_min, q_rem_min, k_move_size;
  double n_plus_nu, minus_n_plus_nu_2, nus2_plus_yy, mu_alpha;
  double nus2;           // nu_sigma2 * s2_sigma2 alone (the direct residual term)
  double g_a, g_b, m_g, tp_type, tp_sum;
  double nu_tau2[4], s2_tau2[4], nus2_tau2[4];
  int32_t allow_terms[4];
  int32_t use_ind;
  double inv_tau2_e0;
  uint64_t tr_cap, th_cap, th_loci_cap;
  uint64_t th_beta_cap;  // save_beta: doubles of thinned beta kept (0: not saved)
  double p_learned;      // probability of a learned swap per iteration
  int32_t dr;            // delayed-rejection move-size limit (0: none)
  int32_t max_size;      // largest model (SNPs) with nonzero prior
  int32_t col_reserve;   // most columns one iteration can add to the current model
};

// Mutable per-chain scalars (device global memory).
struct ChainScal {
  uint32_t rng_c0, rng_c1;
  uint64_t n_iter, n_accepted;
  double alpha, alpha2, inv_tau2_alpha2[4];
  int64_t n_rao_burnin, p_rao_n;
  uint64_t p_proposal_n;
  double p_ms;
  uint64_t n_acpt[4], n_moves[4];   // add/remove, swap, state change, learned swap
  int32_t learn_on;
  // wall-clock ticks per phase: tau2/alpha refresh, add/remove, swap, state
  // change, learned swap, beta/sigma2, pve + thinning, whole iterations;
  // then inside additions: x.y, X^T x, triangular solve, the rest of the
  // update, and the count of additions and removals, time in removals,
  // tau2 draws of the refresh, the downdate and the solve of removals, whole
  // delayed-rejection second stages and their scoring (walk, proposal
  // probabilities, draw); then inside add/remove moves: the proposal draws
  // (sum-tree picks), the removals and additions with their tree updates,
  // the reverse move's weights, and the accept / reject handling (journal,
  // commit or restore)
  uint64_t prof[24];
  uint32_t rb_calls;
  int32_t movesize;
  double pves[3];
  // current model scalars
  int32_t cols, nloci, beta_len, mu_computed;
  int32_t Ns[5];
  double sigma2, syx, ldq, ldl, ll;
  // RB block decided at the end of an RB iteration (src/sampler.cpp:731-811)
  int32_t rb_pending, rb_do_scan, rb_accum_rao, rb_update_prop, rb_rebuild;
  double rao_z1, rao_z2, prop_z1, prop_z2;
  uint32_t rb_call_idx;
  int32_t err;
  uint64_t tr_count, th_count, th_loci_len, th_beta_len;
  double tot_add, tot_rem;   // sum-tree totals
  // capacity: a chain whose model comes within col_reserve columns of the
  // launch's capacity continues in the 256-column kernel (big = 1) from the
  // same iteration; seg_end is the iteration the current segment ends at
  int32_t big;
  uint64_t seg_end;
};

// Per-chain device arrays.  Every per-column array is laid out for
// BMG_KMAXG columns whatever the launch's capacity, so a chain's state is
// read the same way by every compiled capacity.
struct ChainArrays {
  int32_t* model_inds;       // p
  double* q_add; double* q_rem;          // p
  double* tree_add; double* tree_rem;    // tree_size each
  double* p_proposal; double* p_rao; double* p_r;  // p
  double* Lcur;              // packed upper factor, BMG_KMAXG (BMG_KMAXG+1) / 2
  double* Lwork;             // the same size: working factor of the 256-column kernel
  double* xxs;               // BMG_KMAXG x BMG_KMAXG slot matrix
  uint32_t* xcol;            // BMG_KMAXG x W32: packed genotypes of each slot's column
  double* y_hat;             // n_pad
  double* vec;               // 5 x BMG_KMAXG: xy, v, d, beta, mu
  int32_t* idx;              // 8 x BMG_KMAXG: loci, x_ind1, x_ind2, slot, snp, xtype, tind, -
  double* adapt;             // 6 x 256: q_pms, q_pms_nbs, q_pms_nbc, r_ms_sum, r_ms_n, q_p0_ms
  uint8_t* tr_move; uint8_t* tr_jd; uint8_t* tr_ms; uint32_t* tr_size; double* tr_ll;
  uint32_t* th_size; uint32_t* th_loci; double* th_ll; double* th_s2; double* th_alpha;
  double* th_pve;        // 3 per thinned sample (src/model.hpp:345-392)
  double* th_beta;       // save_beta: beta of every thinned sample, concatenated
  double* dr_lp;         // 2^dr second-stage log-probabilities (null: dr == 0)
};

// error flags of a chain; the flags in ERR_STOP stop the chain
enum {
  ERR_CAPACITY = 1,      // a proposal needed more than BMG_KMAXG columns
  ERR_TRACE = 2, ERR_TREE = 4,
  ERR_NONFINITE = 8,     // tau2/alpha refresh left a non-finite log-likelihood
  ERR_SIGMA2 = 16,       // sigma2 draw not finite and positive
  ERR_ALPHA = 32,        // alpha draw did not finish
  ERR_WALK = 64,         // a neighbourhood walk left the genome
  ERR_STOP = ERR_CAPACITY | ERR_NONFINITE | ERR_SIGMA2 | ERR_ALPHA | ERR_WALK
};

}  // namespace bmg

// launch arguments of the chain kernels (bmg_chain.hip)
struct KernelArgs {
  const bmg::DevData* data;
  const bmg::ChainConst* consts;
  bmg::ChainScal* scal;
  const bmg::ChainArrays* arrays;
  int n_chains;
  uint64_t n_iter;     // iterations to run in this launch
  int mode;            // 0 iterate, 1 set-up
  // 0: the only chain kernel of the segment; 1: the launch capacity's kernel,
  // followed by 2: the 256-column kernel, which runs the chains with big = 1
  // up to their seg_end
  int role;
  int small_kmax;      // role 2: the capacity of the role-1 kernel
};

// one Model method on a one-chain state (the drop-in Model API)
struct ModelOpArgs {
  const bmg::DevData* data;
  const bmg::ChainConst* consts;
  bmg::ChainScal* scal;
  const bmg::ChainArrays* arrays;
  int op;              // 0 add, 1 remove, 2 compute_log_likelihood, 3 swapadj,
                       // 4 sample_beta_sigma2, 5 compute_pve
  uint32_t ind;
  int t_ind, model_ind;
  double d0, d1;
};

namespace bmg {

// out[j] = x_j . y for every SNP j of D under `coding`, summed in exactly
// the order the chain kernel's per-add dot uses, so the chain reads the
// same double it would have computed
int launch_xy_all(const DevData& D, int coding, double* out, hipStream_t s);

}  // namespace bmg
