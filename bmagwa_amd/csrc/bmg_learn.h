// bmg_learn.h -- online training of the learned swap proposal of a batch
// (one network and one weight table per dataset).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bmg_chain.h"
#include "bmg_internal.h"

#define LEARN_FEATURES 16

namespace bmg {

struct LearnDS;

struct LearnCfg {
  int H;              // window: partners within H SNPs (<= 32)
  int steps;          // Adam steps per training round
  int batch;          // rows per Adam step
  float lr;           // Adam learning rate
  uint32_t seed;      // minibatch sampling and initial weights
  int swap;           // train and use the learned swap's partner table
  int snp;            // train the per-SNP network and, at the freeze, set
                      // each chain's add/remove proposal table from it
};

// band X^T X (window H + 1), marginal scores and buffers for dataset D
int learn_create(LearnDS** out, bmg_ctx* ctx, const DevData& D, const LearnCfg& cfg, int n_chains,
                 hipStream_t s);
void learn_destroy(LearnDS* L);
float* learn_table(LearnDS* L);   // p x 2H partner weights (device)
int learn_H(const LearnDS* L);
int learn_frozen(const LearnDS* L);
void learn_freeze(LearnDS* L);
uint64_t learn_rounds(const LearnDS* L);

// Swap labels from one RB group's scan: for every model SNP j of every chain
// in list[0..nv) and every SNP k within H of j outside the model,
//   label = lbf(k | gamma - j) - lbf(j | gamma - j)
// with lbf the single-SNP log Bayes factor given sigma2, the shared prior
// precision and the other effects, appended to the dataset's sample ring.
int learn_labels(LearnDS* L, const DevData* d_data, const ChainConst* consts, const ChainScal* scal,
                 const ChainArrays* arrays, const int* d_list, int nv, int kmax, const double* xa,
                 const double* rsum, hipStream_t s);

// One training round: pbar = mean p_proposal of the dataset's chains, Adam
// steps on minibatches drawn from the ring, then the table
// w[j][c] = exp(clamp(net(features(j, j + off(c))), -30, 30) / 2) for every
// pair (0 where the partner is outside the genome).
int learn_round(LearnDS* L, const ChainArrays* d_arrays, const int* chains_host, int n, hipStream_t s);

// Per-SNP proposal samples from one RB group's scan (after the RB
// probabilities, before the p_proposal update): for every chain in
// list[0..nv), the SNPs within the window of each model SNP and LEARN_SRAND
// uniformly drawn SNPs, each with its features in the chain's state and the
// scan's RB probability (clipped logit) as the label.
int learn_snp_labels(LearnDS* L, const ChainScal* scal, const ChainArrays* arrays, const int* d_list, int nv,
                     int kmax, hipStream_t s);
// At the freeze: each chain's table p_j = logistic(net(features of j in the
// chain's state)) becomes its p_proposal (q_add, q_rem and trees follow);
// with adaptation over it stays fixed for the rest of the run.
int learn_snp_apply(LearnDS* L, const DevData* d_data, const ChainConst* d_consts, ChainScal* d_scal,
                    const ChainArrays* d_arrays, const int* chains_host, int n, const uint64_t* lev_len, int top,
                    hipStream_t s);
int learn_snp_applied(const LearnDS* L);
// device-side per-SNP features of SNPs j in chain `chain`'s state, for
// checking: out rows x 16
int learn_snp_features(LearnDS* L, const ChainArrays* d_arrays, int chain, const uint32_t* j_host, int64_t rows,
                       float* out, hipStream_t s);

// device-side features of pairs (j, offset) for checking: out rows x 16
int learn_features(LearnDS* L, const uint32_t* j_host, const int32_t* off_host, int64_t rows, float* out,
                   hipStream_t s);
// the sample ring: copies up to cap samples, returns the number written
int learn_samples(LearnDS* L, uint32_t* j, int32_t* off, float* label, int64_t cap, int64_t* count,
                  hipStream_t s);
// the network parameters (MLP_NPAR floats) and the static per-SNP arrays
int learn_params(LearnDS* L, float* params, hipStream_t s);
int learn_static(LearnDS* L, float* z, float* pbar, int32_t* band, hipStream_t s);

}  // namespace bmg
