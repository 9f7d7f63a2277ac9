/*
 * oracle.h -- CPU restatement of BMAGWA's MCMC step over SNP inclusion.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP engine in bmagwa_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates the reference algorithm (to-mi/bmagwa v2.0) and
 * cites the file:line it follows.  Two things are deliberately *not* the
 * reference's, because the reference's versions cannot be reproduced here:
 *   - the random stream (Boost.Random, unpinned version) is replaced by the
 *     project stream: Philox4x32-10 (Salmon et al., SC'11) with Box-Muller
 *     normals and Marsaglia-Tsang gammas;
 *   - the weighted sampler (threaded binary tree,
 *     src/discrete_distribution.hpp:64-330) is replaced by the project's
 *     64-ary sum tree; same distribution, different uniform->index map.
 * Everything downstream of a random draw is therefore "parity unpinned"
 * against the reference's exact trajectory; it is pinned against the HIP
 * engine (same stream) and statistically against exact enumeration.
 */
#ifndef BMAGWA_ORACLE_H_
#define BMAGWA_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* effect types, src/data_model.hpp:41 */
enum { ORC_A = 0, ORC_H = 1, ORC_D = 2, ORC_R = 3, ORC_AH = 4 };

/* ------------------------------------------------------------------ RNG */
typedef struct { uint32_t key[2]; uint32_t ctr[4]; } orc_rng;

void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                         uint32_t out[4]);
/* stream = (key = {seed, chain}, ctr = {0, 0, c2, c3}); every draw uses one
 * Philox block and then increments the 64-bit (ctr[0], ctr[1]) */
void   orc_rng_init(orc_rng* r, uint32_t seed, uint32_t chain, uint32_t c2,
                    uint32_t c3);
double orc_rng_u01(orc_rng* r);
double orc_rng_normal(orc_rng* r);
double orc_rng_gamma(orc_rng* r, double a);
double orc_rng_sinvchi2(orc_rng* r, double nu, double s2);
/* x = scale * U^-1 z + mu, z standard normal, U upper triangular
 * (column-major, leading dimension ld), src/rand.hpp:144-168 */
void   orc_rng_mvnormal_invchol(orc_rng* r, const double* mu, const double* U, int ld, int d,
                                double scale, double* x);
/* fills out[0..cnt) with draws of one kind: 0 u01, 1 normal, 2 gamma(a),
 * 3 sinvchi2(a, b) */
void   orc_rng_fill(uint32_t seed, uint32_t chain, uint32_t c2, uint32_t c3,
                    int kind, double a, double b, int64_t cnt, double* out);

/* ------------------------------------------------------------ decoding */
/* value of one 2-bit code, src/data.cpp:36,40-54: 00->0 01->-1 10->1 11->2 */
double orc_get_genotype(const uint8_t* body, size_t stride, size_t ind,
                        size_t snp);
/* one SNP, one coding (A,H,D,R), src/data.cpp:56-138 */
void   orc_get_genotypes(const uint8_t* body, size_t stride, uint32_t n,
                         int coding, size_t snp, double* out);

/* ------------------------------------------------------------- options */
/* mirrors the INI keys of src/options.hpp:138-321 */
typedef struct {
  uint32_t n; uint32_t m_e; uint64_t m_g;
  uint64_t do_n_iter;
  uint32_t n_rao, n_rao_burnin, thin, n_sample_tau2_and_missing;
  uint32_t max_move_size, max_SNP_neighborhood_size, delay_rejection;
  uint32_t seed;
  double   p_move_size, p_move_size_nbs, p_move_size_nbc;
  double   p_move_size_acpt_goal;
  int32_t  adapt_p_move_size, adaptation, flat_proposal_dist, save_beta;
  int32_t  sampler_type;          /* 0 = PMV */
  int32_t  n_types; int32_t types[5];
  double   e_qg, var_qg, nu_sigma2, R2mode_sigma2, s2_sigma2, mu_alpha;
  double   inv_tau2_e_const_val, inv_tau2_e_val;
  int32_t  use_individual_tau2;
  double   types_prior[5];
  double   nu_tau2[4], eh_tau2[4], s2_tau2[4];
  int32_t  recode_g_to_minor_allele_count;
  uint32_t chain;                 /* chain index (Philox key word 1) */
} orc_opts;

/* ---------------------------------------------------------------- data */
typedef struct orc_data orc_data;
/* body: p*ceil(n/4) bytes after the 3-byte header; e_cov: n x m_e
 * column-major (NULL if m_e == 0); the constant column is added here as the
 * reference does (src/data.hpp:48,53).  The body is copied. */
orc_data* orc_data_create(uint32_t n, uint64_t m_g, uint32_t m_e,
                          const uint8_t* body, const double* y,
                          const double* e_cov, int recode);
void      orc_data_destroy(orc_data* d);
/* header check + body read, src/data.cpp:245-273; returns 0 on success */
int       orc_read_bed(const char* path, uint32_t n, uint64_t m_g,
                       uint8_t* body_out);
const uint8_t* orc_data_body(const orc_data* d);
void      orc_data_stats(const orc_data* d, double out[4]); /* yy var_y var_x mean_x */
/* missing-genotype index, src/data.cpp:341-376: returns count, fills locs
 * (up to cap) and the cumulative prior counts */
int       orc_data_missing(const orc_data* d, size_t snp, size_t* locs,
                           int cap, double prior_cum[3]);

/* --------------------------------------------------------- data model */
typedef struct orc_datamodel orc_datamodel;
orc_datamodel* orc_datamodel_create(const orc_data* d, const int32_t* types,
                                    int n_types);
void   orc_datamodel_destroy(orc_datamodel* dm);
void   orc_datamodel_get(const orc_datamodel* dm, int coding, size_t snp,
                         double* out);
void   orc_datamodel_set_missing(orc_datamodel* dm, size_t snp, int j,
                                 int val);
int    orc_datamodel_offset(const orc_datamodel* dm, int offset_type[4]);
void   orc_datamodel_update_prexx(const orc_datamodel* dm, size_t snp,
                                  double* pre_xx);
/* PrecomputedSNPCovariances::precompute, one row of `offset` per SNP */
void   orc_self_moments(const orc_datamodel* dm, double* out);

/* band X^T X (no reference counterpart): out[j*W + d] = sum_i x_j(i) *
 * x_{j+d}(i) for d < W, j + d < p, exact int64; missing codes count as 0 */
void   orc_band_xtx(const uint8_t* body, size_t stride, uint32_t n,
                    uint64_t p, uint32_t W, int coding_a, int coding_b,
                    int64_t* out);

/* ---------------------------------------------------------- linalg */
/* upper Cholesky factor in place (column-major, ld); 0 if not pos. def. */
int  orc_chol(double* a, int ld, int m);
/* bordered update, src/symmmatrix.cpp:144-164 */
int  orc_chol_update(double* l, int ld, int m, const double* newcol,
                     double inv_tau2);
/* dchex job=2 + drop last + sign fix, src/symmmatrix.cpp:166-195 */
void orc_chol_downdate(double* l, int ld, int m, int col_rem);
/* adjacent swap, src/symmmatrix.cpp:220-265; v may be NULL */
void orc_chol_swapadj(double* l, int ld, int m, int col, double* v);
void orc_drotg(double* a, double* b, double* c, double* s);

/* ---------------------------------------------------------- sampler */
typedef struct orc_sampler orc_sampler;
orc_sampler* orc_sampler_create(const orc_opts* o, const orc_data* d);
void   orc_sampler_destroy(orc_sampler* s);
void   orc_sampler_init_p_proposal_flat(orc_sampler* s);
/* run do_n_iter iterations (src/sampler.cpp:551-880) */
int    orc_sampler_sample(orc_sampler* s, uint64_t n_iter);
/* run only the start-of-sample() setup (src/sampler.cpp:591-620) */
int    orc_sampler_setup(orc_sampler* s);
/* learned swap: frozen partner weights w[p][2H] (column c <-> offset
 * c < H ? c - H : c - H + 1); a swap drawn by the plain move draw becomes a
 * learned swap with probability p_learned; w = NULL switches it off */
int    orc_sampler_set_learned(orc_sampler* s, const float* w, int H, double p_learned);
/* test knob: probabilities of the add/remove, swap and state-change moves
 * (normally 0.7 / 0.15 / 0.15, src/sampler.cpp:468-476); call after setup */
int    orc_sampler_set_move_mix(orc_sampler* s, const double p[3]);
/* run iterations after setup (the loop body of src/sampler.cpp:626-834) */
int    orc_sampler_iterate(orc_sampler* s, uint64_t n_iter);
const double* orc_sampler_p_rao(const orc_sampler* s);
const double* orc_sampler_p_proposal(const orc_sampler* s);
/* replace p_proposal (each in [0, 1]) and rebuild q_add / q_rem and the
 * proposal trees from it; -1 on a value outside [0, 1] */
int    orc_sampler_set_p_proposal(orc_sampler* s, const double* p);
/* per-iteration trace (length = iterations run): move, jumpdistance,
 * movesize, model size after the move, log-likelihood after the move */
int64_t orc_sampler_trace(const orc_sampler* s, uint8_t* move, uint8_t* jd,
                          uint8_t* ms, uint32_t* msize, double* loglik);
/* thinned trace: model size and loci, flattened */
int64_t orc_sampler_thinned(const orc_sampler* s, uint32_t* sizes,
                            uint32_t* loci, int64_t loci_cap, double* ll,
                            double* sigma2, double* alpha);
/* PVE of every thinned sample (3 values each, src/model.hpp:345-392) */
int64_t orc_sampler_thinned_pve(const orc_sampler* s, double* pve);
/* save_beta: every thinned sample's beta, concatenated (src/sampler.cpp:
 * 710-713); returns the number of doubles kept */
int64_t orc_sampler_thinned_beta(const orc_sampler* s, double* out, int64_t cap);
void   orc_sampler_state(const orc_sampler* s, double out[8]);
int    orc_sampler_loci(const orc_sampler* s, uint32_t* loci);
/* RB scan on the current model with a given y_hat, src/sampler.cpp:32-261 */
void   orc_sampler_rb(orc_sampler* s, const double* y_hat, double* p_r);
/* expose the chain's models for Model-level tests */
typedef struct orc_model orc_model;
orc_model* orc_sampler_current_model(orc_sampler* s);
orc_model* orc_sampler_new_model(orc_sampler* s);
uint64_t   orc_sampler_rng_counter(const orc_sampler* s);

/* ------------------------------------------------------------ model */
int    orc_model_add_term(orc_model* m, uint32_t ind, int t_ind,
                          const double* inv_tau2_alpha2);
void   orc_model_remove_term(orc_model* m, int model_ind);
void   orc_model_compute_log_likelihood(orc_model* m);
double orc_model_log_likelihood(const orc_model* m);
int    orc_model_size(const orc_model* m);
int    orc_model_cols(const orc_model* m);
void   orc_model_copy(orc_model* dst, const orc_model* src);
void   orc_model_get(const orc_model* m, double* xx, double* l, double* xy,
                     double* v, double* d, double scal[6]);
void   orc_model_set_d(orc_model* m, int x_ind, double val);
void   orc_model_sample_beta_sigma2(orc_sampler* s, orc_model* m);
/* N draws of (sigma2, beta) from the sampler's stream on a fixed model:
 * s2_out[N], beta_out[N x k]; returns k */
int64_t orc_model_draw_beta_sigma2(orc_sampler* s, orc_model* m, int64_t N, double* s2_out,
                                   double* beta_out);
/* N independent tau2/alpha refreshes (src/prior.cpp:30-141) of the current
 * model, each from the same starting state: alpha_out[N], and the terms'
 * inv_tau2_alpha2 after each, dv_out[N x k]; returns k.  The state is
 * restored afterwards. */
int64_t orc_sampler_refresh_draws(orc_sampler* s, int64_t N, double* alpha_out, double* dv_out);
void   orc_model_compute_pve(orc_model* m, double pves[3]);
const double* orc_model_beta(const orc_model* m);
const double* orc_model_y_hat(const orc_model* m);
const int32_t* orc_model_model_inds(const orc_model* m);
/* ExhModel + DR helpers, src/model.hpp:585-844, src/sampler.cpp:882-1049 */
void   orc_exhaustive_modelset(orc_model* full, int const_loci, int n_inds,
                               double* log_model_probs, double* max_log);
void   orc_proposal_probs_exh(int use_types, int n_inds,
                              const unsigned char* bit_to_normalized_order,
                              const double* q_add, const double* q_rem,
                              double z_add, double z_rem, size_t const_loci,
                              size_t m_g, const double* log_q_add_types,
                              double* log_prop_probs);

/* ------------------------------------------------------------ prior */
double orc_prior_change_on_add(const orc_sampler* s, const int* Ns,
                               int n_loci_old, int type);
double orc_prior_change_on_rem(const orc_sampler* s, const int* Ns,
                               int n_loci_old, int type);
double orc_prior_change_on_swi(const orc_sampler* s, const int* Ns,
                               int type_add, int type_rem);
double orc_prior_log_model(const orc_sampler* s, const int* Ns);
void   orc_prior_get(const orc_sampler* s, double out[8]);

/* ------------------------------------------------- sum-tree sampler */
typedef struct orc_sumtree orc_sumtree;
orc_sumtree* orc_sumtree_create(const double* w, uint64_t n);
void   orc_sumtree_destroy(orc_sumtree* t);
void   orc_sumtree_zero(orc_sumtree* t, uint64_t i);    /* "adddate" */
void   orc_sumtree_unzero(orc_sumtree* t, uint64_t i);  /* "remdate" */
void   orc_sumtree_set_weights(orc_sumtree* t, const double* w);
double orc_sumtree_total(const orc_sumtree* t);
uint64_t orc_sumtree_pick(const orc_sumtree* t, double u01);
int    orc_sumtree_zeroed(const orc_sumtree* t, uint64_t i);

/* ----------------------------------------------- synthetic datasets */
/* latent AR(1) haplotypes thresholded per SNP (see orc_synth.c); thresh[j]
 * = Phi^-1(MAF_j), rho[j] = AR coefficient linking SNP j to j-1 (0 starts a
 * block); writes p*ceil(n/4) bytes in the reference's code mapping */
void   orc_synth_genotypes(uint32_t seed, uint32_t n, uint64_t p,
                           const double* thresh, const double* rho,
                           uint8_t* body);

/* ------------------------------------------------------- utilities */
void   orc_geometric_dist_cdf(int maxsize, double p, double* values);
/* move-size adaptation by expected jump distance, src/sampler.hpp:1728-1769
 * (updates q_p0[max_ms], returns the new p_move_size) */
double orc_adapt_jd_mb(double p_ms, const double* r_ms, const double* r_n, double* q_p0, int max_ms);
int    orc_sample_discrete_naive(const double* cumsum, int m, double u01);
double orc_ess_geyer(const double* x, int64_t n);

/* ------------------------------------------- learned-proposal network --- */
/* 16 -> 256 -> 256 -> 1 MLP, parameters in one fp32 array:
 * W1[16][256], b1[256], W2[256][256], b2[256], w3[256], b3 */
#define ORC_MLP_IN 16
#define ORC_MLP_HID 256
#define ORC_MLP_OFF_W1 0
#define ORC_MLP_OFF_B1 (ORC_MLP_IN * ORC_MLP_HID)
#define ORC_MLP_OFF_W2 (ORC_MLP_OFF_B1 + ORC_MLP_HID)
#define ORC_MLP_OFF_B2 (ORC_MLP_OFF_W2 + ORC_MLP_HID * ORC_MLP_HID)
#define ORC_MLP_OFF_W3 (ORC_MLP_OFF_B2 + ORC_MLP_HID)
#define ORC_MLP_OFF_B3 (ORC_MLP_OFF_W3 + ORC_MLP_HID)
#define ORC_MLP_NPAR (ORC_MLP_OFF_B3 + 1)
float  orc_bf16_round(float f);
/* learned-proposal inputs (see orc_learn.c / orc_sampler.c) */
void   orc_learn_pair_features(const double* pre, int off, int ot, uint32_t n,
                               const int32_t* band, int H, const float* z, const float* pbar,
                               uint64_t j, int o, float* f);
/* features of SNP j in a chain's state (model_inds, p_proposal) for the
 * per-SNP proposal network */
void   orc_learn_snp_features(const double* pre, int off, int ot, uint32_t n, uint64_t p,
                              const int32_t* band, int H, const float* z, const float* pbar,
                              const int32_t* model_inds, const double* p_proposal, uint64_t j,
                              float* f);
void   orc_sampler_marginal_z(orc_sampler* s, float* z);
int64_t orc_sampler_swap_labels(orc_sampler* s, int H, uint32_t* oj, int32_t* oo, float* olab,
                                int64_t cap);
void   orc_mlp_forward(const float* params, const float* X, int64_t B, float* out);
/* gradients of the mean Huber loss with the device's bf16 rounding points;
 * returns the loss */
double orc_mlp_gradients(const float* params, const float* X, const float* labels, int64_t B,
                         double* grads);
/* the same loss and gradients in double without any rounding */
double orc_mlp_loss_exact(const float* params, const float* X, const float* labels, int64_t B,
                          double* grads);
void   orc_mlp_adam(float* params, float* m, float* v, const float* grads, float lr, int t);

#ifdef __cplusplus
}
#endif
#endif /* BMAGWA_ORACLE_H_ */
