/*
 * bmagwa_hip.h -- C ABI of the MI355X (gfx950) engine for BMAGWA's MCMC step
 * over SNP inclusion.
 *
 * Plain pointers and sizes only.  Pointers named *_dev are device pointers;
 * every other pointer is host memory.  Every call returns a status code
 * (BMG_OK = 0); bmg_last_error() gives the message of the last failure on the
 * calling thread.  A non-positive-definite Cholesky factor is not an error:
 * as in the reference it yields a log-likelihood of -INFINITY
 * (src/model.hpp:210-220,476-484).
 *
 * Handles: a bmg_ctx is one dataset resident in HBM (Data + DataModel +
 * PrecomputedSNPCovariances of the reference); a bmg_model is one FP64 linear
 * model state on the GPU (Model); a bmg_batch is a set of chains, possibly on
 * several datasets, advanced by one persistent launch (Sampler x chains).
 */
#ifndef BMAGWA_HIP_H_
#define BMAGWA_HIP_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BMG_OK = 0,
  BMG_E_ARG = -1,         /* invalid argument */
  BMG_E_HIP = -2,         /* HIP runtime error */
  BMG_E_UNSUPPORTED = -3, /* valid for the reference, not built for the GPU */
  BMG_E_CAPACITY = -4,    /* model larger than the compiled capacity */
  BMG_E_STATE = -5,       /* call out of order */
  BMG_E_NUMERIC = -6      /* a chain's current model became non-finite */
};

/* effect types, src/data_model.hpp:41 */
enum { BMG_A = 0, BMG_H = 1, BMG_D = 2, BMG_R = 3, BMG_AH = 4 };

typedef struct bmg_ctx bmg_ctx;
typedef struct bmg_model bmg_model;
typedef struct bmg_batch bmg_batch;
typedef struct bmg_mlp bmg_mlp;

const char* bmg_last_error(void);
int bmg_device_count(int* count);

/* Sampler options: the INI keys of src/options.hpp:138-321 (same field order
 * and meaning as the [sizes]/[sampler]/[thread]/[model]/[prior] sections). */
typedef struct {
  uint32_t n; uint32_t m_e; uint64_t m_g;
  uint64_t do_n_iter;
  uint32_t n_rao, n_rao_burnin, thin, n_sample_tau2_and_missing;
  uint32_t max_move_size, max_SNP_neighborhood_size, delay_rejection;
  uint32_t seed;
  double   p_move_size, p_move_size_nbs, p_move_size_nbc;
  double   p_move_size_acpt_goal;
  int32_t  adapt_p_move_size, adaptation, flat_proposal_dist, save_beta;
  int32_t  sampler_type;
  int32_t  n_types; int32_t types[5];
  double   e_qg, var_qg, nu_sigma2, R2mode_sigma2, s2_sigma2, mu_alpha;
  double   inv_tau2_e_const_val, inv_tau2_e_val;
  int32_t  use_individual_tau2;
  double   types_prior[5];
  double   nu_tau2[4], eh_tau2[4], s2_tau2[4];
  int32_t  recode_g_to_minor_allele_count;
  uint32_t chain;
} bmg_opts;

/* ------------------------------------------------------------ dataset --- */
/* Replaces Data(n, m_g, m_e, file_fam, file_g, ...) (src/data.hpp:45-72) with
 * the .bed body already read (src/data.cpp:245-273: p * ceil(n/4) bytes after
 * the 3-byte header) and y from the .fam/.y file; DataModel(data, types)
 * (src/data_model.hpp:52-89) and PrecomputedSNPCovariances(data_model)
 * (src/precomputed_snp_covariances.hpp:58-87) are built on the device.
 * The constant column is always present; covariates are added with
 * bmg_ctx_set_covariates.  recode != 0 applies
 * recode_g_to_minor_allele_count (src/data.cpp:324-339). */
int bmg_ctx_create(int device, const uint8_t* bed_body, uint32_t n, uint64_t p,
                   const double* y, const int32_t* types, int n_types, int recode,
                   bmg_ctx** out);
/* Same, genotypes generated on the device by the synthetic-data model of the
 * benchmark (latent AR(1) haplotypes, thresh = Phi^-1(MAF), rho_j links SNP j
 * to j-1, 0 starts a block); y is supplied by the caller afterwards. */
int bmg_ctx_create_synthetic(int device, uint32_t seed, uint32_t n, uint64_t p,
                             const double* thresh, const double* rho,
                             const int32_t* types, int n_types, bmg_ctx** out);
/* A phenotype measured on a subset of a pool's individuals (no counterpart
 * in the reference, where each such phenotype is its own .bed/.fam pair,
 * src/data.hpp:45-72): a new dataset holding the genotypes of
 * individuals[0..n_sub) of the pool (gathered on the device, in that
 * order) with its own y (n_sub) and types; the pool stays usable. */
int bmg_ctx_create_subset(const bmg_ctx* pool, const uint32_t* individuals, uint32_t n_sub,
                          const double* y, const int32_t* types, int n_types, bmg_ctx** out);
int bmg_ctx_set_y(bmg_ctx* ctx, const double* y);
/* The covariates of Data (src/data.hpp:45-60, read_e; the constant column is
 * added by the engine as by the reference): e is n x m_e, column-major, m_e
 * <= 32.  Chains and models on the dataset then carry m_e + 1 fixed columns
 * with precisions inv_tau2_e_const_val / inv_tau2_e_val (src/model.hpp:
 * 92-105, src/prior.hpp:185-190); bmg_opts.m_e must equal m_e. */
int bmg_ctx_set_covariates(bmg_ctx* ctx, const double* e, uint32_t m_e);
int bmg_ctx_destroy(bmg_ctx* ctx);
/* packed body back to the host, reference layout (p * ceil(n/4) bytes) */
int bmg_ctx_read_body(bmg_ctx* ctx, uint8_t* out);
/* yy, var_y, var_x, mean_x (src/data.hpp:67-71, src/data.cpp:403-434) */
int bmg_ctx_stats(bmg_ctx* ctx, double out[4]);
/* number of missing genotype codes in the dataset */
int bmg_ctx_missing_count(bmg_ctx* ctx, uint64_t* out);

/* Replaces DataModel::get_genotypes[type](snp, v) (src/data_model.hpp:44-59,
 * src/data.cpp:56-138); -1 marks a missing genotype as in Data. */
int bmg_decode(bmg_ctx* ctx, uint64_t snp, int coding, double* out_n);
/* many SNPs at once: out[i * n + k] */
int bmg_decode_many(bmg_ctx* ctx, const uint64_t* snps, uint64_t count, int coding,
                    double* out);

/* Replaces PrecomputedSNPCovariances::xx (src/precomputed_snp_covariances.hpp:
 * 95-131): p rows of `offset` doubles (offset from bmg_self_moments_layout),
 * bit-exact (integer sums). */
int bmg_self_moments_layout(bmg_ctx* ctx, int* offset, int offset_type[4]);
int bmg_self_moments(bmg_ctx* ctx, double* out);

/* Band covariances (no reference counterpart; the reference's precomputed
 * covariances are per-SNP only): out[j * W + d] = sum_i x_j(i) x_{j+d}(i),
 * d < W, exact int32, codings ca/cb in {A,H,D,R}, int8 MFMA. */
int bmg_band_xtx(bmg_ctx* ctx, uint32_t W, int coding_a, int coding_b, int32_t* out);
int bmg_band_xtx_dev(bmg_ctx* ctx, uint32_t W, int coding_a, int coding_b,
                     int32_t* out_dev, void* stream);

/* X^T R for nvec residual vectors (the inner product of the RB scan,
 * src/sampler.cpp:171,173,188): r is nvec x n (row-major, host), out is
 * nvec x p.  Exact integer MFMA on the fixed-point residual; relative
 * quantisation error <= 2^-(8*digits-2) of max|r| per product term. */
int bmg_scan_xtr(bmg_ctx* ctx, const double* r, int nvec, int coding, int digits,
                 double* out);
/* device-pointer form for timing: r_dev nvec x n_pad (double), out_dev nvec x p */
int bmg_scan_xtr_dev(bmg_ctx* ctx, const double* r_dev, int nvec, int coding,
                     int digits, double* out_dev, void* stream);
/* No counterpart in the reference: bmg_scan_xtr_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results, bit for
 * bit; the scratch (digit planes, accumulator, scales) is one file-scope
 * device array of 65.0 MiB per GPU, held until the process ends, so it
 * takes p <= 1,047,552 SNPs and n <= 65,536 individuals (BMG_E_CAPACITY
 * beyond).  The scan shape is the one the autotuner picked for this width
 * on the context if it ran, else a fixed shape (no tuning, no events).
 * Calls that overlap on one GPU share the array and race: serialise them
 * (one stream, or events). */
int bmg_scan_xtr_dev_noalloc(bmg_ctx* ctx, const double* r_dev, int nvec, int coding, int digits,
                             double* out_dev, void* stream);
int bmg_ctx_n_pad(bmg_ctx* ctx, uint32_t* n_pad);
/* Timing form of bmg_scan_xtr_dev: prepares the digit planes once, then
 * launches the int8-MFMA scan kernel `reps` times on `stream`, bracketing
 * each launch with HIP events; *ms_per_launch is the mean kernel time. */
int bmg_scan_xtr_timed(bmg_ctx* ctx, const double* r_dev, int nvec, int coding, int digits,
                       double* out_dev, int reps, double* ms_per_launch, void* stream);
/* The scan shape the autotuner picked for nvec * digits digit columns (no
 * counterpart in the reference): out[0] SNP tiles per wave, out[1] LDS tile
 * width in 32-bit words, out[2] individual tiles, out[3] workgroups, out[4]
 * waves per workgroup.  The first scan of each width on a context times
 * every candidate shape and keeps the fastest; all zeros before that. */
int bmg_scan_shape(bmg_ctx* ctx, int nvec, int digits, int32_t* out);
/* Pin the scan shape for nvec * digits columns instead of tuning it:
 * tiles_per_wave 2 or 4, waves 4, 8 or 16, tile_words a multiple of 32
 * whose padded digit tile fits the LDS ((min(tile_words, words per SNP
 * rounded up to 32) * 16 + 16) * nvec * digits <= 160 KiB), workgroups >= 1
 * (clamped to the number of SNP blocks). */
int bmg_scan_set_shape(bmg_ctx* ctx, int nvec, int digits, int tiles_per_wave, int waves,
                       uint32_t tile_words, uint32_t workgroups);
/* Timing form of bmg_band_xtx_dev: `reps` launches, each bracketed by HIP
 * events on `stream`; *ms_per_launch is the mean kernel time. */
int bmg_band_xtx_timed(bmg_ctx* ctx, uint32_t W, int coding_a, int coding_b, int32_t* out_dev,
                       int reps, double* ms_per_launch, void* stream);

/* Measurement helpers (no counterpart in the reference): device buffers for
 * the timing forms above without a second runtime in the process, and a
 * plain 16-byte copy kernel of known traffic that calibrates the HBM
 * counters (FETCH_SIZE, WRITE_SIZE) of a profiler run. */
int bmg_dev_alloc(int device, uint64_t bytes, void** ptr);
int bmg_dev_free(int device, void* ptr);
int bmg_memcpy_h2d(int device, void* dst, const void* src, uint64_t bytes);
/* device-to-host copy after every kernel queued on the device has finished */
int bmg_memcpy_d2h(int device, void* dst, const void* src, uint64_t bytes);
int bmg_copy_timed(int device, const void* src_dev, void* dst_dev, uint64_t bytes, int reps,
                   double* ms_per_launch);
/* Streaming read of `bytes` (no counterpart in the reference): the best
 * mean launch time over a few grid sizes, i.e. the read rate the device
 * reaches, the practical ceiling of a read-bound kernel.  sink_dev: 4
 * writable bytes that are written only in a case that does not occur. */
int bmg_read_timed(int device, const void* src_dev, uint64_t bytes, void* sink_dev, int reps,
                   double* ms_per_launch);

/* -------------------------------------------------------------- model --- */
/* Prior inputs of Model/Prior (src/prior.hpp:45-122), already solved. */
typedef struct {
  double n_plus_nu, minus_n_plus_nu_2, nus2_plus_yy;
  double inv_tau2_e0;          /* d of the constant column */
} bmg_model_prior;

/* Replaces Model(data, data_model, prior) (src/model.hpp:49-113) for data
 * without covariates; kmax = column capacity (<= bmg_model_kmax()). */
int bmg_model_kmax(void);
int bmg_model_create(bmg_ctx* ctx, const bmg_model_prior* prior, bmg_model** out);
/* The Model of one Sampler with its Prior and random stream
 * (src/sampler.hpp:171-203, src/prior.hpp:45-122, src/rand.hpp:36-58): the
 * prior is solved from the options as for a batch chain, and the stream is
 * the chain's (key = seed, chain).  The empty model's log-likelihood is
 * computed. */
int bmg_model_create_sampler(bmg_ctx* ctx, const bmg_opts* opts, bmg_model** out);
int bmg_model_destroy(bmg_model* m);
/* Model::sample_beta_sigma2(rng) (src/model.hpp:314-342) from the model's
 * stream; BMG_E_NUMERIC if the sigma2 draw cannot finish */
int bmg_model_sample_beta_sigma2(bmg_model* m);
/* Model::compute_pve(pves) (src/model.hpp:345-392, no covariates): leaves
 * y_hat = X beta on the device for bmg_model_rb */
int bmg_model_compute_pve(bmg_model* m, double pves[3]);
/* beta (*len values: the constant column first), sigma2 and y_hat (n) of
 * the last draws; any pointer may be NULL */
int bmg_model_beta(bmg_model* m, double* beta, double* sigma2, double* y_hat, int* len);
/* RaoBlackwellizer::p_raoblackwell(Sampler*, y_hat, p_r, p_r_types)
 * (src/sampler.hpp:73-74, src/sampler.cpp:32-261) for a caller-held model:
 * p_r[p] from the model's state and prior and the given y_hat[n]; one
 * effect type */
int bmg_model_rb(bmg_model* m, const double* y_hat, double* p_r);
/* *current_model = *new_model after an accepted move, and *new_model =
 * *current_model after a rejected one (Model::operator=, src/model.hpp:
 * 115-162; src/sampler.cpp:592,634 and src/sampler.hpp:604-608,737-741) */
int bmg_model_commit(bmg_model* current, const bmg_model* proposed);
int bmg_model_rollback(bmg_model* proposed, const bmg_model* current);
/* Model::add_term(ind, t_ind, inv_tau2_alpha2) (src/model.hpp:239-268) */
int bmg_model_add_term(bmg_model* m, uint32_t ind, int t_ind, const double* d);
/* Model::remove_term(model_ind) (src/model.hpp:270-312) */
int bmg_model_remove_term(bmg_model* m, int model_ind);
/* Model::compute_log_likelihood() (src/model.hpp:199-237) */
int bmg_model_compute_log_likelihood(bmg_model* m);
/* Model::log_likelihood(), size() */
int bmg_model_log_likelihood(bmg_model* m, double* out);
int bmg_model_size(bmg_model* m, int* out);
/* Model::operator= (src/model.hpp:115-162) */
int bmg_model_copy(bmg_model* dst, const bmg_model* src);
/* ExhModel support: SymmMatrix::cholesky_swapadj on the model's factor and v
 * (src/symmmatrix.cpp:220-265) */
int bmg_model_swapadj(bmg_model* m, int col);
/* state: xx (k x k col-major, upper), l (same), xy, v, d (k each), scal[6] =
 * loglik, syx_plus_vs2, log_det_invQ, log_det_invQ_plus_xx, sigma2, k */
int bmg_model_get(bmg_model* m, double* xx, double* l, double* xy, double* v,
                  double* d, double scal[6]);
int bmg_model_set_d(bmg_model* m, int x_ind, double val);

/* ------------------------------------------------------ chains (batch) --- */
/* A batch runs n_chains PMV chains (src/sampler.cpp:551-880, moves 0-2 of
 * src/sampler.hpp:899-1669) entirely on the device: one workgroup per chain,
 * FP64 model in LDS, all-SNP RB scans between segments.  chain_ctx[c] picks
 * the dataset of chain c (phenotypes with different individual sets are
 * separate datasets).  opts[c] are the options of chain c (seed = its
 * reference seed).  Delayed rejection (src/sampler.cpp:882-1049) runs on
 * the device for min(delay_rejection, max_move_size) <= 20.  Unsupported on
 * the device: multiple effect types, missing genotypes. */
typedef struct {
  int32_t trace_every_iter;    /* keep per-iteration move/jd/size/loglik */
  int32_t proposal;            /* 0 plain, 1 plain + learned swap (online) */
  int32_t learn_every;         /* learned: training period in iterations
                                  (rounded to RB blocks, burn-in only) */
  int32_t learn_freeze_iter;   /* learned: freeze at the first RB block at or
                                  after this iteration (0: end of burn-in) */
  int32_t scan_digits;         /* fixed-point digits of the RB scan (4..5) */
  int32_t use_graph;           /* after burn-in, replay segments as a captured hipGraph */
  uint64_t trace_capacity;     /* thinned loci entries kept per chain */
  int32_t learn_window;        /* learned: partners within this many SNPs */
  int32_t learn_steps;         /* learned: Adam steps per training round */
  int32_t learn_batch;         /* learned: rows per Adam step */
  int32_t capacity;            /* model columns of the launch: 64, 128, 160
                                  (factor in LDS) or 256 (factor in HBM); 0:
                                  the smallest that holds the prior mean plus
                                  one move (AH: twice).  A chain whose model
                                  comes within one move of it continues in the
                                  256-column kernel from that iteration (same
                                  results); a proposal past 256 columns stops
                                  the chain and bmg_batch_run fails with
                                  BMG_E_CAPACITY (the reference's MAXMODELSIZE
                                  error, src/model.hpp:247-262) */
  double  p_learned;           /* learned: share of the drawn swaps that are learned */
  double  learn_rate;          /* learned: Adam learning rate */
  int32_t max_model_size;      /* SNPs; 0: none.  Larger models get prior
                                  probability 0 (the beta-binomial prior
                                  truncated there), so moves past it are
                                  rejected and RB draws respect it */
  int32_t cfg_pad;
} bmg_batch_cfg;

int bmg_batch_create(bmg_ctx** ctxs, int n_ctx, int n_chains, const int32_t* chain_ctx,
                     const bmg_opts* opts, const bmg_batch_cfg* cfg, bmg_batch** out);
int bmg_batch_destroy(bmg_batch* b);
/* Sampler::initialize_p_proposal_flat + the set-up of sample()
 * (src/sampler.hpp:351-362, src/sampler.cpp:591-620) */
int bmg_batch_setup(bmg_batch* b);
/* n_iter more iterations of every chain (the loop of src/sampler.cpp:626-834) */
int bmg_batch_run(bmg_batch* b, uint64_t n_iter);
int bmg_batch_sync(bmg_batch* b);
/* Learned swap (no counterpart in the reference): frozen partner weights
 * w[p][2H] for the chains of dataset ds (column c <-> SNP offset
 * c < H ? c - H : c - H + 1; negative or non-finite entries count as 0).
 * A swap drawn by the plain move draw (probability 0.15) becomes a learned
 * swap with probability p_learned: remove a uniformly chosen model SNP j,
 * add k ~ w[j][.] over the SNPs within H of j outside the model, accept
 * with the exact Metropolis-Hastings ratio; otherwise the plain swap runs.
 * w = NULL switches it off.  1 <= H <= 32. */
int bmg_batch_set_learned(bmg_batch* b, int ds, const float* w, int H, double p_learned);
/* The online learned proposal (proposal = 1): state of dataset ds's learner
 * (out: training rounds, frozen flag, labels collected, window H), its
 * partner-weight table (p x 2H), label ring, network parameters, static
 * per-SNP inputs (marginal scores z, mean p_proposal, band X^T X p x (H+1))
 * and the features of given (j, offset) pairs (rows x 16). */
int bmg_batch_learn_state(bmg_batch* b, int ds, double out[4]);
int bmg_batch_learn_table(bmg_batch* b, int ds, float* w);
int bmg_batch_learn_samples(bmg_batch* b, int ds, uint32_t* j, int32_t* off, float* label,
                            int64_t cap, int64_t* count);
int bmg_batch_learn_params(bmg_batch* b, int ds, float* params);
int bmg_batch_learn_static(bmg_batch* b, int ds, float* z, float* pbar, int32_t* band);
/* features of SNPs j[0..rows) for the per-SNP proposal network in chain
 * `chain`'s present state (marginal score, running RB means, LD to the
 * model; out rows x 16), for checking against the oracle */
int bmg_batch_learn_snp_features(bmg_batch* b, int ds, int chain, const uint32_t* j, int64_t rows,
                                 float* out);
int bmg_batch_learn_features(bmg_batch* b, int ds, const uint32_t* j, const int32_t* off,
                             int64_t rows, float* out);
/* per-chain state: loglik, size, sigma2, alpha, p_move_size, n_accepted,
 * n_iter, loci_len.  bmg_batch_run fails with BMG_E_NUMERIC when a chain's
 * current model turned non-finite (that chain stops; the others ran on) and
 * with BMG_E_CAPACITY when a thinned-trace buffer filled up. */
int bmg_batch_state(bmg_batch* b, int chain, double out[8]);
int bmg_batch_loci(bmg_batch* b, int chain, uint32_t* loci, int cap, int* k);
/* p_rao (RB PIP estimate, src/sampler.cpp:739-757) and p_proposal */
int bmg_batch_p_rao(bmg_batch* b, int chain, double* out);
int bmg_batch_p_proposal(bmg_batch* b, int chain, double* out);
/* a new source for a chain's add/remove proposal (after bmg_batch_setup):
 * p_proposal <- p (p values in [0, 1]), q_add = max(p, q_add_min) and
 * q_rem = max(1 - p, q_rem_min) with the proposal trees rebuilt, as an
 * adaptation step leaves them (src/sampler.cpp:517-524, 797-806).  Set at
 * the end of burn-in (adaptation off) the table is the proposal for the
 * rest of the run, and the moves keep the posterior for any table; the
 * learned per-SNP proposal sets its network's table through it. */
int bmg_batch_set_p_proposal(bmg_batch* b, int chain, const double* p);
/* per-iteration trace (needs trace_every_iter): move, jd, movesize, size,
 * loglik; returns the number of iterations in *len */
int bmg_batch_trace(bmg_batch* b, int chain, uint8_t* move, uint8_t* jd, uint8_t* ms,
                    uint32_t* size, double* loglik, int64_t cap, int64_t* len);
/* thinned trace (src/sampler.cpp:691-728): model sizes, concatenated loci,
 * loglik, sigma2, alpha */
int bmg_batch_thinned(bmg_batch* b, int chain, uint32_t* sizes, int64_t cap_sizes,
                      uint32_t* loci, int64_t cap_loci, double* loglik, double* sigma2,
                      double* alpha, int64_t* n_samples, int64_t* n_loci);
/* The binary output files of Sampler::sample for one chain
 * (src/sampler.cpp:560-584,687-728,846-858; README "Binary files"), named
 * basename + "_loci.dat", "_modelsize.dat", "_jumpdistance.dat",
 * "_log_likelihood.dat", "_log_prior.dat", "_move_type.dat",
 * "_move_size.dat", "_pve.dat", "_alpha.dat", "_sigma2.dat" and "_rao.dat",
 * in the reference's value types (uint32, uint8, double; native byte order);
 * one effect type, so no "_types.dat" / "_rao_types.dat", as the reference
 * with a single type.  The per-iteration files need trace_every_iter = 1.  A
 * learned swap (no counterpart in the reference) is move code
 * BMG_MOVE_LEARNED.  The basename is the reference's basename plus the
 * chain number. */
#define BMG_MOVE_LEARNED 7
int bmg_batch_write_outputs(bmg_batch* b, int chain, const char* basename);
/* save_beta (src/sampler.cpp:584,710-713): the beta of every thinned sample
 * of one chain, concatenated (the fixed columns first, then the loci's
 * columns in model order); *len = the number of doubles kept.  With
 * bmg_opts.save_beta set, bmg_batch_write_outputs also writes
 * basename + "_beta.dat" in this layout. */
int bmg_batch_thinned_beta(bmg_batch* b, int chain, double* out, int64_t cap, int64_t* len);
/* RB scan of one chain's current state with a given y_hat (the drop-in for
 * RaoBlackwellizer::p_raoblackwell, src/sampler.cpp:32-261) */
int bmg_batch_rb(bmg_batch* b, int chain, const double* y_hat, double* p_r);
/* event timing summed over every bmg_batch_run since set-up: total ms, ms
 * in chain kernels, ms in RB blocks, number of scan launches */
int bmg_batch_timing(bmg_batch* b, double out[4]);
/* The steady-state graph (no counterpart in the reference; use_graph = 1):
 * once every chain is past burn-in and every learner frozen, a segment (the
 * chain kernel over n_rao iterations and the RB block after it) is captured
 * into a hipGraph and replayed.  out: segments replayed, ms in replays (HIP
 * events), graph builds. */
int bmg_batch_graph_stats(bmg_batch* b, double out[3]);
/* milliseconds one chain's workgroup spent, since set-up, in: tau2/alpha
 * refreshes, add/remove moves, swaps, state changes, learned swaps,
 * beta/sigma2 draws, PVE + thinning, whole iterations; then within term
 * additions: x.y, X^T x, the triangular solve, the rest; then the number of
 * additions, the number of removals, ms in removals, ms in the alpha /
 * tau2 draws of the refreshes; then within removals of an inner column: the
 * Cholesky downdate, the triangular solve; then delayed-rejection second
 * stages, and within them the scoring of the sub-models (walk, proposal
 * probabilities, draw); then within add/remove moves: the proposal draws,
 * the removals and additions with their sum-tree updates, the reverse
 * move's weights, the accept / reject handling (device clock) */
int bmg_batch_profile(bmg_batch* b, int chain, double out[24]);

/* --------------------------------------- learned-proposal network ----- */
/* No counterpart in the reference.  The network of the learned proposal:
 * 16 features -> 256 -> 256 (ReLU) -> 1, bf16 weights and activations,
 * fp32 accumulation on the MFMA units, fp32 master weights and Adam.
 * Parameters are one fp32 array of bmg_mlp_nparams() values:
 * W1[16][256], b1[256], W2[256][256], b2[256], w3[256], b3.  Features are
 * rounded to bf16 (nearest even) on the way in.  These entry points expose
 * the kernels the batch trains with, for checking them in isolation. */
int bmg_mlp_nparams(void);
int bmg_mlp_create(int device, bmg_mlp** out);        /* all parameters 0 */
int bmg_mlp_destroy(bmg_mlp* m);
int bmg_mlp_set_params(bmg_mlp* m, const float* params);   /* resets Adam */
int bmg_mlp_get_params(bmg_mlp* m, float* params);
int bmg_mlp_forward(bmg_mlp* m, const float* X, int64_t rows, float* out);
/* gradient of the mean Huber(1) loss of out - labels */
int bmg_mlp_gradients(bmg_mlp* m, const float* X, const float* labels, int64_t rows,
                      float* grads);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient */
int bmg_mlp_adam_step(bmg_mlp* m, const float* X, const float* labels, int64_t rows, float lr);

#ifdef __cplusplus
}
#endif
#endif /* BMAGWA_HIP_H_ */
