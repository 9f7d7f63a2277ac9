/*
 * orc_sampler.c -- the PMV sampler of the reference (moves, Metropolis-
 * Hastings ratios, delayed rejection, schedule, adaptation) and the
 * Rao-Blackwellised all-SNP scan, restated for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "orc_internal.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ============================================================ sum tree */
/*
 * The project's weighted sampler (replaces the threaded binary tree of
 * src/discrete_distribution.hpp:64-330; same distribution: item i is drawn
 * with probability w_i / sum of non-zeroed w).  Level 0 holds the leaf
 * values (w_i, or 0 when zeroed), padded with zeros to a multiple of 64;
 * each entry of level h+1 is the sum of one 64-entry group of level h, and
 * the top level is a single group.  A group's sum is the last element of its
 * inclusive Hillis-Steele scan (offsets 1,2,4,...,32), the order a 64-lane
 * wavefront computes it in.  Drawing with u in [0,1): r = u * total; at each
 * level take the first child i with value > 0 and r < carry + scan[i] (else
 * the last child with value > 0), then carry += scan[i-1].
 */
#define ST_MAXLEV 8
struct orc_sumtree {
  uint64_t n;
  int top;
  uint64_t len[ST_MAXLEV];
  double* lev[ST_MAXLEV];
  double* w;
  unsigned char* zeroed;
  double total;
};

static void scan64(const double* c, double* s)
{
  double t[64];
  for (int i = 0; i < 64; ++i) s[i] = c[i];
  for (int d = 1; d < 64; d <<= 1) {
    for (int i = 0; i < 64; ++i) t[i] = (i >= d) ? s[i] + s[i - d] : s[i];
    for (int i = 0; i < 64; ++i) s[i] = t[i];
  }
}

static double group_sum(const double* c)
{
  double s[64];
  scan64(c, s);
  return s[63];
}

static void st_rebuild(orc_sumtree* t)
{
  for (uint64_t i = 0; i < t->len[0]; ++i)
    t->lev[0][i] = (i < t->n && !t->zeroed[i]) ? t->w[i] : 0.0;
  for (int h = 1; h <= t->top; ++h) {
    for (uint64_t g = 0; g < t->len[h]; ++g)
      t->lev[h][g] = (g < t->len[h - 1] / 64) ? group_sum(t->lev[h - 1] + 64 * g) : 0.0;
  }
  t->total = group_sum(t->lev[t->top]);
}

orc_sumtree* orc_sumtree_create(const double* w, uint64_t n)
{
  orc_sumtree* t = (orc_sumtree*)calloc(1, sizeof(orc_sumtree));
  t->n = n;
  uint64_t len = (n + 63) / 64 * 64;
  int h = 0;
  for (;;) {
    t->len[h] = len;
    t->lev[h] = (double*)calloc(len, sizeof(double));
    if (len == 64) break;
    len = ((len / 64) + 63) / 64 * 64;
    ++h;
  }
  t->top = h;
  t->w = (double*)malloc(sizeof(double) * n);
  memcpy(t->w, w, sizeof(double) * n);
  t->zeroed = (unsigned char*)calloc(n, 1);
  st_rebuild(t);
  return t;
}

void orc_sumtree_destroy(orc_sumtree* t)
{
  if (!t) return;
  for (int h = 0; h <= t->top; ++h) free(t->lev[h]);
  free(t->w); free(t->zeroed); free(t);
}

static void st_path(orc_sumtree* t, uint64_t i)
{
  for (int h = 1; h <= t->top; ++h) {
    uint64_t g = i >> (6 * h);
    t->lev[h][g] = group_sum(t->lev[h - 1] + 64 * g);
  }
  t->total = group_sum(t->lev[t->top]);
}

void orc_sumtree_zero(orc_sumtree* t, uint64_t i)
{
  t->zeroed[i] = 1; t->lev[0][i] = 0.0; st_path(t, i);
}

void orc_sumtree_unzero(orc_sumtree* t, uint64_t i)
{
  t->zeroed[i] = 0; t->lev[0][i] = t->w[i]; st_path(t, i);
}

void orc_sumtree_set_weights(orc_sumtree* t, const double* w)
{
  memcpy(t->w, w, sizeof(double) * t->n);
  st_rebuild(t);
}

double orc_sumtree_total(const orc_sumtree* t) { return t->total; }
int orc_sumtree_zeroed(const orc_sumtree* t, uint64_t i) { return t->zeroed[i]; }

uint64_t orc_sumtree_pick(const orc_sumtree* t, double u01)
{
  double r = u01 * t->total;
  double carry = 0.0;
  uint64_t g = 0;
  for (int h = t->top; h >= 0; --h) {
    const double* c = t->lev[h] + 64 * g;
    double s[64];
    scan64(c, s);
    int pick = -1, lastnz = -1;
    for (int i = 0; i < 64; ++i) {
      if (c[i] > 0.0) {
        lastnz = i;
        if (pick < 0 && r < carry + s[i]) pick = i;
      }
    }
    if (pick < 0) pick = lastnz;
    if (pick < 0) return UINT64_MAX;
    if (pick > 0) carry = carry + s[pick - 1];
    g = 64 * g + (uint64_t)pick;
  }
  return g;
}

/* ============================================================ utilities */
/* src/utils.cpp:144-152 */
void orc_geometric_dist_cdf(int maxsize, double p, double* values)
{
  double q = 1 - p, qp = q;
  for (int i = 0; i < maxsize; ++i) { values[i] = 1 - q; q *= qp; }
}

/* src/utils.cpp:46-57 with the uniform passed in */
int orc_sample_discrete_naive(const double* cumsum, int m, double u01)
{
  double r = u01 * cumsum[m - 1];
  for (int i = 0; i < m; ++i) if (r < cumsum[i]) return i;
  return -1;
}

static int sample_discrete_naive(const double* cumsum, int m, orc_rng* rng)
{
  return orc_sample_discrete_naive(cumsum, m, orc_rng_u01(rng));
}

/* src/utils.cpp:73-101 */
static size_t sample_discrete(const double* cumsum, size_t m, int level,
                              orc_rng* rng)
{
  size_t a = 0, b = m - 1, c;
  double r = orc_rng_u01(rng) * cumsum[b];
  while (level > 0) {
    c = (a + b) / 2;
    if (r < cumsum[c]) b = c; else a = c + 1;
    --level;
  }
  for (size_t i = a; i <= b; ++i) if (r < cumsum[i]) return i;
  return m - 1;
}

/* Geyer (1992) initial monotone sequence estimator of the effective sample
 * size of one series (the reference defines no ESS; this is the build's). */
double orc_ess_geyer(const double* x, int64_t n)
{
  if (n < 4) return (double)n;
  double mean = 0;
  for (int64_t i = 0; i < n; ++i) mean += x[i];
  mean /= n;
  double c0 = 0;
  for (int64_t i = 0; i < n; ++i) c0 += (x[i] - mean) * (x[i] - mean);
  c0 /= n;
  if (c0 <= 0) return (double)n;
  double sum = 0, prev = INFINITY;
  for (int64_t k = 0; 2 * k + 1 < n; ++k) {
    double g = 0;
    for (int lag = 2 * (int)k; lag <= 2 * (int)k + 1; ++lag) {
      double a = 0;
      for (int64_t i = 0; i + lag < n; ++i) a += (x[i] - mean) * (x[i + lag] - mean);
      g += a / n / c0;
    }
    if (g <= 0) break;
    if (g > prev) g = prev;
    prev = g;
    sum += g;
  }
  double tau = -1.0 + 2.0 * sum;
  if (tau < 1.0 / n) tau = 1.0 / n;
  return n / tau;
}

/* ============================================================== sampler */
#define NMOVES 7

struct orc_sampler {
  orc_opts o;
  const orc_data* d;
  orc_datamodel* dm;
  orc_prior pr;
  orc_model *cur, *nw, *exh;
  orc_rng rng;
  uint64_t n_iter, n_accepted;
  int64_t n_rao_burnin;
  size_t m_g; int n_types, use_types;
  double* p_proposal; size_t p_proposal_n;
  double* p_rao; int p_rao_n;
  double p_moves[NMOVES], p_moves_cumsum[NMOVES];
  double *q_add, q_add_min, *q_rem, q_rem_min;
  orc_sumtree *dd_add, *dd_rem;
  double* p_r;
  int max_ms; size_t max_nbh;
  double p_ms, p_ms_nbs, p_ms_nbc; int adapt_p_ms;
  double *q_pms, *q_pms_nbs, *q_pms_nbc;
  double *r_ms_sum, *r_ms_n, *q_p0_ms;
  size_t n_acpt[NMOVES], n_moves[NMOVES];
  double acpt_goal;
  size_t *mi_add, *mi_rem, *mi; int64_t* mi_map; int* mi_isadd;
  int movesize;
  int dr; double *dr_probs, *dr_q_add, *dr_q_rem, *dr_lqt; unsigned char* dr_b2n;
  uint32_t rb_calls;
  double* pre_xx; int pre_offset; int pre_ot[4];
  double pves[3];
  double k_move_size;
  int setup_done;
  /* learned swap: frozen proposal weights W[j][2H] (see do_learned_swap) */
  float* lw; int lH; double p_learned; int learn_on;
  /* traces */
  int64_t tr_len, tr_cap;
  uint8_t *tr_move, *tr_jd, *tr_ms; uint32_t* tr_size; double* tr_ll;
  int64_t th_len, th_cap, th_loci_len, th_loci_cap;
  uint32_t* th_size; uint32_t* th_loci; double *th_ll, *th_s2, *th_alpha, *th_pve;
  double* th_beta; int64_t th_beta_len, th_beta_cap;   /* save_beta */
};

/* src/sampler.cpp:455-515 (PMV only) */
static void compute_p_moves(orc_sampler* s)
{
  double* p = s->p_moves;
  for (int i = 0; i < NMOVES; ++i) p[i] = 0.0;
  if (s->use_types) { p[0] = 0.8; p[1] = 0.1; p[3] = 0.1; }
  else { p[0] = 0.7; p[1] = 0.15; p[2] = 0.15; }
  s->p_moves_cumsum[0] = p[0];
  for (int i = 1; i < NMOVES; ++i) s->p_moves_cumsum[i] = s->p_moves_cumsum[i - 1] + p[i];
}

/* src/sampler.cpp:517-524 and 597 */
static void compute_proposal_dist(orc_sampler* s)
{
  for (size_t i = 0; i < s->m_g; ++i) {
    s->q_add[i] = fmax(s->p_proposal[i], s->q_add_min);
    s->q_rem[i] = fmax(1 - s->p_proposal[i], s->q_rem_min);
  }
}

/* src/sampler.hpp:107-282 */
orc_sampler* orc_sampler_create(const orc_opts* o, const orc_data* d)
{
  if (o->sampler_type != 0) return NULL;
  orc_sampler* s = (orc_sampler*)calloc(1, sizeof(orc_sampler));
  s->o = *o; s->d = d;
  s->m_g = orc__data_m_g(d);
  s->n_types = o->n_types;
  s->use_types = o->n_types > 1;
  s->n_rao_burnin = o->n_rao_burnin;
  orc_rng_init(&s->rng, o->seed, o->chain, 0, 0);
  s->p_proposal_n = 1;
  s->q_add_min = 1.0 / (s->m_g - o->e_qg);
  s->q_rem_min = 1.0 / o->e_qg;
  size_t mm = s->m_g < 255 ? s->m_g : 255;
  s->max_ms = (int)(mm < o->max_move_size ? mm : o->max_move_size);
  s->max_nbh = o->max_SNP_neighborhood_size;
  s->p_ms = o->p_move_size; s->p_ms_nbs = o->p_move_size_nbs;
  s->p_ms_nbc = o->p_move_size_nbc;
  s->adapt_p_ms = o->adapt_p_move_size && s->max_ms > 1;
  s->acpt_goal = o->p_move_size_acpt_goal;
  s->dr = (int)(o->delay_rejection < (uint32_t)s->max_ms ? o->delay_rejection : (uint32_t)s->max_ms);

  s->dm = orc_datamodel_create(d, o->types, o->n_types);
  double st[4]; orc_data_stats(d, st);
  double s2_sigma2 = o->s2_sigma2 > 0 ? o->s2_sigma2
      : st[1] * (1 - o->R2mode_sigma2) * (o->nu_sigma2 + 2) / o->nu_sigma2;
  double s2_tau2[4] = {0, 0, 0, 0};
  for (int t = 0; t < 4; ++t) {
    if (!orc__dm_allow_terms(s->dm, t)) continue;
    if (o->s2_tau2[t] > 0) s2_tau2[t] = o->s2_tau2[t];
    else s2_tau2[t] = o->eh_tau2[t] * (o->nu_tau2[t] - 2)
                      / (o->nu_tau2[t] * (st[2] + st[3] * st[3]) * (1 - o->R2mode_sigma2));
  }
  double inv_tau2_e[64];
  uint32_t me = orc__data_m_e(d);
  for (uint32_t i = 0; i < me && i < 64; ++i) inv_tau2_e[i] = o->inv_tau2_e_val;
  inv_tau2_e[0] = o->inv_tau2_e_const_val;
  if (orc__prior_init(&s->pr, d, s->dm, o->types_prior, o->e_qg, o->var_qg,
                      inv_tau2_e, o->nu_sigma2, s2_sigma2, o->nu_tau2,
                      s2_tau2, o->mu_alpha, o->use_individual_tau2)) {
    orc_datamodel_destroy(s->dm); free(s); return NULL;
  }
  int kmax = 1024;
  /* src/main.cpp:62-67: per-SNP moments once, with missing values at 0 */
  s->pre_offset = orc_datamodel_offset(s->dm, s->pre_ot);
  s->pre_xx = (double*)malloc(sizeof(double) * s->pre_offset * s->m_g);
  orc_self_moments(s->dm, s->pre_xx);
  s->cur = orc__model_create(s, d, s->dm, &s->pr, kmax);
  s->nw = orc__model_create(s, d, s->dm, &s->pr, kmax);
  s->exh = orc__model_create(s, d, s->dm, &s->pr, kmax);

  s->p_proposal = (double*)calloc(s->m_g, sizeof(double));
  s->p_rao = (double*)calloc(s->m_g, sizeof(double));
  s->q_add = (double*)calloc(s->m_g, sizeof(double));
  s->q_rem = (double*)calloc(s->m_g, sizeof(double));
  s->p_r = (double*)calloc(s->m_g, sizeof(double));
  int M = s->max_ms > 0 ? s->max_ms : 1;
  s->mi_add = (size_t*)calloc(M, sizeof(size_t));
  s->mi_rem = (size_t*)calloc(M, sizeof(size_t));
  s->mi = (size_t*)calloc(M, sizeof(size_t));
  s->mi_map = (int64_t*)calloc(M, sizeof(int64_t));
  s->mi_isadd = (int*)calloc(M, sizeof(int));
  for (int i = 0; i < 3; ++i) { s->n_acpt[i] = 1; s->n_moves[i] = 2; }
  s->q_pms = (double*)calloc(M, sizeof(double));
  s->r_ms_sum = (double*)calloc(M, sizeof(double));
  s->r_ms_n = (double*)calloc(M, sizeof(double));
  s->q_p0_ms = (double*)calloc(M, sizeof(double));
  int M2 = (s->max_ms + 1) / 2; if (M2 < 1) M2 = 1;
  s->q_pms_nbs = (double*)calloc(M2, sizeof(double));
  s->q_pms_nbc = (double*)calloc(M, sizeof(double));
  orc_geometric_dist_cdf(s->max_ms, s->p_ms, s->q_pms);
  orc_geometric_dist_cdf((s->max_ms + 1) / 2, s->p_ms_nbs, s->q_pms_nbs);
  orc_geometric_dist_cdf(s->max_ms, s->p_ms_nbc, s->q_pms_nbc);
  if (s->dr > 0) {
    s->dr_probs = (double*)calloc((size_t)1 << s->dr, sizeof(double));
    s->dr_q_add = (double*)calloc(s->dr, sizeof(double));
    s->dr_q_rem = (double*)calloc(s->dr, sizeof(double));
    s->dr_lqt = (double*)calloc(s->dr, sizeof(double));
    s->dr_b2n = (unsigned char*)calloc(s->dr, 1);
  }
  return s;
}

void orc_sampler_destroy(orc_sampler* s)
{
  if (!s) return;
  orc__model_destroy(s->cur); orc__model_destroy(s->nw); orc__model_destroy(s->exh);
  orc_datamodel_destroy(s->dm);
  orc_sumtree_destroy(s->dd_add); orc_sumtree_destroy(s->dd_rem);
  free(s->pre_xx); free(s->lw);
  free(s->p_proposal); free(s->p_rao); free(s->q_add); free(s->q_rem); free(s->p_r);
  free(s->mi_add); free(s->mi_rem); free(s->mi); free(s->mi_map); free(s->mi_isadd);
  free(s->q_pms); free(s->r_ms_sum); free(s->r_ms_n); free(s->q_p0_ms);
  free(s->q_pms_nbs); free(s->q_pms_nbc);
  free(s->dr_probs); free(s->dr_q_add); free(s->dr_q_rem); free(s->dr_lqt); free(s->dr_b2n);
  free(s->tr_move); free(s->tr_jd); free(s->tr_ms); free(s->tr_size); free(s->tr_ll);
  free(s->th_size); free(s->th_loci); free(s->th_ll); free(s->th_s2); free(s->th_alpha); free(s->th_pve);
  free(s->th_beta);
  free(s);
}

/* src/sampler.hpp:351-362 */
void orc_sampler_init_p_proposal_flat(orc_sampler* s)
{
  double e_g = s->pr.g_a / (s->pr.g_a + s->pr.g_b) * s->m_g;
  double val = e_g / (double)s->m_g;
  for (size_t i = 0; i < s->m_g; ++i) s->p_proposal[i] = val;
}

/* src/sampler.hpp:500-515 (missing values only exist in the fixtures) */
static void data_model_sample_missing_single(orc_sampler* s, size_t snp)
{
  const size_t* ml = orc__data_miss_loc(s->d, snp);
  if (!ml) return;
  char* mv = orc__dm_miss_val(s->dm, snp);
  const double* mp = orc__data_miss_prior(s->d, snp);
  for (size_t i = 1; i <= ml[0]; ++i)
    mv[i] = (char)sample_discrete_naive(mp, 3, &s->rng);
}

static void prepare_add_new_term(orc_sampler* s, size_t ind, int type,
                                 double* d2, int sample_missing)
{
  if (sample_missing) data_model_sample_missing_single(s, ind);
  if (type == ORC_AH) {
    d2[0] = orc__prior_draw_inv_tau2(&s->pr, ORC_A, &s->rng);
    d2[1] = orc__prior_draw_inv_tau2(&s->pr, ORC_H, &s->rng);
  } else {
    d2[0] = orc__prior_draw_inv_tau2(&s->pr, type, &s->rng);
  }
}

/* src/data_model.cpp:78-90 */
static void data_model_sample_missing(orc_sampler* s, const int32_t* model_inds)
{
  for (size_t snp = 0; snp < s->m_g; ++snp) {
    if (!orc__dm_miss_val(s->dm, snp) || model_inds[snp] >= 0) continue;
    data_model_sample_missing_single(s, snp);
  }
}

static int data_has_missing(const orc_sampler* s)
{
  for (size_t snp = 0; snp < s->m_g; ++snp)
    if (orc__data_miss_loc(s->d, snp)) return 1;
  return 0;
}

/* src/sampler.cpp:264-453 */
static void sampler_sample_missing(orc_sampler* s)
{
  if (!data_has_missing(s)) { s->cur->mu_beta_computed = 0; return; }
  orc_model* cm = s->cur; orc_model* nm = s->nw;
  size_t n = cm->n; int ld = cm->kmax;
  const double* y = orc__data_y(s->d);
  double s2x2 = cm->sigma2 * 2;
  double* y_hat = (double*)calloc(n, sizeof(double));
  double* res = (double*)calloc(n, sizeof(double));
  for (int j = 0; j < cm->beta_len; ++j)
    for (size_t i = 0; i < n; ++i) y_hat[i] += cm->x[(size_t)j * n + i] * cm->beta[j];
  double r2 = 0;
  for (size_t i = 0; i < n; ++i) { res[i] = y[i] - y_hat[i]; r2 += res[i] * res[i]; }
  const double* b = cm->beta;
  for (int i = 0; i < cm->nloci; ++i) {
    uint32_t ind = cm->loci[i];
    const size_t* ml = orc__data_miss_loc(s->d, ind);
    if (!ml) continue;
    int type = orc__dm_type(s->dm, cm->type_inds[i]);
    unsigned xi1 = cm->x_ind1[i];
    unsigned xi2 = (type == ORC_AH) ? cm->x_ind2[i] : 0;
    double* t1 = cm->x + (size_t)xi1 * n;
    double* t2 = cm->x + (size_t)xi2 * n;
    const double* mp = orc__data_miss_prior(s->d, ind);
    double lp[3] = {log(mp[0]), log(mp[1] - mp[0]), log(mp[2] - mp[1])};
    char* mv = orc__dm_miss_val(s->dm, ind);
    for (size_t j = 1; j <= ml[0]; ++j) {
      size_t im = ml[j];
      double old2 = res[im] * res[im];
      double lk[3];
      for (int k = 0; k < 3; ++k) {
        double nt;
        switch (type) {
          case ORC_A: nt = res[im] - b[xi1] * ((double)k - t1[im]); break;
          case ORC_H: nt = res[im] - b[xi1] * ((double)(k == 1) - t1[im]); break;
          case ORC_R: nt = res[im] - b[xi1] * ((double)(k == 2) - t1[im]); break;
          case ORC_D: nt = res[im] - b[xi1] * ((double)(k > 0) - t1[im]); break;
          default:
            nt = res[im] - b[xi1] * ((double)k - t1[im]);
            nt -= b[xi2] * ((double)(k == 1) - t2[im]);
        }
        lk[k] = lp[k] - (r2 + nt * nt - old2) / s2x2;
      }
      lk[1] -= lk[0]; lk[2] -= lk[0]; lk[0] = 1;
      lk[1] = exp(lk[1]) + lk[0];
      lk[2] = exp(lk[2]) + lk[1];
      int k = sample_discrete_naive(lk, 3, &s->rng);
      mv[j] = (char)k;
      switch (type) {
        case ORC_A: y_hat[im] += b[xi1] * ((double)k - t1[im]); t1[im] = k; break;
        case ORC_H: k = (k == 1); y_hat[im] += b[xi1] * ((double)k - t1[im]); t1[im] = k; break;
        case ORC_R: k = (k == 2); y_hat[im] += b[xi1] * ((double)k - t1[im]); t1[im] = k; break;
        case ORC_D: k = (k > 0); y_hat[im] += b[xi1] * ((double)k - t1[im]); t1[im] = k; break;
        default:
          y_hat[im] += b[xi1] * ((double)k - t1[im]); t1[im] = k;
          k = (k == 1);
          y_hat[im] += b[xi2] * ((double)k - t2[im]); t2[im] = k;
      }
      res[im] = y[im] - y_hat[im];
      r2 += res[im] * res[im] - old2;
    }
  }
  /* x^T x and x^T y fix-up (src/sampler.cpp:392-450) */
  const uint8_t* body = orc_data_body(s->d);
  size_t stride = orc__data_stride(s->d);
  for (int i = 0; i < cm->nloci; ++i) {
    uint32_t ind = cm->loci[i];
    const size_t* ml = orc__data_miss_loc(s->d, ind);
    if (!ml) continue;
    int type = orc__dm_type(s->dm, cm->type_inds[i]);
    for (int term = 0; term < (type == ORC_AH ? 2 : 1); ++term) {
      unsigned xi = term == 0 ? cm->x_ind1[i] : cm->x_ind2[i];
      double* xn = cm->x + (size_t)xi * n;
      double* xo = nm->x + (size_t)xi * n;
      for (size_t j = 1; j <= ml[0]; ++j) {
        size_t im = ml[j];
        cm->xy[xi] += y[im] * (xn[im] - xo[im]);
        unsigned k = 0;
        for (; k < xi; ++k) {
          if ((int)k < cm->m_e ||
              !(orc_get_genotype(body, stride, im, cm->x_ind_to_g_ind[k]) < 0))
            cm->xx[(size_t)xi * ld + k] += xn[im] * cm->x[(size_t)k * n + im]
                                           - xo[im] * nm->x[(size_t)k * n + im];
        }
        for (; (int)k < cm->cols; ++k)
          cm->xx[(size_t)k * ld + xi] += xn[im] * cm->x[(size_t)k * n + im]
                                         - xo[im] * nm->x[(size_t)k * n + im];
      }
    }
  }
  cm->mu_beta_computed = 0;
  free(y_hat); free(res);
}

/* src/prior.cpp:71-114 */
static void prior_sample_tau2(orc_sampler* s, orc_model* m)
{
  orc_prior* p = &s->pr;
  double a2s2 = p->alpha2 * m->sigma2;
  if (p->use_individual_tau2) {
    for (int i = m->m_e; i < m->beta_len; ++i) {
      int t = m->x_types[i];
      double nu = p->nu_tau2[t] + 1.0;
      double s2 = (p->nus2_tau2[t] + m->beta[i] * m->beta[i] / a2s2) / nu;
      m->dv[i] = 1.0 / (p->alpha2 * orc_rng_sinvchi2(&s->rng, nu, s2));
    }
  } else {
    double b2[4] = {0, 0, 0, 0}; size_t mt[4] = {0, 0, 0, 0};
    for (int i = m->m_e; i < m->beta_len; ++i) {
      b2[m->x_types[i]] += m->beta[i] * m->beta[i];
      ++mt[m->x_types[i]];
    }
    for (int t = 0; t < 4; ++t) {
      if (!p->allow_terms[t]) continue;
      double nu = p->nu_tau2[t] + mt[t];
      double s2 = (p->nus2_tau2[t] + b2[t] / a2s2) / nu;
      p->inv_tau2_alpha2[t] = 1.0 / (p->alpha2 * orc_rng_sinvchi2(&s->rng, nu, s2));
    }
    for (int i = m->m_e; i < m->beta_len; ++i) m->dv[i] = p->inv_tau2_alpha2[m->x_types[i]];
  }
  m->mu_beta_computed = 0;
}

/* src/prior.cpp:30-69 */
static double prior_sample_alpha(orc_sampler* s, orc_model* m)
{
  orc_prior* p = &s->pr;
  int nterms = m->beta_len - m->m_e;
  size_t n = m->n;
  double a = p->alpha;
  const double* y = orc__data_y(s->d);
  double* xb = NULL; double* eb = NULL;
  if (nterms > 0) { xb = (double*)malloc(sizeof(double) * n); eb = (double*)malloc(sizeof(double) * n); }
  do {
    if (nterms > 0) {
      if ((a <= 0.0) || log(orc_rng_u01(&s->rng)) <= -2.0 * p->mu_alpha * a) a = -a;
      for (size_t i = 0; i < n; ++i) { xb[i] = 0; eb[i] = 0; }
      for (int j = 0; j < nterms; ++j) {
        double bj = m->beta[m->m_e + j];
        const double* col = m->x + (size_t)(m->m_e + j) * n;
        if (bj != 0.0) for (size_t i = 0; i < n; ++i) xb[i] += bj * col[i];
      }
      for (int j = 0; j < m->m_e; ++j) {
        double bj = m->beta[j];
        const double* col = m->x + (size_t)j * n;
        if (bj != 0.0) for (size_t i = 0; i < n; ++i) eb[i] += bj * col[i];
      }
      for (size_t i = 0; i < n; ++i) eb[i] -= y[i];
      double as2 = a * m->sigma2;
      double xbxb = 0, ebxb = 0;
      for (size_t i = 0; i < n; ++i) { xbxb += xb[i] * xb[i]; ebxb += eb[i] * xb[i]; }
      double var = 1.0 / (1.0 + xbxb / (a * as2));
      double mu = var * (p->mu_alpha - ebxb / as2);
      a = sqrt(var) * orc_rng_normal(&s->rng) + mu;
    } else {
      a = orc_rng_normal(&s->rng) + p->mu_alpha;
    }
  } while (a == 0 || a * a == 0);
  m->mu_beta_computed = 0;
  free(xb); free(eb);
  return a;
}

/* src/prior.cpp:116-141 */
static void prior_set_alpha(orc_sampler* s, double val, orc_model* m)
{
  orc_prior* p = &s->pr;
  double old = p->alpha2;
  p->alpha = val; p->alpha2 = val * val;
  if (p->use_individual_tau2) {
    for (int i = m->m_e; i < m->beta_len; ++i) m->dv[i] = m->dv[i] * old / p->alpha2;
  } else {
    for (int t = 0; t < 4; ++t)
      if (p->allow_terms[t]) p->inv_tau2_alpha2[t] = p->inv_tau2_alpha2[t] * old / p->alpha2;
    for (int i = m->m_e; i < m->beta_len; ++i) m->dv[i] = p->inv_tau2_alpha2[m->x_types[i]];
  }
}

static void prior_sample_alpha_and_tau2(orc_sampler* s, orc_model* m)
{
  prior_sample_tau2(s, m);
  prior_set_alpha(s, prior_sample_alpha(s, m), m);
}

/* ------------------------------------------------------------ RB scan */
/* src/sampler.cpp:32-261.  The per-SNP tau2 draws (use_individual_tau2)
 * come from the project's RB substream: key {seed, chain}, counter
 * {draw, 0, snp, 0x80000000 | rb_call}; terms are drawn in A,H,D,R order. */
void orc_sampler_rb(orc_sampler* s, const double* y_hat, double* p_r)
{
  const orc_model* cm = s->cur;
  const orc_prior* pr = &s->pr;
  size_t n = cm->n;
  const double* y = orc__data_y(s->d);
  double s2x2 = 2 * cm->sigma2;
  int n_types = orc__dm_n_types(s->dm);
  double* rcm = (double*)malloc(sizeof(double) * n);
  double* rom = (double*)malloc(sizeof(double) * n);
  double* x = (double*)malloc(sizeof(double) * n);
  double mr_cm = 0;
  for (size_t i = 0; i < n; ++i) { rcm[i] = y[i] - y_hat[i]; mr_cm += rcm[i]; }
  mr_cm /= n;
  int Ns[5]; memcpy(Ns, cm->Ns, sizeof(Ns));
  int n_loci = cm->nloci;
  double lmp_add[5] = {0}, lmp_rem[5][5];
  memset(lmp_rem, 0, sizeof(lmp_rem));
  for (int t = 0; t < 5; ++t) {
    if (!orc__dm_allow_types(s->dm, t)) continue;
    if ((size_t)(Ns[t] + 1) <= s->m_g) lmp_add[t] = orc__prior_add(pr, Ns, n_loci, t);
    Ns[t]--; n_loci--;
    if (Ns[t] >= 0)
      for (int j = 0; j < 5; ++j)
        if (orc__dm_allow_types(s->dm, j)) lmp_rem[t][j] = orc__prior_add(pr, Ns, n_loci, j);
    Ns[t]++; n_loci++;
  }
  double d_vals[4] = {0, 0, 0, 0};
  if (!pr->use_individual_tau2)
    for (int t = 0; t < 4; ++t)
      if (orc__dm_allow_terms(s->dm, t)) d_vals[t] = pr->inv_tau2_alpha2[t];
  const int* ot = s->pre_ot;
  int offset = s->pre_offset;
  double* pre = (double*)malloc(sizeof(double) * offset);
  for (size_t i = 0; i < s->m_g; ++i) {
    const double* res; double mr; const double* lmp;
    int32_t mind = cm->model_inds[i];
    if (pr->use_individual_tau2) {
      orc_rng rr;
      orc_rng_init(&rr, s->o.seed, s->o.chain, (uint32_t)i, 0x80000000u | s->rb_calls);
      for (int t = 0; t < 4; ++t)
        if (orc__dm_allow_terms(s->dm, t))
          d_vals[t] = 1.0 / (pr->alpha2 * orc_rng_sinvchi2(&rr, pr->nu_tau2[t], pr->s2_tau2[t]));
    }
    if (mind < 0) {
      res = rcm; mr = mr_cm; lmp = lmp_add;
    } else {
      unsigned xi = cm->x_ind1[mind];
      int type = orc__dm_type(s->dm, cm->type_inds[mind]);
      if (pr->use_individual_tau2) {
        if (type == ORC_AH) {
          d_vals[ORC_A] = cm->dv[xi];
          d_vals[ORC_H] = cm->dv[cm->x_ind2[mind]];
        } else {
          d_vals[type] = cm->dv[xi];
        }
      }
      for (size_t k = 0; k < n; ++k) rom[k] = rcm[k] + cm->beta[xi] * cm->x[(size_t)xi * n + k];
      if (type == ORC_AH) {
        xi = cm->x_ind2[mind];
        for (size_t k = 0; k < n; ++k) rom[k] += cm->beta[xi] * cm->x[(size_t)xi * n + k];
      }
      res = rom;
      mr = 0; for (size_t k = 0; k < n; ++k) mr += rom[k];
      mr /= n;
      lmp = lmp_rem[type];
    }
    memcpy(pre, s->pre_xx + i * offset, sizeof(double) * offset);
    orc_datamodel_update_prexx(s->dm, i, pre);
    double maxt = -INFINITY;
    double pt[5];
    for (int t = 0; t < n_types; ++t) {
      int type = orc__dm_type(s->dm, t);
      double det, exp_term, slq;
      if (type == ORC_AH) {
        double sa = pre[0], va = pre[1] + d_vals[ORC_A];
        double sh = pre[2], vh = pre[3] + d_vals[ORC_H];
        double vah = pre[offset - 1];
        slq = -log(d_vals[ORC_A]) - log(d_vals[ORC_H]);
        orc_datamodel_get(s->dm, ORC_A, i, x);
        double rxa = 0; for (size_t k = 0; k < n; ++k) rxa += x[k] * res[k];
        rxa -= sa * mr;
        orc_datamodel_get(s->dm, ORC_H, i, x);
        double rxh = 0; for (size_t k = 0; k < n; ++k) rxh += x[k] * res[k];
        rxh -= sh * mr;
        det = va * vh - vah * vah;
        exp_term = rxa * rxa * vh - 2.0 * rxa * rxh * vah + rxh * rxh * va;
        exp_term /= det;
      } else {
        int off = ot[type];
        double sx = pre[off];
        det = pre[off + 1] + d_vals[type];
        slq = -log(d_vals[type]);
        orc_datamodel_get(s->dm, type, i, x);
        double rx = 0; for (size_t k = 0; k < n; ++k) rx += x[k] * res[k];
        rx -= sx * mr;
        exp_term = rx * rx / det;
      }
      double p = exp_term / s2x2 - 0.5 * (log(det) + slq) + lmp[type];
      if (s->use_types) {
        if (p > maxt) maxt = p;
        pt[t] = p;
      } else {
        p = exp(p);
        p_r[i] = isfinite(p) ? p / (1 + p) : 1;
      }
    }
    if (s->use_types) {
      if (!isfinite(maxt)) { p_r[i] = maxt > 0 ? 1 : 0; continue; }
      double sum = 0;
      for (int t = 0; t < n_types; ++t) { pt[t] = exp(pt[t] - maxt); sum += pt[t]; }
      sum = exp(log(sum) + maxt);
      p_r[i] = isfinite(sum) ? sum / (1 + sum) : 1;
    }
  }
  s->rb_calls++;
  free(rcm); free(rom); free(x); free(pre);
}

/* -------------------------------------------------------------- moves */
static void undo_dd(orc_sampler* s, int ms)
{
  for (int i = 0; i < ms; ++i) {
    size_t ind = s->mi[i];
    if (s->mi_isadd[i]) { orc_sumtree_unzero(s->dd_add, ind); orc_sumtree_zero(s->dd_rem, ind); }
    else { orc_sumtree_zero(s->dd_add, ind); orc_sumtree_unzero(s->dd_rem, ind); }
  }
}

/* src/sampler.hpp:1126-1165 */
static void prepare_addrem(orc_sampler* s, double* lqf, int* n_rem, int ms)
{
  size_t max_adds = s->m_g - s->cur->nloci, max_rems = s->cur->nloci;
  *n_rem = 0;
  for (int i = 0; i < ms; ++i) {
    int add;
    if (max_rems == 0) add = 1;
    else if (max_adds == 0) add = 0;
    else { *lqf += ORC_LOG_HALF; add = (orc_rng_u01(&s->rng) < 0.5); }
    if (add) {
      --max_adds;
      s->mi_isadd[i] = 1;
      size_t ind = orc_sumtree_pick(s->dd_add, orc_rng_u01(&s->rng));
      s->mi[i] = ind; s->mi_map[i] = (int64_t)ind;
      *lqf += log(s->q_add[ind]) - log(orc_sumtree_total(s->dd_add));
      orc_sumtree_zero(s->dd_add, ind);
    } else {
      ++*n_rem; --max_rems;
      s->mi_isadd[i] = 0;
      size_t ind = orc_sumtree_pick(s->dd_rem, orc_rng_u01(&s->rng));
      s->mi[i] = ind; s->mi_map[i] = -1;
      *lqf += log(s->q_rem[ind]) - log(orc_sumtree_total(s->dd_rem));
      orc_sumtree_zero(s->dd_rem, ind);
    }
  }
}

/* src/sampler.hpp:1168-1202 */
static void backward_prepare_addrem(orc_sampler* s, double* lqb, int ms)
{
  size_t max_adds = s->m_g - s->nw->nloci, max_rems = s->nw->nloci;
  double w_add = orc_sumtree_total(s->dd_add), w_rem = orc_sumtree_total(s->dd_rem);
  int last = ms - 1;
  for (int i = 0; i < ms; ++i) {
    if (max_adds > 0 && max_rems > 0) *lqb += ORC_LOG_HALF;
    size_t ind;
    if (s->mi_isadd[i]) {
      --max_rems;
      while (!s->mi_isadd[last]) --last;
      ind = (size_t)s->mi_map[last];
      --last;
      *lqb += log(s->q_rem[ind]) - log(w_rem);
      w_rem -= s->q_rem[ind];
    } else {
      --max_adds;
      ind = (size_t)s->mi_map[i];
      *lqb += log(s->q_add[ind]) - log(w_add);
      w_add -= s->q_add[ind];
    }
  }
}

/* src/sampler.hpp:1204-1268 (single effect type: use_types is false for
 * every configuration of the engine; the type draws are kept for parity) */
static void do_addrem(orc_sampler* s, double* lqf, double* lqb, double* lmpc, int ms)
{
  (void)lqf; (void)lqb;
  int last = ms - 1;
  orc_model* nm = s->nw;
  for (int i = 0; i < ms; ++i) {
    if (s->mi_isadd[i]) continue;
    size_t ind = s->mi[i];
    orc_sumtree_unzero(s->dd_add, ind);
    while (s->mi_map[last] >= 0) --last;
    s->mi_map[last] = (int64_t)ind;
    int mind = nm->model_inds[ind];
    int t_ind = nm->type_inds[mind];
    int type = orc__dm_type(s->dm, t_ind);
    *lmpc += orc__prior_rem(&s->pr, nm->Ns, nm->nloci, type);
    orc_model_remove_term(nm, mind);
  }
  double d2[2];
  for (int i = 0; i < ms; ++i) {
    if (!s->mi_isadd[i]) continue;
    size_t ind = s->mi[i];
    orc_sumtree_unzero(s->dd_rem, ind);
    int t_ind = 0;
    int type = orc__dm_type(s->dm, t_ind);
    *lmpc += orc__prior_add(&s->pr, nm->Ns, nm->nloci, type);
    prepare_add_new_term(s, ind, type, d2, 1);
    orc_model_add_term(nm, (uint32_t)ind, t_ind, d2);
  }
}

static void readd_removed_for_dr(orc_sampler* s, size_t ind)
{
  orc_model* cm = s->cur;
  int mi = cm->model_inds[ind];
  int t_ind = cm->type_inds[mi];
  int type = orc__dm_type(s->dm, t_ind);
  double d2[2];
  d2[0] = cm->dv[cm->x_ind1[mi]];
  if (type == ORC_AH) d2[1] = cm->dv[cm->x_ind2[mi]];
  orc_model_add_term(s->nw, (uint32_t)ind, t_ind, d2);
}

/* second-stage draw shared by both DR branches, src/sampler.cpp-equivalent
 * code at src/sampler.hpp:1008-1028 and 1592-1612 */
static unsigned long dr_second_stage(orc_sampler* s, int ms, double max_log)
{
  unsigned long nmodels = 1ul << ms, mask = nmodels - 1;
  double* p = s->dr_probs;
  double sum = 0.0;
  for (unsigned long i = 0; i < nmodels / 2; ++i) {
    unsigned long j = (~i) & mask;
    double a = p[j] - p[i];
    if (a > 0) { p[i] = exp(p[j] - max_log) * (1 - exp(-a)); p[j] = -1.0; }
    else { p[i] = exp(p[i] - max_log) * (1 - exp(a)); p[j] = 1.0; }
    p[i] += sum; sum = p[i];
  }
  unsigned long sm = sample_discrete(p, nmodels / 2, ms / 2 - 3, &s->rng);
  if (p[(~sm) & mask] < 0) sm = (~sm) & mask;
  return sm;
}

/* src/sampler.hpp:899-1123 */
static int do_multistep_additions_and_removals(orc_sampler* s)
{
  int ms = sample_discrete_naive(s->q_pms, s->max_ms, &s->rng) + 1;
  s->movesize = ms;
  double lqf = 0, lqb = 0, lmpc = 0;
  int ms_rem = 0;
  prepare_addrem(s, &lqf, &ms_rem, ms);
  do_addrem(s, &lqf, &lqb, &lmpc, ms);
  backward_prepare_addrem(s, &lqb, ms);
  double log_r = lqb - lqf;
  log_r += lmpc + s->nw->loglik - s->cur->loglik;
  if (s->dr == 0) s->r_ms_sum[ms - 1] += (log_r >= 0 ? 1.0 : exp(log_r));
  s->r_ms_n[ms - 1] += 1.0;
  if (log_r >= 0 || log(orc_rng_u01(&s->rng)) <= log_r) {
    orc_model_copy(s->cur, s->nw);
    if (s->dr != 0) s->r_ms_sum[ms - 1] += 1.0;
    return ms;
  }
  if (ms <= s->dr && ms > 1) {
    size_t const_loci = s->cur->nloci - ms_rem;
    int ms_add = ms - ms_rem;
    double z_add = orc_sumtree_total(s->dd_add), z_rem = orc_sumtree_total(s->dd_rem);
    unsigned long nmb = (1ul << ms_add) - 1;
    for (int i = 0; i < ms; ++i) {
      size_t ind = s->mi[i];
      if (s->mi_isadd[i]) { z_add += s->q_add[ind]; z_rem -= s->q_rem[ind]; }
      else readd_removed_for_dr(s, ind);
    }
    for (int i = 0; i < ms; ++i) {
      size_t ind = (size_t)s->mi_map[i];
      size_t mi = s->nw->model_inds[ind];
      s->dr_b2n[mi - const_loci] = (unsigned char)i;
      s->dr_q_add[i] = s->q_add[ind];
      s->dr_q_rem[i] = s->q_rem[ind];
    }
    double max_log;
    orc__exhaustive_modelset(s->exh, s->nw, (int)const_loci, ms, s->dr_probs, &max_log);
    orc_proposal_probs_exh(s->use_types, ms, s->dr_b2n, s->dr_q_add, s->dr_q_rem,
                           z_add, z_rem, const_loci, s->m_g, s->dr_lqt, s->dr_probs);
    unsigned long mask = (1ul << ms) - 1;
    unsigned long sm = dr_second_stage(s, ms, max_log);
    int moved = 0;
    if (sm == ((~nmb) & mask)) {
      orc_model_copy(s->nw, s->cur);
      undo_dd(s, ms);
    } else {
      for (int i = 0; i < ms; ++i) {
        if ((sm >> i) & 1) {
          if (!((nmb >> i) & 1)) {
            size_t ind = (size_t)s->mi_map[s->dr_b2n[i]];
            orc_sumtree_zero(s->dd_add, ind); orc_sumtree_unzero(s->dd_rem, ind);
          } else ++moved;
        } else {
          if ((nmb >> i) & 1) {
            size_t ind = (size_t)s->mi_map[s->dr_b2n[i]];
            orc_sumtree_unzero(s->dd_add, ind); orc_sumtree_zero(s->dd_rem, ind);
          } else ++moved;
        }
      }
      unsigned long removals = mask & sm;
      size_t cl = const_loci;
      for (int i = 0; i < ms; ++i) {
        if ((removals & 1) == 0) { orc_model_remove_term(s->nw, (int)(cl + i)); --cl; }
        removals >>= 1;
      }
      s->r_ms_sum[ms - 1] += (double)moved / (double)ms;
      orc_model_copy(s->cur, s->nw);
    }
    return moved;
  }
  orc_model_copy(s->nw, s->cur);
  undo_dd(s, ms);
  return 0;
}

/* src/sampler.hpp:1271-1418 */
static int do_switch_of_nearby_snps(orc_sampler* s)
{
  orc_model* cm = s->cur; orc_model* nm = s->nw;
  size_t full = s->m_g - cm->nloci;
  size_t mm2 = (s->max_ms + 1) / 2;
  size_t mx = mm2 < (full + 1) / 2 ? mm2 : (full + 1) / 2;
  mx = mx < (size_t)cm->nloci ? mx : (size_t)cm->nloci;
  int ms = sample_discrete_naive(s->q_pms_nbs, (int)mx, &s->rng) + 1;
  s->movesize = ms;
  size_t nbh = (full - ms) / 2;
  if (nbh < 1) nbh = 1;
  if (nbh > s->max_nbh) nbh = s->max_nbh;
  int64_t m_g = (int64_t)s->m_g;
  for (int i = 0; i < ms; ++i) {
    size_t mir = (size_t)floor(orc_rng_u01(&s->rng) * nm->nloci);
    int64_t ir = nm->loci[mir];
    s->mi_rem[i] = (size_t)ir;
    int64_t ia;
    int j = 0;
    int move = (int)floor(orc_rng_u01(&s->rng) * nbh) + 1;
    if (orc_rng_u01(&s->rng) < 0.5) {
      ia = ir; j = 0;
      while (j < move) {
        ++ia;
        if (ia < m_g) { if (cm->model_inds[ia] < 0) ++j; }
        else break;
      }
      if (ia == m_g) {
        ia = ir; move = move + move - j; j = 0;
        while (j < move) {
          --ia;
          while (cm->model_inds[ia] >= 0) --ia;
          ++j;
        }
      }
    } else {
      ia = ir; j = 0;
      if (ia > 0) {
        while (j < move) {
          --ia;
          if (ia > 0) { if (cm->model_inds[ia] < 0) ++j; }
          else break;
        }
      }
      if (j < move) {
        if (cm->model_inds[ia] < 0) ++j;
        if (j < move) {
          ia = ir; move = move + move - j; j = 0;
          while (j < move) {
            ++ia;
            while (cm->model_inds[ia] >= 0) ++ia;
            ++j;
          }
        }
      }
    }
    s->mi_add[i] = (size_t)ia;
    int t_ind = nm->type_inds[mir];
    orc_model_remove_term(nm, (int)mir);
    cm->model_inds[ia] = t_ind;
  }
  double d2[2];
  for (int i = 0; i < ms; ++i) {
    size_t ia = s->mi_add[i];
    int t_ind = cm->model_inds[ia];
    int type = orc__dm_type(s->dm, t_ind);
    prepare_add_new_term(s, ia, type, d2, 1);
    orc_model_add_term(nm, (uint32_t)ia, t_ind, d2);
    cm->model_inds[ia] = -1;
  }
  double log_r = nm->loglik - cm->loglik;
  if (log_r >= 0 || log(orc_rng_u01(&s->rng)) <= log_r) {
    orc_model_copy(cm, nm);
    for (int i = 0; i < ms; ++i) {
      orc_sumtree_unzero(s->dd_add, s->mi_rem[i]); orc_sumtree_zero(s->dd_rem, s->mi_rem[i]);
      orc_sumtree_zero(s->dd_add, s->mi_add[i]); orc_sumtree_unzero(s->dd_rem, s->mi_add[i]);
    }
    return 2 * ms;
  }
  orc_model_copy(nm, cm);
  return 0;
}

/* Offset o in [-H, -1] U [1, H] <-> column of a weight row. */
static int lw_col(int H, int o) { return o < 0 ? o + H : o + H - 1; }
static int lw_off(int H, int c) { return c < H ? c - H : c - H + 1; }

/* Learned swap (no counterpart in the reference; the Metropolis-Hastings
 * form of the swap move of src/sampler.hpp:1271-1418 with a non-uniform,
 * state-dependent partner draw).  A model SNP j is chosen uniformly; its
 * partner k is drawn from the frozen weights W[j][.] restricted to the SNPs
 * within H of j that are not in the current model; gamma' = gamma - j + k.
 * The reverse move chooses k uniformly (same model size) and draws j from
 * W[k][.] restricted to the SNPs within H of k not in gamma'.  Hence
 *   log r = l(gamma') - l(gamma) + log W[k][j] - log Z_r - log W[j][k] + log Z_f
 * with Z_f, Z_r the two normalisers (the prior and the uniform choices
 * cancel).  The sums run sequentially in offset order so that the device
 * reproduces them exactly. */
static int do_learned_swap(orc_sampler* s)
{
  orc_model* cm = s->cur; orc_model* nm = s->nw;
  const int H = s->lH;
  const int64_t m_g = (int64_t)s->m_g;
  s->movesize = 1;
  size_t mir = (size_t)floor(orc_rng_u01(&s->rng) * nm->nloci);
  int64_t j = nm->loci[mir];
  const float* Wj = s->lw + (size_t)j * 2 * H;
  double zf = 0.0;
  for (int c = 0; c < 2 * H; ++c) {
    int64_t k = j + lw_off(H, c);
    if (k >= 0 && k < m_g && cm->model_inds[k] < 0) zf += (double)Wj[c];
  }
  if (!(zf > 0.0)) return 0;
  double u = orc_rng_u01(&s->rng) * zf, acc = 0.0;
  int pick = -1;
  for (int c = 0; c < 2 * H; ++c) {
    int64_t k = j + lw_off(H, c);
    if (k >= 0 && k < m_g && cm->model_inds[k] < 0 && Wj[c] > 0.0f) {
      acc += (double)Wj[c];
      pick = c;
      if (u < acc) break;
    }
  }
  int64_t k = j + lw_off(H, pick);
  const float* Wk = s->lw + (size_t)k * 2 * H;
  double zr = 0.0, wr = 0.0;
  for (int c = 0; c < 2 * H; ++c) {
    int64_t q = k + lw_off(H, c);
    if (q < 0 || q >= m_g) continue;
    if (q == j || cm->model_inds[q] < 0) {
      zr += (double)Wk[c];
      if (q == j) wr = (double)Wk[c];
    }
  }
  if (!(wr > 0.0)) return 0;
  double lq = log(wr) - log(zr) - log((double)Wj[pick]) + log(zf);
  s->mi_rem[0] = (size_t)j; s->mi_add[0] = (size_t)k;
  int t_ind = nm->type_inds[mir];
  orc_model_remove_term(nm, (int)mir);
  double d2[2];
  prepare_add_new_term(s, (size_t)k, orc__dm_type(s->dm, t_ind), d2, 1);
  orc_model_add_term(nm, (uint32_t)k, t_ind, d2);
  double log_r = nm->loglik - cm->loglik + lq;
  if (log_r >= 0 || log(orc_rng_u01(&s->rng)) <= log_r) {
    orc_model_copy(cm, nm);
    orc_sumtree_unzero(s->dd_add, (size_t)j); orc_sumtree_zero(s->dd_rem, (size_t)j);
    orc_sumtree_zero(s->dd_add, (size_t)k); orc_sumtree_unzero(s->dd_rem, (size_t)k);
    return 2;
  }
  orc_model_copy(nm, cm);
  return 0;
}

/* Learned-proposal inputs from the sampler's current state (no reference
 * counterpart).  Marginal scores z_j = corr(x_j, y) sqrt(n), clipped to
 * +-50, from dense double dot products. */
void orc_sampler_marginal_z(orc_sampler* s, float* z)
{
  size_t n = s->cur->n;
  const double* y = orc__data_y(s->d);
  double my = 0, syy = 0;
  for (size_t i = 0; i < n; ++i) my += y[i];
  my /= (double)n;
  for (size_t i = 0; i < n; ++i) syy += (y[i] - my) * (y[i] - my);
  int type = orc__dm_type(s->dm, 0);
  int ot = s->pre_ot[type];
  double* x = (double*)malloc(sizeof(double) * n);
  for (size_t j = 0; j < s->m_g; ++j) {
    orc_datamodel_get(s->dm, type, j, x);
    double rx = 0;
    for (size_t i = 0; i < n; ++i) rx += x[i] * (y[i] - my);
    double css = s->pre_xx[j * s->pre_offset + ot + 1];
    double c = (css > 0 && syy > 0) ? rx / sqrt(css * syy) : 0.0;
    double v = c * sqrt((double)n);
    z[j] = (float)(v > 50 ? 50 : (v < -50 ? -50 : v));
  }
  free(x);
}

/* Swap labels of the current model: for each model SNP j (locus order) and
 * each SNP k = j + o (o = -H..-1, 1..H) inside the genome and outside the
 * model,
 *   label = lbf(k | gamma - j) - lbf(j | gamma - j),
 *   lbf(rx, css) = rx^2 / (2 sigma2 (css + d)) - log((css + d) / d) / 2,
 * where rx is the centred dot product of the SNP with the residual that
 * excludes j's effect (y - y_hat + beta_j x_j, formed densely here) and d is
 * the shared prior precision of the type; clipped to +-30.  Single coding
 * (not AH) only.  Returns the number of labels (at most cap are written). */
int64_t orc_sampler_swap_labels(orc_sampler* s, int H, uint32_t* oj, int32_t* oo, float* olab,
                                int64_t cap)
{
  const orc_model* cm = s->cur;
  size_t n = cm->n;
  const double* y = orc__data_y(s->d);
  double* r = (double*)malloc(sizeof(double) * n);
  double* xk = (double*)malloc(sizeof(double) * n);
  int64_t cnt = 0;
  for (int m = 0; m < cm->nloci; ++m) {
    int type = orc__dm_type(s->dm, cm->type_inds[m]);
    if (type == ORC_AH) continue;
    int64_t j = cm->loci[m];
    unsigned xi = cm->x_ind1[m];
    double bj = cm->beta[xi];
    const double* xj = cm->x + (size_t)xi * n;
    double mr = 0;
    for (size_t i = 0; i < n; ++i) { r[i] = y[i] - cm->y_hat[i] + bj * xj[i]; mr += r[i]; }
    mr /= (double)n;
    double rxj = 0;
    for (size_t i = 0; i < n; ++i) rxj += xj[i] * (r[i] - mr);
    int ot = s->pre_ot[type];
    double cj = s->pre_xx[(size_t)j * s->pre_offset + ot + 1];
    double d = s->pr.inv_tau2_alpha2[type], s2 = cm->sigma2;
    for (int c = 0; c < 2 * H; ++c) {
      int o = c < H ? c - H : c - H + 1;
      int64_t k = j + o;
      if (k < 0 || k >= (int64_t)s->m_g || cm->model_inds[k] >= 0) continue;
      double ck = s->pre_xx[(size_t)k * s->pre_offset + ot + 1];
      if (!(cj > 0) || !(ck > 0) || !(d > 0) || !(s2 > 0)) continue;
      orc_datamodel_get(s->dm, type, (size_t)k, xk);
      double rxk = 0;
      for (size_t i = 0; i < n; ++i) rxk += xk[i] * (r[i] - mr);
      double lk = rxk * rxk / (2 * s2 * (ck + d)) - 0.5 * log((ck + d) / d);
      double lj = rxj * rxj / (2 * s2 * (cj + d)) - 0.5 * log((cj + d) / d);
      double lab = lk - lj;
      if (!isfinite(lab)) continue;
      lab = lab > 30 ? 30 : (lab < -30 ? -30 : lab);
      if (cnt < cap) { oj[cnt] = (uint32_t)j; oo[cnt] = o; olab[cnt] = (float)lab; }
      ++cnt;
    }
  }
  free(r); free(xk);
  return cnt;
}

int orc_sampler_set_move_mix(orc_sampler* s, const double p[3])
{
  if (!s || p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] + p[1] + p[2] <= 0) return -1;
  for (int i = 0; i < NMOVES; ++i) s->p_moves[i] = 0.0;
  for (int i = 0; i < 3; ++i) s->p_moves[i] = p[i];
  s->p_moves_cumsum[0] = s->p_moves[0];
  for (int i = 1; i < NMOVES; ++i) s->p_moves_cumsum[i] = s->p_moves_cumsum[i - 1] + s->p_moves[i];
  return 0;
}

int orc_sampler_set_learned(orc_sampler* s, const float* w, int H, double p_learned)
{
  if (!s || H < 1 || H > 32 || p_learned < 0 || p_learned > 1) return -1;
  free(s->lw);
  s->lw = NULL; s->learn_on = 0;
  if (!w) return 0;
  size_t cnt = s->m_g * (size_t)(2 * H);
  s->lw = (float*)malloc(sizeof(float) * cnt);
  for (size_t i = 0; i < cnt; ++i) s->lw[i] = w[i] > 0.0f ? w[i] : 0.0f;
  s->lH = H; s->p_learned = p_learned; s->learn_on = 1;
  return 0;
}

/* src/sampler.hpp:1421-1669 */
static int do_statechange_of_nearby_snps(orc_sampler* s)
{
  orc_model* cm = s->cur; orc_model* nm = s->nw;
  size_t m_g = s->m_g;
  size_t a = s->max_nbh * 2, b = m_g - 1 < (size_t)s->max_ms ? m_g - 1 : (size_t)s->max_ms;
  size_t mx = a < b ? a : b;
  int ms = sample_discrete_naive(s->q_pms_nbc, (int)mx, &s->rng) + 1;
  s->movesize = ms;
  size_t nbh = s->max_nbh < m_g / 2 ? s->max_nbh : m_g / 2;
  size_t mic = (size_t)floor(orc_rng_u01(&s->rng) * nm->nloci);
  size_t ic = nm->loci[mic];
  size_t* perm = (size_t*)malloc(sizeof(size_t) * 2 * nbh);
  for (size_t i = 0; i < 2 * nbh; ++i) perm[i] = i;
  size_t wrap = ic + nbh;
  if (ic < nbh) wrap += (nbh - ic);
  else if (wrap > m_g - 1) wrap = m_g - 1;
  for (int i = 0; i < ms; ++i) {
    size_t ind = (size_t)floor(orc_rng_u01(&s->rng) * (nbh * 2 - i)) + i;
    size_t t = perm[i]; perm[i] = perm[ind]; perm[ind] = t;
  }
  int ms_rem = 0, ms_add = 0;
  for (int i = 0; i < ms; ++i) {
    size_t ind = perm[i] + ic + 1;
    if (ind > wrap) ind = ic - (ind - wrap);
    if (nm->model_inds[ind] < 0) s->mi_add[ms_add++] = ind;
    else s->mi_rem[ms_rem++] = ind;
  }
  free(perm);
  double lmpc = 0.0;
  for (int i = 0; i < ms_rem; ++i) {
    int mir = nm->model_inds[s->mi_rem[i]];
    int type = orc__dm_type(s->dm, nm->type_inds[mir]);
    lmpc += orc__prior_rem(&s->pr, nm->Ns, nm->nloci, type);
    orc_model_remove_term(nm, mir);
  }
  double d2[2];
  int t_ind = nm->type_inds[mic];
  int type = orc__dm_type(s->dm, t_ind);
  for (int i = 0; i < ms_add; ++i) {
    size_t ia = s->mi_add[i];
    lmpc += orc__prior_add(&s->pr, nm->Ns, nm->nloci, type);
    prepare_add_new_term(s, ia, type, d2, 1);
    orc_model_add_term(nm, (uint32_t)ia, t_ind, d2);
  }
  double log_r = lmpc + nm->loglik - cm->loglik;
  log_r += log((double)cm->nloci) - log((double)nm->nloci);
  if (log_r >= 0 || log(orc_rng_u01(&s->rng)) <= log_r) {
    orc_model_copy(cm, nm);
    for (int i = 0; i < ms_add; ++i) {
      orc_sumtree_zero(s->dd_add, s->mi_add[i]); orc_sumtree_unzero(s->dd_rem, s->mi_add[i]);
    }
    for (int i = 0; i < ms_rem; ++i) {
      orc_sumtree_unzero(s->dd_add, s->mi_rem[i]); orc_sumtree_zero(s->dd_rem, s->mi_rem[i]);
    }
    return ms;
  }
  if (ms <= s->dr && ms > 1) {
    size_t const_loci = cm->nloci - ms_rem;
    unsigned long nmb = (1ul << ms_add) - 1;
    for (int i = 0; i < ms_rem; ++i) readd_removed_for_dr(s, s->mi_rem[i]);
    double max_log;
    orc__exhaustive_modelset(s->exh, nm, (int)const_loci, ms, s->dr_probs, &max_log);
    unsigned long nmodels = 1ul << ms, mask = nmodels - 1;
    for (unsigned long i = 0; i < nmodels; ++i)
      s->dr_probs[i] -= log((double)(const_loci + __builtin_popcountl(i)));
    unsigned long sm = dr_second_stage(s, ms, max_log);
    int moved = 0;
    if (sm == ((~nmb) & mask)) {
      orc_model_copy(nm, cm);
    } else {
      for (int i = 0; i < ms; ++i) {
        if ((sm >> i) & 1) {
          if ((nmb >> i) & 1) {
            ++moved;
            size_t ind = s->mi_add[i];
            orc_sumtree_zero(s->dd_add, ind); orc_sumtree_unzero(s->dd_rem, ind);
          }
        } else {
          if (!((nmb >> i) & 1)) {
            ++moved;
            size_t ind = s->mi_rem[i - ms_add];
            orc_sumtree_unzero(s->dd_add, ind); orc_sumtree_zero(s->dd_rem, ind);
          }
        }
      }
      unsigned long removals = mask & sm;
      size_t cl = const_loci;
      for (int i = 0; i < ms; ++i) {
        if ((removals & 1) == 0) { orc_model_remove_term(nm, (int)(cl + i)); --cl; }
        removals >>= 1;
      }
      orc_model_copy(cm, nm);
    }
    return moved;
  }
  orc_model_copy(nm, cm);
  return 0;
}

/* src/sampler.hpp:1715-1725 */
static void adapt_acptrate(orc_sampler* s, double* p_ms, size_t* na, size_t* nm, double t)
{
  double ar = (double)*na / (double)*nm;
  *p_ms = *p_ms + (s->acpt_goal - ar) / t;
  *p_ms = fmax(fmin(*p_ms, 0.99), 0.01);
  *na = 1; *nm = 2;
}

/* src/sampler.hpp:1728-1769: adds the current truncated geometric move-size
 * distribution to q_p0 and returns the grid point (50 points from 0.01 to
 * 0.99) of largest importance-weighted expected jump distance */
double orc_adapt_jd_mb(double p_ms, const double* r_ms, const double* r_n, double* q_p0, int max_ms)
{
  double q = 1 - p_ms;
  double zp = 1 - pow(q, max_ms);
  double q_p = p_ms / zp;
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
  return p_best;
}

static void adapt_jd_mb(orc_sampler* s, double* p_ms, const double* r_ms,
                        const double* r_n, int max_ms)
{
  *p_ms = orc_adapt_jd_mb(*p_ms, r_ms, r_n, s->q_p0_ms, max_ms);
}

static void trace_push(orc_sampler* s, uint8_t mv, uint8_t jd)
{
  if (s->tr_len == s->tr_cap) {
    s->tr_cap = s->tr_cap ? 2 * s->tr_cap : 1024;
    s->tr_move = (uint8_t*)realloc(s->tr_move, s->tr_cap);
    s->tr_jd = (uint8_t*)realloc(s->tr_jd, s->tr_cap);
    s->tr_ms = (uint8_t*)realloc(s->tr_ms, s->tr_cap);
    s->tr_size = (uint32_t*)realloc(s->tr_size, s->tr_cap * sizeof(uint32_t));
    s->tr_ll = (double*)realloc(s->tr_ll, s->tr_cap * sizeof(double));
  }
  s->tr_move[s->tr_len] = mv; s->tr_jd[s->tr_len] = jd;
  s->tr_ms[s->tr_len] = (uint8_t)s->movesize;
  s->tr_size[s->tr_len] = (uint32_t)s->cur->nloci;
  s->tr_ll[s->tr_len] = s->cur->loglik;
  s->tr_len++;
}

static void thinned_push(orc_sampler* s)
{
  if (s->th_len == s->th_cap) {
    s->th_cap = s->th_cap ? 2 * s->th_cap : 256;
    s->th_size = (uint32_t*)realloc(s->th_size, s->th_cap * sizeof(uint32_t));
    s->th_ll = (double*)realloc(s->th_ll, s->th_cap * sizeof(double));
    s->th_s2 = (double*)realloc(s->th_s2, s->th_cap * sizeof(double));
    s->th_alpha = (double*)realloc(s->th_alpha, s->th_cap * sizeof(double));
    s->th_pve = (double*)realloc(s->th_pve, 3 * s->th_cap * sizeof(double));
  }
  int k = s->cur->nloci;
  while (s->th_loci_len + k > s->th_loci_cap) {
    s->th_loci_cap = s->th_loci_cap ? 2 * s->th_loci_cap : 4096;
    s->th_loci = (uint32_t*)realloc(s->th_loci, s->th_loci_cap * sizeof(uint32_t));
  }
  if (k > 0) memcpy(s->th_loci + s->th_loci_len, s->cur->loci, sizeof(uint32_t) * k);
  s->th_loci_len += k;
  s->th_size[s->th_len] = (uint32_t)k;
  s->th_ll[s->th_len] = s->cur->loglik;
  s->th_s2[s->th_len] = s->cur->sigma2;
  s->th_alpha[s->th_len] = s->pr.alpha;
  memcpy(s->th_pve + 3 * s->th_len, s->pves, 3 * sizeof(double));
  if (s->o.save_beta) {
    /* src/sampler.cpp:710-713 */
    int kb = s->cur->beta_len;
    while (s->th_beta_len + kb > s->th_beta_cap) {
      s->th_beta_cap = s->th_beta_cap ? 2 * s->th_beta_cap : 4096;
      s->th_beta = (double*)realloc(s->th_beta, s->th_beta_cap * sizeof(double));
    }
    if (kb > 0) memcpy(s->th_beta + s->th_beta_len, s->cur->beta, sizeof(double) * kb);
    s->th_beta_len += kb;
  }
  s->th_len++;
}

/* src/sampler.cpp:551-620 */
int orc_sampler_setup(orc_sampler* s)
{
  if (!isfinite(s->cur->loglik)) return -1;
  s->pves[0] = s->pves[1] = s->pves[2] = 0.0;
  s->k_move_size = 1000.0 / (double)s->o.n_rao_burnin;
  orc_model_copy(s->nw, s->cur);
  compute_proposal_dist(s);
  orc_sumtree_destroy(s->dd_add); orc_sumtree_destroy(s->dd_rem);
  s->dd_add = orc_sumtree_create(s->q_add, s->m_g);
  s->dd_rem = orc_sumtree_create(s->q_rem, s->m_g);
  for (size_t i = 0; i < s->m_g; ++i) s->dd_rem->zeroed[i] = 1;
  st_rebuild(s->dd_rem);
  compute_p_moves(s);
  orc__model_sample_beta_sigma2(s->cur, &s->rng);
  sampler_sample_missing(s);
  prior_sample_alpha_and_tau2(s, s->cur);
  orc_model_compute_log_likelihood(s->cur);
  orc_model_copy(s->nw, s->cur);
  s->setup_done = 1;
  return 0;
}

/* src/sampler.cpp:626-834 */
int orc_sampler_iterate(orc_sampler* s, uint64_t n_iter)
{
  if (!s->setup_done) return -1;
  const orc_opts* o = &s->o;
  uint64_t end = s->n_iter + n_iter;
  for (uint64_t iter = s->n_iter; iter < end; ++iter) {
    if ((iter + 1) % o->n_sample_tau2_and_missing == 0) {
      /* src/sampler.cpp:628-635; a refresh whose factorisation fails (the
       * reference would continue with ll = -inf and then draw sigma2 from an
       * infinite scale forever) is undone: the previous tau2 / alpha and the
       * previous model (still in the working copy) are kept */
      double al = s->pr.alpha, al2 = s->pr.alpha2, ita[4];
      memcpy(ita, s->pr.inv_tau2_alpha2, sizeof(ita));
      sampler_sample_missing(s);
      prior_sample_alpha_and_tau2(s, s->cur);
      orc_model_compute_log_likelihood(s->cur);
      if (isfinite(s->cur->loglik)) {
        orc_model_copy(s->nw, s->cur);
      } else {
        orc_model_copy(s->cur, s->nw);
        s->pr.alpha = al; s->pr.alpha2 = al2;
        memcpy(s->pr.inv_tau2_alpha2, ita, sizeof(ita));
      }
    }
    /* with a learned table, a drawn swap is a learned swap with probability
     * p_learned (a state-independent choice between two kernels that each
     * keep the posterior) */
    int move = sample_discrete_naive(s->p_moves_cumsum, NMOVES, &s->rng);
    if (move == 1 && s->learn_on && orc_rng_u01(&s->rng) < s->p_learned) move = 3;
    int jd = 0;
    if (!isfinite(s->cur->loglik)) return -2;
    switch (move) {
      case 3:
        if (s->cur->nloci > 0) jd = do_learned_swap(s);
        break;
      case 0: jd = do_multistep_additions_and_removals(s); break;
      case 1:
        if (s->cur->nloci > 0 && (size_t)s->cur->nloci < s->m_g - 1)
          jd = do_switch_of_nearby_snps(s);
        break;
      case 2:
        if (s->cur->nloci > 0 && s->m_g > 2) jd = do_statechange_of_nearby_snps(s);
        break;
      default: return -3;
    }
    s->n_moves[move]++;
    s->n_acpt[move] += jd > 0;
    s->n_accepted += jd > 0;
    if ((iter + 1) % o->thin == 0 || (iter + 2) % o->n_sample_tau2_and_missing == 0)
      orc__model_sample_beta_sigma2(s->cur, &s->rng);
    trace_push(s, (uint8_t)move, (uint8_t)jd);
    if ((iter + 1) % o->thin == 0) {
      orc_model_compute_pve(s->cur, s->pves);
      thinned_push(s);
    }
    if ((iter + 1) % o->n_rao == 0) {
      if (!o->flat_proposal_dist || s->n_rao_burnin <= 0) {
        data_model_sample_missing(s, s->cur->model_inds);
        orc_sampler_rb(s, s->cur->y_hat, s->p_r);
      }
      if (s->n_rao_burnin <= 0) {
        double z1 = (double)(s->p_rao_n + 1), z2 = (double)s->p_rao_n / z1;
        for (size_t i = 0; i < s->m_g; ++i) s->p_rao[i] = z2 * s->p_rao[i] + s->p_r[i] / z1;
        ++s->p_rao_n;
      } else {
        s->n_rao_burnin--;
      }
      if (o->adaptation || s->n_rao_burnin > 0) {
        if (!o->flat_proposal_dist) {
          double z1 = (double)(s->p_proposal_n + 1), z2 = (double)s->p_proposal_n / z1;
          for (size_t i = 0; i < s->m_g; ++i)
            s->p_proposal[i] = z2 * s->p_proposal[i] + s->p_r[i] / z1;
        }
        if (s->adapt_p_ms) {
          double t = fmax(1.0, s->k_move_size * (double)s->p_proposal_n);
          if (s->acpt_goal > 0)
            adapt_acptrate(s, &s->p_ms, &s->n_acpt[0], &s->n_moves[0], 2.0 + t);
          else
            adapt_jd_mb(s, &s->p_ms, s->r_ms_sum, s->r_ms_n, s->max_ms);
          orc_geometric_dist_cdf(s->max_ms, s->p_ms, s->q_pms);
        }
        ++s->p_proposal_n;
        if (!o->flat_proposal_dist) {
          compute_proposal_dist(s);
          orc_sumtree_set_weights(s->dd_add, s->q_add);
          orc_sumtree_set_weights(s->dd_rem, s->q_rem);
        }
      }
    }
  }
  s->n_iter = end;
  return 0;
}

int orc_sampler_sample(orc_sampler* s, uint64_t n_iter)
{
  int rc = orc_sampler_setup(s);
  if (rc) return rc;
  return orc_sampler_iterate(s, n_iter);
}

const double* orc_sampler_p_rao(const orc_sampler* s) { return s->p_rao; }
const double* orc_sampler_p_proposal(const orc_sampler* s) { return s->p_proposal; }

/* A new source of the add/remove proposal: p_proposal is replaced and the
 * proposal weights follow as after an adaptation step (src/sampler.cpp:
 * 797-806, 517-524): q_add = max(p, q_add_min), q_rem = max(1 - p,
 * q_rem_min), sum trees rebuilt with their zeroed leaves kept.  With
 * adaptation finished the proposal then stays fixed, so the moves keep the
 * posterior for any such table. */
int orc_sampler_set_p_proposal(orc_sampler* s, const double* p)
{
  for (size_t i = 0; i < s->m_g; ++i)
    if (!(p[i] >= 0.0 && p[i] <= 1.0)) return -1;
  memcpy(s->p_proposal, p, sizeof(double) * s->m_g);
  compute_proposal_dist(s);
  orc_sumtree_set_weights(s->dd_add, s->q_add);
  orc_sumtree_set_weights(s->dd_rem, s->q_rem);
  return 0;
}
orc_model* orc_sampler_current_model(orc_sampler* s) { return s->cur; }
orc_model* orc_sampler_new_model(orc_sampler* s) { return s->nw; }
uint64_t orc_sampler_rng_counter(const orc_sampler* s)
{ return (uint64_t)s->rng.ctr[0] | ((uint64_t)s->rng.ctr[1] << 32); }

int64_t orc_sampler_trace(const orc_sampler* s, uint8_t* move, uint8_t* jd,
                          uint8_t* ms, uint32_t* msize, double* loglik)
{
  if (move) memcpy(move, s->tr_move, s->tr_len);
  if (jd) memcpy(jd, s->tr_jd, s->tr_len);
  if (ms) memcpy(ms, s->tr_ms, s->tr_len);
  if (msize) memcpy(msize, s->tr_size, s->tr_len * sizeof(uint32_t));
  if (loglik) memcpy(loglik, s->tr_ll, s->tr_len * sizeof(double));
  return s->tr_len;
}

int64_t orc_sampler_thinned(const orc_sampler* s, uint32_t* sizes,
                            uint32_t* loci, int64_t loci_cap, double* ll,
                            double* sigma2, double* alpha)
{
  if (sizes) memcpy(sizes, s->th_size, s->th_len * sizeof(uint32_t));
  if (loci) {
    int64_t c = s->th_loci_len < loci_cap ? s->th_loci_len : loci_cap;
    memcpy(loci, s->th_loci, c * sizeof(uint32_t));
  }
  if (ll) memcpy(ll, s->th_ll, s->th_len * sizeof(double));
  if (sigma2) memcpy(sigma2, s->th_s2, s->th_len * sizeof(double));
  if (alpha) memcpy(alpha, s->th_alpha, s->th_len * sizeof(double));
  return s->th_len;
}

int64_t orc_sampler_thinned_beta(const orc_sampler* s, double* out, int64_t cap)
{
  if (out && s->th_beta_len) memcpy(out, s->th_beta, sizeof(double) * (size_t)(cap < s->th_beta_len ? cap : s->th_beta_len));
  return s->th_beta_len;
}

int64_t orc_sampler_thinned_pve(const orc_sampler* s, double* pve)
{
  if (pve) memcpy(pve, s->th_pve, 3 * s->th_len * sizeof(double));
  return s->th_len;
}

void orc_sampler_state(const orc_sampler* s, double out[8])
{
  out[0] = s->cur->loglik; out[1] = s->cur->nloci; out[2] = s->cur->sigma2;
  out[3] = s->pr.alpha; out[4] = s->p_ms; out[5] = (double)s->n_accepted;
  out[6] = (double)s->n_iter; out[7] = (double)s->th_loci_len;
}

int orc_sampler_loci(const orc_sampler* s, uint32_t* loci)
{
  memcpy(loci, s->cur->loci, sizeof(uint32_t) * s->cur->nloci);
  return s->cur->nloci;
}

void orc_model_sample_beta_sigma2(orc_sampler* s, orc_model* m)
{ orc__model_sample_beta_sigma2(m, &s->rng); }

int64_t orc_model_draw_beta_sigma2(orc_sampler* s, orc_model* m, int64_t N, double* s2_out,
                                   double* beta_out)
{
  int64_t k = 0;
  for (int64_t i = 0; i < N; ++i) {
    orc__model_sample_beta_sigma2(m, &s->rng);
    k = m->beta_len;
    s2_out[i] = m->sigma2;
    for (int64_t j = 0; j < k; ++j) beta_out[i * k + j] = m->beta[j];
  }
  return k;
}

int64_t orc_sampler_refresh_draws(orc_sampler* s, int64_t N, double* alpha_out, double* dv_out)
{
  orc_model* m = s->cur;
  orc_prior* p = &s->pr;
  const double a0 = p->alpha, a20 = p->alpha2;
  double ita0[4];
  for (int t = 0; t < 4; ++t) ita0[t] = p->inv_tau2_alpha2[t];
  const int k = m->beta_len - m->m_e;
  double* dv0 = (double*)malloc(sizeof(double) * (k > 0 ? k : 1));
  for (int j = 0; j < k; ++j) dv0[j] = m->dv[m->m_e + j];
  for (int64_t i = 0; i < N; ++i) {
    p->alpha = a0; p->alpha2 = a20;
    for (int t = 0; t < 4; ++t) p->inv_tau2_alpha2[t] = ita0[t];
    for (int j = 0; j < k; ++j) m->dv[m->m_e + j] = dv0[j];
    prior_sample_alpha_and_tau2(s, m);
    alpha_out[i] = p->alpha;
    for (int j = 0; j < k; ++j) dv_out[i * k + j] = m->dv[m->m_e + j];
  }
  p->alpha = a0; p->alpha2 = a20;
  for (int t = 0; t < 4; ++t) p->inv_tau2_alpha2[t] = ita0[t];
  for (int j = 0; j < k; ++j) m->dv[m->m_e + j] = dv0[j];
  m->mu_beta_computed = 0;
  free(dv0);
  return k;
}

void orc_exhaustive_modelset(orc_model* full, int const_loci, int n_inds,
                             double* log_model_probs, double* max_log)
{
  orc_sampler* s = full->s;
  orc__exhaustive_modelset(s->exh, full, const_loci, n_inds, log_model_probs, max_log);
}

double orc_prior_change_on_add(const orc_sampler* s, const int* Ns, int n, int t)
{ return orc__prior_add(&s->pr, Ns, n, t); }
double orc_prior_change_on_rem(const orc_sampler* s, const int* Ns, int n, int t)
{ return orc__prior_rem(&s->pr, Ns, n, t); }
double orc_prior_change_on_swi(const orc_sampler* s, const int* Ns, int ta, int tr)
{ return orc__prior_swi(&s->pr, Ns, ta, tr); }
double orc_prior_log_model(const orc_sampler* s, const int* Ns)
{ return orc__prior_log_model(&s->pr, Ns); }

void orc_prior_get(const orc_sampler* s, double out[8])
{
  out[0] = s->pr.g_a; out[1] = s->pr.g_b; out[2] = s->pr.alpha;
  out[3] = s->pr.inv_tau2_alpha2[0]; out[4] = s->pr.inv_tau2_alpha2[1];
  out[5] = s->pr.nus2_plus_yy; out[6] = s->pr.n_plus_nu; out[7] = s->pr.s2_sigma2;
}
