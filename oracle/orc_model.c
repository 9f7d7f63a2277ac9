/*
 * orc_model.c -- Prior, Model and ExhModel of the reference, restated for
 * the CPU oracle (FP64, column-major, capacity kmax like MAXMODELSIZE).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "orc_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define E(a, ld, r, c) ((a)[(size_t)(c) * (ld) + (r)])

/* ---------------------------------------------------------------- Prior */
/* src/prior.hpp:273-290 */
static int solve_betadist_params(orc_prior* p, double e_q, double var_q)
{
  double m_g = (double)p->m_g;
  if (p->m_g > 1) {
    double z = (var_q - e_q * (1 - e_q)) / ((m_g - 1) * e_q);
    p->g_a = (z - 1) / (1 - m_g * z / e_q);
    p->g_b = (m_g / e_q - 1) * p->g_a;
  } else {
    double z = e_q / (1 - e_q);
    p->g_b = z / var_q / pow(1 + z, 3) - 1 / (1 + z);
    p->g_a = z * p->g_b;
  }
  if (p->g_a <= 0 || p->g_b <= 0 || isnan(p->g_a) || isnan(p->g_b)) return -1;
  return 0;
}

/* src/prior.hpp:45-122 */
int orc__prior_init(orc_prior* p, const orc_data* d, const orc_datamodel* dm,
                    const double* types_prior, double e_qg, double var_qg,
                    const double* inv_tau2_e, double nu_sigma2,
                    double s2_sigma2, const double* nu_tau2,
                    const double* s2_tau2, double mu_alpha, int use_ind)
{
  memset(p, 0, sizeof(*p));
  p->n = orc__data_n(d); p->m_g = orc__data_m_g(d); p->m_e = orc__data_m_e(d);
  double st[4]; orc_data_stats(d, st);
  p->n_plus_nu = p->n + nu_sigma2;
  p->minus_n_plus_nu_2 = -0.5 * p->n_plus_nu;
  p->nus2_plus_yy = nu_sigma2 * s2_sigma2 + st[0];
  p->nus2 = nu_sigma2 * s2_sigma2;
  p->mu_alpha = mu_alpha;
  p->use_individual_tau2 = use_ind;
  p->n_types = orc__dm_n_types(dm);
  p->nu_sigma2 = nu_sigma2; p->s2_sigma2 = s2_sigma2;
  p->alpha = mu_alpha > 0.5 ? mu_alpha : 0.5;
  p->alpha2 = p->alpha * p->alpha;
  for (size_t i = 0; i < p->m_e && i < 64; ++i) p->inv_tau2_e[i] = inv_tau2_e[i];
  if (s2_sigma2 <= 0 || nu_sigma2 < 0) return -1;
  if (solve_betadist_params(p, e_qg, var_qg)) return -2;
  p->types_prior_sum = 0;
  for (int i = 0; i < 5; ++i) {
    p->allow_types[i] = orc__dm_allow_types(dm, i);
    if (p->allow_types[i]) {
      p->types_prior[i] = types_prior[i];
      p->types_prior_sum += types_prior[i];
    } else {
      p->types_prior[i] = NAN;
    }
  }
  for (int i = 0; i < 4; ++i) {
    p->allow_terms[i] = orc__dm_allow_terms(dm, i);
    if (p->allow_terms[i]) {
      p->nu_tau2[i] = nu_tau2[i];
      p->s2_tau2[i] = s2_tau2[i];
      if (s2_tau2[i] <= 0 || nu_tau2[i] <= 0) return -3;
      p->nus2_tau2[i] = nu_tau2[i] * s2_tau2[i];
      double tau2_init = p->nus2_tau2[i] / fmax(0.1, nu_tau2[i] - 2);
      p->inv_tau2_alpha2[i] = 1.0 / (p->alpha2 * tau2_init);
    } else {
      p->nu_tau2[i] = p->s2_tau2[i] = p->nus2_tau2[i] = NAN;
      p->inv_tau2_alpha2[i] = NAN;
    }
  }
  p->allow_terms[4] = 0;
  return 0;
}

/* src/prior.hpp:144-167 */
double orc__prior_add(const orc_prior* p, const int* Ns, int n_old, int type)
{
  return (log(p->g_a + n_old) - log(p->g_b + (double)p->m_g - n_old - 1)) +
         (log(p->types_prior[type] + Ns[type]) - log(p->types_prior_sum + n_old));
}

double orc__prior_rem(const orc_prior* p, const int* Ns, int n_old, int type)
{
  return (log(p->g_b + (double)p->m_g - n_old) - log(p->g_a + n_old - 1)) +
         (log(p->types_prior_sum + n_old - 1) - log(p->types_prior[type] + Ns[type] - 1));
}

double orc__prior_swi(const orc_prior* p, const int* Ns, int ta, int tr)
{
  return (ta == tr) ? 0 : log(p->types_prior[ta] + Ns[ta]) -
                          log(p->types_prior[tr] + Ns[tr] - 1);
}

/* src/prior.hpp:169-183 (gammaln -> lgamma) */
double orc__prior_log_model(const orc_prior* p, const int* Ns)
{
  double lp = 0; int n_g = 0;
  for (int i = 0; i < 5; ++i) {
    if (p->allow_types[i]) {
      lp += lgamma(p->types_prior[i] + (double)Ns[i]);
      n_g += Ns[i];
    }
  }
  lp += lgamma(p->g_a + (double)n_g) + lgamma(p->g_b + (double)p->m_g - n_g)
        - lgamma(p->types_prior_sum + (double)n_g);
  return lp;
}

/* src/prior.hpp:192-199 */
double orc__prior_draw_inv_tau2(const orc_prior* p, int type, orc_rng* rng)
{
  if (p->use_individual_tau2)
    return 1.0 / (p->alpha2 * orc_rng_sinvchi2(rng, p->nu_tau2[type], p->s2_tau2[type]));
  return p->inv_tau2_alpha2[type];
}

/* ---------------------------------------------------------------- Model */
static double dot(const double* a, const double* b, size_t n)
{
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* y := A x (NoTrans dgemv, column order), A is n x k with stride ld */
static void gemv_n(const double* a, size_t ld, size_t n, int k,
                   const double* x, double* y)
{
  for (size_t i = 0; i < n; ++i) y[i] = 0.0;
  for (int j = 0; j < k; ++j) {
    double t = x[j];
    if (t != 0.0) {
      const double* col = a + (size_t)j * ld;
      for (size_t i = 0; i < n; ++i) y[i] += t * col[i];
    }
  }
}

orc_model* orc__model_create(orc_sampler* s, const orc_data* d,
                             const orc_datamodel* dm, const orc_prior* pr,
                             int kmax)
{
  orc_model* m = (orc_model*)calloc(1, sizeof(orc_model));
  m->s = s; m->d = d; m->dm = dm; m->pr = pr; m->kmax = kmax;
  m->n = orc__data_n(d); m->m_g = orc__data_m_g(d); m->m_e = (int)orc__data_m_e(d);
  size_t n = m->n;
  m->x = (double*)calloc(n * (size_t)kmax, sizeof(double));
  m->xx = (double*)calloc((size_t)kmax * kmax, sizeof(double));
  m->l = (double*)calloc((size_t)kmax * kmax, sizeof(double));
  m->xy = (double*)calloc(kmax, sizeof(double));
  m->v = (double*)calloc(kmax, sizeof(double));
  m->dv = (double*)calloc(kmax, sizeof(double));
  m->beta = (double*)calloc(kmax, sizeof(double));
  m->mu_beta = (double*)calloc(kmax, sizeof(double));
  m->y_hat = (double*)calloc(n, sizeof(double));
  m->y_hat_tmp_e = (double*)calloc(n, sizeof(double));
  m->loci = (uint32_t*)calloc(kmax, sizeof(uint32_t));
  m->type_inds = (unsigned char*)calloc(kmax, 1);
  m->x_ind1 = (uint32_t*)calloc(kmax, sizeof(uint32_t));
  m->x_ind2 = (uint32_t*)calloc(kmax, sizeof(uint32_t));
  m->x_ind_to_g_ind = (uint32_t*)calloc(kmax, sizeof(uint32_t));
  m->x_types = (unsigned char*)calloc(kmax, 1);
  m->model_inds = (int32_t*)malloc(sizeof(int32_t) * m->m_g);
  for (size_t i = 0; i < m->m_g; ++i) m->model_inds[i] = -1;
  m->loglik = NAN;

  /* src/model.hpp:92-105: covariates (column 0 is the constant) */
  int me = m->m_e;
  const double* e = orc__data_e(d);
  const double* y = orc__data_y(d);
  memcpy(m->x, e, sizeof(double) * n * me);
  m->cols = me;
  for (int c = 0; c < me; ++c)
    for (int r = 0; r <= c; ++r)
      E(m->xx, kmax, r, c) = dot(e + (size_t)r * n, e + (size_t)c * n, n);
  for (int c = 0; c < me; ++c) m->xy[c] = dot(e + (size_t)c * n, y, n);
  for (int c = 0; c < me; ++c) { m->x_types[c] = ORC_AH; m->x_ind_to_g_ind[c] = 0; }
  for (int c = 0; c < me; ++c) m->dv[c] = pr->inv_tau2_e[c];
  orc_model_compute_log_likelihood(m);
  return m;
}

void orc__model_destroy(orc_model* m)
{
  if (!m) return;
  free(m->x); free(m->xx); free(m->l); free(m->xy); free(m->v); free(m->dv);
  free(m->beta); free(m->mu_beta); free(m->y_hat); free(m->y_hat_tmp_e);
  free(m->loci); free(m->type_inds); free(m->x_ind1); free(m->x_ind2);
  free(m->x_ind_to_g_ind); free(m->x_types); free(m->model_inds);
  free(m);
}

/* src/model.hpp:115-162 */
void orc_model_copy(orc_model* dst, const orc_model* src)
{
  if (dst == src) return;
  int k = src->cols, ld = src->kmax;
  memcpy(dst->x, src->x, sizeof(double) * src->n * (size_t)k);
  for (int c = 0; c < k; ++c)
    for (int r = 0; r <= c; ++r) {
      E(dst->xx, ld, r, c) = E(src->xx, ld, r, c);
      E(dst->l, ld, r, c) = E(src->l, ld, r, c);
    }
  memcpy(dst->xy, src->xy, sizeof(double) * k);
  memcpy(dst->v, src->v, sizeof(double) * k);
  memcpy(dst->dv, src->dv, sizeof(double) * k);
  memcpy(dst->beta, src->beta, sizeof(double) * src->beta_len);
  memcpy(dst->mu_beta, src->mu_beta, sizeof(double) * src->mu_len);
  dst->beta_len = src->beta_len; dst->mu_len = src->mu_len;
  dst->cols = k;
  dst->sigma2 = src->sigma2;
  dst->syx_plus_vs2 = src->syx_plus_vs2;
  dst->log_det_invQ = src->log_det_invQ;
  dst->log_det_invQ_plus_xx = src->log_det_invQ_plus_xx;
  memcpy(dst->Ns, src->Ns, sizeof(dst->Ns));
  dst->mu_beta_computed = src->mu_beta_computed;
  dst->loglik = src->loglik;
  dst->nloci = src->nloci;
  memcpy(dst->loci, src->loci, sizeof(uint32_t) * src->nloci);
  memcpy(dst->type_inds, src->type_inds, src->nloci);
  memcpy(dst->x_ind1, src->x_ind1, sizeof(uint32_t) * src->nloci);
  memcpy(dst->x_ind2, src->x_ind2, sizeof(uint32_t) * src->nloci);
  memcpy(dst->x_ind_to_g_ind, src->x_ind_to_g_ind, sizeof(uint32_t) * k);
  memcpy(dst->x_types, src->x_types, k);
  memcpy(dst->model_inds, src->model_inds, sizeof(int32_t) * src->m_g);
}

/* src/model.hpp:558-576 */
static double prepare_l_chol(orc_model* m)
{
  int k = m->cols, ld = m->kmax;
  for (int c = 0; c < k; ++c)
    for (int r = 0; r <= c; ++r) E(m->l, ld, r, c) = E(m->xx, ld, r, c);
  double ls2 = 0;
  E(m->l, ld, 0, 0) += m->dv[0];
  for (int i = 1; i < k; ++i) {
    E(m->l, ld, i, i) += m->dv[i];
    ls2 += log(m->dv[i]);
  }
  return 0.5 * ls2;
}

/* The residual term S = nu s2 + y'y - v'v (src/model.hpp:226,498,545) cancels
 * when the factor is nearly singular; where it has lost its significant
 * digits (below 1e-8 of nu s2 + y'y) it is formed directly, as
 * nu s2 + |y - X mu|^2 + mu'D mu with mu = U^-1 v (equal in exact
 * arithmetic, a sum of non-negative terms).  The reference keeps the
 * cancelled value; a fit that close does not occur away from that case. */
#define ORC_S_TINY 1e-8
static void guard_syx(orc_model* m)
{
  if (m->syx_plus_vs2 >= ORC_S_TINY * m->pr->nus2_plus_yy) return;
  int k = m->cols, ld = m->kmax;
  size_t n = m->n;
  double* mu = (double*)malloc(sizeof(double) * (size_t)k);
  memcpy(mu, m->v, sizeof(double) * (size_t)k);
  orc_trsv_un(m->l, ld, k, mu);
  const double* y = orc__data_y(m->d);
  double rss = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double yh = 0.0;
    for (int c = 0; c < k; ++c) yh += m->x[(size_t)c * n + i] * mu[c];
    double r = y[i] - yh;
    rss += r * r;
  }
  double md = 0.0;
  for (int c = 0; c < k; ++c) md += m->dv[c] * mu[c] * mu[c];
  free(mu);
  m->syx_plus_vs2 = m->pr->nus2 + rss + md;
}

/* src/model.hpp:199-237 */
void orc_model_compute_log_likelihood(orc_model* m)
{
  if (m->cols < 1) { m->loglik = NAN; return; }
  int k = m->cols, ld = m->kmax;
  m->log_det_invQ = prepare_l_chol(m);
  if (!orc_chol(m->l, ld, k)) {
    m->loglik = -INFINITY; m->syx_plus_vs2 = INFINITY;
    m->mu_beta_computed = 0; return;
  }
  memcpy(m->v, m->xy, sizeof(double) * k);
  orc_trsv_ut(m->l, ld, k, m->v);
  m->syx_plus_vs2 = m->pr->nus2_plus_yy - dot(m->v, m->v, k);
  guard_syx(m);
  m->log_det_invQ_plus_xx = 0;
  for (int i = 0; i < k; ++i) m->log_det_invQ_plus_xx += log(E(m->l, ld, i, i));
  m->loglik = orc__finite_ll(m->log_det_invQ - m->log_det_invQ_plus_xx
                             + m->pr->minus_n_plus_nu_2 * log(m->syx_plus_vs2));
  m->mu_beta_computed = 0;
}

/* src/model.hpp:432-506 */
static void update_on_add(orc_model* m, uint32_t ind, int model_ind,
                          int type, int is_x_ind2, double dval, int upd)
{
  int col = m->cols, ld = m->kmax;
  size_t n = m->n;
  if (is_x_ind2) m->x_ind2[model_ind] = col;
  else { m->x_ind1[model_ind] = col; m->x_ind2[model_ind] = 0; }
  m->x_ind_to_g_ind[col] = ind;
  m->x_types[col] = (unsigned char)type;
  double* xcol = m->x + (size_t)col * n;
  orc_datamodel_get(m->dm, type, ind, xcol);
  m->cols = col + 1;
  m->xy[col] = dot(xcol, orc__data_y(m->d), n);
  m->dv[col] = dval;
  /* xxcol = X^T xcol (Trans dgemv, one dot per column) */
  for (int c = 0; c <= col; ++c)
    E(m->xx, ld, c, col) = dot(m->x + (size_t)c * n, xcol, n);
  if (upd) {
    if (!orc_chol_update(m->l, ld, col, &E(m->xx, ld, 0, col), dval)) {
      m->loglik = -INFINITY; m->syx_plus_vs2 = INFINITY;
      m->mu_beta_computed = 0; return;
    }
    double* lcol = &E(m->l, ld, 0, col);
    double vm2 = m->xy[col] - dot(lcol, m->v, col);
    m->v[col] = vm2 / lcol[col];
    vm2 = m->v[col] * m->v[col];
    m->syx_plus_vs2 = m->syx_plus_vs2 - vm2;
    guard_syx(m);
    m->log_det_invQ += 0.5 * log(dval);
    m->log_det_invQ_plus_xx += log(lcol[col]);
    m->loglik = orc__finite_ll(m->log_det_invQ - m->log_det_invQ_plus_xx
                               + m->pr->minus_n_plus_nu_2 * log(m->syx_plus_vs2));
  }
}

/* src/model.hpp:239-268 */
int orc_model_add_term(orc_model* m, uint32_t ind, int t_ind,
                       const double* d2)
{
  int type = orc__dm_type(m->dm, t_ind);
  int is_ah = (type == ORC_AH);
  int model_ind = m->nloci;
  if (m->cols + 1 + is_ah > m->kmax) return -1;
  m->loci[m->nloci] = ind;
  m->type_inds[m->nloci] = (unsigned char)t_ind;
  m->nloci++;
  m->Ns[type]++;
  m->model_inds[ind] = model_ind;
  m->mu_beta_computed = 0;
  if (is_ah) {
    update_on_add(m, ind, model_ind, ORC_A, 0, d2[0], 1);
    if (isfinite(m->syx_plus_vs2))
      update_on_add(m, ind, model_ind, ORC_H, 1, d2[1], 1);
  } else {
    update_on_add(m, ind, model_ind, type, 0, d2[0], 1);
  }
  return 0;
}

/* src/model.hpp:508-556 */
static void update_on_remove(orc_model* m, int x_ind, int upd)
{
  int k = m->cols, ld = m->kmax;
  size_t n = m->n;
  /* x.remove_column, xy.remove, xx.remove_colrow, erase vectors */
  memmove(m->x + (size_t)x_ind * n, m->x + (size_t)(x_ind + 1) * n,
          sizeof(double) * n * (size_t)(k - 1 - x_ind));
  memmove(m->x_ind_to_g_ind + x_ind, m->x_ind_to_g_ind + x_ind + 1,
          sizeof(uint32_t) * (k - 1 - x_ind));
  memmove(m->xy + x_ind, m->xy + x_ind + 1, sizeof(double) * (k - 1 - x_ind));
  orc_remove_colrow(m->xx, ld, k, x_ind);
  memmove(m->x_types + x_ind, m->x_types + x_ind + 1, k - 1 - x_ind);
  double inv_val = m->dv[x_ind];
  memmove(m->dv + x_ind, m->dv + x_ind + 1, sizeof(double) * (k - 1 - x_ind));
  m->cols = k - 1;
  if (upd) {
    if (x_ind == m->cols) {
      m->log_det_invQ_plus_xx -= log(E(m->l, ld, x_ind, x_ind));
      m->syx_plus_vs2 += m->v[x_ind] * m->v[x_ind];
    } else {
      orc_chol_downdate(m->l, ld, k, x_ind);
      memcpy(m->v, m->xy, sizeof(double) * m->cols);
      orc_trsv_ut(m->l, ld, m->cols, m->v);
      m->log_det_invQ_plus_xx = 0;
      for (int i = 0; i < m->cols; ++i)
        m->log_det_invQ_plus_xx += log(E(m->l, ld, i, i));
      m->syx_plus_vs2 = m->pr->nus2_plus_yy - dot(m->v, m->v, m->cols);
      guard_syx(m);
    }
    m->log_det_invQ -= 0.5 * log(inv_val);
    m->loglik = orc__finite_ll(m->log_det_invQ - m->log_det_invQ_plus_xx
                               + m->pr->minus_n_plus_nu_2 * log(m->syx_plus_vs2));
  }
}

/* src/model.hpp:270-312 */
void orc_model_remove_term(orc_model* m, int model_ind)
{
  uint32_t ind = m->loci[model_ind];
  int type = orc__dm_type(m->dm, m->type_inds[model_ind]);
  unsigned xi1 = m->x_ind1[model_ind], xi2 = m->x_ind2[model_ind];
  m->model_inds[ind] = -1;
  m->Ns[type]--;
  int nl = m->nloci;
  memmove(m->loci + model_ind, m->loci + model_ind + 1, sizeof(uint32_t) * (nl - 1 - model_ind));
  memmove(m->type_inds + model_ind, m->type_inds + model_ind + 1, nl - 1 - model_ind);
  memmove(m->x_ind1 + model_ind, m->x_ind1 + model_ind + 1, sizeof(uint32_t) * (nl - 1 - model_ind));
  memmove(m->x_ind2 + model_ind, m->x_ind2 + model_ind + 1, sizeof(uint32_t) * (nl - 1 - model_ind));
  m->nloci = nl - 1;
  int xmove = (type == ORC_AH) ? 2 : 1;
  for (int i = model_ind; i < m->nloci; ++i) {
    m->x_ind1[i] -= xmove;
    m->x_ind2[i] -= xmove;
    m->model_inds[m->loci[i]]--;
  }
  if (type == ORC_AH) {
    if (xi1 > xi2) { update_on_remove(m, xi1, 1); update_on_remove(m, xi2, 1); }
    else { update_on_remove(m, xi2, 1); update_on_remove(m, xi1, 1); }
  } else {
    update_on_remove(m, xi1, 1);
  }
  m->mu_beta_computed = 0;
}

/* src/model.hpp:314-324 */
static void compute_mu_beta(orc_model* m)
{
  if (!m->mu_beta_computed) {
    m->mu_len = m->cols;
    memcpy(m->mu_beta, m->v, sizeof(double) * m->cols);
    orc_trsv_un(m->l, m->kmax, m->cols, m->mu_beta);
    m->mu_beta_computed = 1;
  }
}

/* src/model.hpp:326-342 with src/rand.hpp:61-78,144-168 */
void orc__model_sample_beta_sigma2(orc_model* m, orc_rng* rng)
{
  double nu = m->pr->n_plus_nu;
  double s2 = m->syx_plus_vs2 / m->pr->n_plus_nu;
  double val = -1.0;
  while (val <= 0 || !isfinite(val))
    val = nu * s2 / (2.0 * orc_rng_gamma(rng, 0.5 * nu));
  m->sigma2 = val;
  compute_mu_beta(m);
  m->beta_len = m->mu_len;
  orc_rng_mvnormal_invchol(rng, m->mu_beta, m->l, m->kmax, m->beta_len, sqrt(m->sigma2), m->beta);
}

/* src/rand.hpp:144-168 (Rand::rand_mvnormal_invchol): x = scale * U^-1 z + mu
 * with z standard normal, U upper triangular (column-major, ld), i.e. x ~
 * N(mu, scale^2 (U'U)^-1) */
void orc_rng_mvnormal_invchol(orc_rng* rng, const double* mu, const double* U, int ld, int d,
                              double scale, double* x)
{
  for (int i = 0; i < d; ++i) x[i] = orc_rng_normal(rng);
  orc_trsv_un(U, ld, d, x);
  for (int i = 0; i < d; ++i) x[i] = scale * x[i] + mu[i];
}

static double vvar(const double* x, size_t n)
{
  double sq = 0, s = 0;
  for (size_t i = 0; i < n; ++i) { sq += x[i] * x[i]; s += x[i]; }
  return (sq - s * s / n) / (n - 1);
}

/* src/model.hpp:345-392 */
void orc_model_compute_pve(orc_model* m, double pves[3])
{
  size_t n = m->n; int me = m->m_e;
  int ce = (me > 1);
  if (ce) {
    gemv_n(m->x, n, n, me, m->beta, m->y_hat_tmp_e);
    pves[2] = vvar(m->y_hat_tmp_e, n);
  } else {
    pves[2] = 0.0;
    for (size_t i = 0; i < n; ++i) m->y_hat_tmp_e[i] = m->beta[0];
  }
  int cg = (m->nloci > 0);
  if (cg) {
    gemv_n(m->x + (size_t)me * n, n, n, m->cols - me, m->beta + me, m->y_hat);
    pves[1] = vvar(m->y_hat, n);
  } else {
    pves[1] = 0.0;
  }
  if (ce || cg) {
    for (size_t i = 0; i < n; ++i) m->y_hat[i] += m->y_hat_tmp_e[i];
    if (!ce) pves[0] = pves[1];
    else if (!cg) pves[0] = pves[2];
    else pves[0] = vvar(m->y_hat, n);
  } else {
    memcpy(m->y_hat, m->y_hat_tmp_e, sizeof(double) * n);
    pves[0] = pves[1] = pves[2] = 0; return;
  }
  double z = pves[0] + m->sigma2;
  pves[0] /= z; pves[1] /= z; pves[2] /= z;
}

double orc_model_log_likelihood(const orc_model* m) { return m->loglik; }
int orc_model_size(const orc_model* m) { return m->nloci; }
int orc_model_cols(const orc_model* m) { return m->cols; }
const double* orc_model_beta(const orc_model* m) { return m->beta; }
const double* orc_model_y_hat(const orc_model* m) { return m->y_hat; }
const int32_t* orc_model_model_inds(const orc_model* m) { return m->model_inds; }
void orc_model_set_d(orc_model* m, int x_ind, double val)
{ if (x_ind < m->kmax) m->dv[x_ind] = val; }

void orc_model_get(const orc_model* m, double* xx, double* l, double* xy,
                   double* v, double* d, double scal[6])
{
  int k = m->cols, ld = m->kmax;
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < k; ++r) {
      if (xx) xx[c * k + r] = r <= c ? E(m->xx, ld, r, c) : 0.0;
      if (l) l[c * k + r] = r <= c ? E(m->l, ld, r, c) : 0.0;
    }
  if (xy) memcpy(xy, m->xy, sizeof(double) * k);
  if (v) memcpy(v, m->v, sizeof(double) * k);
  if (d) memcpy(d, m->dv, sizeof(double) * k);
  if (scal) {
    scal[0] = m->loglik; scal[1] = m->syx_plus_vs2; scal[2] = m->log_det_invQ;
    scal[3] = m->log_det_invQ_plus_xx; scal[4] = m->sigma2; scal[5] = k;
  }
}

/* ------------------------------------------------------------- ExhModel */
/* src/model.hpp:585-844.  The exhaustive model reuses orc_model storage and
 * keeps its own bookkeeping in orc_exh. */
typedef struct {
  orc_model* m;
  int model_size, const_loci, v_size;
  double log_model_prior;
} orc_exh;

/* src/model.hpp:602-671 */
static double exh_update_to_model(orc_exh* x, orc_model* em, const orc_model* src,
                                  int const_loci)
{
  int k = src->cols, ld = src->kmax;
  x->m = em;
  for (int c = 0; c < k; ++c)
    for (int r = 0; r <= c; ++r) E(em->l, ld, r, c) = E(src->l, ld, r, c);
  memcpy(em->xy, src->xy, sizeof(double) * k);
  memcpy(em->v, src->v, sizeof(double) * k);
  memcpy(em->dv, src->dv, sizeof(double) * k);
  em->cols = k;
  em->nloci = src->nloci;
  memcpy(em->type_inds, src->type_inds, src->nloci);
  memcpy(em->x_types, src->x_types, k);
  memcpy(em->x_ind1, src->x_ind1, sizeof(uint32_t) * src->nloci);
  memcpy(em->x_ind2, src->x_ind2, sizeof(uint32_t) * src->nloci);
  const orc_prior* pr = src->pr;
  x->const_loci = const_loci;
  x->model_size = const_loci;
  if (x->model_size > 0) {
    if (orc__dm_type(src->dm, em->type_inds[x->model_size - 1]) == ORC_AH)
      x->v_size = em->x_ind2[x->model_size - 1] + 1;
    else
      x->v_size = em->x_ind1[x->model_size - 1] + 1;
  } else {
    x->v_size = src->m_e;
  }
  em->syx_plus_vs2 = pr->nus2_plus_yy - dot(em->v, em->v, x->v_size);
  em->log_det_invQ_plus_xx = 0;
  for (int i = 0; i < x->v_size; ++i) em->log_det_invQ_plus_xx += log(E(em->l, ld, i, i));
  double ls = 0.0;
  for (int i = 1; i < src->m_e; ++i) ls += log(pr->inv_tau2_e[i]);
  em->log_det_invQ = ls;
  for (int i = src->m_e; i < x->v_size; ++i) em->log_det_invQ += log(em->dv[i]);
  em->log_det_invQ *= 0.5;
  em->loglik = em->log_det_invQ - em->log_det_invQ_plus_xx
               + pr->minus_n_plus_nu_2 * log(em->syx_plus_vs2);
  x->log_model_prior = 0.0;
  for (int i = 0; i < 5; ++i) em->Ns[i] = 0;
  for (int i = 0; i < x->model_size; ++i)
    em->Ns[orc__dm_type(src->dm, em->type_inds[i])]++;
  return em->loglik + x->log_model_prior;
}

/* src/model.hpp:673-701 */
static double exh_update_on_add(orc_exh* x)
{
  orc_model* m = x->m; int ld = m->kmax;
  int type = orc__dm_type(m->dm, m->type_inds[x->model_size]);
  m->syx_plus_vs2 -= m->v[x->v_size] * m->v[x->v_size];
  m->log_det_invQ += 0.5 * log(m->dv[x->v_size]);
  m->log_det_invQ_plus_xx += log(E(m->l, ld, x->v_size, x->v_size));
  ++x->v_size;
  if (type == ORC_AH) {
    m->syx_plus_vs2 -= m->v[x->v_size] * m->v[x->v_size];
    m->log_det_invQ += 0.5 * log(m->dv[x->v_size]);
    m->log_det_invQ_plus_xx += log(E(m->l, ld, x->v_size, x->v_size));
    ++x->v_size;
  }
  m->loglik = m->log_det_invQ - m->log_det_invQ_plus_xx
              + m->pr->minus_n_plus_nu_2 * log(m->syx_plus_vs2);
  x->log_model_prior += orc__prior_add(m->pr, m->Ns, x->model_size, type);
  m->Ns[type]++;
  ++x->model_size;
  return m->loglik + x->log_model_prior;
}

#define SWAPD(a, b) do { double _t = (a); (a) = (b); (b) = _t; } while (0)
#define SWAPU(a, b) do { unsigned _t = (a); (a) = (b); (b) = _t; } while (0)

/* src/model.hpp:703-832 */
static double exh_update_on_moveleft(orc_exh* x)
{
  orc_model* m = x->m; int ld = m->kmax;
  const orc_prior* pr = m->pr;
  --x->model_size;
  int ik = x->model_size, ir = x->model_size - 1;
  int tk = orc__dm_type(m->dm, m->type_inds[ik]);
  int tr = orc__dm_type(m->dm, m->type_inds[ir]);
  double* v = m->v; double* xy = m->xy; double* dv = m->dv;
  uint32_t* x1 = m->x_ind1; uint32_t* x2 = m->x_ind2;

  --x->v_size;
  m->syx_plus_vs2 += v[x->v_size] * v[x->v_size];
  m->log_det_invQ_plus_xx -= log(E(m->l, ld, x->v_size, x->v_size));
  if (tk == ORC_AH) {
    --x->v_size;
    m->syx_plus_vs2 += v[x->v_size] * v[x->v_size];
    m->log_det_invQ_plus_xx -= log(E(m->l, ld, x->v_size, x->v_size));
  }
  --x->v_size;
  m->syx_plus_vs2 += v[x->v_size] * v[x->v_size];
  m->log_det_invQ -= 0.5 * log(dv[x->v_size]);
  m->log_det_invQ_plus_xx -= log(E(m->l, ld, x->v_size, x->v_size));
  if (tr == ORC_AH) {
    --x->v_size;
    m->syx_plus_vs2 += v[x->v_size] * v[x->v_size];
    m->log_det_invQ -= 0.5 * log(dv[x->v_size]);
    m->log_det_invQ_plus_xx -= log(E(m->l, ld, x->v_size, x->v_size));
  }
  { unsigned char t = m->type_inds[ik]; m->type_inds[ik] = m->type_inds[ir]; m->type_inds[ir] = t; }
  int k = m->cols;
  if (tr == ORC_AH) {
    if (tk == ORC_AH) {
      int si = x2[ir];
      orc_chol_swapadj(m->l, ld, k, si, v); ++si;
      orc_chol_swapadj(m->l, ld, k, si, v); si -= 2;
      orc_chol_swapadj(m->l, ld, k, si, v); ++si;
      orc_chol_swapadj(m->l, ld, k, si, v);
      SWAPD(xy[x1[ik]], xy[x2[ir]]); SWAPD(xy[x1[ik]], xy[x2[ik]]);
      SWAPD(xy[x1[ir]], xy[x2[ir]]); SWAPD(xy[x1[ik]], xy[x2[ir]]);
      SWAPD(dv[x1[ik]], dv[x2[ir]]); SWAPD(dv[x1[ik]], dv[x2[ik]]);
      SWAPD(dv[x1[ir]], dv[x2[ir]]); SWAPD(dv[x1[ik]], dv[x2[ir]]);
    } else {
      int si = x2[ir];
      orc_chol_swapadj(m->l, ld, k, si, v); --si;
      orc_chol_swapadj(m->l, ld, k, si, v);
      SWAPD(xy[x2[ir]], xy[x1[ik]]); SWAPD(xy[x1[ir]], xy[x2[ir]]);
      SWAPD(dv[x2[ir]], dv[x1[ik]]); SWAPD(dv[x1[ir]], dv[x2[ir]]);
      { unsigned char t = m->x_types[x2[ir]]; m->x_types[x2[ir]] = m->x_types[x1[ik]]; m->x_types[x1[ik]] = t; }
      { unsigned char t = m->x_types[x1[ir]]; m->x_types[x1[ir]] = m->x_types[x2[ir]]; m->x_types[x2[ir]] = t; }
      x2[ik] = x1[ik];
      --x1[ik];
    }
  } else {
    if (tk == ORC_AH) {
      int si = x1[ir];
      orc_chol_swapadj(m->l, ld, k, si, v); ++si;
      orc_chol_swapadj(m->l, ld, k, si, v);
      SWAPD(xy[x1[ir]], xy[x1[ik]]); SWAPD(xy[x1[ik]], xy[x2[ik]]);
      SWAPD(dv[x1[ir]], dv[x1[ik]]); SWAPD(dv[x1[ik]], dv[x2[ik]]);
      { unsigned char t = m->x_types[x1[ir]]; m->x_types[x1[ir]] = m->x_types[x1[ik]]; m->x_types[x1[ik]] = t; }
      { unsigned char t = m->x_types[x1[ik]]; m->x_types[x1[ik]] = m->x_types[x2[ik]]; m->x_types[x2[ik]] = t; }
      x2[ir] = x1[ik];
      x1[ik] = x2[ik];
    } else {
      int si = x1[ir];
      orc_chol_swapadj(m->l, ld, k, si, v);
      SWAPD(xy[x1[ik]], xy[x1[ir]]);
      SWAPD(dv[x1[ik]], dv[x1[ir]]);
      { unsigned char t = m->x_types[x1[ik]]; m->x_types[x1[ik]] = m->x_types[x1[ir]]; m->x_types[x1[ir]] = t; }
    }
  }
  m->syx_plus_vs2 -= v[x->v_size] * v[x->v_size];
  m->log_det_invQ_plus_xx += log(E(m->l, ld, x->v_size, x->v_size));
  ++x->v_size;
  if (tk == ORC_AH) {
    m->syx_plus_vs2 -= v[x->v_size] * v[x->v_size];
    m->log_det_invQ_plus_xx += log(E(m->l, ld, x->v_size, x->v_size));
    ++x->v_size;
  }
  m->loglik = m->log_det_invQ - m->log_det_invQ_plus_xx
              + pr->minus_n_plus_nu_2 * log(m->syx_plus_vs2);
  x->log_model_prior += orc__prior_rem(pr, m->Ns, x->model_size + 1, tr);
  m->Ns[tr]--;
  return m->loglik + x->log_model_prior;
}

static double exh_update_on_twonewswap(orc_exh* x)
{
  exh_update_on_add(x);
  exh_update_on_add(x);
  return exh_update_on_moveleft(x);
}

/* src/sampler.cpp:882-980; `em` is scratch storage for the ExhModel */
void orc__exhaustive_modelset(orc_model* em, const orc_model* full,
                              int const_loci, int n_inds, double* lp,
                              double* max_log)
{
  orc_exh ex;
  exh_update_to_model(&ex, em, full, const_loci);
  size_t cmodel_binary = 0, tmp;
  int model_size = 0;
  size_t* ci = (size_t*)malloc(sizeof(size_t) * n_inds);
  for (int i = 0; i < n_inds; ++i) ci[i] = i;
  lp[0] = em->loglik + ex.log_model_prior;
  *max_log = lp[0];
  for (int i = 0; i < n_inds; ++i) {
    ++model_size;
    if (i > 1) { ++model_size; exh_update_on_add(&ex); }
    cmodel_binary = ((size_t)1 << model_size) - 1;
    lp[cmodel_binary] = exh_update_on_add(&ex);
    if (lp[cmodel_binary] > *max_log) *max_log = lp[cmodel_binary];
    for (int j = 0; j < i; ++j) {
      --model_size;
      tmp = ci[model_size - 1]; ci[model_size - 1] = ci[model_size]; ci[model_size] = tmp;
      cmodel_binary &= ~((size_t)1 << tmp);
      lp[cmodel_binary] = exh_update_on_moveleft(&ex);
      if (lp[cmodel_binary] > *max_log) *max_log = lp[cmodel_binary];
    }
    size_t nmodels = ((size_t)1 << i) - i - 1;
    size_t j = 0, K, nK = 0;
    int do_lefts = 0;
    while (j < nmodels) {
      if (do_lefts < 2) {
        ++model_size;
        tmp = ci[model_size]; ci[model_size] = ci[model_size - 1]; ci[model_size - 1] = tmp;
        cmodel_binary |= ((size_t)1 << tmp);
        lp[cmodel_binary] = exh_update_on_twonewswap(&ex);
        if (lp[cmodel_binary] > *max_log) *max_log = lp[cmodel_binary];
        ++j; ++do_lefts;
      } else {
        ++nK; K = 0;
        while (((nK >> K) & 1) == 0) K++;
        for (size_t kk = 0; kk <= K; ++kk) {
          --model_size;
          tmp = ci[model_size - 1]; ci[model_size - 1] = ci[model_size]; ci[model_size] = tmp;
          cmodel_binary &= ~((size_t)1 << tmp);
          lp[cmodel_binary] = exh_update_on_moveleft(&ex);
          if (lp[cmodel_binary] > *max_log) *max_log = lp[cmodel_binary];
          ++j;
        }
        do_lefts = 0;
      }
    }
  }
  free(ci);
}

/* src/sampler.cpp:982-1049 (log_prop_probs is accumulated into, as in the
 * reference, so the caller passes the log model probabilities) */
void orc_proposal_probs_exh(int use_types, int n_inds,
                            const unsigned char* b2n, const double* q_add,
                            const double* q_rem, double z_add, double z_rem,
                            size_t const_loci, size_t m_g,
                            const double* log_q_add_types,
                            double* log_prop_probs)
{
  int isadd[64];
  unsigned long nmodels = 1ul << n_inds;
  for (unsigned long i = 0; i < nmodels; ++i) {
    size_t max_adds = m_g - const_loci, max_rems = const_loci;
    double za = z_add, zr = z_rem;
    for (int j = 0; j < n_inds; ++j) {
      int bit = (i >> j) & 1;
      int nind = b2n[j];
      if (bit) {
        isadd[nind] = 0;
        za -= q_add[nind]; zr += q_rem[nind];
        ++max_rems; --max_adds;
      } else {
        isadd[nind] = 1;
      }
    }
    int last_rem_pos = n_inds - 1;
    for (int j = 0; j < n_inds; ++j) {
      if (max_adds > 0 && max_rems > 0) log_prop_probs[i] += ORC_LOG_HALF;
      if (isadd[j]) {
        --max_adds;
        log_prop_probs[i] += log(q_add[j]) - log(za);
        za -= q_add[j];
        if (use_types) log_prop_probs[i] += log_q_add_types[j];
      } else {
        --max_rems;
        while (isadd[last_rem_pos]) --last_rem_pos;
        log_prop_probs[i] += log(q_rem[last_rem_pos]) - log(zr);
        zr -= q_rem[last_rem_pos];
        --last_rem_pos;
      }
    }
  }
}
