/*
 * orc_internal.h -- shared structs of the CPU oracle (test infrastructure).
 */
#ifndef BMAGWA_ORC_INTERNAL_H_
#define BMAGWA_ORC_INTERNAL_H_

#include "oracle.h"

/* A log-likelihood that is not finite (+inf when the residual term reaches
 * zero, NaN) can only come from rounding in a near-singular design; it is
 * taken as -inf, so such a proposal is rejected.  The reference keeps the
 * raw value (+inf would be accepted). */
#include <math.h>
static inline double orc__finite_ll(double v) { return isfinite(v) ? v : -INFINITY; }

#define ORC_LOG_HALF (-0.69314718055994528622676398299518041312694549560546875)

/* orc_data.c */
uint32_t orc__data_n(const orc_data*);
uint64_t orc__data_m_g(const orc_data*);
uint32_t orc__data_m_e(const orc_data*);
size_t   orc__data_stride(const orc_data*);
const double* orc__data_y(const orc_data*);
const double* orc__data_e(const orc_data*);
const size_t* orc__data_miss_loc(const orc_data*, size_t);
const double* orc__data_miss_prior(const orc_data*, size_t);
int   orc__dm_allow_types(const orc_datamodel*, int);
int   orc__dm_allow_terms(const orc_datamodel*, int);
int   orc__dm_n_types(const orc_datamodel*);
int   orc__dm_type(const orc_datamodel*, int);
char* orc__dm_miss_val(const orc_datamodel*, size_t);
const orc_data* orc__dm_data(const orc_datamodel*);

/* orc_linalg.c */
void orc_trsv_ut(const double* u, int ld, int m, double* x);
void orc_trsv_un(const double* u, int ld, int m, double* x);
void orc_remove_colrow(double* a, int ld, int m, int col_rem);
void orc_swap_colrow(double* a, int ld, int m, int colrow);

/* src/prior.hpp:38-291 */
typedef struct {
  double n_plus_nu, minus_n_plus_nu_2, nus2_plus_yy, mu_alpha;
  double nus2;   /* nu_sigma2 * s2_sigma2 alone */
  int use_individual_tau2;
  size_t n, m_g, m_e;
  int n_types;
  double g_a, g_b;
  double inv_tau2_e[64];
  double nu_sigma2, s2_sigma2;
  double nu_tau2[4], s2_tau2[4], nus2_tau2[4];
  double alpha, alpha2;
  double inv_tau2_alpha2[4];
  double types_prior[5], types_prior_sum;
  int allow_types[5], allow_terms[5];
} orc_prior;

/* src/model.hpp:46-582 */
struct orc_model {
  orc_sampler* s;
  const orc_data* d; const orc_datamodel* dm; const orc_prior* pr;
  int kmax; uint32_t n; size_t m_g; int m_e;
  int cols;
  double *x, *xx, *l, *xy, *v, *dv, *beta, *mu_beta, *y_hat, *y_hat_tmp_e;
  int beta_len, mu_len;
  double sigma2, syx_plus_vs2, log_det_invQ, log_det_invQ_plus_xx, loglik;
  int Ns[5];
  int mu_beta_computed;
  int nloci;
  uint32_t *loci; unsigned char* type_inds; uint32_t *x_ind1, *x_ind2;
  uint32_t *x_ind_to_g_ind; unsigned char* x_types;
  int32_t* model_inds;
};

int    orc__prior_init(orc_prior* p, const orc_data* d, const orc_datamodel* dm,
                       const double* types_prior, double e_qg, double var_qg,
                       const double* inv_tau2_e, double nu_sigma2,
                       double s2_sigma2, const double* nu_tau2,
                       const double* s2_tau2, double mu_alpha, int use_ind);
double orc__prior_add(const orc_prior* p, const int* Ns, int n_old, int type);
double orc__prior_rem(const orc_prior* p, const int* Ns, int n_old, int type);
double orc__prior_swi(const orc_prior* p, const int* Ns, int ta, int tr);
double orc__prior_log_model(const orc_prior* p, const int* Ns);
double orc__prior_draw_inv_tau2(const orc_prior* p, int type, orc_rng* rng);

orc_model* orc__model_create(orc_sampler* s, const orc_data* d,
                             const orc_datamodel* dm, const orc_prior* pr,
                             int kmax);
void   orc__model_destroy(orc_model* m);
void   orc__model_sample_beta_sigma2(orc_model* m, orc_rng* rng);
void   orc__exhaustive_modelset(orc_model* em, const orc_model* full,
                                int const_loci, int n_inds, double* lp,
                                double* max_log);

#endif
