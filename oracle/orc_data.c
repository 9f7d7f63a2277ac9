/*
 * orc_data.c -- genotype store, decoding, missing-value handling, per-SNP
 * moments and band X^T X, restated for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* src/data.cpp:36 (single) and the four 256x4 tables of
 * src/genotype_tables.hpp, which are the per-field application of these */
static const double CODE_VALUE[4][4] = {
  /* A */ {0.0, -1.0, 1.0, 2.0},
  /* H */ {0.0, -1.0, 1.0, 0.0},
  /* D */ {0.0, -1.0, 1.0, 1.0},
  /* R */ {0.0, -1.0, 0.0, 1.0},
};

/* per byte: number of fields with code 10 (value 1), 11 (value 2) and 01
 * (missing); used to take exact integer sums a byte at a time */
static unsigned char BYTE_C1[256], BYTE_C2[256], BYTE_CM[256];
static int byte_tables_ready = 0;
static void byte_tables(void)
{
  if (byte_tables_ready) return;
  for (int b = 0; b < 256; ++b) {
    int c1 = 0, c2 = 0, cm = 0;
    for (int f = 0; f < 4; ++f) {
      int c = (b >> (2 * f)) & 3;
      c1 += c == 2; c2 += c == 3; cm += c == 1;
    }
    BYTE_C1[b] = (unsigned char)c1; BYTE_C2[b] = (unsigned char)c2; BYTE_CM[b] = (unsigned char)cm;
  }
  byte_tables_ready = 1;
}

/* counts of codes 10, 11, 01 of one SNP (padding fields are 00) */
static void snp_counts(const uint8_t* g, size_t stride, long long* c1, long long* c2,
                       long long* cm)
{
  long long a = 0, b = 0, m = 0;
  for (size_t k = 0; k < stride; ++k) {
    a += BYTE_C1[g[k]]; b += BYTE_C2[g[k]]; m += BYTE_CM[g[k]];
  }
  *c1 = a; *c2 = b; *cm = m;
}

/* src/data.cpp:40-54 */
double orc_get_genotype(const uint8_t* body, size_t stride, size_t ind,
                        size_t snp)
{
  unsigned char c = (body[stride * snp + ind / 4] >> (2 * (ind % 4))) & 0x03;
  return CODE_VALUE[0][c];
}

/* src/data.cpp:56-138: four fields per byte, LSB first, tail byte partial */
void orc_get_genotypes(const uint8_t* body, size_t stride, uint32_t n,
                       int coding, size_t snp, double* out)
{
  const uint8_t* g = body + stride * snp;
  const double* tbl = CODE_VALUE[coding];
  for (uint32_t i = 0; i < n; ++i)
    out[i] = tbl[(g[i >> 2] >> (2 * (i & 3))) & 3];
}

/* ------------------------------------------------------------------ Data */
struct orc_data {
  uint32_t n; uint64_t m_g; uint32_t m_e;  /* m_e includes the constant */
  size_t stride;
  uint8_t* g;
  double* y;
  double* e;            /* n x m_e column-major, column 0 = 1 */
  size_t** miss_loc;    /* miss_loc[snp][0] = count, then individuals */
  double** miss_prior;  /* cumulative counts of 0,1,2 */
  double yy, var_y, var_x, mean_x;
};

int orc_read_bed(const char* path, uint32_t n, uint64_t m_g,
                 uint8_t* body_out)
{
  /* src/data.cpp:245-273 */
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  unsigned char hdr[3];
  if (fread(hdr, 1, 3, f) != 3) { fclose(f); return -2; }
  if (hdr[0] != 0x6C || hdr[1] != 0x1B) { fclose(f); return -3; }
  if (hdr[2] != 0x01) { fclose(f); return -4; }
  size_t len = (size_t)((n + 3) / 4) * m_g;
  size_t got = fread(body_out, 1, len, f);
  fclose(f);
  return got == len ? 0 : -5;
}

static double vec_var(const double* x, size_t n)
{
  /* src/vector.cpp:113-124 */
  double sq = 0, s = 0;
  for (size_t i = 0; i < n; ++i) { sq += x[i] * x[i]; s += x[i]; }
  return (sq - s * s / n) / (n - 1);
}

static void set_genotype(orc_data* d, size_t ind, size_t snp, int geno)
{
  /* src/data.cpp:275-322 */
  static const unsigned char CODE_OF[4] = {0x01, 0x00, 0x02, 0x03};
  unsigned char* byte = d->g + d->stride * snp + ind / 4;
  unsigned char null_mask = (unsigned char)~(0x03 << (2 * (ind % 4)));
  unsigned char mask = (unsigned char)(CODE_OF[geno + 1] << (2 * (ind % 4)));
  *byte = (*byte & null_mask) | mask;
}

orc_data* orc_data_create(uint32_t n, uint64_t m_g, uint32_t m_e,
                          const uint8_t* body, const double* y,
                          const double* e_cov, int recode)
{
  orc_data* d = (orc_data*)calloc(1, sizeof(orc_data));
  d->n = n; d->m_g = m_g; d->m_e = m_e + 1;
  d->stride = (n + 3) / 4;
  d->g = (uint8_t*)malloc(d->stride * m_g);
  memcpy(d->g, body, d->stride * m_g);
  d->y = (double*)malloc(sizeof(double) * n);
  memcpy(d->y, y, sizeof(double) * n);
  d->e = (double*)malloc(sizeof(double) * n * d->m_e);
  for (uint32_t i = 0; i < n; ++i) d->e[i] = 1.0;
  if (m_e > 0) memcpy(d->e + n, e_cov, sizeof(double) * n * m_e);

  if (recode) {
    /* src/data.cpp:324-339 */
    byte_tables();
    for (uint64_t snp = 0; snp < m_g; ++snp) {
      /* allele count and non-missing count are exact integers (the
       * reference sums them as doubles, src/utils.cpp:33-44) */
      long long c1, c2, cm;
      snp_counts(d->g + d->stride * snp, d->stride, &c1, &c2, &cm);
      double mac = (double)(c1 + 2 * c2); size_t ng = n - (size_t)cm;
      if (mac / (2.0 * ng) > 0.5) {
        for (uint32_t i = 0; i < n; ++i) {
          double gv = orc_get_genotype(d->g, d->stride, i, snp);
          if (gv == 0.0) set_genotype(d, i, snp, 2);
          else if (gv == 2.0) set_genotype(d, i, snp, 0);
        }
      }
    }
  }

  /* src/data.cpp:341-376 */
  byte_tables();
  d->miss_loc = (size_t**)calloc(m_g, sizeof(size_t*));
  d->miss_prior = (double**)calloc(m_g, sizeof(double*));
  for (uint64_t snp = 0; snp < m_g; ++snp) {
    long long c1, c2, cm;
    snp_counts(d->g + d->stride * snp, d->stride, &c1, &c2, &cm);
    int n_miss = (int)cm;
    if (n_miss > 0) {
      d->miss_loc[snp] = (size_t*)malloc(sizeof(size_t) * (n_miss + 1));
      d->miss_prior[snp] = (double*)calloc(3, sizeof(double));
      d->miss_loc[snp][0] = n_miss;
      n_miss = 0;
      for (uint32_t i = 0; i < n; ++i) {
        double gv = orc_get_genotype(d->g, d->stride, i, snp);
        if (gv < 0) d->miss_loc[snp][++n_miss] = i;
        else ++d->miss_prior[snp][(int)gv];
      }
      d->miss_prior[snp][1] += d->miss_prior[snp][0];
      d->miss_prior[snp][2] += d->miss_prior[snp][1];
    }
  }

  /* src/data.hpp:67-71, src/data.cpp:403-434 */
  d->yy = 0;
  for (uint32_t i = 0; i < n; ++i) d->yy += d->y[i] * d->y[i];
  d->var_y = vec_var(d->y, n);
  double tv = 0, tm = 0; size_t nsm = 0, nsv = 0;
  for (uint64_t snp = 0; snp < m_g; ++snp) {
    /* the per-SNP sums are exact integers */
    long long c1, c2, cm;
    snp_counts(d->g + d->stride * snp, d->stride, &c1, &c2, &cm);
    double sq = (double)(c1 + 4 * c2), s = (double)(c1 + 2 * c2);
    size_t ng = n - (size_t)cm;
    if (ng > 1) {
      tm += s / ng; tv += (sq - s * s / ng) / (ng - 1); ++nsm; ++nsv;
    } else if (ng == 1) {
      tm += s / ng; ++nsm;
    }
  }
  d->var_x = tv / nsv;
  d->mean_x = tm / nsm;
  return d;
}

void orc_data_destroy(orc_data* d)
{
  if (!d) return;
  for (uint64_t s = 0; s < d->m_g; ++s) {
    free(d->miss_loc[s]); free(d->miss_prior[s]);
  }
  free(d->miss_loc); free(d->miss_prior);
  free(d->g); free(d->y); free(d->e); free(d);
}

const uint8_t* orc_data_body(const orc_data* d) { return d->g; }

void orc_data_stats(const orc_data* d, double out[4])
{
  out[0] = d->yy; out[1] = d->var_y; out[2] = d->var_x; out[3] = d->mean_x;
}

int orc_data_missing(const orc_data* d, size_t snp, size_t* locs, int cap,
                     double prior_cum[3])
{
  if (!d->miss_loc[snp]) return 0;
  int cnt = (int)d->miss_loc[snp][0];
  for (int j = 0; j < cnt && j < cap; ++j) locs[j] = d->miss_loc[snp][j + 1];
  for (int k = 0; k < 3; ++k) prior_cum[k] = d->miss_prior[snp][k];
  return cnt;
}

/* accessors for the sampler (internal linkage across oracle files) */
uint32_t orc__data_n(const orc_data* d) { return d->n; }
uint64_t orc__data_m_g(const orc_data* d) { return d->m_g; }
uint32_t orc__data_m_e(const orc_data* d) { return d->m_e; }
size_t orc__data_stride(const orc_data* d) { return d->stride; }
const double* orc__data_y(const orc_data* d) { return d->y; }
const double* orc__data_e(const orc_data* d) { return d->e; }
const size_t* orc__data_miss_loc(const orc_data* d, size_t s)
{ return d->miss_loc[s]; }
const double* orc__data_miss_prior(const orc_data* d, size_t s)
{ return d->miss_prior[s]; }

/* ------------------------------------------------------------ DataModel */
struct orc_datamodel {
  const orc_data* d;
  uint64_t m_g;         /* the data's SNP count, so destroy does not read d */
  int n_types; int types[5];
  int allow_types[5], allow_terms[5];
  char** miss_val;      /* same indexing as miss_loc (slot 0 unused) */
  int offset; int offset_type[4];
};

orc_datamodel* orc_datamodel_create(const orc_data* d, const int32_t* types,
                                    int n_types)
{
  /* src/data_model.hpp:52-89 */
  orc_datamodel* dm = (orc_datamodel*)calloc(1, sizeof(orc_datamodel));
  dm->d = d; dm->n_types = n_types;
  for (int i = 0; i < n_types; ++i) dm->types[i] = types[i];
  for (int i = 0; i < n_types; ++i) {
    dm->allow_types[types[i]] = 1;
    if (types[i] == ORC_AH) { dm->allow_terms[ORC_A] = 1; dm->allow_terms[ORC_H] = 1; }
    else dm->allow_terms[types[i]] = 1;
  }
  dm->m_g = d->m_g;
  dm->miss_val = (char**)calloc(d->m_g, sizeof(char*));
  for (uint64_t s = 0; s < d->m_g; ++s)
    if (d->miss_loc[s])
      dm->miss_val[s] = (char*)calloc(d->miss_loc[s][0] + 1, 1);
  /* src/precomputed_snp_covariances.hpp:58-83 */
  int ah = dm->allow_types[ORC_AH];
  int both = dm->allow_types[ORC_A] && dm->allow_types[ORC_H];
  int either = dm->allow_types[ORC_A] || dm->allow_types[ORC_H];
  dm->offset = 2 * n_types + (ah ? 1 : 0) * (1 + (both ? -2 : 0) + (either ? 0 : 2));
  int k = 0;
  for (int j = 0; j < 4; ++j) {
    if (dm->allow_terms[j]) dm->offset_type[j] = k++;
    else dm->offset_type[j] = -1;
  }
  return dm;
}

void orc_datamodel_destroy(orc_datamodel* dm)
{
  if (!dm) return;
  for (uint64_t s = 0; s < dm->m_g; ++s) free(dm->miss_val[s]);
  free(dm->miss_val); free(dm);
}

int orc_datamodel_offset(const orc_datamodel* dm, int offset_type[4])
{
  for (int j = 0; j < 4; ++j) offset_type[j] = dm->offset_type[j];
  return dm->offset;
}

/* src/data_model.cpp:30-72 */
void orc_datamodel_get(const orc_datamodel* dm, int coding, size_t snp,
                       double* out)
{
  const orc_data* d = dm->d;
  orc_get_genotypes(d->g, d->stride, d->n, coding, snp, out);
  if (!d->miss_loc[snp]) return;
  for (size_t i = 1; i <= d->miss_loc[snp][0]; ++i) {
    int v = dm->miss_val[snp][i];
    double x;
    switch (coding) {
      case ORC_A: x = (double)v; break;
      case ORC_H: x = (double)(v == 1); break;
      case ORC_D: x = (double)(v > 0); break;
      default:    x = (double)(v == 2); break;
    }
    out[d->miss_loc[snp][i]] = x;
  }
}

void orc_datamodel_set_missing(orc_datamodel* dm, size_t snp, int j, int val)
{
  if (dm->miss_val[snp]) dm->miss_val[snp][j] = (char)val;
}

int orc__dm_allow_types(const orc_datamodel* dm, int t) { return dm->allow_types[t]; }
int orc__dm_allow_terms(const orc_datamodel* dm, int t) { return dm->allow_terms[t]; }
int orc__dm_n_types(const orc_datamodel* dm) { return dm->n_types; }
int orc__dm_type(const orc_datamodel* dm, int i) { return dm->types[i]; }
char* orc__dm_miss_val(const orc_datamodel* dm, size_t s) { return dm->miss_val[s]; }
const orc_data* orc__dm_data(const orc_datamodel* dm) { return dm->d; }

/* src/data_model.cpp:105-167 */
void orc_datamodel_update_prexx(const orc_datamodel* dm, size_t snp,
                                double* pre_xx)
{
  const orc_data* d = dm->d;
  if (!dm->miss_val[snp]) return;
  double* xxloc = pre_xx;
  const char* mv = dm->miss_val[snp];
  size_t cnt = d->miss_loc[snp][0];
  double sum_a_times_sum_h = 0.0;
  if (dm->allow_types[ORC_AH]) sum_a_times_sum_h = pre_xx[0] * pre_xx[2];
  for (int t = 0; t < 4; ++t) {
    if (!dm->allow_terms[t]) continue;
    double sg = 0.0, sg2 = 0.0;
    for (size_t i = 1; i <= cnt; ++i) {
      int v = mv[i];
      double vg;
      switch (t) {
        case ORC_A: vg = (double)v; break;
        case ORC_H: vg = (double)(v == 1); break;
        case ORC_D: vg = (double)(v > 0); break;
        default:    vg = (double)(v == 2); break;
      }
      sg += vg; sg2 += vg * vg;
    }
    double old_sum = *xxloc;
    *xxloc += sg; xxloc++;
    *xxloc += sg2 - sg * (old_sum * 2.0 + sg) / (double)d->n; xxloc++;
  }
  if (dm->allow_types[ORC_AH]) {
    *xxloc += (sum_a_times_sum_h - pre_xx[0] * pre_xx[2]) / (double)d->n;
    for (size_t i = 1; i <= cnt; ++i)
      if (mv[i] == 1) *xxloc += 1.0;
  }
}

/* src/precomputed_snp_covariances.hpp:95-131 */
void orc_self_moments(const orc_datamodel* dm, double* out)
{
  const orc_data* d = dm->d;
  uint32_t n = d->n;
  double* x = (double*)malloc(sizeof(double) * n);
  double* xah = (double*)malloc(sizeof(double) * n);
  double* xxloc = out;
  byte_tables();
  for (uint64_t i = 0; i < d->m_g; ++i) {
    if (!d->miss_loc[i]) {
      /* no missing value: the sums of the loop below are exact integer
       * functions of the code counts, taken a byte at a time */
      long long c1, c2, cm;
      snp_counts(d->g + d->stride * i, d->stride, &c1, &c2, &cm);
      const long long sum[4] = {c1 + 2 * c2, c1, c1 + c2, c2};
      const long long sq[4] = {c1 + 4 * c2, c1, c1 + c2, c2};
      for (int t = 0; t < 4; ++t) {
        if (!dm->allow_terms[t]) continue;
        double xsum = (double)sum[t];
        *xxloc++ = xsum;
        *xxloc++ = (double)sq[t] - xsum * xsum / (double)n;
      }
      if (dm->allow_types[ORC_AH]) {
        double as = (double)sum[0], hs = (double)sum[1];
        *xxloc++ = (double)c1 - as * hs / (double)n;
      }
      continue;
    }
    for (int t = 0; t < 4; ++t) {
      if (!dm->allow_terms[t]) continue;
      orc_datamodel_get(dm, t, i, x);
      double xsum = 0, xx = 0;
      for (uint32_t k = 0; k < n; ++k) xsum += x[k];
      for (uint32_t k = 0; k < n; ++k) xx += x[k] * x[k];
      *xxloc++ = xsum;
      *xxloc++ = xx - xsum * xsum / (double)n;
    }
    if (dm->allow_types[ORC_AH]) {
      orc_datamodel_get(dm, ORC_A, i, x);
      orc_datamodel_get(dm, ORC_H, i, xah);
      double as = 0, hs = 0, ah = 0;
      for (uint32_t k = 0; k < n; ++k) as += x[k];
      for (uint32_t k = 0; k < n; ++k) hs += xah[k];
      for (uint32_t k = 0; k < n; ++k) ah += x[k] * xah[k];
      *xxloc++ = ah - as * hs / (double)n;
    }
  }
  free(x); free(xah);
}

/* Band X^T X by brute force in int64.  Missing codes count as 0, which is
 * the value the reference's precompute sees for them (missing values are
 * zero until sampled, src/precomputed_snp_covariances.hpp:38-40). */
void orc_band_xtx(const uint8_t* body, size_t stride, uint32_t n,
                  uint64_t p, uint32_t W, int coding_a, int coding_b,
                  int64_t* out)
{
  static const int IVAL[4][4] = {
    {0, 0, 1, 2}, {0, 0, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 1}};
  int* xa = (int*)malloc(sizeof(int) * n);
  int* xb = (int*)malloc(sizeof(int) * n);
  for (uint64_t j = 0; j < p; ++j) {
    const uint8_t* gj = body + stride * j;
    for (uint32_t i = 0; i < n; ++i)
      xa[i] = IVAL[coding_a][(gj[i >> 2] >> (2 * (i & 3))) & 3];
    for (uint32_t dlt = 0; dlt < W; ++dlt) {
      if (j + dlt >= p) { out[j * W + dlt] = 0; continue; }
      const uint8_t* gk = body + stride * (j + dlt);
      for (uint32_t i = 0; i < n; ++i)
        xb[i] = IVAL[coding_b][(gk[i >> 2] >> (2 * (i & 3))) & 3];
      int64_t acc = 0;
      for (uint32_t i = 0; i < n; ++i) acc += (int64_t)xa[i] * xb[i];
      out[j * W + dlt] = acc;
    }
  }
  free(xa); free(xb);
}
