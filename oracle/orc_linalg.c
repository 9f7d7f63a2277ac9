/*
 * orc_linalg.c -- the dense FP64 kernels of the reference's Model, restated
 * as textbook loops for the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Storage follows the reference: column-major, element (r, c) at c*ld + r,
 * only the upper triangle (r <= c) is meaningful (src/symmmatrix.hpp:46-60).
 * The reference calls LAPACK dpotrf/BLAS dtrsv/drotg and LINPACK dchex
 * (src/symmmatrix.hpp:33-38, src/dchex.f).  Their summation order is
 * unspecified by the reference (ATLAS or OpenBLAS, README.markdown:39-40),
 * so the loops below use the unblocked textbook order of the netlib
 * reference implementations (dpotf2 'U', dtrsv, drotg, dchex).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define E(a, ld, r, c) ((a)[(size_t)(c) * (ld) + (r)])

/* dpotf2 'U' order; returns 0 when a leading minor is not positive */
int orc_chol(double* a, int ld, int m)
{
  for (int j = 0; j < m; ++j) {
    double ajj = E(a, ld, j, j);
    for (int k = 0; k < j; ++k) ajj -= E(a, ld, k, j) * E(a, ld, k, j);
    if (!(ajj > 0.0)) return 0;
    ajj = sqrt(ajj);
    E(a, ld, j, j) = ajj;
    for (int i = j + 1; i < m; ++i) {
      double t = E(a, ld, j, i);
      for (int k = 0; k < j; ++k) t -= E(a, ld, k, j) * E(a, ld, k, i);
      E(a, ld, j, i) = t / ajj;
    }
  }
  return 1;
}

/* dtrsv Upper/Trans/NonUnit: x := U^-T x (src/vector.cpp:165-178) */
void orc_trsv_ut(const double* u, int ld, int m, double* x)
{
  for (int j = 0; j < m; ++j) {
    double t = x[j];
    for (int i = 0; i < j; ++i) t -= E(u, ld, i, j) * x[i];
    x[j] = t / E(u, ld, j, j);
  }
}

/* dtrsv Upper/NoTrans/NonUnit: x := U^-1 x */
void orc_trsv_un(const double* u, int ld, int m, double* x)
{
  for (int j = m - 1; j >= 0; --j) {
    if (x[j] != 0.0) {
      x[j] /= E(u, ld, j, j);
      double t = x[j];
      for (int i = j - 1; i >= 0; --i) x[i] -= t * E(u, ld, i, j);
    }
  }
}

/* src/symmmatrix.cpp:144-164: append column `m` (newcol has m+1 entries) */
int orc_chol_update(double* l, int ld, int m, const double* newcol,
                    double inv_tau2)
{
  double* lcol = &E(l, ld, 0, m);
  for (int i = 0; i < m; ++i) lcol[i] = newcol[i];
  orc_trsv_ut(l, ld, m, lcol);
  double dot = 0.0;
  for (int i = 0; i < m; ++i) dot += lcol[i] * lcol[i];
  double lmm2 = newcol[m] + inv_tau2 - dot;
  if (lmm2 <= 0) return 0;
  E(l, ld, m, m) = sqrt(lmm2);
  return 1;
}

/* netlib drotg (BLAS level 1, Lawson, Hanson, Kincaid & Krogh 1979) */
void orc_drotg(double* a, double* b, double* c, double* s)
{
  double sa = *a, sb = *b;
  double roe = (fabs(sa) > fabs(sb)) ? sa : sb;
  double scale = fabs(sa) + fabs(sb);
  double r, z;
  if (scale == 0.0) {
    *c = 1.0; *s = 0.0; r = 0.0; z = 0.0;
  } else {
    double as = sa / scale, bs = sb / scale;
    r = scale * sqrt(as * as + bs * bs);
    r = (roe >= 0.0 ? 1.0 : -1.0) * r;
    *c = sa / r; *s = sb / r;
    z = 1.0;
    if (fabs(sa) > fabs(sb)) z = *s;
    if (fabs(sb) >= fabs(sa) && *c != 0.0) z = 1.0 / *c;
  }
  *a = r; *b = z;
}

/*
 * src/symmmatrix.cpp:166-195 -> LINPACK dchex job = 2 (src/dchex.f:186-231)
 * with k = col_rem + 1, l = p = m (1-based): left circular shift of column
 * col_rem to the end, re-triangularisation by l - k Givens rotations, then
 * drop the last row/column and flip rows whose diagonal became negative.
 */
void orc_chol_downdate(double* l, int ld, int m, int col_rem)
{
  int k = col_rem;            /* 0-based first permuted column */
  int last = m - 1;           /* 0-based last column */
  int lmk = last - k;
  double* s = (double*)malloc(sizeof(double) * (m + 1));
  double* c = (double*)malloc(sizeof(double) * (m + 1));
  /* save column k (rows 0..k) into s[lmk + i] (dchex.f:192-195) */
  for (int i = 0; i <= k; ++i) s[lmk + i] = E(l, ld, i, k);
  /* shift columns left; remember the displaced diagonals (dchex.f:196-202) */
  for (int j = k; j < last; ++j) {
    for (int i = 0; i <= j; ++i) E(l, ld, i, j) = E(l, ld, i, j + 1);
    s[j - k] = E(l, ld, j + 1, j + 1);
  }
  /* moved column into the last position (dchex.f:203-209) */
  for (int i = 0; i <= k; ++i) E(l, ld, i, last) = s[lmk + i];
  for (int i = k + 1; i <= last; ++i) E(l, ld, i, last) = 0.0;
  /* reduction loop (dchex.f:213-231) */
  for (int j = k; j <= last; ++j) {
    if (j != k) {
      int iu = (j - 1 < last - 1) ? j - 1 : last - 1;
      for (int i = k; i <= iu; ++i) {
        int ii = i - k;
        double t = c[ii] * E(l, ld, i, j) + s[ii] * E(l, ld, i + 1, j);
        E(l, ld, i + 1, j) = c[ii] * E(l, ld, i + 1, j) - s[ii] * E(l, ld, i, j);
        E(l, ld, i, j) = t;
      }
    }
    if (j < last) {
      int jj = j - k;
      double t = s[jj];
      orc_drotg(&E(l, ld, j, j), &t, &c[jj], &s[jj]);
    }
  }
  /* drop the last row/column; sign fix (src/symmmatrix.cpp:188-194) */
  int mm = m - 1;
  for (int i = col_rem; i < mm; ++i) {
    if (E(l, ld, i, i) < 0) {
      for (int j = i; j < mm; ++j) E(l, ld, i, j) *= -1;
    }
  }
  free(s); free(c);
}

/* src/symmmatrix.cpp:220-265 */
void orc_chol_swapadj(double* l, int ld, int m, int col, double* v)
{
  double* c1 = l + (size_t)col * ld;
  double* c2 = c1 + ld;
  for (int i = 0; i < col + 2; ++i) {
    double t = c1[i]; c1[i] = c2[i]; c2[i] = t;
  }
  c2[col + 1] = 0.0;
  double* a = &E(l, ld, col, col);
  double* b = a + 1;
  double c, s;
  orc_drotg(a, b, &c, &s);
  if (*a < 0) { *a *= -1.0; s *= -1.0; c *= -1.0; }
  for (int i = col + 1; i < m; ++i) {
    double* x1 = &E(l, ld, col, i);
    double* x2 = x1 + 1;
    double tmp = *x1 * s - *x2 * c;
    *x1 = *x2 * s + *x1 * c;
    *x2 = tmp;
  }
  if (v) {
    double* x1 = &v[col];
    double* x2 = &v[col + 1];
    double tmp = *x1 * s - *x2 * c;
    *x1 = *x2 * s + *x1 * c;
    *x2 = tmp;
  }
}

/* src/symmmatrix.cpp:126-142 */
void orc_remove_colrow(double* a, int ld, int m, int col_rem)
{
  for (int i = col_rem; i < m - 1; ++i) {
    for (int j = 0; j <= i; ++j) {
      if (j < col_rem) E(a, ld, j, i) = E(a, ld, j, i + 1);
      else E(a, ld, j, i) = E(a, ld, j + 1, i + 1);
    }
  }
}

/* src/symmmatrix.cpp:197-218 */
void orc_swap_colrow(double* a, int ld, int m, int colrow)
{
  double* c1 = a + (size_t)colrow * ld;
  double* c2 = c1 + ld;
  for (int i = 0; i < colrow; ++i) { double t = c1[i]; c1[i] = c2[i]; c2[i] = t; }
  double t = c1[colrow]; c1[colrow] = c2[colrow + 1]; c2[colrow + 1] = t;
  for (int i = colrow + 2; i < m; ++i) {
    double* x1 = &E(a, ld, colrow, i);
    double tt = x1[0]; x1[0] = x1[1]; x1[1] = tt;
  }
}
