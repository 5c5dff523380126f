/*
 * mthqr_shim.cc -- extern "C" entry points over the reference's own
 * mth/mthQR.cc (Householder QR used by spr::recoverField,
 * spr/sprRecoverField.cc:283-309).  Compiled from the reference sources
 * where they lie by oracle/Makefile into oracle/_ref/; test infrastructure
 * only.  The shim copies plain arrays into mth::Matrix / mth::Vector and
 * back; all arithmetic is the reference's.
 */
#include "mthQR.h"
#include "mth_def.h"

extern "C" {

int ref_decompose_qr(int m, int n, const double* a, double* q, double* r)
{
  mth::Matrix<double> A(m, n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) A(i, j) = a[i * n + j];
  mth::Matrix<double> Q, R;
  unsigned rank = mth::decomposeQR(A, Q, R);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) q[i * m + j] = Q(i, j);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) r[i * n + j] = R(i, j);
  return (int)rank;
}

void ref_solve_from_qr(int m, int n, const double* q, const double* r,
                       const double* b, double* x)
{
  mth::Matrix<double> Q(m, m), R(m, n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Q(i, j) = q[i * m + j];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) R(i, j) = r[i * n + j];
  mth::Vector<double> B(m), X;
  for (int i = 0; i < m; ++i) B(i) = b[i];
  mth::solveFromQR(Q, R, B, X);
  for (int i = 0; i < n; ++i) x[i] = X(i);
}

int ref_solve_qr(int m, int n, const double* a, const double* b, double* x)
{
  mth::Matrix<double> A(m, n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) A(i, j) = a[i * n + j];
  mth::Vector<double> B(m), X;
  for (int i = 0; i < m; ++i) B(i) = b[i];
  bool ok = mth::solveQR(A, B, X);
  if (ok)
    for (int i = 0; i < n; ++i) x[i] = X(i);
  return ok ? 1 : 0;
}

}
