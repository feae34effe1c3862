// Dense FP64 linear algebra for the CPU oracle (test infrastructure only).
//
// Column-major storage throughout, leading dimensions explicit. These routines
// restate the arithmetic that the reference obtains from Eigen 3.4.0
// (pinned by Makefile:2 / scripts/Dockerfile:4 of the reference):
//   - HouseholderQR          (h2matrix.cpp:78-79, 102-105; solver.cpp:190)
//   - ColPivHouseholderQR    (h2matrix.cpp:84-89), including its rank() rule
//   - PartialPivLU           (h2matrix.cpp:435-441)
//   - triangular solveInPlace(h2matrix.cpp:94)
// The Householder vector convention is Eigen's makeHouseholderInPlace:
//   beta = -sign(x0) * ||x||, v = [1; x_tail / (x0 - beta)], tau = (beta - x0) / beta.
// The column-norm downdating in cpqr() is the LAPACK xGEQPF / LAWN 176 rule that
// Eigen's ColPivHouseholderQR uses (recompute when the downdated norm has lost
// more than sqrt(eps) of relative accuracy).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

// C[m x n] = alpha * op(A) * op(B) + beta * C, op = 'N' or 'T'.
void gemm(char ta, char tb, long m, long n, long k, double alpha, const double* A, long lda,
          const double* B, long ldb, double beta, double* C, long ldc);

// y = alpha * op(A) x + beta * y
void gemv(char ta, long m, long n, double alpha, const double* A, long lda, const double* x,
          double beta, double* y);

// Householder vector for x[0..n): on exit x[1..n) holds the essential part,
// returns beta (new x[0]) and writes tau.
inline double make_householder(long n, double* x, double* tau) {
  double c0 = x[0];
  double tail = 0.;
  for (long i = 1; i < n; i++) tail += x[i] * x[i];
  const double tol = std::numeric_limits<double>::min();
  if (tail <= tol) {
    *tau = 0.;
    for (long i = 1; i < n; i++) x[i] = 0.;
    return c0;
  }
  double beta = std::sqrt(c0 * c0 + tail);
  if (c0 >= 0.) beta = -beta;
  double s = 1. / (c0 - beta);
  for (long i = 1; i < n; i++) x[i] *= s;
  *tau = (beta - c0) / beta;
  return beta;
}

// Apply H = I - tau [1;v][1;v]^T from the left to the m x n block C (ldc),
// where v has m-1 entries.
inline void apply_householder_left(long m, long n, const double* v, double tau, double* C, long ldc) {
  if (m == 1) {
    for (long j = 0; j < n; j++) C[j * ldc] *= (1. - tau);
    return;
  }
  if (tau == 0.) return;
  for (long j = 0; j < n; j++) {
    double* c = C + j * ldc;
    double w = c[0];
    for (long i = 1; i < m; i++) w += v[i - 1] * c[i];
    w *= tau;
    c[0] -= w;
    for (long i = 1; i < m; i++) c[i] -= w * v[i - 1];
  }
}

// Unpivoted Householder QR in place (reflectors below the diagonal, R on/above).
inline void householder_qr(long m, long n, double* A, long lda, double* tau) {
  long size = std::min(m, n);
  for (long k = 0; k < size; k++) {
    double* col = A + k + k * lda;
    double beta = make_householder(m - k, col, &tau[k]);
    col[0] = beta;
    apply_householder_left(m - k, n - k - 1, col + 1, tau[k], col + lda, lda);
  }
}

// Explicit Q (m x m) from the reflectors stored in A (m x n, size = min(m,n)).
inline void form_q(long m, long n, const double* A, long lda, const double* tau, double* Q, long ldq) {
  long size = std::min(m, n);
  for (long j = 0; j < m; j++)
    for (long i = 0; i < m; i++) Q[i + j * ldq] = (i == j) ? 1. : 0.;
  for (long k = size - 1; k >= 0; k--)
    apply_householder_left(m - k, m - k, A + (k + 1) + k * lda, tau[k], Q + k + k * ldq, ldq);
}

struct CPQRInfo {
  std::vector<long> perm;   // perm[j] = original column now at position j
  double maxpivot = 0.;
  long nonzero_pivots = 0;
};

// Column-pivoted Householder QR in place (Eigen ColPivHouseholderQR::computeInPlace).
// Only the first `steps` pivots are computed when steps < min(m, n); the norm
// bookkeeping (nonzero_pivots, maxpivot) then covers those steps only.
inline CPQRInfo cpqr(long m, long n, double* A, long lda, long steps = -1) {
  CPQRInfo info;
  long size = std::min(m, n);
  if (steps < 0 || steps > size) steps = size;
  info.perm.resize(n);
  for (long j = 0; j < n; j++) info.perm[j] = j;
  std::vector<double> upd(n), dir(n), tau(size);
  double maxnorm = 0.;
  for (long j = 0; j < n; j++) {
    double s = 0.;
    for (long i = 0; i < m; i++) s += A[i + j * lda] * A[i + j * lda];
    dir[j] = upd[j] = std::sqrt(s);
    maxnorm = std::max(maxnorm, upd[j]);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  double threshold_helper = (maxnorm * eps) * (maxnorm * eps) / (double)m;
  double downdate_threshold = std::sqrt(eps);
  info.nonzero_pivots = size;
  info.maxpivot = 0.;
  for (long k = 0; k < steps; k++) {
    long big = k;
    for (long j = k + 1; j < n; j++)
      if (upd[j] > upd[big]) big = j;
    double bsq = upd[big] * upd[big];
    if (info.nonzero_pivots == size && bsq < threshold_helper * (double)(m - k)) info.nonzero_pivots = k;
    if (big != k) {
      for (long i = 0; i < m; i++) std::swap(A[i + k * lda], A[i + big * lda]);
      std::swap(upd[k], upd[big]);
      std::swap(dir[k], dir[big]);
      std::swap(info.perm[k], info.perm[big]);
    }
    double* col = A + k + k * lda;
    double beta = make_householder(m - k, col, &tau[k]);
    col[0] = beta;
    if (std::fabs(beta) > info.maxpivot) info.maxpivot = std::fabs(beta);
    apply_householder_left(m - k, n - k - 1, col + 1, tau[k], col + lda, lda);
    for (long j = k + 1; j < n; j++) {
      if (upd[j] != 0.) {
        double t = std::fabs(A[k + j * lda]) / upd[j];
        t = (1. + t) * (1. - t);
        if (t < 0.) t = 0.;
        double r = upd[j] / dir[j];
        double t2 = t * r * r;
        if (t2 <= downdate_threshold) {
          double s = 0.;
          for (long i = k + 1; i < m; i++) s += A[i + j * lda] * A[i + j * lda];
          dir[j] = upd[j] = std::sqrt(s);
        } else {
          upd[j] *= std::sqrt(t);
        }
      }
    }
  }
  if (steps < size && info.nonzero_pivots > steps) info.nonzero_pivots = steps;
  return info;
}

// Eigen ColPivHouseholderQR::rank() with a prescribed threshold.
inline long cpqr_rank(const CPQRInfo& info, const double* A, long lda, double threshold) {
  double pt = std::fabs(info.maxpivot) * threshold;
  long r = 0;
  for (long i = 0; i < info.nonzero_pivots; i++) r += (std::fabs(A[i + i * lda]) > pt);
  return r;
}

// Solve U X = B in place, U upper triangular n x n, B n x nrhs.
inline void trsm_upper(long n, long nrhs, const double* U, long ldu, double* B, long ldb) {
  for (long c = 0; c < nrhs; c++) {
    double* b = B + c * ldb;
    for (long i = n - 1; i >= 0; i--) {
      double s = b[i];
      for (long j = i + 1; j < n; j++) s -= U[i + j * ldu] * b[j];
      b[i] = s / U[i + i * ldu];
    }
  }
}

// LU with partial pivoting in place (Eigen PartialPivLU / LAPACK getrf semantics).
// Returns log|det| (-inf if exactly singular).
inline double lu_factor(long n, double* A, long lda, long* ipiv) {
  double logdet = 0.;
  for (long k = 0; k < n; k++) {
    long p = k;
    double pmax = std::fabs(A[k + k * lda]);
    for (long i = k + 1; i < n; i++)
      if (std::fabs(A[i + k * lda]) > pmax) { pmax = std::fabs(A[i + k * lda]); p = i; }
    ipiv[k] = p;
    if (p != k)
      for (long j = 0; j < n; j++) std::swap(A[k + j * lda], A[p + j * lda]);
    double d = A[k + k * lda];
    if (d == 0.) { logdet = -std::numeric_limits<double>::infinity(); continue; }
    logdet += std::log(std::fabs(d));
    double inv = 1. / d;
    for (long i = k + 1; i < n; i++) A[i + k * lda] *= inv;
    for (long j = k + 1; j < n; j++) {
      double a = A[k + j * lda];
      if (a == 0.) continue;
      double* cj = A + j * lda;
      const double* ck = A + k * lda;
      for (long i = k + 1; i < n; i++) cj[i] -= ck[i] * a;
    }
  }
  return logdet;
}

// Solve (P L U) X = B for nrhs columns.
inline void lu_solve(long n, long nrhs, const double* LU, long lda, const long* ipiv, double* B, long ldb) {
  for (long c = 0; c < nrhs; c++) {
    double* b = B + c * ldb;
    for (long k = 0; k < n; k++)
      if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
    for (long k = 0; k < n; k++) {
      double bk = b[k];
      if (bk == 0.) continue;
      for (long i = k + 1; i < n; i++) b[i] -= LU[i + k * lda] * bk;
    }
    for (long i = n - 1; i >= 0; i--) {
      b[i] /= LU[i + i * lda];
      double bi = b[i];
      for (long k = 0; k < i; k++) b[k] -= LU[k + i * lda] * bi;
    }
  }
}

}  // namespace orc
