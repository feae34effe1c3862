// Blocked FP64 GEMM/GEMV for the CPU oracle (test infrastructure only).
// Packs op(A) into 8-row panels and op(B) into 4-column panels and runs an
// 8x4 register-tiled micro-kernel written with GCC vector extensions, so the
// oracle's factorization runs at a reasonable fraction of a host core's FMA
// rate. Results are ordinary IEEE FP64 sums; the summation order differs from
// Eigen's, which is within the tolerance the parity tests use.
#include "dense.hpp"

namespace orc {

typedef double v4d __attribute__((vector_size(32)));

static const long MR = 8, NR = 4, KC = 256, MC = 128, NC = 1024;

static inline double getA(char ta, const double* A, long lda, long i, long p) {
  return ta == 'N' ? A[i + p * lda] : A[p + i * lda];
}
static inline double getB(char tb, const double* B, long ldb, long p, long j) {
  return tb == 'N' ? B[p + j * ldb] : B[j + p * ldb];
}

static void pack_a(char ta, const double* A, long lda, long i0, long mc, long p0, long kc, double* buf) {
  for (long ib = 0; ib < mc; ib += MR) {
    long mr = std::min(MR, mc - ib);
    for (long p = 0; p < kc; p++) {
      for (long i = 0; i < mr; i++) buf[i] = getA(ta, A, lda, i0 + ib + i, p0 + p);
      for (long i = mr; i < MR; i++) buf[i] = 0.;
      buf += MR;
    }
  }
}

static void pack_b(char tb, const double* B, long ldb, long p0, long kc, long j0, long nc, double* buf) {
  for (long jb = 0; jb < nc; jb += NR) {
    long nr = std::min(NR, nc - jb);
    for (long p = 0; p < kc; p++) {
      for (long j = 0; j < nr; j++) buf[j] = getB(tb, B, ldb, p0 + p, j0 + jb + j);
      for (long j = nr; j < NR; j++) buf[j] = 0.;
      buf += NR;
    }
  }
}

static inline void micro(long kc, const double* a, const double* b, double alpha, double* C, long ldc, long mr, long nr) {
  v4d c00 = {0, 0, 0, 0}, c01 = c00, c02 = c00, c03 = c00;
  v4d c10 = c00, c11 = c00, c12 = c00, c13 = c00;
  for (long p = 0; p < kc; p++) {
    v4d a0, a1;
    std::memcpy(&a0, a + p * MR, 32);
    std::memcpy(&a1, a + p * MR + 4, 32);
    const double* bp = b + p * NR;
    v4d b0 = {bp[0], bp[0], bp[0], bp[0]};
    v4d b1 = {bp[1], bp[1], bp[1], bp[1]};
    v4d b2 = {bp[2], bp[2], bp[2], bp[2]};
    v4d b3 = {bp[3], bp[3], bp[3], bp[3]};
    c00 += a0 * b0; c10 += a1 * b0;
    c01 += a0 * b1; c11 += a1 * b1;
    c02 += a0 * b2; c12 += a1 * b2;
    c03 += a0 * b3; c13 += a1 * b3;
  }
  double t[MR * NR];
  std::memcpy(t + 0, &c00, 32); std::memcpy(t + 4, &c10, 32);
  std::memcpy(t + 8, &c01, 32); std::memcpy(t + 12, &c11, 32);
  std::memcpy(t + 16, &c02, 32); std::memcpy(t + 20, &c12, 32);
  std::memcpy(t + 24, &c03, 32); std::memcpy(t + 28, &c13, 32);
  for (long j = 0; j < nr; j++)
    for (long i = 0; i < mr; i++) C[i + j * ldc] += alpha * t[i + j * MR];
}

void gemm(char ta, char tb, long m, long n, long k, double alpha, const double* A, long lda,
          const double* B, long ldb, double beta, double* C, long ldc) {
  if (m <= 0 || n <= 0) return;
  if (beta != 1.)
    for (long j = 0; j < n; j++)
      for (long i = 0; i < m; i++) C[i + j * ldc] = (beta == 0.) ? 0. : beta * C[i + j * ldc];
  if (k <= 0 || alpha == 0.) return;
  std::vector<double> pa(MC * KC + MR * KC), pb(NC * KC + NR * KC);
  for (long j0 = 0; j0 < n; j0 += NC) {
    long nc = std::min(NC, n - j0);
    for (long p0 = 0; p0 < k; p0 += KC) {
      long kc = std::min(KC, k - p0);
      pack_b(tb, B, ldb, p0, kc, j0, nc, pb.data());
      for (long i0 = 0; i0 < m; i0 += MC) {
        long mc = std::min(MC, m - i0);
        pack_a(ta, A, lda, i0, mc, p0, kc, pa.data());
        for (long jb = 0; jb < nc; jb += NR)
          for (long ib = 0; ib < mc; ib += MR)
            micro(kc, pa.data() + ib * kc, pb.data() + jb * kc, alpha, C + (i0 + ib) + (j0 + jb) * ldc, ldc,
                  std::min(MR, mc - ib), std::min(NR, nc - jb));
      }
    }
  }
}

void gemv(char ta, long m, long n, double alpha, const double* A, long lda, const double* x, double beta, double* y) {
  if (ta == 'N') {
    for (long i = 0; i < m; i++) y[i] = (beta == 0.) ? 0. : beta * y[i];
    for (long j = 0; j < n; j++) {
      double a = alpha * x[j];
      const double* c = A + j * lda;
      for (long i = 0; i < m; i++) y[i] += c[i] * a;
    }
  } else {
    for (long j = 0; j < n; j++) {
      const double* c = A + j * lda;
      double s = 0.;
      for (long i = 0; i < m; i++) s += c[i] * x[i];
      y[j] = alpha * s + ((beta == 0.) ? 0. : beta * y[j]);
    }
  }
}

}  // namespace orc
