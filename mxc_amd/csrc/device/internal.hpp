// Internal interfaces shared by the device-layer translation units.
#pragma once

#include "common.hpp"

namespace mxg {

// ---- kernel-block generation (kgen.hip) ----
// Block out[r + c*ld] = K(|p(row r) - p(col c)|) with rows/cols given either
// by int32 body-index lists or by contiguous ranges starting at rbeg/cbeg.
struct GenTask {
  const int* ri;   // row body indices (nullptr -> contiguous from rbeg)
  const int* ci;   // column body indices (nullptr -> contiguous from cbeg)
  long long rbeg, cbeg;
  void* out;       // double* or float*
  int m, n, ld;
  int pad;
};
// Tile shape of the generator: 64 rows x 32 columns per workgroup.
constexpr int GEN_TM = 64, GEN_TN = 32;
template <class T>
void gen_tasks(const std::vector<GenTask>& tasks, const double* d_bodies, mxg_kernel_t k, hipStream_t s);

// ---- batched GEMM with K segments (gemm.hip) ----
// C[m x n] = alpha * sum_s op(A_s) op(B_s) + beta * C. op = transpose when
// ta/tb = 1. amask/bmask = 1 zeroes stored entries below the diagonal
// (the operand is the upper triangle of a stored square matrix).
template <class T>
struct GemmSeg {
  const T* A;
  const T* B;
  int lda, ldb, k;
  unsigned char ta, tb, amask, bmask;
};
template <class T>
struct GemmTask {
  T* C;
  int ldc, m, n;
  int seg0, nseg;
  double alpha, beta;
};
// bf16: FP32 tasks whose inputs are rounded to bfloat16 and multiplied on the
// bfloat16 matrix cores with FP32 accumulation. With a stage, the descriptor
// arrays are staged there and the call returns without waiting for the
// launch; without one it waits, and the caller may free what it passed.
template <class T>
void gemm_batched(const std::vector<GemmTask<T>>& tasks, const std::vector<GemmSeg<T>>& segs, hipStream_t s,
                  bool bf16 = false, Stage* st = nullptr);

// ---- dense factorizations, one workgroup per matrix (lapack.hip) ----
// Householder QR of an m x n column-major matrix in place (R on/above the
// diagonal, reflectors below), Eigen's reflector convention.
struct QrTask {
  double* A;
  double* tau;  // min(m, n) entries
  int m, n, lda;
  int zero_lower;  // 1: clear the reflectors within the top min(m, n) rows (leaves triu R there)
};
void qr_batched(const std::vector<QrTask>& tasks, hipStream_t s);

// Explicit M x M Q from the r reflectors of a QR-factored M x r matrix P
// (LAPACK org2r order), and R = triu(P[:r, :r]) into an M x M block (zeroed
// outside its leading r x r).
struct FormQTask {
  const double* P;
  const double* tau;
  double* Q;
  double* R;
  int M, r, ldp;
  int pad;
};
void formq_batched(const std::vector<FormQTask>& tasks, hipStream_t s);

// Column-pivoted QR with Eigen ColPivHouseholderQR semantics on K x M
// matrices, stopping after `steps` pivots (or earlier once the threshold rank
// is settled). Writes the rank and the column permutation.
struct CpqrTask {
  double* A;      // K x M, ld = lda
  int* perm;      // M entries: perm[j] = original column at position j
  int* rank_out;  // 1 int
  double* work;   // 3*M doubles
  int K, M, lda, steps;
  double tol;     // > 0: threshold rank, else all `steps` pivots count
  int maxrank;    // > 0: cap
  int pad;
};
void cpqr_batched(const std::vector<CpqrTask>& tasks, hipStream_t s);

// LU with partial pivoting and explicit inverse of an n x n block:
// inv = A^{-1} (ld n); info = 1 when log|det| <= log(DBL_MIN).
template <class T>
struct LuTask {
  const T* A;  // input block, ld lda (not modified)
  T* inv;      // output n x n, ld n
  T* work;     // n*n + 2n scratch
  int* info;
  int n, lda;
};
template <class T>
void lu_inverse_batched(const std::vector<LuTask<T>>& tasks, hipStream_t s, Stage* st = nullptr);

// ---- Rademacher sketch of far-field blocks (sketch.hip) ----
// Y (M x M, ld M) = Omega G, G(t, j) = K(|body(rows[t]) - body(cols[j])|),
// t < nf, Omega(i, t) = -1 when bit (i mod 64) of the word mixed from
// (t, i / 64) is set, else +1. Boxes with more than SKETCH_ROWS far rows are
// summed from slices of SKETCH_ROWS rows held in a partial buffer of
// sketch_partial_size(M, nf) doubles per box (in box order).
constexpr int SKETCH_ROWS = 8192;
struct SketchBox {
  const int* rows;
  const int* cols;
  double* Y;
  long long nf;
  int M;
  int pad;
};
struct SketchTask {
  const int* rows;
  const int* cols;
  double* out;
  int t0, t1, M, i0, j0;
  int pad;
};
struct SliceSum {
  const double* src;
  double* dst;
  long long len;
  int nsplit;
  int pad;
};
void sketch_rows(mxg_ctx_s* ctx, const std::vector<SketchBox>& boxes, const double* d_bodies, mxg_kernel_t k,
                 double* d_partial, hipStream_t s);
size_t sketch_partial_size(long long M, long long nf);

// ---- misc kernels (h2.hip) ----
template <class Tdst, class Tsrc>
void convert(Tdst* dst, const Tsrc* src, size_t n, hipStream_t s);

}  // namespace mxg
