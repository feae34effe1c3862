// Internal layout of an H^2 matrix resident on the device.
//
// Per level l (boxes i = 0..nb-1, tree order):
//   dims[i]  = M_i   box size (leaf: bodies; upper: sum of the children's ranks)
//   ranks[i] = k_i   skeleton rank
//   offX/offZ        prefix sums of M and k: vector layouts. The upper level's
//                    X is the lower level's Z and the lower level's W is the
//                    upper level's Y, so the passes hand vectors over in place.
//   offQ             prefix sum of M^2: Q (basis, M x M) and R (M x M) per box
//   near CSR + offA  near blocks A_ij (M_i x M_j, column major)
//   far CSR + offC   coupling blocks C_ij (k_i x k_j, column major)
//   S                skeleton body indices (int32, M per box, permuted so the
//                    k_i skeleton points come first)
#pragma once

#include "internal.hpp"

namespace mxg {

// One dense product term t = B v of a block-sparse pass: B is an n x m
// column-major (sub-)block with leading dimension ld; b, v and out are
// element offsets into the matrix, input and output arrays of its pass.
struct PairTask {
  long long b, v, out;
  int n, m, ld;
  int pad;
};
// The tasks with their rows cut into chunks of at most `rows` (row r0 of a
// column-major block starts at b + r0): more, shorter jobs for a level whose
// few blocks would otherwise leave the GPU idle. Every output row is still
// written by exactly one job.
inline std::vector<PairTask> split_rows(const std::vector<PairTask>& tasks, int rows) {
  std::vector<PairTask> out;
  for (const PairTask& t : tasks)
    for (int r0 = 0; r0 < std::max(t.n, 1); r0 += rows)
      out.push_back(PairTask{t.b + r0, t.v, t.out + r0, std::min(rows, t.n - r0), t.m, t.ld, t.pad});
  return out;
}
// levels with fewer blocks than this get their rows split (chunks of 32)
constexpr size_t SPLIT_BELOW = 1024;
constexpr int SPLIT_ROWS = 32;
inline std::vector<PairTask> spread(const std::vector<PairTask>& tasks, size_t below = SPLIT_BELOW,
                                    int rows = SPLIT_ROWS) {
  return tasks.size() < below ? split_rows(tasks, rows) : tasks;
}

// Launch shapes of one precision's matvec, chosen by mxg_h2_autotune: the
// symmetric (read-once) or the general block products for the leaf near
// blocks and for the couplings, and the row chunk of the upward products of
// levels with fewer than 1024 blocks.
struct MatvecTuning {
  int near_sym = 0, far_sym = 0;
  int split_rows = SPLIT_ROWS;
  int matrix_cores = 0;  // general near and coupling products of aligned blocks on the matrix cores
};

// The device form of a task: absolute addresses, and vec = 1 when the block
// admits 16-byte column loads (address, rows and ld multiples of 16 bytes).
struct PairJob {
  const void* B;
  const void* v;
  void* out;
  long long ldv, ldo;  // strides (elements) between right-hand sides of v and out
  int n, m, ld;
  int vec;   // 1: B 16-byte aligned with ld and n multiples of the vector width
  int wave;  // 1: small enough for one wave (handled four to a workgroup)
  int mx;    // 1: runs on the matrix cores (k_pair_gemv_mx), else on the vector units
};
// right-hand sides one launch of the solve-phase kernels carries at most
constexpr int MAX_RHS = 8;
// right-hand sides of one launch out of nr remaining: 8, 4, 2 or 1
inline int batch_width(int nr) { return nr >= 8 ? 8 : nr >= 4 ? 4 : nr >= 2 ? 2 : 1; }

// A stored near block A_ij (i <= j) of a symmetric kernel matrix and the two
// terms it gives: o1 = A_ij x_j (n rows) and, for j != i, o2 = A_ij^T x_i =
// A_ji x_i (m rows; -1 for a diagonal block). Offsets into the matrix, the
// input vector (xi, xj) and the term array of the pass.
struct SymTask {
  long long b, xi, xj, o1, o2;
  int n, m, ld;
  int pad;
};
// Device form: absolute addresses; out2 = nullptr for a diagonal block; vec =
// 1 when the block admits 16-byte column loads.
struct SymJob {
  const void* B;
  const void* xi;
  const void* xj;
  void* out1;
  void* out2;
  long long ldv, ldo;  // strides between right-hand sides of x and of the terms
  int n, m, ld;
  int vec;
};
struct SymList {
  DevBuf<SymJob> jobs;
  long long count = 0;
};
// Jobs of one list: those for the matrix cores first (workgroup-sized, then
// wave-sized), then those for the vector units (workgroup-sized, then
// wave-sized); each group is one launch.
struct PairList {
  DevBuf<PairJob> jobs;
  long long count = 0;                   // all jobs
  long long count_mx = 0, nbig_mx = 0;   // jobs [0, count_mx), the first nbig_mx workgroup-sized
  long long nbig = 0;                    // workgroup-sized jobs among [count_mx, count)
};

struct Level {
  long long nb = 0;
  bool built = false;
  std::vector<long long> dims, ranks, nrp, ncol, frp, fcol;
  std::vector<long long> offX, offZ, offQ, offA, offC, diag;
  long long lenX = 0, lenZ = 0, maxM = 0, maxK = 0;
  DevBuf<double> A, Q, R, C;
  DevBuf<int> S;
  DevBuf<long long> d_offX, d_offZ, d_offQ, d_offA, d_offC, d_nrp, d_frp;
  DevBuf<int> d_dims, d_ranks, d_ncol, d_fcol;
  // Schur hand-off to the parent: destination offset in the parent's near
  // blocks and its leading dimension, per near pair of this level (-1: none).
  std::vector<long long> up_off;
  std::vector<int> up_ld;
  // FP32 copies of the operator data (inner matvec of mixed precision GMRES).
  // C32 pads every coupling block's rows to a multiple of 4 with zeros
  // (offsets offC32), so that its columns admit 16-byte loads.
  DevBuf<float> A32, Q32, C32;
  std::vector<long long> offC32;
  // matvec tasks: up z_i = QT_i x_i (QT_i = Q_i[:, :k]^T, k x M), far
  // t_p = C_p z_j, leaf near u_p = A_p x_j; toffF / toffN: term offsets
  // within the level's share of the term scratch (tbaseF / tbaseN; the FP32
  // far terms have the padded row counts of C32: far_t32, toffF32)
  std::vector<long long> offQT, toffF, toffN, toffF32;
  std::vector<PairTask> up_t, far_t, near_t, far_t32;
  // leaf near blocks with i <= j, each giving both of its transposed terms
  std::vector<SymTask> near_sym_t;
  // couplings generated as K(S_i, S_j) on the skeleton points (interpolative
  // bases), so C_ji = C_ij^T bit for bit: each far pair i < j once, FP64 and
  // padded FP32 (offC32 / toffF32)
  bool sym_far = false;
  std::vector<SymTask> far_sym_t, far_sym_t32;
  DevBuf<double> QT;
  DevBuf<float> QT32;
  PairList up_pl[2];  // [0] FP64, [1] FP32
  DevBuf<long long> d_toffF, d_toffN, d_toffF32;
  long long lenTF = 0, lenTN = 0, lenTF32 = 0;
  long long tbaseF[2] = {0, 0}, tbaseN[2] = {0, 0};  // [0] FP64, [1] FP32
  bool vecQ[2] = {false, false};  // every M_i admits 16-byte loads of Q columns (FP64 / FP32)
};

}  // namespace mxg

struct mxg_h2_s {
  mxg_ctx_t ctx = nullptr;
  mxg_kernel_t k{};
  long long n = 0, L = 0;
  int flags = 0;
  mxg::DevBuf<double> bodies;
  std::vector<mxg::Level> lv;
  std::vector<mxg::DevBuf<double>> Z, W;
  std::vector<mxg::DevBuf<float>> Z32, W32;
  mxg::DevBuf<float> x32, y32;
  bool shadow32 = false;
  // matvec: input copy, product terms of every level, the launch of every
  // level's far terms and the launch of the leaf near terms ([0] FP64, [1] FP32)
  mxg::DevBuf<double> x64, t64;
  mxg::DevBuf<float> t32;
  mxg::PairList prod_pl[2];  // coupling products of every level
  mxg::PairList near_pl[2];  // leaf near products
  // the leaf near blocks come from the kernel on one point set (A_ji = A_ij^T
  // bit for bit): the near products read each stored pair once
  // (near_sym); false once near data are defined or uploaded by the caller
  bool sym_near = true;
  mxg::SymList near_sym[2];
  mxg::SymList prod_sym[2];  // coupling products of the levels with sym_far
  mxg::MatvecTuning tune[2];  // [0] FP64, [1] FP32
  bool jobs_ready[2] = {false, false};
  long long tlen[2] = {1, 1};  // term count of the matvec per precision (the stride between right-hand sides)
};

namespace mxg {
// One block copy dst = src (m x n, column major), or dst = src^T with trans.
template <class T>
struct Copy2d {
  const T* src;
  T* dst;
  int m, n, lds, ldd;
  int trans, pad;
};
// every copy of the list in one launch (staged as gemm_batched's)
template <class T>
void copy2d_run(const std::vector<Copy2d<T>>& tasks, hipStream_t s, Stage* st = nullptr);
// launch helpers used by ulv.hip / gmres.hip
template <class T>
void copy2d_batched(const std::vector<const T*>& src, const std::vector<T*>& dst, const std::vector<int>& m,
                    const std::vector<int>& n, const std::vector<int>& lds, const std::vector<int>& ldd, hipStream_t s);
// y = A x for nrhs <= MAX_RHS right-hand sides: x is the internal staging
// buffer of the precision (n x MAX_RHS), y has leading dimension ldy
template <class T>
void h2_matvec_levels(mxg_h2_s* h, T* y, long long ldy, int nrhs);
// Y = A X (n x nrhs, leading dimensions ldx / ldy, FP64) on the context stream
// without a host synchronization; MXG_FP32 applies the FP32 copy of the operator
void h2_apply(mxg_h2_s* h, const double* x, long long ldx, double* y, long long ldy, long long nrhs, int precision);
// dst[r + c ldd] = src[r + c lds], r < n, c < ncols (with conversion)
template <class Td, class Ts>
void convert2d(Td* dst, long long ldd, const Ts* src, long long lds, long long n, int ncols, hipStream_t s);
// X <- M^{-1} X (n x nrhs, leading dimension ldx) without a host synchronization
void ulv_apply(mxg_ulv_s* f, double* x, long long ldx, long long nrhs);

// ---- block-sparse products (bgemv.hip) ----
// bfloat16 storage: the upper half of the FP32 bit pattern
typedef unsigned short bf16_t;
// Appends the jobs of tasks over the arrays B (nB elements, stored as S), v
// (nv per right-hand side, ldv apart) and out (nout per right-hand side, ldo
// apart), after checking every task against those sizes. With matrix_cores,
// blocks stored in the compute type that admit 16-byte column loads go to the
// matrix-core kernel.
template <class T, class S>
void pair_jobs_append(std::vector<PairJob>& jobs, const std::vector<PairTask>& tasks, const S* B, long long nB, const T* v,
                      long long nv, long long ldv, T* out, long long nout, long long ldo, bool matrix_cores = false);
void pair_list_upload(PairList& pl, const std::vector<PairJob>& jobs, hipStream_t s);
// out[r + k ldo] = sum_c B[r + c*ld] v[c + k ldv] for every job and the
// right-hand sides k < nrhs (<= MAX_RHS), one workgroup per job, blocks
// stored as S, arithmetic in T
template <class T, class S>
void pair_gemv(const PairList& pl, int nrhs, hipStream_t s);
// Symmetric near products (symv.hip): jobs of tasks over the matrix B (nB
// elements), the input x (nx per right-hand side, ldv apart) and the terms
// (nout per right-hand side, ldo apart), checked against those sizes; and the
// launch for nrhs <= MAX_RHS right-hand sides.
template <class T>
void sym_jobs_append(std::vector<SymJob>& jobs, const std::vector<SymTask>& tasks, const T* B, long long nB, const T* x,
                     long long nx, long long ldv, T* out, long long nout, long long ldo);
void sym_list_upload(SymList& sl, const std::vector<SymJob>& jobs, hipStream_t s);
template <class T>
void sym_gemv(const SymList& sl, int nrhs, hipStream_t s);
// FP32 -> bfloat16, round to nearest even
void to_bf16(bf16_t* dst, const float* src, size_t n, hipStream_t s);
// w_i = w_in_i + sum_far tF, y_i = Q_i[:, :k] w_i (+ sum_near tN when tN !=
// nullptr); ld*: strides between right-hand sides (tF and tN share ldt)
template <class T>
void comb_down(const Level& Lv, bool vec, const long long* d_toffF, const long long* d_toffN, const T* Q, const T* w_in,
               long long ldw, const T* tF, const T* tN, long long ldt, T* y, long long ldy, int nrhs, hipStream_t s);
// z_i = x_i[:s] - sum_near t
template <class T>
void comb_fwd(const Level& Lv, const long long* d_toff, const T* x, long long ldx, const T* t, long long ldt, T* z,
              long long ldz, int nrhs, hipStream_t s);
// v_i = [w_i ; x_i[s:] - sum_near t] (v at the offsets of x, stride ldv)
template <class T>
void bwd_gather(const Level& Lv, const long long* d_toff, const T* x, long long ldx, const T* w, long long ldw, const T* t,
                long long ldt, T* v, long long ldv, int nrhs, hipStream_t s);
// y_i = U_i [w_i ; x_i[s:] - sum_near t], U stored as S
template <class T, class S>
void comb_bwd(const Level& Lv, bool vec, const long long* d_toff, const S* U, const T* x, long long ldx, const T* w,
              long long ldw, const T* t, long long ldt, T* y, long long ldy, int nrhs, hipStream_t s);
}  // namespace mxg
