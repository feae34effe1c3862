// Dense factorizations, one 256-thread workgroup per matrix, variable sizes.
//
//  * qr:   Householder QR (Eigen HouseholderQR reflector convention,
//          h2matrix.cpp:78-79,102-105): beta = -sign(x0)||x||,
//          v = [1; x_tail/(x0 - beta)], tau = (beta - x0)/beta.
//  * cpqr: column-pivoted QR with Eigen ColPivHouseholderQR semantics
//          (h2matrix.cpp:84-89): first-index argmax of the updated column
//          norms, LAPACK xGEQPF norm downdating with recomputation below
//          sqrt(eps), the "nonzero pivots" cut and rank() = #{|R_ii| > tol *
//          max pivot}. Stops as soon as the rank is settled.
//  * lu_inverse: LU with partial pivoting (Eigen PartialPivLU,
//          h2matrix.cpp:435-441) followed by the explicit inverse from the
//          triangular factors; flags |det| <= DBL_MIN through log|det|.
// Each column operation maps a wave to a column with lanes down the rows, so
// global accesses are coalesced; workgroup reductions go through LDS.
#include "internal.hpp"

#include <cfloat>

namespace mxg {
namespace {

struct Red {
  double v[4];
  int i[4];
};

__device__ inline double wg_sum(double v, Red& r) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) r.v[wave] = v;
  __syncthreads();
  return (r.v[0] + r.v[1]) + (r.v[2] + r.v[3]);
}

// max of v over the workgroup; ties resolved to the smallest index.
__device__ inline void wg_argmax(double v, int idx, Red& r, double& bv, int& bi) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, 64);
    int oi = __shfl_xor(idx, o, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) { r.v[wave] = v; r.i[wave] = idx; }
  __syncthreads();
  bv = r.v[0];
  bi = r.i[0];
  for (int w = 1; w < 4; w++)
    if (r.v[w] > bv || (r.v[w] == bv && r.i[w] < bi)) { bv = r.v[w]; bi = r.i[w]; }
}

// Householder reflector of x[0..len) (stride 1) in place; returns beta, writes tau.
// All threads participate; x[0] is left for the caller to overwrite with beta.
__device__ inline double make_reflector(double* x, int len, Red& r, double& tau) {
  double s = 0.;
  for (int i = 1 + threadIdx.x; i < len; i += 256) s += x[i] * x[i];
  s = wg_sum(s, r);
  const double c0 = x[0];
  double beta;
  if (s <= DBL_MIN) {
    tau = 0.;
    beta = c0;
    for (int i = 1 + threadIdx.x; i < len; i += 256) x[i] = 0.;
  } else {
    beta = sqrt(c0 * c0 + s);
    if (c0 >= 0.) beta = -beta;
    const double sc = 1. / (c0 - beta);
    for (int i = 1 + threadIdx.x; i < len; i += 256) x[i] *= sc;
    tau = (beta - c0) / beta;
  }
  return beta;
}

// Apply H = I - tau [1; v][1; v]^T (v = x[1..len)) to columns c of the block
// starting at row 0 of `base` (ld) for c in [c0, c1). Wave per column.
__device__ inline void apply_reflector(const double* x, int len, double tau, double* base, long long ld, int c0, int c1) {
  if (tau == 0.) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = c0 + wave; c < c1; c += 4) {
    double* col = base + (long long)c * ld;
    double w = (lane == 0) ? col[0] : 0.;
    for (int i = 1 + lane; i < len; i += 64) w += x[i] * col[i];
    w = wave_sum(w) * tau;
    if (lane == 0) col[0] -= w;
    for (int i = 1 + lane; i < len; i += 64) col[i] -= w * x[i];
  }
}

__global__ void __launch_bounds__(256) k_qr(const QrTask* __restrict__ tasks) {
  __shared__ Red red;
  __shared__ double sh_tau;
  const QrTask t = tasks[blockIdx.x];
  const int size = min(t.m, t.n);
  for (int k = 0; k < size; k++) {
    double* col = t.A + k + (long long)k * t.lda;
    const int len = t.m - k;
    double tau;
    const double beta = make_reflector(col, len, red, tau);
    __syncthreads();
    if (threadIdx.x == 0) {
      col[0] = beta;
      t.tau[k] = tau;
      sh_tau = tau;
    }
    __syncthreads();
    apply_reflector(col, len, sh_tau, t.A + k, t.lda, k + 1, t.n);
    __syncthreads();
  }
  if (t.zero_lower)
    for (long long e = threadIdx.x; e < (long long)size * size; e += 256) {
      const int i = (int)(e % size), j = (int)(e / size);
      if (i > j) t.A[i + (long long)j * t.lda] = 0.;
    }
}

__global__ void __launch_bounds__(256) k_cpqr(const CpqrTask* __restrict__ tasks) {
  __shared__ Red red;
  __shared__ double sh_tau, sh_maxpivot;
  const CpqrTask t = tasks[blockIdx.x];
  const int K = t.K, M = t.M, size = min(K, M);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* upd = t.work;
  double* dir = t.work + M;
  for (int j = threadIdx.x; j < M; j += 256) t.perm[j] = j;
  for (int j = wave; j < M; j += 4) {
    const double* col = t.A + (long long)j * t.lda;
    double s = 0.;
    for (int i = lane; i < K; i += 64) s += col[i] * col[i];
    s = wave_sum(s);
    if (lane == 0) upd[j] = dir[j] = sqrt(s);
  }
  __syncthreads();
  double mx = 0.;
  for (int j = threadIdx.x; j < M; j += 256) mx = fmax(mx, upd[j]);
  double maxnorm;
  int dummy;
  wg_argmax(mx, 0, red, maxnorm, dummy);
  const double eps = DBL_EPSILON;
  const double th_helper = (maxnorm * eps) * (maxnorm * eps) / (double)K;
  const double downdate_th = sqrt(eps);
  int nonzero = size, done = 0;
  double maxpivot = 0.;
  for (int k = 0; k < t.steps; k++) {
    double bv = -1.;
    int bi = M;
    for (int j = k + threadIdx.x; j < M; j += 256)
      if (upd[j] > bv) { bv = upd[j]; bi = j; }
    double gv;
    int gi;
    wg_argmax(bv, bi, red, gv, gi);
    if (nonzero == size && gv * gv < th_helper * (double)(K - k)) nonzero = k;
    if (t.tol > 0. && nonzero < size) break;                      // rank counts i < nonzero only
    if (t.tol > 0. && k > 0 && gv <= t.tol * maxpivot) break;    // every later |R_ii| is below the threshold
    if (gi != k) {
      double* ck = t.A + (long long)k * t.lda;
      double* cb = t.A + (long long)gi * t.lda;
      for (int i = threadIdx.x; i < K; i += 256) {
        double a = ck[i];
        ck[i] = cb[i];
        cb[i] = a;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double a = upd[k]; upd[k] = upd[gi]; upd[gi] = a;
        a = dir[k]; dir[k] = dir[gi]; dir[gi] = a;
        int p = t.perm[k]; t.perm[k] = t.perm[gi]; t.perm[gi] = p;
      }
    }
    __syncthreads();
    double* col = t.A + k + (long long)k * t.lda;
    const int len = K - k;
    double tau;
    const double beta = make_reflector(col, len, red, tau);
    __syncthreads();
    if (threadIdx.x == 0) {
      col[0] = beta;
      sh_tau = tau;
      sh_maxpivot = fmax(maxpivot, fabs(beta));
    }
    __syncthreads();
    maxpivot = sh_maxpivot;
    apply_reflector(col, len, sh_tau, t.A + k, t.lda, k + 1, M);
    __syncthreads();
    for (int j = k + 1 + threadIdx.x; j < M; j += 256) {
      if (upd[j] != 0.) {
        double q = fabs(t.A[k + (long long)j * t.lda]) / upd[j];
        q = (1. + q) * (1. - q);
        q = q < 0. ? 0. : q;
        const double r = upd[j] / dir[j];
        if (q * r * r <= downdate_th) {
          const double* cj = t.A + (long long)j * t.lda;
          double s = 0.;
          for (int i = k + 1; i < K; i++) s += cj[i] * cj[i];
          dir[j] = upd[j] = sqrt(s);
        } else {
          upd[j] *= sqrt(q);
        }
      }
    }
    __syncthreads();
    done = k + 1;
  }
  if (threadIdx.x == 0) {
    int rank;
    if (t.tol > 0.) {
      const double th = maxpivot * t.tol;
      const int lim = min(nonzero, done);
      rank = 0;
      for (int i = 0; i < lim; i++) rank += fabs(t.A[i + (long long)i * t.lda]) > th;
    } else {
      rank = done;
    }
    if (t.maxrank > 0) rank = min(rank, t.maxrank);
    *t.rank_out = rank;
  }
}

__global__ void __launch_bounds__(256) k_formq(const FormQTask* __restrict__ tasks) {
  const FormQTask t = tasks[blockIdx.x];
  const int M = t.M;
  for (long long e = threadIdx.x; e < (long long)M * M; e += 256) {
    const int i = (int)(e % M), j = (int)(e / M);
    t.Q[e] = (i == j) ? 1. : 0.;
    t.R[e] = (j < t.r && i <= j) ? t.P[i + (long long)j * t.ldp] : 0.;
  }
  __syncthreads();
  for (int k = t.r - 1; k >= 0; k--) {
    apply_reflector(t.P + k + (long long)k * t.ldp, M - k, t.tau[k], t.Q + k, M, k, M);
    __syncthreads();
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_lu_inverse(const LuTask<T>* __restrict__ tasks) {
  __shared__ Red red;
  __shared__ int ipiv[1024];
  const LuTask<T> t = tasks[blockIdx.x];
  const int n = t.n;
  T* W = t.work;
  T* X = t.inv;
  for (long long e = threadIdx.x; e < (long long)n * n; e += 256) {
    const int i = (int)(e % n), j = (int)(e / n);
    W[e] = t.A[i + (long long)j * t.lda];
  }
  __syncthreads();
  double logdet = 0.;
  int singular = 0;
  for (int k = 0; k < n; k++) {
    double bv = -1.;
    int bi = n;
    for (int i = k + threadIdx.x; i < n; i += 256) {
      const double a = fabs((double)W[i + (long long)k * n]);
      if (a > bv) { bv = a; bi = i; }
    }
    double gv;
    int gi;
    wg_argmax(bv, bi, red, gv, gi);
    if (threadIdx.x == 0) ipiv[k] = gi;
    if (gi != k)
      for (int c = threadIdx.x; c < n; c += 256) {
        T a = W[k + (long long)c * n];
        W[k + (long long)c * n] = W[gi + (long long)c * n];
        W[gi + (long long)c * n] = a;
      }
    __syncthreads();
    const T d = W[k + (long long)k * n];
    if (d == T(0)) {
      singular = 1;
      continue;
    }
    logdet += log(fabs((double)d));
    const T inv = T(1) / d;
    for (int i = k + 1 + threadIdx.x; i < n; i += 256) W[i + (long long)k * n] *= inv;
    __syncthreads();
    const int r = n - k - 1;
    for (long long e = threadIdx.x; e < (long long)r * r; e += 256) {
      const int i = k + 1 + (int)(e % r), j = k + 1 + (int)(e / r);
      W[i + (long long)j * n] -= W[i + (long long)k * n] * W[k + (long long)j * n];
    }
    __syncthreads();
  }
  // X = P I, then L^{-1}, then U^{-1}
  for (long long e = threadIdx.x; e < (long long)n * n; e += 256) X[e] = ((e % n) == (e / n)) ? T(1) : T(0);
  __syncthreads();
  for (int k = 0; k < n; k++) {
    const int p = ipiv[k];
    if (p != k)
      for (int c = threadIdx.x; c < n; c += 256) {
        T a = X[k + (long long)c * n];
        X[k + (long long)c * n] = X[p + (long long)c * n];
        X[p + (long long)c * n] = a;
      }
    __syncthreads();
  }
  for (int k = 0; k < n - 1; k++) {
    const int r = n - k - 1;
    for (long long e = threadIdx.x; e < (long long)r * n; e += 256) {
      const int i = k + 1 + (int)(e % r), c = (int)(e / r);
      X[i + (long long)c * n] -= W[i + (long long)k * n] * X[k + (long long)c * n];
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; k--) {
    const T d = W[k + (long long)k * n];
    for (int c = threadIdx.x; c < n; c += 256) X[k + (long long)c * n] /= d;
    __syncthreads();
    for (long long e = threadIdx.x; e < (long long)k * n; e += 256) {
      const int i = (int)(e % k), c = (int)(e / k);
      X[i + (long long)c * n] -= W[i + (long long)k * n] * X[k + (long long)c * n];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *t.info = (singular || logdet <= log(DBL_MIN)) ? 1 : 0;
}

// The same LU with partial pivoting and explicit inverse, with the matrix in
// registers: 512 threads form a 16 x 32 grid and thread (ti, tj) = (t % 16,
// t / 16) owns the TR x TC tile of rows ti TR .. and columns tj TC .. (n <=
// 16 TR = 32 TC); at two waves per SIMD a thread may hold 256 registers, so
// a 192 x 192 FP64 block (TR 12, TC 6: 144 registers of tile) stays on
// chip. Every step of the LU (pivot search, row swap, scaling of column k,
// rank-1 update) and of the two triangular sweeps of the inverse touches LDS
// only for the pivot column and the row it broadcasts, double-buffered by
// the parity of the step, so each step costs one or two workgroup barriers;
// W goes to global memory once, between the LU and the sweeps, which read
// its columns. The arithmetic per element and its order are those of
// k_lu_inverse. Tile indices into the register arrays are compile-time; no
// register is written under a condition on a runtime row index (that would
// turn the tile into an indexed array in scratch memory), so row swaps are
// logical (pos) and the U sweep divides a row when it is written out. The
// FP32 192-row shape (72 registers of tile) is held to 128 registers a
// lane, four waves per SIMD, so that two blocks share a CU and one block's
// barriers overlap the other's arithmetic; a few values spill to scratch
// (measured at N = 2^20: 4096 blocks of 192 x 192, FP32 factorization
// 0.095 -> 0.077 s).
template <class T, int TR, int TC>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && TR == 12 ? 4 : 1)))
k_lu_inverse_reg(const LuTask<T>* __restrict__ tasks) {
  constexpr int NP = 16 * TR;
  static_assert(16 * TR == 32 * TC, "square coverage");
  __shared__ T colb[2][NP];   // column k
  __shared__ T rowp[2][NP];   // the pivot row (the new row k)
  __shared__ int pos[2][NP];  // logical row of each stored row, before / after step k's swap
  __shared__ int perm[NP];
  __shared__ T diag[NP];
  const LuTask<T> task = tasks[blockIdx.x];
  const int n = task.n;
  const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
  const int r0 = ti * TR, c0 = tj * TC;
  const int lane = tid & 63;
  T a[TR][TC];
#pragma unroll
  for (int i = 0; i < TR; i++)
#pragma unroll
    for (int j = 0; j < TC; j++)
      a[i][j] = (r0 + i < n && c0 + j < n) ? task.A[(r0 + i) + (long long)(c0 + j) * task.lda] : T(0);
  double logdet = 0.;
  int singular = 0;
  for (int i = tid; i < NP; i += 512) pos[0][i] = i;
  // Rows never move: a swap of logical rows k and p exchanges their logical
  // positions (pos) instead, and a stored row takes part in step k's update
  // while its logical position is above k.
  for (int k = 0; k < n; k++) {
    const int b = k & 1;
    if (tj == k / TC)
#pragma unroll
      for (int i = 0; i < TR; i++)
#pragma unroll
        for (int j = 0; j < TC; j++)
          if (c0 + j == k) colb[b][r0 + i] = a[i][j];
    __syncthreads();
    // pivot: the largest |a| among rows at logical position >= k, ties to the
    // smallest logical position (every wave finds it)
    double bv = -1.;
    int bl = n, br = n;
    for (int r = lane; r < n; r += 64) {
      const int l = pos[b][r];
      if (l < k) continue;
      const double v = fabs((double)colb[b][r]);
      if (v > bv || (v == bv && l < bl)) { bv = v; bl = l; br = r; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int ol = __shfl_xor(bl, o, 64), orr = __shfl_xor(br, o, 64);
      if (ov > bv || (ov == bv && ol < bl)) { bv = ov; bl = ol; br = orr; }
    }
    const int pr = br;  // the stored row that becomes logical row k
    // the stored row now at logical k moves to the pivot's position bl
    for (int r = tid; r < NP; r += 512) {
      const int l = pos[b][r];
      pos[b ^ 1][r] = (r == pr) ? k : (l == k ? bl : l);
    }
#pragma unroll
    for (int i = 0; i < TR; i++)
#pragma unroll
      for (int j = 0; j < TC; j++)
        if (r0 + i == pr) rowp[b][c0 + j] = a[i][j];
    __syncthreads();
    const T d = rowp[b][k];
    const bool upd = d != T(0);
    if (!upd) singular = 1;
    else logdet += log(fabs((double)d));
    const T inv = upd ? T(1) / d : T(0);
#pragma unroll
    for (int i = 0; i < TR; i++) {
      const int r = r0 + i;
      const bool below = upd && r < n && pos[b ^ 1][r] > k;  // rows the step updates
      const T l = colb[b][r] * inv;
#pragma unroll
      for (int j = 0; j < TC; j++) {
        const int c = c0 + j;
        T v = a[i][j];
        if (below) {
          if (c == k) v = l;
          else if (c > k && c < n) v -= l * rowp[b][c];
        }
        a[i][j] = v;
      }
    }
  }
  // W in logical row order to the work buffer (the sweeps read its columns),
  // its diagonal and the row permutation P to LDS
  const int* fpos = pos[n & 1];
  for (int r = tid; r < n; r += 512) perm[fpos[r]] = r;
#pragma unroll
  for (int i = 0; i < TR; i++)
#pragma unroll
    for (int j = 0; j < TC; j++)
      if (r0 + i < n && c0 + j < n) {
        const int l = fpos[r0 + i];
        task.work[l + (long long)(c0 + j) * n] = a[i][j];
        if (l == c0 + j) diag[l] = a[i][j];
      }
  __syncthreads();
  // X = P I, then L^{-1}, then U^{-1}; x reuses a's registers
  auto& x = a;
#pragma unroll
  for (int i = 0; i < TR; i++)
#pragma unroll
    for (int j = 0; j < TC; j++) x[i][j] = (r0 + i < n && c0 + j == perm[r0 + i]) ? T(1) : T(0);
  const T* W = task.work;
  for (int k = 0; k < n - 1; k++) {
    const int b = k & 1;
#pragma unroll
    for (int i = 0; i < TR; i++)
#pragma unroll
      for (int j = 0; j < TC; j++)
        if (r0 + i == k) colb[b][c0 + j] = x[i][j];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TR; i++) {
      const int r = r0 + i;
      if (r > k && r < n) {
        const T l = W[r + (long long)k * n];
#pragma unroll
        for (int j = 0; j < TC; j++) x[i][j] -= l * colb[b][c0 + j];
      }
    }
  }
  // U^{-1}: step k divides row k by U_kk and subtracts it from the rows
  // above. Row k no longer changes after step k, so its registers keep the
  // undivided values and are divided once when written out; the divided row
  // goes to the other threads through LDS.
  for (int k = n - 1; k >= 0; k--) {
    const int b = k & 1;
    const T dk = diag[k];
#pragma unroll
    for (int i = 0; i < TR; i++)
#pragma unroll
      for (int j = 0; j < TC; j++)
        if (r0 + i == k) rowp[b][c0 + j] = x[i][j] / dk;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TR; i++) {
      const int r = r0 + i;
      if (r < k) {
        const T u = W[r + (long long)k * n];
#pragma unroll
        for (int j = 0; j < TC; j++) x[i][j] -= u * rowp[b][c0 + j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < TR; i++)
#pragma unroll
    for (int j = 0; j < TC; j++)
      if (r0 + i < n && c0 + j < n) task.inv[(r0 + i) + (long long)(c0 + j) * n] = x[i][j] / diag[r0 + i];
  if (tid == 0) *task.info = (singular || logdet <= log(DBL_MIN)) ? 1 : 0;
}

template <class Task, class Kern>
void launch_tasks(const std::vector<Task>& tasks, Kern kern, hipStream_t s, Stage* st = nullptr) {
  if (tasks.empty()) return;
  MXG_REQUIRE(tasks.size() < (1u << 31), "too many tasks");
  DevBuf<Task> dt;
  const Task* pt;
  if (st) {
    pt = st->put(tasks, s);
  } else {
    dt.upload(tasks, s);
    pt = dt.p;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)tasks.size()), dim3(256), 0, s, pt);
  MXG_HIP(hipGetLastError());
  if (!st) MXG_HIP(hipStreamSynchronize(s));
}

}  // namespace

void qr_batched(const std::vector<QrTask>& tasks, hipStream_t s) { launch_tasks(tasks, k_qr, s); }

void formq_batched(const std::vector<FormQTask>& tasks, hipStream_t s) { launch_tasks(tasks, k_formq, s); }

void cpqr_batched(const std::vector<CpqrTask>& tasks, hipStream_t s) {
  for (const CpqrTask& t : tasks) MXG_REQUIRE(t.steps <= std::min(t.K, t.M), "cpqr steps exceed min(K, M)");
  launch_tasks(tasks, k_cpqr, s);
}

// Tile shapes of the register LU (512 threads, n <= 16 TR): FP64 up to
// 192 rows (144 registers of tile per thread), FP32 up to 256.
template <class T>
constexpr int lu_max_rows() {
  return sizeof(T) == 8 ? 192 : 256;
}

template <class Task, class Kern>
void launch_tasks_512(const std::vector<Task>& tasks, Kern kern, hipStream_t s, Stage* st = nullptr) {
  if (tasks.empty()) return;
  MXG_REQUIRE(tasks.size() < (1u << 31), "too many tasks");
  DevBuf<Task> dt;
  const Task* pt;
  if (st) {
    pt = st->put(tasks, s);
  } else {
    dt.upload(tasks, s);
    pt = dt.p;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)tasks.size()), dim3(512), 0, s, pt);
  MXG_HIP(hipGetLastError());
  if (!st) MXG_HIP(hipStreamSynchronize(s));
}

template <class T>
void lu_inverse_batched(const std::vector<LuTask<T>>& tasks, hipStream_t s, Stage* st) {
  for (const LuTask<T>& t : tasks) MXG_REQUIRE(t.n >= 0 && t.n <= 1024, "LU block larger than 1024");
  // by size: the register kernel for each tile shape in use, the
  // global-memory kernel above lu_max_rows
  std::vector<LuTask<T>> by[4], big;
  for (const LuTask<T>& t : tasks) {
    if (t.n <= 64) by[0].push_back(t);
    else if (t.n <= 128) by[1].push_back(t);
    else if (t.n <= 192 && t.n <= lu_max_rows<T>()) by[2].push_back(t);
    else if (t.n <= lu_max_rows<T>()) by[3].push_back(t);
    else big.push_back(t);
  }
  launch_tasks_512(by[0], k_lu_inverse_reg<T, 4, 2>, s, st);
  launch_tasks_512(by[1], k_lu_inverse_reg<T, 8, 4>, s, st);
  launch_tasks_512(by[2], k_lu_inverse_reg<T, 12, 6>, s, st);
  if constexpr (lu_max_rows<T>() >= 256) launch_tasks_512(by[3], k_lu_inverse_reg<T, 16, 8>, s, st);
  launch_tasks(big, k_lu_inverse<T>, s, st);
}

template void lu_inverse_batched<double>(const std::vector<LuTask<double>>&, hipStream_t, Stage*);
template void lu_inverse_batched<float>(const std::vector<LuTask<float>>&, hipStream_t, Stage*);

}  // namespace mxg
