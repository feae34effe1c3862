// Restarted, left-preconditioned GMRES on the device (replaces
// solveDeviceGMRES, device_gmres.cu:26-115; algorithm of solver.cpp:144-201),
// for a batch of right-hand sides solved in lockstep.
//
// Vectors (Krylov bases, residuals, iterates) are FP64 in HBM, except the
// Krylov basis of FP32 inner iterations (CGS2 mode), which is stored in FP32:
// those iterations apply FP32 operators to it anyway, and the Arnoldi inner
// products still accumulate in FP64 against the FP64 vector w. Inner products
// use a two-stage reduction (per-block partial sums, then a fixed-order sum),
// so results are bitwise reproducible run to run. Each cycle starts from the
// true residual b - A x with the FP64 H^2 operator, which makes the restarts
// an iterative refinement: the inner operator applications may run in FP32
// (MXG_FP32) against an FP32 or bfloat16-stored preconditioner while the
// residual stays FP64.
//
// Batches: every right-hand side has its own Krylov basis, Hessenberg matrix
// and stopping test; the operator and preconditioner applications of a step
// run for all of them at once (one pass over the H^2 and ULV data per step),
// and the Arnoldi inner products of all of them share one launch. A right-hand
// side that meets its cycle target stops updating its least-squares problem;
// one that meets the tolerance drops out of later cycles.
//
// Modes:
//   MXG_GMRES_REFERENCE  CGS1 Arnoldi, exactly `restart` inner iterations,
//                        least squares by Householder QR (solver.cpp:172-191).
//   MXG_GMRES_CGS2       CGS2 Arnoldi, Givens rotations, inner exit once the
//                        preconditioned residual estimate meets the cycle's
//                        target (or FP32 accuracy for FP32 inner products).
#include <cmath>
#include <vector>

#include "h2.hpp"

namespace mxg {
namespace {

constexpr int DOT_ROWS = 2048;  // rows per block of the partial sums
constexpr int DOT_MAXC = 32;    // columns per pass of k_dots

// Right-hand side q = blockIdx.y: part[(q * nblk + blk) * ncols + c] = sum
// over the block's rows of V_q[:, c] * w_q (w read once per row).
template <class B>
__global__ void __launch_bounds__(256) k_dots(const B* __restrict__ V, int ncols, long long ldv, long long ldq,
                                              long long n, const double* __restrict__ w, long long ldw,
                                              double* __restrict__ part) {
  __shared__ double red[4][DOT_MAXC];
  const int q = blockIdx.y;
  const B* Vq = V + (long long)q * ldq;
  const double* wq = w + (long long)q * ldw;
  double* pq = part + ((long long)q * gridDim.x + blockIdx.x) * ncols;
  const long long r0 = (long long)blockIdx.x * DOT_ROWS;
  const long long r1 = min(r0 + DOT_ROWS, n);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < ncols; c0 += DOT_MAXC) {
    const int nc = min(DOT_MAXC, ncols - c0);
    double s[DOT_MAXC];
#pragma unroll
    for (int c = 0; c < DOT_MAXC; c++) s[c] = 0.;
    for (long long r = r0 + threadIdx.x; r < r1; r += 256) {
      const double wr = wq[r];
#pragma unroll
      for (int c = 0; c < DOT_MAXC; c++)
        if (c < nc) s[c] += (double)Vq[r + (long long)(c0 + c) * ldv] * wr;
    }
#pragma unroll
    for (int c = 0; c < DOT_MAXC; c++) {
      if (c < nc) {
        const double t = wave_sum(s[c]);
        if (lane == 0) red[wave][c] = t;
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < nc)
      pq[c0 + threadIdx.x] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    __syncthreads();
  }
}

// out[q * ldo + c] = sum_b part[(q * nblk + b) * ncols + c]: one wave per
// column, lanes over the blocks in a fixed stride, a fixed-shape reduction
__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int nblk, int ncols,
                                                   double* __restrict__ out, long long ldo) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int q = blockIdx.y;
  if (c >= ncols) return;
  const double* pq = part + (long long)q * nblk * ncols;
  double s = 0.;
  for (int b = lane; b < nblk; b += 64) s += pq[(long long)b * ncols + c];
  s = wave_sum(s);
  if (lane == 0) out[(long long)q * ldo + c] = s;
}

// y_q = x_q / sqrt(nrm2[q * ldn]) (zero when nrm2 is not positive), stored
// as B; y2_q (when given) receives the stored value in FP64
template <class B>
__global__ void k_scale_rsqrt(const double* __restrict__ x, long long ldx, const double* __restrict__ nrm2, long long ldn,
                              long long n, B* __restrict__ y, long long ldy, double* __restrict__ y2, long long ldy2) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  const double v = nrm2[(long long)q * ldn];
  if (r >= n) return;
  const B b = (B)(v > 0. ? x[r + q * ldx] / sqrt(v) : 0.);
  y[r + q * ldy] = b;
  if (y2) y2[r + q * ldy2] = (double)b;
}

// w_q -= V_q[:, :ncols] h_q
template <class B>
__global__ void k_sub_vh(const B* __restrict__ V, int ncols, long long ldv, long long ldq, long long n,
                         const double* __restrict__ h, long long ldh, double* __restrict__ w, long long ldw) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (r >= n) return;
  const B* Vq = V + (long long)q * ldq;
  const double* hq = h + (long long)q * ldh;
  double s = w[r + q * ldw];
  for (int c = 0; c < ncols; c++) s -= (double)Vq[r + (long long)c * ldv] * hq[c];
  w[r + q * ldw] = s;
}

// x_q += V_q[:, :k_q] y_q
template <class B>
__global__ void k_add_vy(const B* __restrict__ V, long long ldv, long long ldq, long long n,
                         const double* __restrict__ y, long long ldy, const int* __restrict__ k, double* __restrict__ x,
                         long long ldx) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (r >= n) return;
  const B* Vq = V + (long long)q * ldq;
  const double* yq = y + (long long)q * ldy;
  double s = x[r + q * ldx];
  for (int c = 0; c < k[q]; c++) s += (double)Vq[r + (long long)c * ldv] * yq[c];
  x[r + q * ldx] = s;
}

// r_q = b_q - r_q
__global__ void k_b_minus(const double* __restrict__ b, long long ldb, long long n, double* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i < n) r[i + q * n] = b[i + q * ldb] - r[i + q * n];
}

struct Gmres {
  mxg_h2_s* A;
  mxg_ulv_t M;
  long long n;
  int m, nr;
  hipStream_t s;
  int nblk;
  long long ldq, ldh;  // distance between right-hand sides: Krylov bases and per-step scalars
  bool f32;            // Krylov basis in FP32 (V32, with the newest column also in vcur)
  // workspace of the context (grown on demand, kept between calls)
  struct Buf {
    double* p;
  };
  Buf V, w, r, part, hd, yd, vcur;
  struct Buf32 {
    float* p;
  } V32;
  struct BufI {
    int* p;
  } kd;
  double* hpin = nullptr;  // pinned host copy of one Arnoldi step, ldh per right-hand side

  Gmres(mxg_h2_s* a, mxg_ulv_t mm, int restart, int nrhs, bool basis32)
      : A(a), M(mm), n(a->n), m(restart), nr(nrhs), s(a->ctx->stream), f32(basis32) {
    nblk = (int)ceil_div(n, DOT_ROWS);
    ldq = n * (long long)(m + 1);
    ldh = 2 * (m + 1) + 1;
    auto& g = a->ctx->gmres;
    // a buffer that grows may still be read by the previous call's kernels
    MXG_HIP(hipStreamSynchronize(s));
    V = {nullptr};
    V32 = {nullptr};
    vcur = {nullptr};
    if (f32) {
      g.V32.reserve((size_t)ldq * nr);
      g.vcur.reserve((size_t)n * nr);
      V32.p = g.V32.p;
      vcur.p = g.vcur.p;
    } else {
      g.V.reserve((size_t)ldq * nr);
      V.p = g.V.p;
    }
    g.w.reserve((size_t)n * nr);
    g.r.reserve((size_t)n * nr);
    g.part.reserve((size_t)nblk * (m + 1) * nr);
    g.hd.reserve((size_t)ldh * nr);
    g.yd.reserve((size_t)(m + 1) * nr);
    g.kd.reserve(nr);
    w.p = g.w.p;
    r.p = g.r.p;
    part.p = g.part.p;
    hd.p = g.hd.p;
    yd.p = g.yd.p;
    kd.p = g.kd.p;
    const size_t hn = (size_t)ldh * nr;
    if (g.hpin_n < hn) {
      if (g.hpin) MXG_HIP(hipHostFree(g.hpin));
      g.hpin = nullptr;
      g.hpin_n = 0;
      MXG_HIP(hipHostMalloc((void**)&g.hpin, hn * sizeof(double), hipHostMallocDefault));
      g.hpin_n = hn;
    }
    hpin = g.hpin;
  }
  Gmres(const Gmres&) = delete;
  Gmres& operator=(const Gmres&) = delete;

  dim3 grid() const { return dim3((unsigned)ceil_div(n, 256), (unsigned)nr); }

  // out_q[0..ncols) = V_q[:, :ncols]^T w_q on the device (out_q = out + q ldh)
  template <class B>
  void dots_dev(const B* Vp, int ncols, long long ldvq, const double* wp, long long ldw, double* out) {
    hipLaunchKernelGGL(k_dots<B>, dim3(nblk, nr), dim3(256), 0, s, Vp, ncols, n, ldvq, n, wp, ldw, part.p);
    MXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_parts, dim3((ncols + 3) / 4, nr), dim3(256), 0, s, part.p, nblk, ncols, out, ldh);
    MXG_HIP(hipGetLastError());
  }

  // |x_q|^2 of every right-hand side (x_q = x + q ldx) into hpin[q ldh]
  void norms2(const double* x, long long ldx) {
    dots_dev(x, 1, ldx, x, ldx, hd.p);
    MXG_HIP(hipMemcpyAsync(hpin, hd.p, (size_t)ldh * nr * sizeof(double), hipMemcpyDeviceToHost, s));
    MXG_HIP(hipStreamSynchronize(s));
  }

  // the Krylov basis: out = V[:, :ncols]^T w; w -= V[:, :ncols] h; column col
  // = x / sqrt(nrm2); x += V[:, :k] y (yd, kd); and column i as a matvec input
  void basis_dots(int ncols, double* out) {
    if (f32)
      dots_dev<float>(V32.p, ncols, ldq, w.p, n, out);
    else
      dots_dev<double>(V.p, ncols, ldq, w.p, n, out);
  }
  void basis_sub(int ncols, const double* h) {
    if (f32)
      hipLaunchKernelGGL(k_sub_vh<float>, grid(), dim3(256), 0, s, V32.p, ncols, n, ldq, n, h, ldh, w.p, n);
    else
      hipLaunchKernelGGL(k_sub_vh<double>, grid(), dim3(256), 0, s, V.p, ncols, n, ldq, n, h, ldh, w.p, n);
    MXG_HIP(hipGetLastError());
  }
  void basis_set(int col, const double* x, long long ldx, const double* nrm2) {
    if (f32)
      hipLaunchKernelGGL(k_scale_rsqrt<float>, grid(), dim3(256), 0, s, x, ldx, nrm2, ldh, n, V32.p + (size_t)col * n, ldq,
                         vcur.p, n);
    else
      hipLaunchKernelGGL(k_scale_rsqrt<double>, grid(), dim3(256), 0, s, x, ldx, nrm2, ldh, n, V.p + (size_t)col * n, ldq,
                         (double*)nullptr, (long long)0);
    MXG_HIP(hipGetLastError());
  }
  void basis_add(double* x, long long ldx) {
    if (f32)
      hipLaunchKernelGGL(k_add_vy<float>, grid(), dim3(256), 0, s, V32.p, n, ldq, n, yd.p, (long long)(m + 1), kd.p, x, ldx);
    else
      hipLaunchKernelGGL(k_add_vy<double>, grid(), dim3(256), 0, s, V.p, n, ldq, n, yd.p, (long long)(m + 1), kd.p, x, ldx);
    MXG_HIP(hipGetLastError());
  }
  // column i (the newest one) as an FP64 operator input, and its stride
  const double* basis_col(int i, long long* ld) const {
    *ld = f32 ? n : ldq;
    return f32 ? vcur.p : V.p + (size_t)i * n;
  }

  void precondition(double* x) { ulv_apply(M, x, n, nr); }

  void matvec(const double* x, long long ldx, double* y, int precision, mxg_h2_s* op = nullptr) {
    h2_apply(op ? op : A, x, ldx, y, n, nr, precision);
  }
};

// Householder least squares min ||s - H y|| for H ((k+1) x k, ld), as the
// reference's H.householderQr().solve(s) (solver.cpp:190).
void householder_ls(int k, int ld, std::vector<double> H, std::vector<double> s, std::vector<double>& y) {
  const int rows = k + 1;
  for (int c = 0; c < k; c++) {
    double* x = &H[c + (size_t)c * ld];
    const int len = rows - c;
    double tail = 0.;
    for (int i = 1; i < len; i++) tail += x[i] * x[i];
    double tau = 0., beta = x[0];
    if (tail > 2.2250738585072014e-308) {
      beta = std::sqrt(x[0] * x[0] + tail);
      if (x[0] >= 0.) beta = -beta;
      const double sc = 1. / (x[0] - beta);
      for (int i = 1; i < len; i++) x[i] *= sc;
      tau = (beta - x[0]) / beta;
    } else {
      for (int i = 1; i < len; i++) x[i] = 0.;
    }
    x[0] = beta;
    auto apply = [&](double* col) {
      double wv = col[0];
      for (int i = 1; i < len; i++) wv += x[i] * col[i];
      wv *= tau;
      col[0] -= wv;
      for (int i = 1; i < len; i++) col[i] -= wv * x[i];
    };
    for (int j = c + 1; j < k; j++) apply(&H[c + (size_t)j * ld]);
    apply(&s[c]);
  }
  y.assign(k, 0.);
  for (int i = k - 1; i >= 0; i--) {
    double v = s[i];
    for (int j = i + 1; j < k; j++) v -= H[i + (size_t)j * ld] * y[j];
    y[i] = v / H[i + (size_t)i * ld];
  }
}

// Per right-hand side state of the cycle loop.
struct Rhs {
  std::vector<double> H, gv, cs, sn, y;
  double beta = 0., normb = 1., target = 0.;
  int k = 0;            // Arnoldi steps that enter this cycle's update
  bool active = false;  // takes part in the current cycle
  bool open = false;    // still extending its least-squares problem in this cycle
  long long cycles = 0, inner = 0;
};

void gmres_batch(mxg_h2_s* A, mxg_ulv_t M, const mxg_gmres_opts_t* o, int nr, const double* d_b, long long ldb,
                 double* d_x, long long ldx, double* resid, long long* iters_out, long long* inner_out) {
  mxg_h2_s* Ain = o->inner_operator;
  Gmres g(A, M, o->restart, nr, o->mode == MXG_GMRES_CGS2 && o->inner_precision == MXG_FP32);
  const long long n = g.n;
  const int m = o->restart;
  const long long R = o->max_cycles + 1;
  const bool ref = (o->mode == MXG_GMRES_REFERENCE);
  std::vector<Rhs> st(nr);
  for (Rhs& q : st) {
    q.H.assign((size_t)(m + 1) * m, 0.);
    q.gv.assign(m + 1, 0.);
    q.cs.assign(m, 0.);
    q.sn.assign(m, 0.);
  }
  // |b| and |x0|
  g.norms2(d_b, ldb);
  for (int q = 0; q < nr; q++) st[q].normb = g.hpin[q * g.ldh] > 0. ? std::sqrt(g.hpin[q * g.ldh]) : 1.;
  g.norms2(d_x, ldx);
  bool zero_guess = true;
  for (int q = 0; q < nr; q++) zero_guess = zero_guess && g.hpin[q * g.ldh] == 0.;
  // r = b - A x (r = b for a zero initial guess)
  if (zero_guess) {
    convert2d(g.r.p, n, d_b, ldb, n, nr, g.s);
  } else {
    g.matvec(d_x, ldx, g.r.p, MXG_FP64);
    hipLaunchKernelGGL(k_b_minus, g.grid(), dim3(256), 0, g.s, d_b, ldb, n, g.r.p);
    MXG_HIP(hipGetLastError());
  }
  g.norms2(g.r.p, n);
  for (int q = 0; q < nr; q++) {
    resid[q * R] = std::sqrt(std::max(g.hpin[q * g.ldh], 0.)) / st[q].normb;
    for (long long c = 1; c < R; c++) resid[q * R + c] = 0.;
  }
  std::vector<int> kh(nr);
  std::vector<double> yh((size_t)(m + 1) * nr);
  for (long long cyc = 0; cyc < o->max_cycles; cyc++) {
    bool any = false;
    for (int q = 0; q < nr; q++) {
      st[q].active = o->tol <= resid[q * R + st[q].cycles];
      any = any || st[q].active;
    }
    if (!any) break;
    g.precondition(g.r.p);
    g.norms2(g.r.p, n);
    for (int q = 0; q < nr; q++) {
      Rhs& S = st[q];
      S.beta = std::sqrt(std::max(g.hpin[q * g.ldh], 0.));
      if (!(S.beta > 0.)) S.active = false;
      S.open = S.active;
      S.k = 0;
      std::fill(S.H.begin(), S.H.end(), 0.);
      std::fill(S.gv.begin(), S.gv.end(), 0.);
      S.gv[0] = S.beta;
      // inner target: reach tol on the true residual in this cycle, with margin
      S.target = S.active ? 0.3 * o->tol / resid[q * R + S.cycles] : 0.;
      if (o->inner_precision == MXG_FP32) S.target = std::max(S.target, 3e-6);
    }
    any = false;
    for (int q = 0; q < nr; q++) any = any || st[q].active;
    if (!any) break;
    // V_q[:, 0] = r_q / beta_q (|r_q|^2 is still in hd)
    g.basis_set(0, g.r.p, n, g.hd.p);
    for (int i = 0; i < m; i++) {
      bool open = false;
      for (int q = 0; q < nr; q++) open = open || st[q].open;
      if (!open) break;
      long long ldvi = 0;
      const double* vi = g.basis_col(i, &ldvi);
      double* h1d = g.hd.p;
      double* h2d = g.hd.p + (m + 1);
      double* nrd = g.hd.p + 2 * (m + 1);
      g.matvec(vi, ldvi, g.w.p, o->inner_precision, Ain);
      g.precondition(g.w.p);
      g.basis_dots(i + 1, h1d);
      g.basis_sub(i + 1, h1d);
      if (!ref) {
        g.basis_dots(i + 1, h2d);
        g.basis_sub(i + 1, h2d);
      }
      g.dots_dev<double>(g.w.p, 1, n, g.w.p, n, nrd);
      g.basis_set(i + 1, g.w.p, n, nrd);
      // one transfer per step: h, h2 and |w|^2 of every right-hand side
      MXG_HIP(hipMemcpyAsync(g.hpin, g.hd.p, (size_t)g.ldh * nr * sizeof(double), hipMemcpyDeviceToHost, g.s));
      MXG_HIP(hipStreamSynchronize(g.s));
      for (int q = 0; q < nr; q++) {
        Rhs& S = st[q];
        if (!S.open) continue;
        const double* hp = g.hpin + (size_t)q * g.ldh;
        double* col = &S.H[(size_t)i * (m + 1)];
        for (int c = 0; c <= i; c++) col[c] = hp[c] + (ref ? 0. : hp[(m + 1) + c]);
        const double hn = std::sqrt(std::max(hp[2 * (m + 1)], 0.));
        col[i + 1] = hn;
        S.k = i + 1;
        S.inner++;
        if (!ref) {
          // Givens update of column i and of the residual estimate
          for (int c = 0; c < i; c++) {
            const double t = S.cs[c] * col[c] + S.sn[c] * col[c + 1];
            col[c + 1] = -S.sn[c] * col[c] + S.cs[c] * col[c + 1];
            col[c] = t;
          }
          const double a = col[i], b = col[i + 1];
          const double rr = std::hypot(a, b);
          S.cs[i] = rr > 0. ? a / rr : 1.;
          S.sn[i] = rr > 0. ? b / rr : 0.;
          col[i] = rr;
          col[i + 1] = 0.;
          S.gv[i + 1] = -S.sn[i] * S.gv[i];
          S.gv[i] = S.cs[i] * S.gv[i];
          if (std::fabs(S.gv[i + 1]) <= S.target * S.beta) S.open = false;
        }
        if (!(hn > 0.)) S.open = false;  // happy breakdown
      }
    }
    // least squares per right-hand side, then x_q += V_q y_q
    for (int q = 0; q < nr; q++) {
      Rhs& S = st[q];
      kh[q] = S.active ? S.k : 0;
      std::fill(yh.begin() + (size_t)q * (m + 1), yh.begin() + (size_t)(q + 1) * (m + 1), 0.);
      if (!S.active || S.k == 0) continue;
      const int k = S.k;
      if (ref) {
        std::vector<double> sv(k + 1, 0.);
        sv[0] = S.beta;
        std::vector<double> Hk((size_t)(k + 1) * k);
        for (int c = 0; c < k; c++)
          for (int rr = 0; rr <= k; rr++) Hk[rr + (size_t)c * (k + 1)] = S.H[rr + (size_t)c * (m + 1)];
        householder_ls(k, k + 1, Hk, sv, S.y);
      } else {
        S.y.assign(k, 0.);
        for (int i2 = k - 1; i2 >= 0; i2--) {
          double v = S.gv[i2];
          for (int j = i2 + 1; j < k; j++) v -= S.H[i2 + (size_t)j * (m + 1)] * S.y[j];
          S.y[i2] = v / S.H[i2 + (size_t)i2 * (m + 1)];
        }
      }
      std::copy(S.y.begin(), S.y.end(), yh.begin() + (size_t)q * (m + 1));
    }
    MXG_HIP(hipMemcpyAsync(g.yd.p, yh.data(), yh.size() * sizeof(double), hipMemcpyHostToDevice, g.s));
    MXG_HIP(hipMemcpyAsync(g.kd.p, kh.data(), kh.size() * sizeof(int), hipMemcpyHostToDevice, g.s));
    g.basis_add(d_x, ldx);
    // true residual of every right-hand side
    g.matvec(d_x, ldx, g.r.p, MXG_FP64);
    hipLaunchKernelGGL(k_b_minus, g.grid(), dim3(256), 0, g.s, d_b, ldb, n, g.r.p);
    MXG_HIP(hipGetLastError());
    g.norms2(g.r.p, n);
    for (int q = 0; q < nr; q++) {
      Rhs& S = st[q];
      if (!S.active) continue;
      S.cycles++;
      resid[q * R + S.cycles] = std::sqrt(std::max(g.hpin[q * g.ldh], 0.)) / S.normb;
    }
  }
  MXG_HIP(hipStreamSynchronize(g.s));
  for (int q = 0; q < nr; q++) {
    if (iters_out) iters_out[q] = st[q].cycles;
    if (inner_out) inner_out[q] = st[q].inner;
  }
}

}  // namespace
}  // namespace mxg

using namespace mxg;

namespace {
void check_opts(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* o) {
  MXG_REQUIRE(A && M && o, "null argument");
  MXG_REQUIRE(o->restart >= 1 && o->restart <= 1000 && o->max_cycles >= 0, "bad restart/max_cycles");
  MXG_REQUIRE(o->mode == MXG_GMRES_REFERENCE || o->mode == MXG_GMRES_CGS2, "unknown GMRES mode");
  MXG_REQUIRE(o->inner_precision == MXG_FP64 || o->inner_precision == MXG_FP32, "unknown inner precision");
  mxg_h2_s* Ain = o->inner_operator;
  MXG_REQUIRE(!Ain || (Ain->n == A->n && Ain->ctx == A->ctx), "inner operator must act on the same bodies and device");
}
}  // namespace

extern "C" int mxg_gmres(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* o, const double* d_b, double* d_x, double* resid,
                         long long* iters_out, long long* inner_out) {
  return guarded([&] {
    check_opts(A, M, o);
    MXG_REQUIRE(d_b && d_x && resid, "null argument");
    MXG_HIP(hipSetDevice(A->ctx->device));
    gmres_batch(A, M, o, 1, d_b, A->n, d_x, A->n, resid, iters_out, inner_out);
  });
}

extern "C" int mxg_gmres_batch(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* o, long long nrhs, const double* d_b,
                               long long ldb, double* d_x, long long ldx, double* resid, long long* iters_out,
                               long long* inner_out) {
  return guarded([&] {
    check_opts(A, M, o);
    MXG_REQUIRE(d_b && d_x && resid, "null argument");
    MXG_REQUIRE(nrhs >= 1 && nrhs <= 64, "nrhs out of range (1..64)");
    MXG_REQUIRE(ldb >= A->n && ldx >= A->n, "bad leading dimensions");
    MXG_HIP(hipSetDevice(A->ctx->device));
    gmres_batch(A, M, o, (int)nrhs, d_b, ldb, d_x, ldx, resid, iters_out, inner_out);
  });
}
