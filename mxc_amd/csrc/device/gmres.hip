// Restarted, left-preconditioned GMRES on the device (replaces
// solveDeviceGMRES, device_gmres.cu:26-115; algorithm of solver.cpp:144-201).
//
// Vectors (Krylov basis, residual, iterate) are FP64 in HBM. Inner products
// use a two-stage reduction (per-block partial sums, then a fixed-order sum),
// so results are bitwise reproducible run to run. Each cycle starts from the
// true residual b - A x with the FP64 H^2 operator, which makes the restarts
// an iterative refinement: the inner operator applications may run in FP32
// (MXG_FP32) against an FP32 preconditioner while the residual stays FP64.
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

// part[blk * ncols + c] = sum over the block's rows of V[:, c] * w
// (w read once per row; columns in passes of DOT_MAXC)
__global__ void __launch_bounds__(256) k_dots(const double* __restrict__ V, int ncols, long long ldv, long long n,
                                              const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double red[4][DOT_MAXC];
  const long long r0 = (long long)blockIdx.x * DOT_ROWS;
  const long long r1 = min(r0 + DOT_ROWS, n);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c0 = 0; c0 < ncols; c0 += DOT_MAXC) {
    const int nc = min(DOT_MAXC, ncols - c0);
    double s[DOT_MAXC];
#pragma unroll
    for (int c = 0; c < DOT_MAXC; c++) s[c] = 0.;
    for (long long r = r0 + threadIdx.x; r < r1; r += 256) {
      const double wr = w[r];
#pragma unroll
      for (int c = 0; c < DOT_MAXC; c++)
        if (c < nc) s[c] += V[r + (long long)(c0 + c) * ldv] * wr;
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
      part[(long long)blockIdx.x * ncols + c0 + threadIdx.x] =
          (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    __syncthreads();
  }
}

// out[c] = sum_b part[b * ncols + c]: one wave per column, lanes over the
// blocks in a fixed stride, then a fixed-shape wave reduction
__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int nblk, int ncols,
                                                   double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncols) return;
  double s = 0.;
  for (int b = lane; b < nblk; b += 64) s += part[(long long)b * ncols + c];
  s = wave_sum(s);
  if (lane == 0) out[c] = s;
}

// y = x / sqrt(nrm2[0]) (zero when nrm2 is not positive)
__global__ void k_scale_rsqrt(const double* __restrict__ x, const double* __restrict__ nrm2, long long n,
                              double* __restrict__ y) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double v = nrm2[0];
  if (r < n) y[r] = v > 0. ? x[r] / sqrt(v) : 0.;
}

// w -= V h (h on the device, ncols entries)
__global__ void k_sub_vh(const double* __restrict__ V, int ncols, long long ldv, long long n, const double* __restrict__ h,
                         double* __restrict__ w) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = w[r];
  for (int c = 0; c < ncols; c++) s -= V[r + (long long)c * ldv] * h[c];
  w[r] = s;
}

// y = alpha * x
__global__ void k_scale(const double* __restrict__ x, double alpha, long long n, double* __restrict__ y) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) y[r] = alpha * x[r];
}

// x += V y
__global__ void k_add_vy(const double* __restrict__ V, int ncols, long long ldv, long long n, const double* __restrict__ y,
                         double* __restrict__ x) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = x[r];
  for (int c = 0; c < ncols; c++) s += V[r + (long long)c * ldv] * y[c];
  x[r] = s;
}

// r = b - r
__global__ void k_b_minus(const double* __restrict__ b, long long n, double* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

struct Gmres {
  mxg_h2_s* A;
  mxg_ulv_t M;
  long long n;
  int m;
  hipStream_t s;
  int nblk;
  DevBuf<double> V, w, r, part, hd, yd;
  double* hpin = nullptr;  // pinned host copy of one Arnoldi step: h, h2, |w|^2

  Gmres(mxg_h2_s* a, mxg_ulv_t mm, int restart) : A(a), M(mm), n(a->n), m(restart), s(a->ctx->stream) {
    nblk = (int)ceil_div(n, DOT_ROWS);
    V.alloc((size_t)n * (m + 1));
    w.alloc(n);
    r.alloc(n);
    part.alloc((size_t)nblk * (m + 1));
    hd.alloc(2 * (m + 1) + 1);
    yd.alloc(m + 1);
    MXG_HIP(hipHostMalloc((void**)&hpin, (2 * (m + 1) + 1) * sizeof(double), hipHostMallocDefault));
  }
  ~Gmres() {
    if (hpin) (void)hipHostFree(hpin);
  }
  Gmres(const Gmres&) = delete;
  Gmres& operator=(const Gmres&) = delete;

  unsigned grid() const { return ceil_div(n, 256); }

  // out[0..ncols) = V[:, :ncols]^T w on the device
  void dots_dev(const double* Vp, int ncols, const double* wp, double* out) {
    hipLaunchKernelGGL(k_dots, dim3(nblk), dim3(256), 0, s, Vp, ncols, n, n, wp, part.p);
    MXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_parts, dim3((ncols + 3) / 4), dim3(256), 0, s, part.p, nblk, ncols, out);
    MXG_HIP(hipGetLastError());
  }

  double norm(const double* x) {
    dots_dev(x, 1, x, hd.p);
    MXG_HIP(hipMemcpyAsync(hpin, hd.p, sizeof(double), hipMemcpyDeviceToHost, s));
    MXG_HIP(hipStreamSynchronize(s));
    return std::sqrt(hpin[0]);
  }

  // w -= V[:, :ncols] h, h on the device
  void sub_vh_dev(int ncols, const double* h) {
    hipLaunchKernelGGL(k_sub_vh, dim3(grid()), dim3(256), 0, s, V.p, ncols, n, n, h, w.p);
    MXG_HIP(hipGetLastError());
  }

  void precondition(double* x) { ulv_apply(M, x); }

  void matvec(const double* x, double* y, int precision, mxg_h2_s* op = nullptr) { h2_apply(op ? op : A, x, y, precision); }
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

}  // namespace
}  // namespace mxg

using namespace mxg;

extern "C" int mxg_gmres(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* o, const double* d_b, double* d_x, double* resid,
                         long long* iters_out, long long* inner_out) {
  return guarded([&] {
    MXG_REQUIRE(A && M && o && d_b && d_x && resid, "null argument");
    MXG_REQUIRE(o->restart >= 1 && o->restart <= 1000 && o->max_cycles >= 0, "bad restart/max_cycles");
    MXG_REQUIRE(o->mode == MXG_GMRES_REFERENCE || o->mode == MXG_GMRES_CGS2, "unknown GMRES mode");
    MXG_REQUIRE(o->inner_precision == MXG_FP64 || o->inner_precision == MXG_FP32, "unknown inner precision");
    mxg_h2_s* Ain = o->inner_operator;
    MXG_REQUIRE(!Ain || (Ain->n == A->n && Ain->ctx == A->ctx), "inner operator must act on the same bodies and device");
    MXG_HIP(hipSetDevice(A->ctx->device));
    Gmres g(A, M, o->restart);
    const long long n = g.n;
    const int m = o->restart;
    const bool ref = (o->mode == MXG_GMRES_REFERENCE);
    double normb = g.norm(d_b);
    if (normb == 0.) normb = 1.;
    // r = b - A x (r = b for a zero initial guess)
    if (g.norm(d_x) == 0.) {
      MXG_HIP(hipMemcpyAsync(g.r.p, d_b, n * sizeof(double), hipMemcpyDeviceToDevice, g.s));
    } else {
      g.matvec(d_x, g.r.p, MXG_FP64);
      hipLaunchKernelGGL(k_b_minus, dim3(g.grid()), dim3(256), 0, g.s, d_b, n, g.r.p);
      MXG_HIP(hipGetLastError());
    }
    resid[0] = g.norm(g.r.p) / normb;
    long long iters = 0, inner_total = 0;
    std::vector<double> H((size_t)(m + 1) * m), h(m + 1), gv(m + 1), cs(m), sn(m), y;
    while (iters < o->max_cycles && o->tol <= resid[iters]) {
      g.precondition(g.r.p);
      const double beta = g.norm(g.r.p);
      if (!(beta > 0.)) break;
      hipLaunchKernelGGL(k_scale, dim3(g.grid()), dim3(256), 0, g.s, g.r.p, 1. / beta, n, g.V.p);
      MXG_HIP(hipGetLastError());
      std::fill(H.begin(), H.end(), 0.);
      std::fill(gv.begin(), gv.end(), 0.);
      gv[0] = beta;
      // inner target: reach tol on the true residual in this cycle, with margin
      double target = 0.3 * o->tol / resid[iters];
      if (o->inner_precision == MXG_FP32) target = std::max(target, 3e-6);
      int k = 0;
      for (int i = 0; i < m; i++) {
        double* vi = g.V.p + (size_t)i * n;
        double* h1d = g.hd.p;
        double* h2d = g.hd.p + (m + 1);
        double* nrd = g.hd.p + 2 * (m + 1);
        g.matvec(vi, g.w.p, o->inner_precision, Ain);
        g.precondition(g.w.p);
        g.dots_dev(g.V.p, i + 1, g.w.p, h1d);
        g.sub_vh_dev(i + 1, h1d);
        if (!ref) {
          g.dots_dev(g.V.p, i + 1, g.w.p, h2d);
          g.sub_vh_dev(i + 1, h2d);
        }
        g.dots_dev(g.w.p, 1, g.w.p, nrd);
        hipLaunchKernelGGL(k_scale_rsqrt, dim3(g.grid()), dim3(256), 0, g.s, g.w.p, nrd, n, g.V.p + (size_t)(i + 1) * n);
        MXG_HIP(hipGetLastError());
        // one transfer per step: h, h2 and |w|^2
        MXG_HIP(hipMemcpyAsync(g.hpin, g.hd.p, (2 * (m + 1) + 1) * sizeof(double), hipMemcpyDeviceToHost, g.s));
        MXG_HIP(hipStreamSynchronize(g.s));
        for (int c = 0; c <= i; c++) h[c] = g.hpin[c] + (ref ? 0. : g.hpin[(m + 1) + c]);
        const double hn = std::sqrt(std::max(g.hpin[2 * (m + 1)], 0.));
        for (int c = 0; c <= i; c++) H[c + (size_t)i * (m + 1)] = h[c];
        H[(i + 1) + (size_t)i * (m + 1)] = hn;
        k = i + 1;
        inner_total++;
        if (!ref) {
          // Givens update of column i and of the residual estimate
          double* col = &H[(size_t)i * (m + 1)];
          for (int c = 0; c < i; c++) {
            const double t = cs[c] * col[c] + sn[c] * col[c + 1];
            col[c + 1] = -sn[c] * col[c] + cs[c] * col[c + 1];
            col[c] = t;
          }
          const double a = col[i], b = col[i + 1];
          const double rr = std::hypot(a, b);
          cs[i] = rr > 0. ? a / rr : 1.;
          sn[i] = rr > 0. ? b / rr : 0.;
          col[i] = rr;
          col[i + 1] = 0.;
          gv[i + 1] = -sn[i] * gv[i];
          gv[i] = cs[i] * gv[i];
          if (std::fabs(gv[i + 1]) <= target * beta) break;
        }
        if (!(hn > 0.)) break;  // happy breakdown
      }
      if (ref) {
        std::vector<double> s(m + 1, 0.);
        s[0] = beta;
        std::vector<double> Hk((size_t)(k + 1) * k);
        for (int c = 0; c < k; c++)
          for (int rr = 0; rr <= k; rr++) Hk[rr + (size_t)c * (k + 1)] = H[rr + (size_t)c * (m + 1)];
        s.resize(k + 1);
        householder_ls(k, k + 1, Hk, s, y);
      } else {
        y.assign(k, 0.);
        for (int i2 = k - 1; i2 >= 0; i2--) {
          double v = gv[i2];
          for (int j = i2 + 1; j < k; j++) v -= H[i2 + (size_t)j * (m + 1)] * y[j];
          y[i2] = v / H[i2 + (size_t)i2 * (m + 1)];
        }
      }
      MXG_HIP(hipMemcpyAsync(g.yd.p, y.data(), k * sizeof(double), hipMemcpyHostToDevice, g.s));
      hipLaunchKernelGGL(k_add_vy, dim3(g.grid()), dim3(256), 0, g.s, g.V.p, k, n, n, g.yd.p, d_x);
      MXG_HIP(hipGetLastError());
      g.matvec(d_x, g.r.p, MXG_FP64);
      hipLaunchKernelGGL(k_b_minus, dim3(g.grid()), dim3(256), 0, g.s, d_b, n, g.r.p);
      MXG_HIP(hipGetLastError());
      resid[++iters] = g.norm(g.r.p) / normb;
    }
    MXG_HIP(hipStreamSynchronize(g.s));
    if (iters_out) *iters_out = iters;
    if (inner_out) *inner_out = inner_total;
  });
}
