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

// ---- the small Arnoldi bookkeeping, on the device in FP64 ----
// Per right-hand side q (one thread each): the Hessenberg matrix H_q
// ((m+1) x m, column major), the Givens rotations cs_q / sn_q, the rotated
// right-hand side g_q (m+1), the inner exit threshold tb_q = target * beta,
// and the flags open_q (still extending its least-squares problem), k_q
// (steps in this cycle) and inner_q (steps over all cycles).

// Step i of every open right-hand side: column i of H_q from this step's
// projections (h1 + h2 with CGS2) and |w|; with CGS2 the previous rotations
// are applied to it, a new one zeroes its subdiagonal, g_q is rotated and the
// inner exit tested (|g_{i+1}| <= target * beta); a zero |w| (happy breakdown)
// closes the right-hand side as well.
__global__ void k_arnoldi_step(int i, int m, int cgs2, int nr, const double* __restrict__ hd, long long ldh,
                               double* __restrict__ H, double* __restrict__ cs, double* __restrict__ sn,
                               double* __restrict__ gv, const double* __restrict__ tb, int* __restrict__ open,
                               int* __restrict__ kq, int* __restrict__ inner) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nr || !open[q]) return;
  const double* hp = hd + (long long)q * ldh;
  double* col = H + (long long)q * (m + 1) * m + (long long)i * (m + 1);
  for (int c = 0; c <= i; c++) col[c] = hp[c] + (cgs2 ? hp[(m + 1) + c] : 0.);
  const double hn = sqrt(fmax(hp[2 * (m + 1)], 0.));
  col[i + 1] = hn;
  kq[q] = i + 1;
  inner[q]++;
  if (cgs2) {
    double* c_ = cs + (long long)q * m;
    double* s_ = sn + (long long)q * m;
    double* g = gv + (long long)q * (m + 1);
    for (int c = 0; c < i; c++) {
      const double t = c_[c] * col[c] + s_[c] * col[c + 1];
      col[c + 1] = -s_[c] * col[c] + c_[c] * col[c + 1];
      col[c] = t;
    }
    const double a = col[i], b = col[i + 1];
    const double rr = hypot(a, b);
    c_[i] = rr > 0. ? a / rr : 1.;
    s_[i] = rr > 0. ? b / rr : 0.;
    col[i] = rr;
    col[i + 1] = 0.;
    g[i + 1] = -s_[i] * g[i];
    g[i] = c_[i] * g[i];
    if (fabs(g[i + 1]) <= tb[q]) open[q] = 0;
  }
  if (!(hn > 0.)) open[q] = 0;
}

// End of a cycle: y_q minimizing |beta_q e_1 - H_q y| over the k_q steps of
// an active right-hand side (kd_q = k_q; 0 for an inactive one, which then
// keeps its iterate). CGS2: back substitution with the rotated H_q and g_q.
// Reference mode: Householder QR of H_q (its copy in W) and of beta e_1, then
// back substitution, as the reference's H.householderQr().solve(s)
// (solver.cpp:187-191).
__global__ void k_cycle_lsq(int m, int cgs2, int nr, const double* __restrict__ H, const double* __restrict__ gv,
                            const double* __restrict__ beta, const int* __restrict__ kq, const int* __restrict__ active,
                            double* __restrict__ W, double* __restrict__ y, int* __restrict__ kd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nr) return;
  const int ld = m + 1;
  double* yq = y + (long long)q * ld;
  for (int t = 0; t <= m; t++) yq[t] = 0.;
  const int k = active[q] ? kq[q] : 0;
  kd[q] = k;
  if (k == 0) return;
  const double* Hq = H + (long long)q * ld * m;
  if (cgs2) {
    const double* g = gv + (long long)q * ld;
    for (int i2 = k - 1; i2 >= 0; i2--) {
      double v = g[i2];
      for (int j = i2 + 1; j < k; j++) v -= Hq[i2 + (long long)j * ld] * yq[j];
      yq[i2] = v / Hq[i2 + (long long)i2 * ld];
    }
    return;
  }
  double* A = W + (long long)q * ld * (m + 1);  // H_q copy, then the right-hand side in column m
  for (int c = 0; c < k; c++)
    for (int r = 0; r <= k; r++) A[r + (long long)c * ld] = Hq[r + (long long)c * ld];
  double* sv = A + (long long)m * ld;
  for (int r = 0; r <= k; r++) sv[r] = r == 0 ? beta[q] : 0.;
  const int rows = k + 1;
  for (int c = 0; c < k; c++) {
    double* x = A + c + (long long)c * ld;
    const int len = rows - c;
    double tail = 0.;
    for (int i = 1; i < len; i++) tail += x[i] * x[i];
    double tau = 0., bet = x[0];
    if (tail > 2.2250738585072014e-308) {
      bet = sqrt(x[0] * x[0] + tail);
      if (x[0] >= 0.) bet = -bet;
      const double sc = 1. / (x[0] - bet);
      for (int i = 1; i < len; i++) x[i] *= sc;
      tau = (bet - x[0]) / bet;
    } else {
      for (int i = 1; i < len; i++) x[i] = 0.;
    }
    x[0] = bet;
    for (int j = c + 1; j <= k; j++) {
      double* colj = (j < k) ? A + c + (long long)j * ld : sv + c;
      double wv = colj[0];
      for (int i = 1; i < len; i++) wv += x[i] * colj[i];
      wv *= tau;
      colj[0] -= wv;
      for (int i = 1; i < len; i++) colj[i] -= wv * x[i];
    }
  }
  for (int i = k - 1; i >= 0; i--) {
    double v = sv[i];
    for (int j = i + 1; j < k; j++) v -= A[i + (long long)j * ld] * yq[j];
    yq[i] = v / A[i + (long long)i * ld];
  }
}

// One HIP graph of an inner step's launches (operator and preconditioner),
// captured once per call on the context stream (the matvec's second stream
// joins the capture through its fork/join events) and replayed for the later
// steps. The graph holds the addresses of the call's buffers, so it never
// outlives the call.
struct StepGraph {
  bool enabled = false, ready = false;
  hipGraphExec_t exec = nullptr;
  template <class F>
  void capture(hipStream_t s, F launches) {
    hipGraph_t graph = nullptr;
    MXG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    try {
      launches();
    } catch (...) {
      (void)hipStreamEndCapture(s, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    MXG_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MXG_HIP(e);
    ready = true;
  }
  StepGraph() = default;
  StepGraph(const StepGraph&) = delete;
  StepGraph& operator=(const StepGraph&) = delete;
  ~StepGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

// Per right-hand side state kept on the host: the cycle bookkeeping.
struct Rhs {
  double beta = 0., normb = 1., target = 0.;
  bool active = false;  // takes part in the current cycle
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
  // device state of the small problems (context workspace)
  auto& ws = A->ctx->gmres;
  ws.H.reserve((size_t)(m + 1) * m * nr);
  ws.W.reserve((size_t)(m + 1) * (m + 1) * nr);
  ws.gv.reserve((size_t)(m + 1) * nr);
  ws.cs.reserve((size_t)m * nr);
  ws.sn.reserve((size_t)m * nr);
  ws.tb.reserve((size_t)2 * nr);  // tb, then beta
  ws.flags.reserve((size_t)4 * nr);  // open, k, inner, active
  double* dH = ws.H.p;
  double* dgv = ws.gv.p;
  double* dtb = ws.tb.p;
  double* dbeta = ws.tb.p + nr;
  int* dopen = ws.flags.p;
  int* dk = ws.flags.p + nr;
  int* dinner = ws.flags.p + 2 * nr;
  int* dactive = ws.flags.p + 3 * nr;
  int* hflags = reinterpret_cast<int*>(g.hpin);  // 4 nr ints fit the pinned buffer (ldh >= 5 doubles per rhs)
  const dim3 small((unsigned)ceil_div(nr, 64)), lanes(64);
  MXG_HIP(hipMemsetAsync(dinner, 0, (size_t)nr * sizeof(int), g.s));
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
  // graph replay of the inner step: its input is the fixed staging column of
  // the FP32 basis; events of the kernel timer cannot live in a graph
  StepGraph step_graph;
  step_graph.enabled = o->graph != 0 && g.f32 && !A->ctx->profiling;
  std::vector<double> h_state((size_t)(m + 1) * nr + 2 * nr);
  std::vector<int> h_open(2 * (size_t)nr);
  for (long long cyc = 0; cyc < o->max_cycles; cyc++) {
    bool any = false;
    for (int q = 0; q < nr; q++) {
      st[q].active = o->tol <= resid[q * R + st[q].cycles];
      any = any || st[q].active;
    }
    if (!any) break;
    g.precondition(g.r.p);
    g.norms2(g.r.p, n);
    // cycle start: beta_q, the inner target, g_q = beta_q e_1, H_q = 0
    double* hg = h_state.data();
    double* htb = hg + (size_t)(m + 1) * nr;
    double* hbeta = htb + nr;
    std::fill(h_state.begin(), h_state.end(), 0.);
    for (int q = 0; q < nr; q++) {
      Rhs& S = st[q];
      S.beta = std::sqrt(std::max(g.hpin[q * g.ldh], 0.));
      if (!(S.beta > 0.)) S.active = false;
      // inner target: reach tol on the true residual in this cycle, with margin
      S.target = S.active ? 0.3 * o->tol / resid[q * R + S.cycles] : 0.;
      if (o->inner_precision == MXG_FP32) S.target = std::max(S.target, 3e-6);
      hg[(size_t)q * (m + 1)] = S.beta;
      htb[q] = S.target * S.beta;
      hbeta[q] = S.beta;
      h_open[q] = S.active ? 1 : 0;
      h_open[nr + q] = S.active ? 1 : 0;
    }
    any = false;
    for (int q = 0; q < nr; q++) any = any || st[q].active;
    if (!any) break;
    MXG_HIP(hipMemcpyAsync(dgv, hg, (size_t)(m + 1) * nr * sizeof(double), hipMemcpyHostToDevice, g.s));
    MXG_HIP(hipMemcpyAsync(dtb, htb, (size_t)2 * nr * sizeof(double), hipMemcpyHostToDevice, g.s));
    MXG_HIP(hipMemcpyAsync(dopen, h_open.data(), (size_t)nr * sizeof(int), hipMemcpyHostToDevice, g.s));
    MXG_HIP(hipMemcpyAsync(dactive, h_open.data() + nr, (size_t)nr * sizeof(int), hipMemcpyHostToDevice, g.s));
    MXG_HIP(hipMemsetAsync(dk, 0, (size_t)nr * sizeof(int), g.s));
    MXG_HIP(hipMemsetAsync(dH, 0, (size_t)(m + 1) * m * nr * sizeof(double), g.s));
    // V_q[:, 0] = r_q / beta_q (|r_q|^2 is still in hd)
    g.basis_set(0, g.r.p, n, g.hd.p);
    for (int i = 0; i < m; i++) {
      long long ldvi = 0;
      const double* vi = g.basis_col(i, &ldvi);
      double* h1d = g.hd.p;
      double* h2d = g.hd.p + (m + 1);
      double* nrd = g.hd.p + 2 * (m + 1);
      if (step_graph.enabled && step_graph.ready) {
        MXG_HIP(hipGraphLaunch(step_graph.exec, g.s));
      } else {
        // the first step of a graph-enabled call runs plainly (it finishes any
        // set-up of the operator and factors) and then once more under capture
        g.matvec(vi, ldvi, g.w.p, o->inner_precision, Ain);
        g.precondition(g.w.p);
        if (step_graph.enabled) step_graph.capture(g.s, [&] {
          g.matvec(vi, ldvi, g.w.p, o->inner_precision, Ain);
          g.precondition(g.w.p);
        });
      }
      g.basis_dots(i + 1, h1d);
      g.basis_sub(i + 1, h1d);
      if (!ref) {
        g.basis_dots(i + 1, h2d);
        g.basis_sub(i + 1, h2d);
      }
      g.dots_dev<double>(g.w.p, 1, n, g.w.p, n, nrd);
      g.basis_set(i + 1, g.w.p, n, nrd);
      hipLaunchKernelGGL(k_arnoldi_step, small, lanes, 0, g.s, i, m, ref ? 0 : 1, nr, g.hd.p, g.ldh, dH, ws.cs.p, ws.sn.p,
                         dgv, dtb, dopen, dk, dinner);
      MXG_HIP(hipGetLastError());
      // one transfer per step: which right-hand sides still extend their basis
      MXG_HIP(hipMemcpyAsync(hflags, dopen, (size_t)nr * sizeof(int), hipMemcpyDeviceToHost, g.s));
      MXG_HIP(hipStreamSynchronize(g.s));
      bool open = false;
      for (int q = 0; q < nr; q++) open = open || hflags[q] != 0;
      if (!open) break;
    }
    // least squares per right-hand side, then x_q += V_q y_q
    hipLaunchKernelGGL(k_cycle_lsq, small, lanes, 0, g.s, m, ref ? 0 : 1, nr, dH, dgv, dbeta, dk, dactive, ws.W.p, g.yd.p,
                       g.kd.p);
    MXG_HIP(hipGetLastError());
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
  MXG_HIP(hipMemcpyAsync(hflags, dinner, (size_t)nr * sizeof(int), hipMemcpyDeviceToHost, g.s));
  MXG_HIP(hipStreamSynchronize(g.s));
  for (int q = 0; q < nr; q++) {
    if (iters_out) iters_out[q] = st[q].cycles;
    if (inner_out) inner_out[q] = hflags[q];
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
