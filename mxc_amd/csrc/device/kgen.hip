// Kernel-block generation and the matrix-free dense check.
//
// gen_matrix (kernel.cpp:71-85) evaluated on the device: one workgroup fills a
// 64 x 32 tile of a column-major block; lane r owns a row (coalesced stores
// down a column), the four waves take interleaved columns. Tasks of different
// shapes share one launch through a prefix sum of their tile counts.
#include <mutex>

#include "internal.hpp"

namespace mxg {
namespace {

template <class T>
__global__ void __launch_bounds__(256) k_gen(const GenTask* __restrict__ tasks, const long long* __restrict__ prefix,
                                             int ntasks, const double* __restrict__ bodies, int kid, double diag,
                                             double alpha) {
  long long tile = blockIdx.x;
  int t = find_task(prefix, ntasks, tile);
  GenTask g = tasks[t];
  long long local = tile - prefix[t];
  int tiles_m = (g.m + GEN_TM - 1) / GEN_TM;
  int r0 = (int)(local % tiles_m) * GEN_TM;
  int c0 = (int)(local / tiles_m) * GEN_TN;
  int r = r0 + (threadIdx.x & 63);
  int cw = threadIdx.x >> 6;
  if (r >= g.m) return;
  long long br = g.ri ? g.ri[r] : g.rbeg + r;
  double p[3] = {bodies[3 * br], bodies[3 * br + 1], bodies[3 * br + 2]};
  T* out = reinterpret_cast<T*>(g.out);
  int cend = min(c0 + GEN_TN, g.n);
  for (int c = c0 + cw; c < cend; c += 4) {
    long long bc = g.ci ? g.ci[c] : g.cbeg + c;
    double q[3] = {bodies[3 * bc], bodies[3 * bc + 1], bodies[3 * bc + 2]};
    out[r + (long long)c * g.ld] = (T)kernel_eval(kid, diag, alpha, dist2(p, q));
  }
}

// Dense check: partial sums of Y = K X over column chunks, then a
// fixed-order reduction (deterministic; no atomics). Each kernel entry is
// evaluated once and applied to the NR right-hand sides of the launch.
constexpr int DM_ROWS = 256, DM_CHUNK = 8192;

// Rows row0 + blockIdx.x * DM_ROWS + t of the launch; partial sums of chunk c
// and right-hand side q at part[(c * NR + q) * ldp + (row - row0)].
template <int NR>
__global__ void __launch_bounds__(256) k_dense_partial(const double* __restrict__ bodies, const double* __restrict__ x,
                                                       long long ldx, long long n, int kid, double diag, double alpha,
                                                       long long row0, long long rows, long long ldp,
                                                       double* __restrict__ part) {
  __shared__ double sp[256 * 3];
  __shared__ double sx[NR][256];
  const long long li = (long long)blockIdx.x * DM_ROWS + threadIdx.x;
  const long long i = row0 + li;
  const bool live = li < rows && i < n;
  long long j0 = (long long)blockIdx.y * DM_CHUNK;
  long long j1 = min(j0 + DM_CHUNK, n);
  double p[3] = {0., 0., 0.};
  if (live) { p[0] = bodies[3 * i]; p[1] = bodies[3 * i + 1]; p[2] = bodies[3 * i + 2]; }
  double acc[NR];
#pragma unroll
  for (int q = 0; q < NR; q++) acc[q] = 0.;
  for (long long jt = j0; jt < j1; jt += 256) {
    int cnt = (int)min((long long)256, j1 - jt);
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      long long j = jt + threadIdx.x;
      sp[3 * threadIdx.x] = bodies[3 * j];
      sp[3 * threadIdx.x + 1] = bodies[3 * j + 1];
      sp[3 * threadIdx.x + 2] = bodies[3 * j + 2];
#pragma unroll
      for (int q = 0; q < NR; q++) sx[q][threadIdx.x] = x[j + q * ldx];
    }
    __syncthreads();
    for (int c = 0; c < cnt; c++) {
      const double kv = kernel_eval(kid, diag, alpha, dist2(p, &sp[3 * c]));
#pragma unroll
      for (int q = 0; q < NR; q++) acc[q] += kv * sx[q][c];
    }
  }
  if (live)
#pragma unroll
    for (int q = 0; q < NR; q++) part[((long long)blockIdx.y * NR + q) * ldp + li] = acc[q];
}

// y[row0 + li] = sum over the chunks, in order, of the partial sums of the row
__global__ void k_dense_reduce(const double* __restrict__ part, long long row0, long long rows, long long ldp, int nchunks,
                               int nr, double* __restrict__ y, long long ldy) {
  const long long li = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= rows) return;
  for (int q = 0; q < nr; q++) {
    double s = 0.;
    for (int c = 0; c < nchunks; c++) s += part[((long long)c * nr + q) * ldp + li];
    y[row0 + li + q * ldy] = s;
  }
}

// Y[:, 0:nr] = K X[:, 0:nr] for rows [row0, row0 + rows), nr in {1, 2, 4, 8},
// partial sums in part (ldp >= rows per chunk and right-hand side); each row
// is summed in the same chunk order whatever the slab of rows it falls in and
// whatever nr is, so a column's result does not depend on either.
void dense_group(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n, const double* x, long long ldx,
                 double* y, long long ldy, int nr, long long row0, long long rows, double* part, long long ldp) {
  const int nchunks = (int)ceil_div(n, DM_CHUNK);
  const dim3 g(ceil_div(rows, DM_ROWS), nchunks), b(256);
  switch (nr) {
    case 8: hipLaunchKernelGGL(k_dense_partial<8>, g, b, 0, ctx->stream, d_bodies, x, ldx, n, k.id, k.diag, k.alpha, row0, rows, ldp, part); break;
    case 4: hipLaunchKernelGGL(k_dense_partial<4>, g, b, 0, ctx->stream, d_bodies, x, ldx, n, k.id, k.diag, k.alpha, row0, rows, ldp, part); break;
    case 2: hipLaunchKernelGGL(k_dense_partial<2>, g, b, 0, ctx->stream, d_bodies, x, ldx, n, k.id, k.diag, k.alpha, row0, rows, ldp, part); break;
    default: hipLaunchKernelGGL(k_dense_partial<1>, g, b, 0, ctx->stream, d_bodies, x, ldx, n, k.id, k.diag, k.alpha, row0, rows, ldp, part); break;
  }
  MXG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dense_reduce, dim3(ceil_div(rows, 256)), dim3(256), 0, ctx->stream, part, row0, rows, ldp, nchunks, nr,
                     y, ldy);
  MXG_HIP(hipGetLastError());
}

// widest group (8, 4, 2 or 1 columns) of the remaining ones
inline int group_width(long long remaining, int cap) {
  const int nr = remaining >= 8 ? 8 : remaining >= 4 ? 4 : remaining >= 2 ? 2 : 1;
  return nr < cap ? nr : cap;
}

// ---- scratch of the allocation-free dense check ----
// The partial sums of mxg_dense_matvec_batch_noalloc live in this array: one
// copy per GPU, loaded with the code object and held until the process ends.
// Host code never uses the array's name as an address: each device's address
// is looked up once, at context set-up, and kept in scratch_ptr.
constexpr long long DENSE_SCRATCH_DOUBLES = (64ll << 20) / (long long)sizeof(double);  // 64 MiB
static __device__ __align__(256) double scratch[DENSE_SCRATCH_DOUBLES];
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

double* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(scratch_mu);
  return (device >= 0 && device < MAX_DEVICES) ? scratch_ptr[device] : nullptr;
}

}  // namespace

// Looks up this device's copy of the scratch array (the device must be
// current); called by mxg_init, once per device.
void dense_scratch_setup(int device) {
  MXG_REQUIRE(device >= 0 && device < MAX_DEVICES, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(MXG_ERR_HIP, std::string("device address of the dense-check scratch: ") + hipGetErrorString(e));
  scratch_ptr[device] = static_cast<double*>(p);
}

namespace {

}  // namespace

template <class T>
void gen_tasks(const std::vector<GenTask>& tasks, const double* d_bodies, mxg_kernel_t k, hipStream_t s) {
  if (tasks.empty()) return;
  std::vector<long long> prefix(tasks.size() + 1, 0);
  for (size_t t = 0; t < tasks.size(); t++) {
    long long tm = (tasks[t].m + GEN_TM - 1) / GEN_TM, tn = (tasks[t].n + GEN_TN - 1) / GEN_TN;
    prefix[t + 1] = prefix[t] + ((tasks[t].m > 0 && tasks[t].n > 0) ? tm * tn : 0);
  }
  long long total = prefix.back();
  if (total == 0) return;
  MXG_REQUIRE(total < (1ll << 31), "generation grid too large");
  DevBuf<GenTask> dt;
  DevBuf<long long> dp;
  dt.upload(tasks, s);
  dp.upload(prefix, s);
  hipLaunchKernelGGL(k_gen<T>, dim3((unsigned)total), dim3(256), 0, s, dt.p, dp.p, (int)tasks.size(), d_bodies, k.id, k.diag,
                     k.alpha);
  MXG_HIP(hipGetLastError());
  MXG_HIP(hipStreamSynchronize(s));  // task arrays are freed on return
}

template void gen_tasks<double>(const std::vector<GenTask>&, const double*, mxg_kernel_t, hipStream_t);
template void gen_tasks<float>(const std::vector<GenTask>&, const double*, mxg_kernel_t, hipStream_t);

}  // namespace mxg

using namespace mxg;

extern "C" {

int mxg_gen_block(mxg_ctx_t ctx, mxg_kernel_t k, long long m, long long n, const double* d_bi, const double* d_bj,
                  double* d_A, long long lda) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_REQUIRE(m >= 0 && n >= 0 && lda >= (m > 0 ? m : 1), "bad block shape");
    MXG_REQUIRE(m < (1ll << 31) && n < (1ll << 31) && lda < (1ll << 31), "block too large");
    if (m == 0 || n == 0) return;
    MXG_REQUIRE(d_bi && d_bj && d_A, "null pointer");
    // Points of the two sides are separate arrays: stage them as one body array.
    DevBuf<double> pts;
    pts.alloc(3 * (size_t)(m + n));
    MXG_HIP(hipMemcpyAsync(pts.p, d_bi, 3 * m * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    MXG_HIP(hipMemcpyAsync(pts.p + 3 * m, d_bj, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    GenTask g{nullptr, nullptr, 0, m, d_A, (int)m, (int)n, (int)lda, 0};
    gen_tasks<double>({g}, pts.p, k, ctx->stream);
  });
}

int mxg_dense_matvec_batch(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n, const double* d_X,
                           long long ldx, double* d_Y, long long ldy, long long nrhs) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_REQUIRE(n >= 0 && nrhs >= 0, "negative size");
    if (n == 0 || nrhs == 0) return;
    MXG_REQUIRE(d_bodies && d_X && d_Y, "null pointer");
    MXG_REQUIRE(ldx >= n && ldy >= n, "bad leading dimensions");
    const int nchunks = (int)ceil_div(n, DM_CHUNK);
    MXG_REQUIRE(ceil_div(n, DM_ROWS) < (1u << 31) && nchunks < 65536, "dense check size too large");
    DevBuf<double> part;
    part.alloc((size_t)nchunks * n * std::min<long long>(nrhs, 8));
    for (long long c0 = 0; c0 < nrhs;) {
      const int nr = group_width(nrhs - c0, 8);
      dense_group(ctx, k, d_bodies, n, d_X + c0 * ldx, ldx, d_Y + c0 * ldy, ldy, nr, 0, n, part.p, n);
      c0 += nr;
    }
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_dense_matvec_batch_noalloc(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n,
                                   const double* d_X, long long ldx, double* d_Y, long long ldy, long long nrhs) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_REQUIRE(n >= 0 && nrhs >= 0, "negative size");
    if (n == 0 || nrhs == 0) return;
    MXG_REQUIRE(d_bodies && d_X && d_Y, "null pointer");
    MXG_REQUIRE(ldx >= n && ldy >= n, "bad leading dimensions");
    const long long nchunks = ceil_div(n, DM_CHUNK);
    // one slab of DM_ROWS rows of one column must fit: n <= 2^28
    MXG_REQUIRE(nchunks * DM_ROWS <= DENSE_SCRATCH_DOUBLES,
                "n = " + std::to_string(n) + " exceeds the allocation-free dense check's limit of " +
                    std::to_string(DENSE_SCRATCH_DOUBLES / DM_ROWS * DM_CHUNK) + " points (64 MiB of scratch)");
    double* part = scratch_of(ctx->device);
    MXG_REQUIRE(part, "the dense-check scratch of this device was not set up");
    const int cap = (int)std::min<long long>(8, DENSE_SCRATCH_DOUBLES / (nchunks * DM_ROWS));
    for (long long c0 = 0; c0 < nrhs;) {
      const int nr = group_width(nrhs - c0, cap == 8 ? 8 : cap >= 4 ? 4 : cap >= 2 ? 2 : 1);
      // rows per slab: as many whole DM_ROWS tiles as the scratch holds
      const long long slab = std::min<long long>(DENSE_SCRATCH_DOUBLES / (nchunks * nr) / DM_ROWS * DM_ROWS, n);
      for (long long row0 = 0; row0 < n; row0 += slab) {
        const long long rows = std::min(slab, n - row0);
        dense_group(ctx, k, d_bodies, n, d_X + c0 * ldx, ldx, d_Y + c0 * ldy, ldy, nr, row0, rows, part, slab);
      }
      c0 += nr;
    }
    MXG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int mxg_dense_matvec(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n, const double* d_x, double* d_y) {
  return mxg_dense_matvec_batch(ctx, k, d_bodies, n, d_x, n, d_y, n, 1);
}

}  // extern "C"
