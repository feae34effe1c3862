// Row reduction of the far-field block by a Rademacher sketch.
//
// For a box with M columns (its bodies or skeleton) and nf > M far rows, the
// shared-basis construction needs an M x M matrix with the row space of the
// far block G(t, j) = K(|far_t - col_j|). The reference takes the R factor of
// a Householder QR of G (h2matrix.cpp:78-79); here it is the sketch
// Y = Omega G with Omega(i, t) = +-1, which has the same column relations in
// exact arithmetic up to the subspace embedding (Halko, Martinsson, Tropp
// 2011) and costs one streaming pass: G is never stored. Both operands are
// generated in registers/LDS: the signs from a 64-bit mix of (t, i / 64), the
// kernel entries from the body coordinates.
//
// Work decomposition: one 256-thread workgroup owns a (64 RT sketch rows) x
// 64 (columns) tile of Y and a contiguous slice of far rows; slices of one
// box are written as separate partial tiles and summed in a fixed order, so
// the result is deterministic. RT, the 16-row MFMA tiles per wave, follows
// the box size (the fewest 64-row groups that cover M, up to 6, then 8), so
// a box of up to 512 columns is one tile high and its kernel entries are
// evaluated once, and up to M = 384 the tile is less than 64 rows taller than
// the box. Per step of 16 far rows the workgroup evaluates a 16 x 64 block of
// G into LDS (4 entries per thread), computes the sign words of its rows, and
// each wave with rows in the box runs all RT x 4 f64 MFMA 16x16x4 tiles over
// its 16 RT sketch rows back to back (v_mfma_f64_16x16x4f64, 16 RT
// accumulators of 4 doubles per lane), tiles past the box included: their
// sums are not stored, and no test separates the MFMAs. Every entry of Y sums
// the far rows in the same order for every RT.
#include <map>

#include "internal.hpp"

namespace mxg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int SK_TJ = 64, SK_BK = 16;

__device__ inline unsigned long long sketch_word(unsigned long long t, unsigned long long b) {
  unsigned long long z = t * 0x9E3779B97F4A7C15ull + b * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Tiles of up to 192 rows are held to three waves per SIMD (168 registers a
// lane; RT = 3 then keeps a few values of the entry evaluation in scratch),
// where the compiler would settle for two, so that more workgroups' entry
// evaluations overlap other workgroups' MFMAs.
template <int RT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RT <= 3 ? 3 : 1)))
k_sketch(const SketchTask* __restrict__ tasks, const double* __restrict__ bodies, int kid, double diag, double alpha) {
  constexpr int WR = 16 * RT;  // sketch rows per wave
  constexpr int NG = RT;       // 64-row groups of the tile (4 WR / 64)
  __shared__ double Gs[SK_BK][SK_TJ + 1];
  __shared__ unsigned long long Ws[SK_BK][NG];
  __shared__ double Xs[SK_TJ][3];
  __shared__ double Xr[SK_BK][3];  // coordinates of the step's far rows
  const SketchTask tk = tasks[blockIdx.x];
  // the wave index as a wave-uniform value, so that the tile-extent tests
  // below are scalar branches rather than exec masks around every MFMA
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lk = lane >> 4;

  if (tid < SK_TJ) {
    const int j = tk.j0 + tid;
    double x0 = 0., x1 = 0., x2 = 0.;
    if (j < tk.M) {
      const long long c = tk.cols[j];
      x0 = bodies[3 * c];
      x1 = bodies[3 * c + 1];
      x2 = bodies[3 * c + 2];
    }
    Xs[tid][0] = x0;
    Xs[tid][1] = x1;
    Xs[tid][2] = x2;
  }
  // The far rows' coordinates reach the evaluation through LDS, gathered by
  // 3 SK_BK threads (row tid / 3, coordinate tid % 3) one step ahead, with
  // the row indices two steps ahead, so that the gathers are in flight while
  // the matrix cores run and no step waits on a dependent pair of loads.
  constexpr int NX = 3 * SK_BK;
  const int xr = tid / 3, xc = tid - 3 * (tid / 3);
  int idx_next = -1;  // row index of this thread's entry one step ahead
  if (tid < NX) {
    const int t = tk.t0 + xr;
    Xr[xr][xc] = t < tk.t1 ? bodies[3 * (long long)tk.rows[t] + xc] : 0.;
    const int t1 = t + SK_BK;
    idx_next = t1 < tk.t1 ? tk.rows[t1] : -1;
  }

  d4 acc[RT][4];
#pragma unroll
  for (int a = 0; a < RT; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = d4{0., 0., 0., 0.};

  const int irow0 = tk.i0 + wave * WR;
  const bool active = irow0 < tk.M;
  const int jj = tid & 63;
  const bool jvalid = tk.j0 + jj < tk.M;

  for (int tb = tk.t0; tb < tk.t1; tb += SK_BK) {
    __syncthreads();  // Xr holds this step's rows; the previous step's reads of Gs are done
    double xnext = 0.;
    int idx_after = -1;
    if (tid < NX) {
      if (idx_next >= 0) xnext = bodies[3 * (long long)idx_next + xc];
      const int t2 = tb + 2 * SK_BK + xr;
      idx_after = t2 < tk.t1 ? tk.rows[t2] : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tt = (tid >> 6) + 4 * q;
      const int t = tb + tt;
      double v = 0.;
      if (t < tk.t1 && jvalid) v = kernel_eval(kid, diag, alpha, dist2(Xr[tt], Xs[jj]));
      Gs[tt][jj] = v;
    }
    if (tid < NG * SK_BK) {
      const int tt = tid / NG, b = tid - NG * (tid / NG);
      Ws[tt][b] = sketch_word((unsigned long long)(tb + tt), (unsigned long long)((tk.i0 >> 6) + b));
    }
    __syncthreads();
    // every row and column tile of the wave, also those past the box (their
    // sums are not stored): the RT x 4 MFMAs of a step run back to back
    if (active) {
#pragma unroll
      for (int kk = 0; kk < SK_BK; kk += 4) {
        double g[4], a[RT];
#pragma unroll
        for (int bj = 0; bj < 4; bj++) g[bj] = Gs[kk + lk][bj * 16 + lr];
        // sketch row (within the tile) x + lr, x = wave WR + 16 bi: bit (row mod
        // 64) of word row / 64; x is a multiple of 16, so row / 64 = x / 64
#pragma unroll
        for (int bi = 0; bi < RT; bi++) {
          const int x = wave * WR + bi * 16;
          a[bi] = ((Ws[kk + lk][x >> 6] >> ((x & 63) + lr)) & 1ull) ? -1. : 1.;
        }
#pragma unroll
        for (int bi = 0; bi < RT; bi++)
#pragma unroll
          for (int bj = 0; bj < 4; bj++)
            acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], g[bj], acc[bi][bj], 0, 0, 0);
      }
    }
    // every read of Xr for this step came before the barrier above
    if (tid < NX) {
      Xr[xr][xc] = xnext;
      idx_next = idx_after;
    }
  }

  if (!active) return;
#pragma unroll
  for (int bi = 0; bi < RT; bi++)
#pragma unroll
    for (int bj = 0; bj < 4; bj++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = irow0 + bi * 16 + lk + 4 * q;
        const int col = tk.j0 + bj * 16 + lr;
        if (row < tk.M && col < tk.M) tk.out[row + (long long)col * tk.M] = acc[bi][bj][q];
      }
}

// dst[e] = sum_{s < nsplit} src[e + s * len], in order of s.
__global__ void __launch_bounds__(256) k_sum_slices(const SliceSum* __restrict__ tasks, const long long* __restrict__ prefix,
                                                    int ntasks) {
  const long long blk = blockIdx.x;
  const int t = find_task(prefix, ntasks, blk);
  const SliceSum tk = tasks[t];
  const long long e = (blk - prefix[t]) * 256 + threadIdx.x;
  if (e >= tk.len) return;
  double s = 0.;
  for (int k = 0; k < tk.nsplit; k++) s += tk.src[e + (long long)k * tk.len];
  tk.dst[e] = s;
}

// row tiles of 16 per wave for a box of M columns: one tile high up to 512,
// and every wave holds sketch rows of the box while 64 RT >= M
int sketch_row_tiles(int M) { return M <= 384 ? std::max(1, (M + 63) / 64) : 8; }

}  // namespace

void sketch_rows(mxg_ctx_s* ctx, const std::vector<SketchBox>& boxes, const double* d_bodies, mxg_kernel_t k,
                 double* d_partial, hipStream_t s) {
  std::map<int, std::vector<SketchTask>> tasks;  // by row tiles per wave
  std::vector<SliceSum> sums;
  std::vector<long long> prefix{0};
  size_t poff = 0;
  double flops = 0.;
  for (const SketchBox& b : boxes) {
    if (b.M <= 0 || b.nf <= 0) continue;
    flops += 2. * b.M * (double)b.M * (double)b.nf;
    const int nsplit = (int)((b.nf + SKETCH_ROWS - 1) / SKETCH_ROWS);
    const long long len = (long long)b.M * b.M;
    double* base = nsplit == 1 ? b.Y : d_partial + poff;
    const int rt = sketch_row_tiles(b.M), ti = 64 * rt;
    for (int sp = 0; sp < nsplit; sp++) {
      const int t0 = sp * SKETCH_ROWS, t1 = (int)std::min<long long>(b.nf, (long long)(sp + 1) * SKETCH_ROWS);
      for (int i0 = 0; i0 < b.M; i0 += ti)
        for (int j0 = 0; j0 < b.M; j0 += SK_TJ)
          tasks[rt].push_back(SketchTask{b.rows, b.cols, base + sp * len, t0, t1, b.M, i0, j0});
    }
    if (nsplit > 1) {
      sums.push_back(SliceSum{base, b.Y, len, nsplit});
      prefix.push_back(prefix.back() + (len + 255) / 256);
      poff += (size_t)nsplit * len;
    }
  }
  if (tasks.empty()) return;
  std::vector<DevBuf<SketchTask>> dts(tasks.size());
  size_t q = 0;
  for (const auto& kv : tasks) {
    MXG_REQUIRE(kv.second.size() < (1u << 31), "sketch grid too large");
    dts[q++].upload(kv.second, s);
  }
  ProfScope prof(ctx, MXG_PROF_SKETCH, s, flops);
  q = 0;
  for (const auto& kv : tasks) {
    const dim3 g((unsigned)kv.second.size()), blk(256);
    const SketchTask* tp = dts[q++].p;
    switch (kv.first) {
      case 1: hipLaunchKernelGGL(k_sketch<1>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      case 2: hipLaunchKernelGGL(k_sketch<2>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      case 3: hipLaunchKernelGGL(k_sketch<3>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      case 4: hipLaunchKernelGGL(k_sketch<4>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      case 5: hipLaunchKernelGGL(k_sketch<5>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      case 6: hipLaunchKernelGGL(k_sketch<6>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
      default: hipLaunchKernelGGL(k_sketch<8>, g, blk, 0, s, tp, d_bodies, (int)k.id, k.diag, k.alpha); break;
    }
    MXG_HIP(hipGetLastError());
  }
  DevBuf<SliceSum> ds;
  DevBuf<long long> dp;
  if (!sums.empty()) {
    ds.upload(sums, s);
    dp.upload(prefix, s);
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)prefix.back()), dim3(256), 0, s, ds.p, dp.p, (int)sums.size());
    MXG_HIP(hipGetLastError());
  }
  prof.end();
  MXG_HIP(hipStreamSynchronize(s));
}

size_t sketch_partial_size(long long M, long long nf) {
  if (M <= 0 || nf <= 0) return 0;
  const long long nsplit = (nf + SKETCH_ROWS - 1) / SKETCH_ROWS;
  return nsplit > 1 ? (size_t)nsplit * M * M : 0;
}

}  // namespace mxg
