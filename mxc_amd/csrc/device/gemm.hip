// Variable-size batched GEMM with K segments on CDNA4 matrix cores.
//
// Each task owns one output block C (m x n, column major) and a list of K
// segments; the result is C = alpha * sum_s op(A_s) op(B_s) + beta * C. This
// single primitive covers every dense contraction of the ULV factorization
// (two-sided basis rotations, Schur complements whose K dimension runs over a
// row of near blocks) and of the basis construction (triangular products).
//
// Tiling: one 256-thread workgroup computes a 64 x 64 tile; the four waves
// form a 2 x 2 grid of 32 x 32 sub-tiles, each a 2 x 2 set of 16 x 16 MFMA
// accumulators (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32). K advances
// in slabs of 16 staged through LDS; op(A) and op(B) are loaded with the
// contiguous index fastest, zero-padded at the edges and masked to the upper
// triangle when requested. Tiles of all tasks share one launch: blockIdx.x is
// mapped to (task, tile) through a prefix sum of tile counts.
//
// MFMA 16x16x4 operand layout (wave64): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] of the 16x4 and 4x16 operands. Accumulator entry q
// (q = 0..3) of lane l is C[4 * (l / 16) + q][l % 16] for f32 and
// C[l / 16 + 4 * q][l % 16] for f64 (checked on gfx950 by the GEMM tests).
#include "internal.hpp"

namespace mxg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

template <class T> struct Acc;
template <> struct Acc<double> { typedef d4 type; };
template <> struct Acc<float> { typedef f4 type; };

constexpr int BM = 64, BN = 64, BK = 16, PAD = 1;

template <class T>
__global__ void __launch_bounds__(256) k_bgemm(const GemmTask<T>* __restrict__ tasks, const GemmSeg<T>* __restrict__ segs,
                                               const long long* __restrict__ prefix, int ntasks) {
  typedef typename Acc<T>::type V;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];
  const long long tile = blockIdx.x;
  const int t = find_task(prefix, ntasks, tile);
  const GemmTask<T> task = tasks[t];
  const long long local = tile - prefix[t];
  const int tiles_m = (task.m + BM - 1) / BM;
  const int m0 = (int)(local % tiles_m) * BM;
  const int n0 = (int)(local / tiles_m) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  V acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = V{0, 0, 0, 0};

  // Slabs of BK columns of op(A) / rows of op(B) over the task's segments in
  // order; the next slab's global loads are issued before the current one's
  // MFMAs, so their latency overlaps the matrix cores. The products run over
  // k in the same order as without the prefetch.
  constexpr int LA = BM * BK / 256, LB = BK * BN / 256;
  T ra[LA], rb[LB];
  int cs = task.seg0, ck = 0;  // the slab to load next
  const int send = task.seg0 + task.nseg;
  auto skip_empty = [&]() {
    while (cs < send && ck >= segs[cs].k) {
      cs++;
      ck = 0;
    }
  };
  auto load = [&]() {
    const GemmSeg<T> g = segs[cs];
#pragma unroll
    for (int u = 0; u < LA; u++) {
      const int e = u * 256 + tid;
      int i, p;
      if (!g.ta) { i = e % BM; p = e / BM; } else { p = e % BK; i = e / BK; }
      const int gi = m0 + i, gp = ck + p;
      T v = T(0);
      if (gi < task.m && gp < g.k) {
        const int sr = g.ta ? gp : gi, sc = g.ta ? gi : gp;
        if (!(g.amask && sr > sc)) v = g.A[sr + (long long)sc * g.lda];
      }
      ra[u] = v;
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
      const int e = u * 256 + tid;
      int p, j;
      if (!g.tb) { p = e % BK; j = e / BK; } else { j = e % BN; p = e / BN; }
      const int gp = ck + p, gj = n0 + j;
      T v = T(0);
      if (gp < g.k && gj < task.n) {
        const int sr = g.tb ? gj : gp, sc = g.tb ? gp : gj;
        if (!(g.bmask && sr > sc)) v = g.B[sr + (long long)sc * g.ldb];
      }
      rb[u] = v;
    }
    ck += BK;
  };
  auto store = [&](const GemmSeg<T>& g) {
#pragma unroll
    for (int u = 0; u < LA; u++) {
      const int e = u * 256 + tid;
      int i, p;
      if (!g.ta) { i = e % BM; p = e / BM; } else { p = e % BK; i = e / BK; }
      As[p][i] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
      const int e = u * 256 + tid;
      int p, j;
      if (!g.tb) { p = e % BK; j = e / BK; } else { j = e % BN; p = e / BN; }
      Bs[p][j] = rb[u];
    }
  };
  skip_empty();
  if (cs < send) {
    int ls = cs;  // segment of the slab held in ra / rb
    load();
    for (;;) {
      store(segs[ls]);
      __syncthreads();
      skip_empty();
      const bool more = cs < send;
      ls = cs;
      if (more) load();
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        const T a0 = As[kk + lk][wr * 32 + lr];
        const T a1 = As[kk + lk][wr * 32 + 16 + lr];
        const T b0 = Bs[kk + lk][wc * 32 + lr];
        const T b1 = Bs[kk + lk][wc * 32 + 16 + lr];
        acc[0][0] = mfma(a0, b0, acc[0][0]);
        acc[0][1] = mfma(a0, b1, acc[0][1]);
        acc[1][0] = mfma(a1, b0, acc[1][0]);
        acc[1][1] = mfma(a1, b1, acc[1][1]);
      }
      __syncthreads();
      if (!more) break;
    }
  }

  const T alpha = (T)task.alpha, beta = (T)task.beta;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = m0 + wr * 32 + i * 16 + (sizeof(T) == 8 ? lk + 4 * q : 4 * lk + q);
        const int col = n0 + wc * 32 + j * 16 + lr;
        if (row < task.m && col < task.n) {
          T* c = task.C + row + (long long)col * task.ldc;
          T v = alpha * acc[i][j][q];
          if (task.beta != 0.) v += beta * (*c);
          *c = v;
        }
      }
}

// The FP32 tasks with bfloat16 inputs: op(A) and op(B) are rounded to
// bfloat16 (nearest even) as they are staged, and each 16-deep slab is one
// v_mfma_f32_16x16x16_bf16 per 16 x 16 accumulator (FP32 accumulation). Lane l
// supplies A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16], e < 4; the
// accumulator layout is that of the FP32 instruction.
typedef short s4_t __attribute__((ext_vector_type(4)));
constexpr int PADH = 2;

__device__ inline unsigned short bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__global__ void __launch_bounds__(256) k_bgemm_bf16(const GemmTask<float>* __restrict__ tasks,
                                                    const GemmSeg<float>* __restrict__ segs,
                                                    const long long* __restrict__ prefix, int ntasks) {
  __shared__ unsigned short As[BK][BM + PADH];
  __shared__ unsigned short Bs[BK][BN + PADH];
  const long long tile = blockIdx.x;
  const int t = find_task(prefix, ntasks, tile);
  const GemmTask<float> task = tasks[t];
  const long long local = tile - prefix[t];
  const int tiles_m = (task.m + BM - 1) / BM;
  const int m0 = (int)(local % tiles_m) * BM;
  const int n0 = (int)(local / tiles_m) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f4{0, 0, 0, 0};

  for (int s = task.seg0; s < task.seg0 + task.nseg; s++) {
    const GemmSeg<float> g = segs[s];
    for (int k0 = 0; k0 < g.k; k0 += BK) {
#pragma unroll
      for (int e0 = 0; e0 < BM * BK; e0 += 256) {
        const int e = e0 + tid;
        int i, p;
        if (!g.ta) { i = e % BM; p = e / BM; } else { p = e % BK; i = e / BK; }
        const int gi = m0 + i, gp = k0 + p;
        float v = 0.f;
        if (gi < task.m && gp < g.k) {
          const int sr = g.ta ? gp : gi, sc = g.ta ? gi : gp;
          if (!(g.amask && sr > sc)) v = g.A[sr + (long long)sc * g.lda];
        }
        As[p][i] = bf16_bits(v);
      }
#pragma unroll
      for (int e0 = 0; e0 < BK * BN; e0 += 256) {
        const int e = e0 + tid;
        int p, j;
        if (!g.tb) { p = e % BK; j = e / BK; } else { j = e % BN; p = e / BN; }
        const int gp = k0 + p, gj = n0 + j;
        float v = 0.f;
        if (gp < g.k && gj < task.n) {
          const int sr = g.tb ? gj : gp, sc = g.tb ? gp : gj;
          if (!(g.bmask && sr > sc)) v = g.B[sr + (long long)sc * g.ldb];
        }
        Bs[p][j] = bf16_bits(v);
      }
      __syncthreads();
      s4_t a[2], b[2];
#pragma unroll
      for (int e = 0; e < 4; e++) {
#pragma unroll
        for (int i = 0; i < 2; i++) a[i][e] = (short)As[4 * lk + e][wr * 32 + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < 2; j++) b[j][e] = (short)Bs[4 * lk + e][wc * 32 + j * 16 + lr];
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
      __syncthreads();
    }
  }

  const float alpha = (float)task.alpha, beta = (float)task.beta;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = m0 + wr * 32 + i * 16 + 4 * lk + q;
        const int col = n0 + wc * 32 + j * 16 + lr;
        if (row < task.m && col < task.n) {
          float* c = task.C + row + (long long)col * task.ldc;
          float v = alpha * acc[i][j][q];
          if (task.beta != 0.) v += beta * (*c);
          *c = v;
        }
      }
}

}  // namespace

template <class T>
void gemm_batched(const std::vector<GemmTask<T>>& tasks, const std::vector<GemmSeg<T>>& segs, hipStream_t s, bool bf16,
                  Stage* st) {
  if (tasks.empty()) return;
  std::vector<long long> prefix(tasks.size() + 1, 0);
  for (size_t t = 0; t < tasks.size(); t++) {
    const GemmTask<T>& g = tasks[t];
    MXG_REQUIRE(g.seg0 >= 0 && g.nseg >= 0 && (size_t)(g.seg0 + g.nseg) <= segs.size(), "gemm task segment range");
    long long n = (g.m > 0 && g.n > 0) ? (long long)((g.m + BM - 1) / BM) * ((g.n + BN - 1) / BN) : 0;
    prefix[t + 1] = prefix[t] + n;
  }
  long long total = prefix.back();
  if (total == 0) return;
  MXG_REQUIRE(total < (1ll << 31), "gemm grid too large");
  DevBuf<GemmTask<T>> dt;
  DevBuf<GemmSeg<T>> ds;
  DevBuf<long long> dp;
  const GemmTask<T>* pt;
  const GemmSeg<T>* ps;
  const long long* pp;
  if (st) {
    pt = st->put(tasks, s);
    ps = st->put(segs, s);
    pp = st->put(prefix, s);
  } else {
    dt.upload(tasks, s);
    ds.upload(segs, s);
    dp.upload(prefix, s);
    pt = dt.p;
    ps = ds.p;
    pp = dp.p;
  }
  if constexpr (std::is_same<T, float>::value) {
    if (bf16)
      hipLaunchKernelGGL(k_bgemm_bf16, dim3((unsigned)total), dim3(256), 0, s, pt, ps, pp, (int)tasks.size());
    else
      hipLaunchKernelGGL(k_bgemm<T>, dim3((unsigned)total), dim3(256), 0, s, pt, ps, pp, (int)tasks.size());
  } else {
    MXG_REQUIRE(!bf16, "bfloat16 inputs are for FP32 products");
    hipLaunchKernelGGL(k_bgemm<T>, dim3((unsigned)total), dim3(256), 0, s, pt, ps, pp, (int)tasks.size());
  }
  MXG_HIP(hipGetLastError());
  if (!st) MXG_HIP(hipStreamSynchronize(s));  // descriptor arrays are freed on return
}

template void gemm_batched<double>(const std::vector<GemmTask<double>>&, const std::vector<GemmSeg<double>>&, hipStream_t,
                                   bool, Stage*);
template void gemm_batched<float>(const std::vector<GemmTask<float>>&, const std::vector<GemmSeg<float>>&, hipStream_t, bool,
                                  Stage*);

}  // namespace mxg

using namespace mxg;

namespace {
template <class T>
void one_gemm(mxg_ctx_t ctx, int ta, int tb, long long m, long long n, long long k, double alpha, const void* A, long long lda,
              const void* B, long long ldb, double beta, void* C, long long ldc, bool bf16 = false) {
  std::vector<GemmSeg<T>> segs{GemmSeg<T>{static_cast<const T*>(A), static_cast<const T*>(B), (int)lda, (int)ldb, (int)k,
                                          (unsigned char)(ta != 0), (unsigned char)(tb != 0), 0, 0}};
  std::vector<GemmTask<T>> tasks{GemmTask<T>{static_cast<T*>(C), (int)ldc, (int)m, (int)n, 0, k > 0 ? 1 : 0, alpha, beta}};
  gemm_batched<T>(tasks, segs, ctx->stream, bf16);
}
}  // namespace

extern "C" int mxg_gemm(mxg_ctx_t ctx, int precision, int ta, int tb, long long m, long long n, long long k, double alpha,
                        const void* d_A, long long lda, const void* d_B, long long ldb, double beta, void* d_C, long long ldc) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_REQUIRE(m >= 0 && n >= 0 && k >= 0, "negative size");
    MXG_REQUIRE(m < (1ll << 31) && n < (1ll << 31) && k < (1ll << 31) && lda < (1ll << 31) && ldb < (1ll << 31) &&
                    ldc < (1ll << 31), "size exceeds 2^31");
    MXG_REQUIRE(ldc >= std::max<long long>(m, 1), "ldc < m");
    MXG_REQUIRE(lda >= std::max<long long>(ta ? k : m, 1) && ldb >= std::max<long long>(tb ? n : k, 1), "lda/ldb too small");
    if (m == 0 || n == 0) return;
    MXG_REQUIRE(d_C && (k == 0 || (d_A && d_B)), "null pointer");
    if (precision == MXG_FP64)
      one_gemm<double>(ctx, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
    else if (precision == MXG_FP32)
      one_gemm<float>(ctx, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc);
    else if (precision == MXG_BF16)
      one_gemm<float>(ctx, ta, tb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc, true);
    else
      throw Error(MXG_ERR_ARG, "unknown precision");
  });
}
