// Block-sparse matrix-vector products of the solve phase, for 1..MAX_RHS
// right-hand sides at once.
//
// Every pass of the H^2 matvec and of the ULV substitution is a sum of small
// dense products t = B v over a list of blocks (a box's basis, its coupling
// or near blocks, or sub-blocks of them), followed by a per-box sum of the
// terms that land on the same box. The terms are computed pair-parallel: one
// 256-thread workgroup per block, so a level's work spreads over as many
// workgroups as it has blocks, whatever its number of boxes. Within a block,
// lanes run down the rows (16-byte vector loads of a column when the block
// is aligned) and the four waves take every fourth column; short columns
// share a wave. Each loaded column is applied to NR right-hand sides, so a
// block is streamed from HBM once per batch. The partial rows are summed
// through LDS in a fixed order, and the per-box sums read the terms back in
// CSR order: every result is deterministic and does not depend on NR.
//
// Blocks are stored in the compute type T (double or float) or, for the
// FP32 substitution of a preconditioner, in bfloat16 (the upper 16 bits of
// the FP32 pattern, rounded to nearest even); products always accumulate in T.
#include "h2.hpp"

namespace mxg {
namespace {

template <class T, int V> struct Vec { typedef T type __attribute__((ext_vector_type(V))); };
template <> struct Vec<float, 1> { typedef float type; };
template <> struct Vec<double, 1> { typedef double type; };

// V elements of storage type S at p (16-byte aligned when V > 1) as T.
template <class T, class S, int V>
__device__ inline typename Vec<T, V>::type load_vec(const S* p) {
  if constexpr (std::is_same<T, S>::value) {
    return *reinterpret_cast<const typename Vec<T, V>::type*>(p);
  } else {
    static_assert(std::is_same<T, float>::value && std::is_same<S, bf16_t>::value, "bf16 storage needs float compute");
    typename Vec<T, V>::type out;
    if constexpr (V == 1) {
      out = __uint_as_float((unsigned)p[0] << 16);
    } else {
      typedef unsigned short U __attribute__((ext_vector_type(V)));
      const U u = *reinterpret_cast<const U*>(p);
#pragma unroll
      for (int e = 0; e < V; e++) out[e] = __uint_as_float((unsigned)u[e] << 16);
    }
    return out;
  }
}

// row groups of one pass of a workgroup (64 lanes x V rows x QN groups):
// 256 rows for V <= 2 (512 for V = 4 with one right-hand side), 512 for V = 8
template <int V, int NR>
constexpr int qn() {
  return V == 1 ? 4 : V == 2 ? 2 : (V == 4 && NR == 1) ? 2 : 1;
}
// columns in flight per wave (the same for every NR, so that a row's sum is
// formed in the same order whatever the batch width)
template <int NR>
constexpr int unroll() {
  return 4;
}

// Partial rows of B[r0 : r0 + 64 V QN, :] v_k for this wave's columns (c =
// wave, wave + 4, ...) and k < NR, summed over the waves through red, then
// handed to emit(r, k, value) by thread r (and r + 256, ...).
template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv_pass(const S* __restrict__ B, long long ld, int n, int m, int r0, VF vfun, EF emit,
                                    T* red) {
  typedef typename Vec<T, V>::type VT;
  constexpr int QN = qn<V, NR>();
  constexpr int U = unroll<NR>();
  constexpr int ROWS = 64 * V * QN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  VT acc[QN][NR];
#pragma unroll
  for (int q = 0; q < QN; q++)
#pragma unroll
    for (int k = 0; k < NR; k++) acc[q][k] = VT(0);
  const int nrows = min(ROWS, n - r0);
  int c = wave;
  for (; c + 4 * (U - 1) < m; c += 4 * U) {
    T vv[U][NR];
    VT a[U][QN];
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int k = 0; k < NR; k++) vv[u][k] = vfun(c + 4 * u, k);
      const S* col = B + (long long)(c + 4 * u) * ld + r0;
#pragma unroll
      for (int q = 0; q < QN; q++) {
        const int r = (q * 64 + lane) * V;
        a[u][q] = (r < nrows) ? load_vec<T, S, V>(col + r) : VT(0);
      }
    }
#pragma unroll
    for (int q = 0; q < QN; q++)
#pragma unroll
      for (int k = 0; k < NR; k++) {
        if constexpr (U == 4)
          acc[q][k] += (a[0][q] * vv[0][k] + a[1][q] * vv[1][k]) + (a[2][q] * vv[2][k] + a[3][q] * vv[3][k]);
        else
          acc[q][k] += a[0][q] * vv[0][k] + a[1][q] * vv[1][k];
      }
  }
  for (; c < m; c += 4) {
    const S* col0 = B + (long long)c * ld + r0;
    T v0[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) v0[k] = vfun(c, k);
#pragma unroll
    for (int q = 0; q < QN; q++) {
      const int r = (q * 64 + lane) * V;
      if (r < nrows) {
        const VT a0 = load_vec<T, S, V>(col0 + r);
#pragma unroll
        for (int k = 0; k < NR; k++) acc[q][k] += a0 * v0[k];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NR; k++) {
#pragma unroll
    for (int q = 0; q < QN; q++) {
      const int r = (q * 64 + lane) * V;
      if (r < nrows) *reinterpret_cast<VT*>(red + wave * ROWS + r) = acc[q][k];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < nrows; r += 256)
      emit(r0 + r, k, red[r] + red[ROWS + r] + red[2 * ROWS + r] + red[3 * ROWS + r]);
    __syncthreads();
  }
}

// Short columns (n <= 64 V): a column takes lpc = 2^ceil(log2(n / V)) lanes
// and each wave works on 64 / lpc columns at once, so no lane idles for want
// of rows. The 4 * 64 / lpc partial columns are summed through LDS.
template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv_narrow(const S* __restrict__ B, long long ld, int n, int m, VF vfun, EF emit, T* red) {
  typedef typename Vec<T, V>::type VT;
  constexpr int U = unroll<NR>();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nv = (n + V - 1) / V;
  int lpc = 1;
  while (lpc < nv) lpc <<= 1;
  const int G = 64 / lpc, g = lane / lpc, lr = lane % lpc;
  const int r = lr * V;
  const bool active = lr < nv;
  const int stride = 4 * G;
  VT acc[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) acc[k] = VT(0);
  int c = wave * G + g;
  for (; c + (U - 1) * stride < m; c += U * stride) {
    T vv[U][NR];
    VT a[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int k = 0; k < NR; k++) vv[u][k] = vfun(c + u * stride, k);
      a[u] = active ? load_vec<T, S, V>(B + (long long)(c + u * stride) * ld + r) : VT(0);
    }
#pragma unroll
    for (int k = 0; k < NR; k++) {
      if constexpr (U == 4)
        acc[k] += (a[0] * vv[0][k] + a[1] * vv[1][k]) + (a[2] * vv[2][k] + a[3] * vv[3][k]);
      else
        acc[k] += a[0] * vv[0][k] + a[1] * vv[1][k];
    }
  }
  for (; c < m; c += stride) {
    if (active) {
      const VT a0 = load_vec<T, S, V>(B + (long long)c * ld + r);
#pragma unroll
      for (int k = 0; k < NR; k++) acc[k] += a0 * vfun(c, k);
    }
  }
  const int W = lpc * V;  // row stride of the partial columns in LDS
  const int slots = 4 * G;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    *reinterpret_cast<VT*>(red + (wave * G + g) * W + r) = active ? acc[k] : VT(0);
    __syncthreads();
    for (int rr = threadIdx.x; rr < n; rr += 256) {
      T sum = T(0);
      for (int s = 0; s < slots; s++) sum += red[s * W + rr];
      emit(rr, k, sum);
    }
    __syncthreads();
  }
}

// LDS elements the block product needs with V rows per lane
template <int V>
constexpr int red_elems() {
  return 4 * 64 * V * qn<V, 1>();
}

template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv(const S* __restrict__ B, long long ld, int n, int m, VF vfun, EF emit, T* red) {
  if (n <= 64 * V) {
    wg_gemv_narrow<T, S, V, NR>(B, ld, n, m, vfun, emit, red);
    return;
  }
  for (int r0 = 0; r0 < n; r0 += 64 * V * qn<V, NR>()) wg_gemv_pass<T, S, V, NR>(B, ld, n, m, r0, vfun, emit, red);
}

template <class S>
constexpr int vwidth() {
  return 16 / (int)sizeof(S);
}

// Right-hand sides c0 .. c0 + NR - 1 of every job.
template <class T, class S, int NR>
__global__ void __launch_bounds__(256) k_pair_gemv(const PairJob* __restrict__ jobs, int c0) {
  constexpr int VW = vwidth<S>();
  __shared__ __align__(16) T red[red_elems<VW>()];
  const PairJob t = jobs[blockIdx.x];
  const S* B = static_cast<const S*>(t.B);
  const T* vb = static_cast<const T*>(t.v) + (long long)c0 * t.ldv;
  T* ob = static_cast<T*>(t.out) + (long long)c0 * t.ldo;
  if (t.m == 0) {
    for (int k = 0; k < NR; k++)
      for (int r = threadIdx.x; r < t.n; r += 256) ob[r + (long long)k * t.ldo] = T(0);
    return;
  }
  auto vf = [&](int c, int k) { return vb[c + (long long)k * t.ldv]; };
  auto ef = [&](int r, int k, T s) { ob[r + (long long)k * t.ldo] = s; };
  if (t.vec)
    wg_gemv<T, S, VW, NR>(B, t.ld, t.n, t.m, vf, ef, red);
  else
    wg_gemv<T, S, 1, NR>(B, t.ld, t.n, t.m, vf, ef, red);
}

// Strides (elements) between right-hand sides of the vectors of a per-box sum.
struct Strides {
  long long a, b, c, d;
};

// Downward pass of one level: w_i = w_in_i + sum_far t_p (k_i entries), then
// y_i = Q_i[:, :k] w_i (+ sum_near u_p at the leaf). Strides: a = w_in,
// b = terms, d = y.
template <class T, int V, int NR>
__global__ void __launch_bounds__(256) k_comb_down(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                   const long long* __restrict__ offQ, const long long* __restrict__ offX,
                                                   const long long* __restrict__ offZ, const long long* __restrict__ frp,
                                                   const long long* __restrict__ toffF, const long long* __restrict__ nrp,
                                                   const long long* __restrict__ toffN, const T* __restrict__ Q,
                                                   const T* __restrict__ w_in, const T* __restrict__ tF,
                                                   const T* __restrict__ tN, T* __restrict__ y, Strides ld, int kmax) {
  __shared__ __align__(16) T red[red_elems<V>()];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* w = reinterpret_cast<T*>(dsm);  // kmax x NR
  const int i = blockIdx.x;
  const int M = dims[i], k = ranks[i];
  for (int e = threadIdx.x; e < k * NR; e += 256) {
    const int r = e % k, q = e / k;
    T s = w_in[offZ[i] + r + q * ld.a];
    for (long long p = frp[i]; p < frp[i + 1]; p++) s += tF[toffF[p] + r + q * ld.b];
    w[r + q * kmax] = s;
  }
  __syncthreads();
  T* yi = y + offX[i];
  const long long n0 = tN ? nrp[i] : 0, n1 = tN ? nrp[i + 1] : 0;
  auto emit = [&](int r, int q, T s) {
    for (long long p = n0; p < n1; p++) s += tN[toffN[p] + r + q * ld.b];
    yi[r + q * ld.d] = s;
  };
  if (k == 0) {
    for (int q = 0; q < NR; q++)
      for (int r = threadIdx.x; r < M; r += 256) emit(r, q, T(0));
    return;
  }
  wg_gemv<T, T, V, NR>(Q + offQ[i], M, M, k, [&](int c, int q) { return w[c + q * kmax]; }, emit, red);
}

// Forward substitution, skeleton part: z_i = x_i[:s] - sum_near t_p.
// Strides: a = x, b = t, c = z.
template <class T>
__global__ void __launch_bounds__(256) k_comb_fwd(const int* __restrict__ ranks, const long long* __restrict__ offX,
                                                  const long long* __restrict__ offZ, const long long* __restrict__ nrp,
                                                  const long long* __restrict__ toff, const T* __restrict__ x,
                                                  const T* __restrict__ t, T* __restrict__ z, Strides ld, int nr) {
  const int i = blockIdx.x;
  const int s = ranks[i];
  for (int e = threadIdx.x; e < s * nr; e += 256) {
    const int r = e % s, q = e / s;
    T acc = x[offX[i] + r + q * ld.a];
    for (long long p = nrp[i]; p < nrp[i + 1]; p++) acc -= t[toff[p] + r + q * ld.b];
    z[offZ[i] + r + q * ld.c] = acc;
  }
}

// Backward substitution: v_i = [w_i ; x_i[s:] - sum_near t_p], y_i = U_i v_i.
// Strides: a = x, b = w, c = t, d = y.
template <class T, class S, int V, int NR>
__global__ void __launch_bounds__(256) k_comb_bwd(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                  const long long* __restrict__ offQ, const long long* __restrict__ offX,
                                                  const long long* __restrict__ offZ, const long long* __restrict__ nrp,
                                                  const long long* __restrict__ toff, const S* __restrict__ U,
                                                  const T* __restrict__ x, const T* __restrict__ w,
                                                  const T* __restrict__ t, T* __restrict__ y, Strides ld, int mmax) {
  __shared__ __align__(16) T red[red_elems<V>()];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* xv = reinterpret_cast<T*>(dsm);  // mmax x NR
  const int i = blockIdx.x;
  const int M = dims[i], s = ranks[i];
  for (int e = threadIdx.x; e < M * NR; e += 256) {
    const int r = e % M, q = e / M;
    T acc;
    if (r < s) {
      acc = w[offZ[i] + r + q * ld.b];
    } else {
      acc = x[offX[i] + r + q * ld.a];
      for (long long p = nrp[i]; p < nrp[i + 1]; p++) acc -= t[toff[p] + (r - s) + q * ld.c];
    }
    xv[r + q * mmax] = acc;
  }
  __syncthreads();
  T* yi = y + offX[i];
  wg_gemv<T, S, V, NR>(U + offQ[i], M, M, M, [&](int c, int q) { return xv[c + q * mmax]; },
                       [&](int r, int q, T v) { yi[r + q * ld.d] = v; }, red);
}

__global__ void k_to_bf16(bf16_t* __restrict__ dst, const float* __restrict__ src, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const unsigned b = __float_as_uint(src[e]);
    dst[e] = (bf16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
  }
}

// largest supported batch width not above nr
inline int batch_width(int nr) { return nr >= 8 ? 8 : nr >= 4 ? 4 : nr >= 2 ? 2 : 1; }

}  // namespace

template <class T, class S>
void pair_jobs_append(std::vector<PairJob>& jobs, const std::vector<PairTask>& tasks, const S* B, long long nB, const T* v,
                      long long nv, long long ldv, T* out, long long nout, long long ldo) {
  MXG_REQUIRE(ldv >= nv && ldo >= nout, "right-hand-side strides shorter than the vectors");
  constexpr long long W = vwidth<S>();
  for (const PairTask& t : tasks) {
    MXG_REQUIRE(t.n >= 0 && t.m >= 0 && t.ld >= std::max(t.n, 1) && t.b >= 0 && t.v >= 0 && t.out >= 0, "bad pair task");
    MXG_REQUIRE(t.out + t.n <= nout && t.v + t.m <= nv, "pair task outside its vectors");
    MXG_REQUIRE(t.n == 0 || t.m == 0 || t.b + (long long)(t.m - 1) * t.ld + t.n <= nB, "pair task outside its matrix");
    const S* b = B + t.b;
    const int vec = ((reinterpret_cast<uintptr_t>(b) % 16) == 0 && t.ld % W == 0 && t.n % W == 0) ? 1 : 0;
    jobs.push_back(PairJob{b, v + t.v, out + t.out, ldv, ldo, t.n, t.m, t.ld, vec});
  }
}
template void pair_jobs_append<float, float>(std::vector<PairJob>&, const std::vector<PairTask>&, const float*, long long,
                                             const float*, long long, long long, float*, long long, long long);
template void pair_jobs_append<double, double>(std::vector<PairJob>&, const std::vector<PairTask>&, const double*,
                                               long long, const double*, long long, long long, double*, long long,
                                               long long);
template void pair_jobs_append<float, bf16_t>(std::vector<PairJob>&, const std::vector<PairTask>&, const bf16_t*,
                                              long long, const float*, long long, long long, float*, long long, long long);

void pair_list_upload(PairList& pl, const std::vector<PairJob>& jobs, hipStream_t s) {
  MXG_REQUIRE(jobs.size() < (1ull << 31), "pair list too long");
  pl.count = (long long)jobs.size();
  if (pl.count) pl.jobs.upload(jobs, s);
}

template <class T, class S>
void pair_gemv(const PairList& pl, int nrhs, hipStream_t s) {
  if (pl.count == 0) return;
  MXG_REQUIRE(nrhs >= 1 && nrhs <= MAX_RHS, "right-hand sides out of range");
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(nrhs - c0);
    const dim3 g((unsigned)pl.count), b(256);
    switch (nr) {
      case 8: hipLaunchKernelGGL((k_pair_gemv<T, S, 8>), g, b, 0, s, pl.jobs.p, c0); break;
      case 4: hipLaunchKernelGGL((k_pair_gemv<T, S, 4>), g, b, 0, s, pl.jobs.p, c0); break;
      case 2: hipLaunchKernelGGL((k_pair_gemv<T, S, 2>), g, b, 0, s, pl.jobs.p, c0); break;
      default: hipLaunchKernelGGL((k_pair_gemv<T, S, 1>), g, b, 0, s, pl.jobs.p, c0); break;
    }
    MXG_HIP(hipGetLastError());
    c0 += nr;
  }
}
template void pair_gemv<float, float>(const PairList&, int, hipStream_t);
template void pair_gemv<double, double>(const PairList&, int, hipStream_t);
template void pair_gemv<float, bf16_t>(const PairList&, int, hipStream_t);

template <class T>
void comb_down(const Level& Lv, bool vec, const long long* d_toffF, const long long* d_toffN, const T* Q, const T* w_in,
               long long ldw, const T* tF, const T* tN, long long ldt, T* y, long long ldy, int nrhs, hipStream_t s) {
  if (Lv.nb == 0) return;
  const int kmax = (int)std::max<long long>(Lv.maxK, 1);
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(nrhs - c0);
    const size_t shm = (size_t)kmax * nr * sizeof(T);
    MXG_REQUIRE(shm <= 64 * 1024, "rank too large for the downward pass");
    const Strides ld{ldw, ldt, 0, ldy};
    const T* wi = w_in + c0 * ldw;
    const T* f = tF + c0 * ldt;
    const T* nn = tN ? tN + c0 * ldt : nullptr;
    T* yo = y + c0 * ldy;
    const dim3 g((unsigned)Lv.nb), b(256);
#define MXG_COMB_DOWN(VV, NN)                                                                                         \
  hipLaunchKernelGGL((k_comb_down<T, VV, NN>), g, b, shm, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p, Lv.d_offX.p,    \
                     Lv.d_offZ.p, Lv.d_frp.p, d_toffF, Lv.d_nrp.p, d_toffN, Q, wi, f, nn, yo, ld, kmax)
    if (vec) {
      switch (nr) {
        case 8: MXG_COMB_DOWN(vwidth<T>(), 8); break;
        case 4: MXG_COMB_DOWN(vwidth<T>(), 4); break;
        case 2: MXG_COMB_DOWN(vwidth<T>(), 2); break;
        default: MXG_COMB_DOWN(vwidth<T>(), 1); break;
      }
    } else {
      switch (nr) {
        case 8: MXG_COMB_DOWN(1, 8); break;
        case 4: MXG_COMB_DOWN(1, 4); break;
        case 2: MXG_COMB_DOWN(1, 2); break;
        default: MXG_COMB_DOWN(1, 1); break;
      }
    }
#undef MXG_COMB_DOWN
    MXG_HIP(hipGetLastError());
    c0 += nr;
  }
}
template void comb_down<float>(const Level&, bool, const long long*, const long long*, const float*, const float*,
                               long long, const float*, const float*, long long, float*, long long, int, hipStream_t);
template void comb_down<double>(const Level&, bool, const long long*, const long long*, const double*, const double*,
                                long long, const double*, const double*, long long, double*, long long, int, hipStream_t);

template <class T>
void comb_fwd(const Level& Lv, const long long* d_toff, const T* x, long long ldx, const T* t, long long ldt, T* z,
              long long ldz, int nrhs, hipStream_t s) {
  if (Lv.nb == 0 || Lv.lenZ == 0) return;
  hipLaunchKernelGGL(k_comb_fwd<T>, dim3((unsigned)Lv.nb), dim3(256), 0, s, Lv.d_ranks.p, Lv.d_offX.p, Lv.d_offZ.p,
                     Lv.d_nrp.p, d_toff, x, t, z, Strides{ldx, ldt, ldz, 0}, nrhs);
  MXG_HIP(hipGetLastError());
}
template void comb_fwd<float>(const Level&, const long long*, const float*, long long, const float*, long long, float*,
                              long long, int, hipStream_t);
template void comb_fwd<double>(const Level&, const long long*, const double*, long long, const double*, long long,
                               double*, long long, int, hipStream_t);

template <class T, class S>
void comb_bwd(const Level& Lv, bool vec, const long long* d_toff, const S* U, const T* x, long long ldx, const T* w,
              long long ldw, const T* t, long long ldt, T* y, long long ldy, int nrhs, hipStream_t s) {
  if (Lv.nb == 0) return;
  const int mmax = (int)std::max<long long>(Lv.maxM, 1);
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(nrhs - c0);
    const size_t shm = (size_t)mmax * nr * sizeof(T);
    MXG_REQUIRE(shm <= 64 * 1024, "box too large for the backward pass");
    const Strides ld{ldx, ldw, ldt, ldy};
    const T* xi = x + c0 * ldx;
    const T* wi = w + c0 * ldw;
    const T* ti = t + c0 * ldt;
    T* yo = y + c0 * ldy;
    const dim3 g((unsigned)Lv.nb), b(256);
#define MXG_COMB_BWD(VV, NN)                                                                                          \
  hipLaunchKernelGGL((k_comb_bwd<T, S, VV, NN>), g, b, shm, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p, Lv.d_offX.p,  \
                     Lv.d_offZ.p, Lv.d_nrp.p, d_toff, U, xi, wi, ti, yo, ld, mmax)
    if (vec) {
      switch (nr) {
        case 8: MXG_COMB_BWD(vwidth<S>(), 8); break;
        case 4: MXG_COMB_BWD(vwidth<S>(), 4); break;
        case 2: MXG_COMB_BWD(vwidth<S>(), 2); break;
        default: MXG_COMB_BWD(vwidth<S>(), 1); break;
      }
    } else {
      switch (nr) {
        case 8: MXG_COMB_BWD(1, 8); break;
        case 4: MXG_COMB_BWD(1, 4); break;
        case 2: MXG_COMB_BWD(1, 2); break;
        default: MXG_COMB_BWD(1, 1); break;
      }
    }
#undef MXG_COMB_BWD
    MXG_HIP(hipGetLastError());
    c0 += nr;
  }
}
template void comb_bwd<float, float>(const Level&, bool, const long long*, const float*, const float*, long long,
                                     const float*, long long, const float*, long long, float*, long long, int, hipStream_t);
template void comb_bwd<double, double>(const Level&, bool, const long long*, const double*, const double*, long long,
                                       const double*, long long, const double*, long long, double*, long long, int,
                                       hipStream_t);
template void comb_bwd<float, bf16_t>(const Level&, bool, const long long*, const bf16_t*, const float*, long long,
                                      const float*, long long, const float*, long long, float*, long long, int,
                                      hipStream_t);

void to_bf16(bf16_t* dst, const float* src, size_t n, hipStream_t s) {
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks), dim3(256), 0, s, dst, src, n);
  MXG_HIP(hipGetLastError());
}

}  // namespace mxg
