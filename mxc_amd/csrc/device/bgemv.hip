// Block-sparse matrix-vector products of the solve phase.
//
// Every pass of the H^2 matvec and of the ULV substitution is a sum of small
// dense products t = B v over a list of blocks (a box's basis, its coupling
// or near blocks, or sub-blocks of them), followed by a per-box sum of the
// terms that land on the same box. The terms are computed pair-parallel: one
// 256-thread workgroup per block, so a level's work spreads over as many
// workgroups as it has blocks, whatever its number of boxes. Within a block,
// lanes run down the rows (16-byte vector loads of a column when the block
// is aligned: 4 floats or 2 doubles per lane) and the four waves take every
// fourth column; the waves' partial rows are summed through LDS in a fixed
// order. The per-box sums then read the terms back in CSR order, so every
// result is deterministic.
#include "h2.hpp"

namespace mxg {
namespace {

template <class T, int V> struct Vec;
template <> struct Vec<float, 1> { typedef float type; };
template <> struct Vec<double, 1> { typedef double type; };
template <> struct Vec<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };

// rows covered by one pass of a workgroup: 64 lanes x V rows x QN groups
constexpr int QN = 4;

// Partial rows of B[r0 : r0 + 64 V QN, :] v for this wave's columns (c = wave,
// wave + 4, ...) into red[wave][...], then the sum over the waves for rows
// r < nrows is handed to emit(r, value) by thread r (and r + 256, ...).
template <class T, int V, class VF, class EF>
__device__ inline void wg_gemv_pass(const T* __restrict__ B, long long ld, int n, int m, int r0, VF vfun, EF emit,
                                    T* red) {
  typedef typename Vec<T, V>::type VT;
  constexpr int ROWS = 64 * V * QN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  VT acc[QN];
#pragma unroll
  for (int q = 0; q < QN; q++) acc[q] = VT(0);
  const int nrows = min(ROWS, n - r0);
  int c = wave;
  for (; c + 4 < m; c += 8) {
    const T v0 = vfun(c), v1 = vfun(c + 4);
    const T* col0 = B + (long long)c * ld + r0;
    const T* col1 = col0 + 4 * ld;
    VT a0[QN], a1[QN];
#pragma unroll
    for (int q = 0; q < QN; q++) {
      const int r = (q * 64 + lane) * V;
      if (r < nrows) {
        a0[q] = *reinterpret_cast<const VT*>(col0 + r);
        a1[q] = *reinterpret_cast<const VT*>(col1 + r);
      } else {
        a0[q] = VT(0);
        a1[q] = VT(0);
      }
    }
#pragma unroll
    for (int q = 0; q < QN; q++) acc[q] += a0[q] * v0 + a1[q] * v1;
  }
  for (; c < m; c += 4) {
    const T v0 = vfun(c);
    const T* col0 = B + (long long)c * ld + r0;
#pragma unroll
    for (int q = 0; q < QN; q++) {
      const int r = (q * 64 + lane) * V;
      if (r < nrows) acc[q] += *reinterpret_cast<const VT*>(col0 + r) * v0;
    }
  }
#pragma unroll
  for (int q = 0; q < QN; q++) {
    const int r = (q * 64 + lane) * V;
    if (r < nrows) *reinterpret_cast<VT*>(red + wave * ROWS + r) = acc[q];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrows; r += 256) emit(r0 + r, red[r] + red[ROWS + r] + red[2 * ROWS + r] + red[3 * ROWS + r]);
  __syncthreads();
}

template <class T, int V, class VF, class EF>
__device__ inline void wg_gemv(const T* __restrict__ B, long long ld, int n, int m, VF vfun, EF emit, T* red) {
  for (int r0 = 0; r0 < n; r0 += 64 * V * QN) wg_gemv_pass<T, V>(B, ld, n, m, r0, vfun, emit, red);
}

template <class T, int V>
__global__ void __launch_bounds__(256) k_pair_gemv(const PairTask* __restrict__ tasks, const T* __restrict__ B,
                                                   const T* __restrict__ v, T* __restrict__ out) {
  __shared__ __align__(16) T red[4 * 64 * V * QN];
  const PairTask t = tasks[blockIdx.x];
  const T* vb = v + t.v;
  T* ob = out + t.out;
  if (t.m == 0) {
    for (int r = threadIdx.x; r < t.n; r += 256) ob[r] = T(0);
    return;
  }
  wg_gemv<T, V>(B + t.b, t.ld, t.n, t.m, [&](int c) { return vb[c]; }, [&](int r, T s) { ob[r] = s; }, red);
}

// Downward pass of one level: w_i = w_in_i + sum_far t_p (k_i entries), then
// y_i = Q_i[:, :k] w_i (+ sum_near u_p at the leaf).
template <class T, int V>
__global__ void __launch_bounds__(256) k_comb_down(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                   const long long* __restrict__ offQ, const long long* __restrict__ offX,
                                                   const long long* __restrict__ offZ, const long long* __restrict__ frp,
                                                   const long long* __restrict__ toffF, const long long* __restrict__ nrp,
                                                   const long long* __restrict__ toffN, const T* __restrict__ Q,
                                                   const T* __restrict__ w_in, const T* __restrict__ tF,
                                                   const T* __restrict__ tN, T* __restrict__ y) {
  __shared__ __align__(16) T red[4 * 64 * V * QN];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* w = reinterpret_cast<T*>(dsm);
  const int i = blockIdx.x;
  const int M = dims[i], k = ranks[i];
  for (int r = threadIdx.x; r < k; r += 256) {
    T s = w_in[offZ[i] + r];
    for (long long p = frp[i]; p < frp[i + 1]; p++) s += tF[toffF[p] + r];
    w[r] = s;
  }
  __syncthreads();
  T* yi = y + offX[i];
  const long long n0 = tN ? nrp[i] : 0, n1 = tN ? nrp[i + 1] : 0;
  auto emit = [&](int r, T s) {
    for (long long p = n0; p < n1; p++) s += tN[toffN[p] + r];
    yi[r] = s;
  };
  if (k == 0) {
    for (int r = threadIdx.x; r < M; r += 256) emit(r, T(0));
    return;
  }
  wg_gemv<T, V>(Q + offQ[i], M, M, k, [&](int c) { return w[c]; }, emit, red);
}

// Forward substitution, skeleton part: z_i = x_i[:s] - sum_near t_p.
template <class T>
__global__ void __launch_bounds__(256) k_comb_fwd(const int* __restrict__ ranks, const long long* __restrict__ offX,
                                                  const long long* __restrict__ offZ, const long long* __restrict__ nrp,
                                                  const long long* __restrict__ toff, const T* __restrict__ x,
                                                  const T* __restrict__ t, T* __restrict__ z) {
  const int i = blockIdx.x;
  const int s = ranks[i];
  for (int r = threadIdx.x; r < s; r += 256) {
    T acc = x[offX[i] + r];
    for (long long p = nrp[i]; p < nrp[i + 1]; p++) acc -= t[toff[p] + r];
    z[offZ[i] + r] = acc;
  }
}

// Backward substitution: v_i = [w_i ; x_i[s:] - sum_near t_p], y_i = U_i v_i.
template <class T, int V>
__global__ void __launch_bounds__(256) k_comb_bwd(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                  const long long* __restrict__ offQ, const long long* __restrict__ offX,
                                                  const long long* __restrict__ offZ, const long long* __restrict__ nrp,
                                                  const long long* __restrict__ toff, const T* __restrict__ U,
                                                  const T* __restrict__ x, const T* __restrict__ w,
                                                  const T* __restrict__ t, T* __restrict__ y) {
  __shared__ __align__(16) T red[4 * 64 * V * QN];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* xv = reinterpret_cast<T*>(dsm);
  const int i = blockIdx.x;
  const int M = dims[i], s = ranks[i];
  for (int r = threadIdx.x; r < M; r += 256) {
    T acc;
    if (r < s) {
      acc = w[offZ[i] + r];
    } else {
      acc = x[offX[i] + r];
      for (long long p = nrp[i]; p < nrp[i + 1]; p++) acc -= t[toff[p] + (r - s)];
    }
    xv[r] = acc;
  }
  __syncthreads();
  T* yi = y + offX[i];
  wg_gemv<T, V>(U + offQ[i], M, M, M, [&](int c) { return xv[c]; }, [&](int r, T v) { yi[r] = v; }, red);
}

template <class T>
constexpr int vwidth() {
  return 16 / (int)sizeof(T);
}

}  // namespace

template <class T>
bool pair_tasks_vectorizable(const std::vector<PairTask>& tasks) {
  constexpr long long W = vwidth<T>();
  for (const PairTask& t : tasks)
    if (t.n % W || t.ld % W || t.b % W) return false;
  return true;
}
template bool pair_tasks_vectorizable<float>(const std::vector<PairTask>&);
template bool pair_tasks_vectorizable<double>(const std::vector<PairTask>&);

void pair_list_build(PairList& pl, const std::vector<PairTask>& tasks, long long nB, long long nv, long long nout,
                     hipStream_t s) {
  for (const PairTask& t : tasks) {
    MXG_REQUIRE(t.n >= 0 && t.m >= 0 && t.ld >= std::max(t.n, 1) && t.b >= 0 && t.v >= 0 && t.out >= 0, "bad pair task");
    MXG_REQUIRE(t.out + t.n <= nout && t.v + t.m <= nv, "pair task outside its vectors");
    MXG_REQUIRE(t.n == 0 || t.m == 0 || t.b + (long long)(t.m - 1) * t.ld + t.n <= nB, "pair task outside its matrix");
  }
  pl.count = (long long)tasks.size();
  pl.vec[0] = pair_tasks_vectorizable<double>(tasks);
  pl.vec[1] = pair_tasks_vectorizable<float>(tasks);
  if (pl.count) pl.tasks.upload(tasks, s);
}

template <class T>
void pair_gemv(const PairList& pl, const T* B, const T* v, T* out, hipStream_t s) {
  if (pl.count == 0) return;
  MXG_REQUIRE(pl.count < (1ll << 31), "pair list too long");
  if (pl.vec[sizeof(T) == 4 ? 1 : 0])
    hipLaunchKernelGGL((k_pair_gemv<T, vwidth<T>()>), dim3((unsigned)pl.count), dim3(256), 0, s, pl.tasks.p, B, v, out);
  else
    hipLaunchKernelGGL((k_pair_gemv<T, 1>), dim3((unsigned)pl.count), dim3(256), 0, s, pl.tasks.p, B, v, out);
  MXG_HIP(hipGetLastError());
}
template void pair_gemv<float>(const PairList&, const float*, const float*, float*, hipStream_t);
template void pair_gemv<double>(const PairList&, const double*, const double*, double*, hipStream_t);

template <class T>
void comb_down(const Level& Lv, bool vec, const long long* d_toffF, const long long* d_toffN, const T* Q, const T* w_in,
               const T* tF, const T* tN, T* y, hipStream_t s) {
  if (Lv.nb == 0) return;
  const size_t shm = std::max<long long>(Lv.maxK, 1) * sizeof(T);
  MXG_REQUIRE(shm <= 64 * 1024, "rank too large for the downward pass");
  if (vec)
    hipLaunchKernelGGL((k_comb_down<T, vwidth<T>()>), dim3((unsigned)Lv.nb), dim3(256), shm, s, Lv.d_dims.p, Lv.d_ranks.p,
                       Lv.d_offQ.p, Lv.d_offX.p, Lv.d_offZ.p, Lv.d_frp.p, d_toffF, Lv.d_nrp.p, d_toffN, Q, w_in, tF, tN, y);
  else
    hipLaunchKernelGGL((k_comb_down<T, 1>), dim3((unsigned)Lv.nb), dim3(256), shm, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p,
                       Lv.d_offX.p, Lv.d_offZ.p, Lv.d_frp.p, d_toffF, Lv.d_nrp.p, d_toffN, Q, w_in, tF, tN, y);
  MXG_HIP(hipGetLastError());
}
template void comb_down<float>(const Level&, bool, const long long*, const long long*, const float*, const float*,
                               const float*, const float*, float*, hipStream_t);
template void comb_down<double>(const Level&, bool, const long long*, const long long*, const double*, const double*,
                                const double*, const double*, double*, hipStream_t);

template <class T>
void comb_fwd(const Level& Lv, const long long* d_toff, const T* x, const T* t, T* z, hipStream_t s) {
  if (Lv.nb == 0 || Lv.lenZ == 0) return;
  hipLaunchKernelGGL(k_comb_fwd<T>, dim3((unsigned)Lv.nb), dim3(256), 0, s, Lv.d_ranks.p, Lv.d_offX.p, Lv.d_offZ.p,
                     Lv.d_nrp.p, d_toff, x, t, z);
  MXG_HIP(hipGetLastError());
}
template void comb_fwd<float>(const Level&, const long long*, const float*, const float*, float*, hipStream_t);
template void comb_fwd<double>(const Level&, const long long*, const double*, const double*, double*, hipStream_t);

template <class T>
void comb_bwd(const Level& Lv, bool vec, const long long* d_toff, const T* U, const T* x, const T* w, const T* t, T* y,
              hipStream_t s) {
  if (Lv.nb == 0) return;
  const size_t shm = std::max<long long>(Lv.maxM, 1) * sizeof(T);
  MXG_REQUIRE(shm <= 64 * 1024, "box too large for the backward pass");
  if (vec)
    hipLaunchKernelGGL((k_comb_bwd<T, vwidth<T>()>), dim3((unsigned)Lv.nb), dim3(256), shm, s, Lv.d_dims.p, Lv.d_ranks.p,
                       Lv.d_offQ.p, Lv.d_offX.p, Lv.d_offZ.p, Lv.d_nrp.p, d_toff, U, x, w, t, y);
  else
    hipLaunchKernelGGL((k_comb_bwd<T, 1>), dim3((unsigned)Lv.nb), dim3(256), shm, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p,
                       Lv.d_offX.p, Lv.d_offZ.p, Lv.d_nrp.p, d_toff, U, x, w, t, y);
  MXG_HIP(hipGetLastError());
}
template void comb_bwd<float>(const Level&, bool, const long long*, const float*, const float*, const float*,
                              const float*, float*, hipStream_t);
template void comb_bwd<double>(const Level&, bool, const long long*, const double*, const double*, const double*,
                               const double*, double*, hipStream_t);

}  // namespace mxg
