// Block-sparse matrix-vector products of the solve phase, for 1..MAX_RHS
// right-hand sides at once.
//
// Every pass of the H^2 matvec and of the ULV substitution is a sum of small
// dense products t = B v over a list of blocks (a box's basis, its coupling
// or near blocks, or sub-blocks of them), followed by a per-box sum of the
// terms that land on the same box. The terms are computed pair-parallel: one
// 256-thread workgroup per block above 64 KiB, one wave per smaller block
// (four to a workgroup), so a level's work spreads over as many workgroups
// as it has blocks, whatever its number of boxes. Within a block, lanes run
// down the rows (16-byte vector loads of a column when the block is aligned,
// 8-byte for bfloat16) and the waves take every fourth column; short columns
// share a wave. Each loaded column is applied to NR right-hand sides, so a
// block is streamed from HBM once per batch; the NR values of a column sit
// side by side in LDS and reach every lane in one broadcast read. Each row
// is a chain of fused multiply-adds down the wave's columns in order, the
// four waves' partial rows are summed through LDS in a fixed order (a lane
// butterfly within a one-wave job), and the per-box sums read the terms back
// in CSR order: every result is deterministic and does not depend on NR.
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

// V elements of storage type S as loaded from memory (16 bytes when V > 1),
// and their conversion to V elements of T: bfloat16 is the upper half of the
// FP32 pattern. Loads stay packed until their column is used, so that more of
// them fit in flight.
template <class T, class S, int V> struct Raw { typedef typename Vec<T, V>::type type; };
template <int V> struct Raw<float, bf16_t, V> { typedef unsigned short type __attribute__((ext_vector_type(V))); };
template <> struct Raw<float, bf16_t, 1> { typedef unsigned short type; };

// Block entries are read once per pass: non-temporal loads keep them from
// displacing the right-hand sides and terms that other jobs reuse from the
// caches (measured at N = 2^20: FP32 near launch 1.92 -> 1.87 ms, leaf
// substitution sweeps 0.40 -> 0.36 ms).
template <class T, class S, int V>
__device__ inline typename Raw<T, S, V>::type load_raw(const S* p) {
  typedef typename Raw<T, S, V>::type R;
  return __builtin_nontemporal_load(reinterpret_cast<const R*>(p));
}

template <class T, class S, int V>
__device__ inline typename Vec<T, V>::type cvt(const typename Raw<T, S, V>::type& u) {
  if constexpr (std::is_same<T, S>::value) {
    return u;
  } else {
    static_assert(std::is_same<T, float>::value && std::is_same<S, bf16_t>::value, "bf16 storage needs float compute");
    if constexpr (V == 1) {
      return __uint_as_float((unsigned)u << 16);
    } else {
      typename Vec<T, V>::type out;
#pragma unroll
      for (int e = 0; e < V; e++) out[e] = __uint_as_float((unsigned)u[e] << 16);
      return out;
    }
  }
}

// LDS of a block product: the staged right-hand-side values and the cross-wave
// row sums, RED_BYTES each.
constexpr int RED_BYTES = 16384;
template <class T>
constexpr int lds_elems() {
  return RED_BYTES / (int)sizeof(T);
}

// row groups of one pass of a workgroup (64 lanes x V rows x QN groups):
// 256 rows for V <= 2 (512 for V = 4 with one right-hand side), 512 for V = 8
template <int V, int NR>
constexpr int qn() {
  return V == 1 ? 4 : V == 2 ? 2 : (V == 4 && NR == 1) ? 2 : 1;
}
// columns whose loads a wave keeps in flight; the sums run down the columns in
// order whatever this is, so it may depend on NR without changing a result
template <int V, int NR>
constexpr int unroll() {
  return 4;
}

__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Stage columns [cb, cb + cn) of the NR right-hand sides as vs[c NR + k], so
// that a column's NR values are one broadcast LDS read.
template <class T, int NR, class VF>
__device__ inline void stage_v(T* vs, int cb, int cn, VF vfun) {
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; k++)
    for (int c = threadIdx.x; c < cn; c += 256) vs[c * NR + k] = vfun(cb + c, k);
  __syncthreads();
}

template <class T, int NR>
__device__ inline typename Vec<T, NR>::type lds_row(const T* p) {
  return *reinterpret_cast<const typename Vec<T, NR>::type*>(p);
}

template <class T, int NR, class VT, class WT>
__device__ inline void fma_col(VT* acc, const VT& a, const WT& w) {
  if constexpr (NR == 1) {
    acc[0] += a * w;
  } else {
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] += a * w[k];
  }
}

// Partial rows of B[r0 : r0 + 64 V QN, :] v_k for this wave's columns (c =
// wave, wave + 4, ...) and k < NR, summed over the waves through red, then
// handed to emit(r, k, value) by thread r (and r + 256, ...).
template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv_pass(const S* __restrict__ B, long long ld, int n, int m, int r0, VF vfun, EF emit,
                                    T* red, T* vs) {
  typedef typename Vec<T, V>::type VT;
  typedef typename Vec<T, NR>::type WT;
  typedef typename Raw<T, S, V>::type RT;
  constexpr int QN = qn<V, NR>();
  constexpr int U = unroll<V, NR>();
  constexpr int ROWS = 64 * V * QN;
  constexpr int CW = lds_elems<T>() / NR;
  static_assert(CW % 256 == 0, "column chunks must keep every wave's column sequence");
  const int lane = threadIdx.x & 63, wave = wave_id();
  VT acc[QN][NR];
#pragma unroll
  for (int q = 0; q < QN; q++)
#pragma unroll
    for (int k = 0; k < NR; k++) acc[q][k] = VT(0);
  const int nrows = min(ROWS, n - r0);
  for (int cb = 0; cb < m; cb += CW) {
    const int cn = min(CW, m - cb);
    stage_v<T, NR>(vs, cb, cn, vfun);
    const S* Bc = B + (long long)cb * ld + r0;
    int c = wave;
    for (; c + 4 * (U - 1) < cn; c += 4 * U) {
      RT a[U][QN];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const S* col = Bc + (long long)(c + 4 * u) * ld;
#pragma unroll
        for (int q = 0; q < QN; q++) {
          const int r = (q * 64 + lane) * V;
          a[u][q] = (r < nrows) ? load_raw<T, S, V>(col + r) : RT(0);
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const WT w = lds_row<T, NR>(vs + (c + 4 * u) * NR);
#pragma unroll
        for (int q = 0; q < QN; q++) fma_col<T, NR>(acc[q], cvt<T, S, V>(a[u][q]), w);
      }
    }
    for (; c < cn; c += 4) {
      const S* col = Bc + (long long)c * ld;
      const WT w = lds_row<T, NR>(vs + c * NR);
#pragma unroll
      for (int q = 0; q < QN; q++) {
        const int r = (q * 64 + lane) * V;
        if (r < nrows) fma_col<T, NR>(acc[q], cvt<T, S, V>(load_raw<T, S, V>(col + r)), w);
      }
    }
  }
  constexpr int KR = (lds_elems<T>() / (4 * ROWS)) < NR ? (lds_elems<T>() / (4 * ROWS)) : NR;
  static_assert(KR >= 1 && NR % KR == 0, "reduction rounds");
#pragma unroll
  for (int k0 = 0; k0 < NR; k0 += KR) {
#pragma unroll
    for (int kk = 0; kk < KR; kk++)
#pragma unroll
      for (int q = 0; q < QN; q++) {
        const int r = (q * 64 + lane) * V;
        if (r < nrows) *reinterpret_cast<VT*>(red + (kk * 4 + wave) * ROWS + r) = acc[q][k0 + kk];
      }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KR; kk++) {
      const T* rk = red + kk * 4 * ROWS;
      for (int r = threadIdx.x; r < nrows; r += 256)
        emit(r0 + r, k0 + kk, rk[r] + rk[ROWS + r] + rk[2 * ROWS + r] + rk[3 * ROWS + r]);
    }
    __syncthreads();
  }
}

// Short columns (n <= 64 V): a column takes lpc = 2^ceil(log2(n / V)) lanes
// and each wave works on 64 / lpc columns at once, so no lane idles for want
// of rows. The 4 * 64 / lpc partial columns are summed through LDS.
template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv_narrow(const S* __restrict__ B, long long ld, int n, int m, VF vfun, EF emit, T* red,
                                      T* vs) {
  typedef typename Vec<T, V>::type VT;
  typedef typename Vec<T, NR>::type WT;
  typedef typename Raw<T, S, V>::type RT;
  constexpr int U = unroll<V, NR>();
  constexpr int CW = lds_elems<T>() / NR;
  static_assert(CW % 256 == 0, "column chunks must keep every wave's column sequence");
  const int lane = threadIdx.x & 63, wave = wave_id();
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
  for (int cb = 0; cb < m; cb += CW) {
    const int cn = min(CW, m - cb);
    stage_v<T, NR>(vs, cb, cn, vfun);
    const S* Bc = B + (long long)cb * ld + r;
    int c = wave * G + g;
    for (; c + (U - 1) * stride < cn; c += U * stride) {
      RT a[U];
#pragma unroll
      for (int u = 0; u < U; u++) a[u] = active ? load_raw<T, S, V>(Bc + (long long)(c + u * stride) * ld) : RT(0);
#pragma unroll
      for (int u = 0; u < U; u++) fma_col<T, NR>(acc, cvt<T, S, V>(a[u]), lds_row<T, NR>(vs + (c + u * stride) * NR));
    }
    for (; c < cn; c += stride)
      if (active)
        fma_col<T, NR>(acc, cvt<T, S, V>(load_raw<T, S, V>(Bc + (long long)c * ld)), lds_row<T, NR>(vs + c * NR));
  }
  const int W = lpc * V;  // row stride of the partial columns in LDS
  const int slots = 4 * G;
  constexpr int PK = 256 * V;  // LDS elements of one right-hand side's partial columns
  constexpr int KR = (lds_elems<T>() / PK) < NR ? (lds_elems<T>() / PK) : NR;
  static_assert(KR >= 1 && NR % KR == 0, "reduction rounds");
#pragma unroll
  for (int k0 = 0; k0 < NR; k0 += KR) {
#pragma unroll
    for (int kk = 0; kk < KR; kk++)
      *reinterpret_cast<VT*>(red + kk * PK + (wave * G + g) * W + r) = active ? acc[k0 + kk] : VT(0);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KR; kk++) {
      const T* rk = red + kk * PK;
      for (int rr = threadIdx.x; rr < n; rr += 256) {
        T sum = T(0);
        for (int s = 0; s < slots; s++) sum += rk[s * W + rr];
        emit(rr, k0 + kk, sum);
      }
    }
    __syncthreads();
  }
}

template <class T, class S, int V, int NR, class VF, class EF>
__device__ inline void wg_gemv(const S* __restrict__ B, long long ld, int n, int m, VF vfun, EF emit, T* red, T* vs) {
  if (n <= 64 * V) {
    wg_gemv_narrow<T, S, V, NR>(B, ld, n, m, vfun, emit, red, vs);
    return;
  }
  for (int r0 = 0; r0 < n; r0 += 64 * V * qn<V, NR>()) wg_gemv_pass<T, S, V, NR>(B, ld, n, m, r0, vfun, emit, red, vs);
}

template <class S>
constexpr int vwidth() {
  return std::is_same<S, bf16_t>::value ? 4 : 16 / (int)sizeof(S);
}

// Orders this wave's LDS accesses: those before complete before those after.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A whole small block by one wave (n <= 64 V): a column takes lpc =
// 2^ceil(log2(n / V)) lanes, the wave's G = 64 / lpc lane groups take every
// G-th column in order, and the G partial columns are summed by a butterfly
// over the lane groups. No workgroup barrier: the four waves of a workgroup
// work on four independent blocks, each staging its right-hand-side values
// in its own quarter of vs.
template <class T, class S, int V, int NR>
__device__ inline void wave_gemv(const PairJob& t, int c0, T* __restrict__ vw) {
  typedef typename Vec<T, V>::type VT;
  typedef typename Vec<T, NR>::type WT;
  typedef typename Raw<T, S, V>::type RT;
  constexpr int U = unroll<V, NR>();
  constexpr int CW = lds_elems<T>() / 4 / NR;  // columns per staged chunk
  static_assert(CW % 64 == 0, "column chunks must keep every lane group's column sequence");
  const S* B = static_cast<const S*>(t.B);
  const T* vb = static_cast<const T*>(t.v) + (long long)c0 * t.ldv;
  T* ob = static_cast<T*>(t.out) + (long long)c0 * t.ldo;
  const int lane = threadIdx.x & 63;
  const int n = t.n, m = t.m;
  const int nv = (n + V - 1) / V;
  int lpc = 1;
  while (lpc < nv) lpc <<= 1;
  const int G = 64 / lpc, g = lane / lpc, r = (lane % lpc) * V;
  const bool active = r < n;
  VT acc[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) acc[k] = VT(0);
  for (int cb = 0; cb < m; cb += CW) {
    const int cn = min(CW, m - cb);
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < NR; k++)
      for (int c = lane; c < cn; c += 64) vw[c * NR + k] = vb[cb + c + (long long)k * t.ldv];
    wave_lds_sync();
    // lanes past the column's end load its first V rows; their sums are dropped
    const S* Bc = B + (long long)cb * t.ld + (active ? r : 0);
    int c = g;
    for (; c + (U - 1) * G < cn; c += U * G) {
      RT a[U];
#pragma unroll
      for (int u = 0; u < U; u++) a[u] = load_raw<T, S, V>(Bc + (long long)(c + u * G) * t.ld);
#pragma unroll
      for (int u = 0; u < U; u++) fma_col<T, NR>(acc, cvt<T, S, V>(a[u]), lds_row<T, NR>(vw + (c + u * G) * NR));
    }
    for (; c < cn; c += G)
      fma_col<T, NR>(acc, cvt<T, S, V>(load_raw<T, S, V>(Bc + (long long)c * t.ld)), lds_row<T, NR>(vw + c * NR));
  }
  for (int sh = lpc; sh < 64; sh <<= 1)
#pragma unroll
    for (int k = 0; k < NR; k++) {
      if constexpr (V == 1) {
        acc[k] += __shfl_xor(acc[k], sh);
      } else {
#pragma unroll
        for (int e = 0; e < V; e++) acc[k][e] += __shfl_xor(acc[k][e], sh);
      }
    }
  if (g == 0 && active)
#pragma unroll
    for (int k = 0; k < NR; k++) {
      T* o = ob + (long long)k * t.ldo + r;
      if constexpr (V == 1) {
        o[0] = acc[k];
      } else {
#pragma unroll
        for (int e = 0; e < V; e++) o[e] = acc[k][e];
      }
    }
}

// ---- workgroup jobs on the matrix cores (blocks stored in the compute type) ----
typedef float mf4_t __attribute__((ext_vector_type(4)));
typedef double md4_t __attribute__((ext_vector_type(4)));
template <class T> struct MAcc;
template <> struct MAcc<float> { typedef mf4_t type; };
template <> struct MAcc<double> { typedef md4_t type; };
__device__ inline mf4_t mfma16(float a, float b, mf4_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline md4_t mfma16(double a, double b, md4_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// B[0:n, 0:m] v for NR right-hand sides on v_mfma_{f32_16x16x4f32,
// f64_16x16x4f64}: out (16 rows x 16 right-hand sides) += B (16 rows x 4
// columns) * v (4 columns x 16 right-hand sides, zero past NR). Wave w owns
// rows [64 w, 64 w + 64) of each 256-row chunk; lane l loads the 16-byte
// vectors of rows 16 V h + V (l % 16) + [0, V) of column c + l / 16 (V = 4
// floats or 2 doubles, h < 4 / V), so a wave reads 64 rows of 4 columns per
// step and element e of vector h is the A operand of row tile V h + e, whose
// rows are 16 V h + V p + e, p = 0..15. The right-hand sides of the staged
// columns sit in LDS as [c][16]; one read per lane gives the B operand of all
// four tiles. No cross-wave sum: every row is one wave's chain of MFMAs over
// the columns in order, so a result does not depend on NR.
constexpr int MF_U = 4;  // column steps (4 columns each) whose loads are in flight
template <class T, int NR>
__device__ inline void wg_gemv_mfma(const T* __restrict__ B, long long ld, int n, int m, const T* __restrict__ vb,
                                    long long ldv, T* __restrict__ ob, long long ldo, T* xs) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int H = 4 / V;  // vectors per lane and column step
  constexpr int CW = lds_elems<T>() / 16;  // staged columns per chunk
  typedef T VT __attribute__((ext_vector_type(V)));
  typedef typename MAcc<T>::type AT;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int lc = lane >> 4, lr = lane & 15;
  for (int r0 = 0; r0 < n; r0 += 256) {
    const int rbase = r0 + 64 * wave;
    AT acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = AT(0);
    for (int cb = 0; cb < m; cb += CW) {
      const int cn = min(CW, m - cb);
      const int cn4 = (cn + 3) & ~3;
      __syncthreads();  // the previous chunk's reads of xs are done
      for (int c = threadIdx.x; c < cn4; c += 256) {
        T row[16];
#pragma unroll
        for (int k = 0; k < 16; k++) row[k] = (k < NR && c < cn) ? vb[cb + c + (long long)k * ldv] : T(0);
#pragma unroll
        for (int k = 0; k < 16; k += V) {
          VT v;
#pragma unroll
          for (int e = 0; e < V; e++) v[e] = row[k + e];
          *reinterpret_cast<VT*>(xs + c * 16 + k) = v;
        }
      }
      __syncthreads();
      if (rbase < n) {  // wave-uniform
        const T* Bc = B + (long long)cb * ld + rbase + V * lr;
        int c = 0;
        for (; c + 4 * (MF_U - 1) < cn4; c += 4 * MF_U) {
          VT a[MF_U][H];
#pragma unroll
          for (int u = 0; u < MF_U; u++) {
            const int col = c + 4 * u + lc;
#pragma unroll
            for (int h = 0; h < H; h++) {
              const int row = rbase + 16 * V * h + V * lr;
              a[u][h] = (col < cn && row < n) ? *reinterpret_cast<const VT*>(Bc + (long long)col * ld + 16 * V * h)
                                              : VT(0);
            }
          }
#pragma unroll
          for (int u = 0; u < MF_U; u++) {
            const T bop = xs[(c + 4 * u + lc) * 16 + lr];
#pragma unroll
            for (int h = 0; h < H; h++)
#pragma unroll
              for (int e = 0; e < V; e++) acc[V * h + e] = mfma16(a[u][h][e], bop, acc[V * h + e]);
          }
        }
        for (; c < cn4; c += 4) {
          const int col = c + lc;
          VT a[H];
#pragma unroll
          for (int h = 0; h < H; h++) {
            const int row = rbase + 16 * V * h + V * lr;
            a[h] = (col < cn && row < n) ? *reinterpret_cast<const VT*>(Bc + (long long)col * ld + 16 * V * h) : VT(0);
          }
          const T bop = xs[col * 16 + lr];
#pragma unroll
          for (int h = 0; h < H; h++)
#pragma unroll
            for (int e = 0; e < V; e++) acc[V * h + e] = mfma16(a[h][e], bop, acc[V * h + e]);
        }
      }
    }
    if (rbase < n) {
      // entry q of tile V h + e in lane l: tile row p = 4 (l / 16) + q (FP32) or
      // l / 16 + 4 q (FP64), right-hand side l % 16
#pragma unroll
      for (int h = 0; h < H; h++)
#pragma unroll
        for (int e = 0; e < V; e++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int p = sizeof(T) == 4 ? 4 * lc + q : lc + 4 * q;
            const int row = rbase + 16 * V * h + V * p + e;
            if (lr < NR && row < n) ob[row + (long long)lr * ldo] = acc[V * h + e][q];
          }
    }
  }
}

// A wave-sized job on the matrix cores: the same operand layout as
// wg_gemv_mfma, with the whole block (n <= 64 V rows) on one wave in row
// groups of 64, and blocks stored as S converted to T on load (bfloat16 is
// the upper half of the FP32 pattern, so the conversion is exact). The
// right-hand sides of up to CW columns are staged in the wave's quarter of
// LDS as [c][16].
template <class T, class S, int NR>
__device__ inline void wave_gemv_mfma(const PairJob& t, int c0, T* __restrict__ vw) {
  constexpr int V = std::is_same<S, bf16_t>::value ? 4 : 16 / (int)sizeof(T);  // elements per vector load
  constexpr int H = 4 / V;
  constexpr int G = sizeof(T) == 4 ? 4 : 2;  // row groups of 64 rows: a wave job has n <= 64 V
  constexpr int CW = lds_elems<T>() / 4 / 16;
  typedef typename Raw<T, S, V>::type RT;
  typedef typename MAcc<T>::type AT;
  static_assert(H * V == 4, "a 16-lane group covers 16 V rows; H vectors make 64");
  const S* B = static_cast<const S*>(t.B);
  const T* vb = static_cast<const T*>(t.v) + (long long)c0 * t.ldv;
  T* ob = static_cast<T*>(t.out) + (long long)c0 * t.ldo;
  const int lane = threadIdx.x & 63, lc = lane >> 4, lr = lane & 15;
  const int n = t.n, m = t.m;
  const int ng = (n + 63) / 64;  // wave-uniform
  AT acc[G][4];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[g][q] = AT(0);
  for (int cb = 0; cb < m; cb += CW) {
    const int cn = min(CW, m - cb);
    const int cn4 = (cn + 3) & ~3;
    wave_lds_sync();
    for (int c = lane; c < cn4; c += 64)
#pragma unroll
      for (int k = 0; k < 16; k++) vw[c * 16 + k] = (k < NR && c < cn) ? vb[cb + c + (long long)k * t.ldv] : T(0);
    wave_lds_sync();
#pragma unroll 4
    for (int c = 0; c < cn4; c += 4) {
      const int col = c + lc;
      const T bop = vw[col * 16 + lr];
      const S* Bc = B + (long long)(cb + col) * t.ld + V * lr;
#pragma unroll
      for (int g = 0; g < G; g++) {
        if (g < ng) {
#pragma unroll
          for (int h = 0; h < H; h++) {
            const int row = 64 * g + 16 * V * h + V * lr;
            const auto a = cvt<T, S, V>((col < cn && row < n) ? load_raw<T, S, V>(Bc + 64 * g + 16 * V * h) : RT(0));
#pragma unroll
            for (int e = 0; e < V; e++) acc[g][V * h + e] = mfma16(a[e], bop, acc[g][V * h + e]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; g++)
    if (g < ng)
#pragma unroll
      for (int h = 0; h < H; h++)
#pragma unroll
        for (int e = 0; e < V; e++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int p = sizeof(T) == 4 ? 4 * lc + q : lc + 4 * q;
            const int row = 64 * g + 16 * V * h + V * p + e;
            if (lr < NR && row < n) ob[row + (long long)lr * t.ldo] = acc[g][V * h + e][q];
          }
}

// The matrix-core jobs (mx = 1) of a list: right-hand sides c0 .. c0 + NR - 1;
// workgroups below nbig take one workgroup-sized job each, the others four
// wave-sized jobs. A kernel of its own, so that the vector-unit paths' register
// needs do not lower this one's occupancy.
template <class T, class S, int NR>
__global__ void __launch_bounds__(256) k_pair_gemv_mx(const PairJob* __restrict__ jobs, long long nbig, long long count,
                                                      int c0) {
  __shared__ __align__(16) T vs[lds_elems<T>()];
  if ((long long)blockIdx.x >= nbig) {
    const int wave = wave_id();
    const long long j = nbig + ((long long)blockIdx.x - nbig) * 4 + wave;
    if (j >= count) return;
    wave_gemv_mfma<T, S, NR>(jobs[j], c0, vs + wave * (lds_elems<T>() / 4));
    return;
  }
  if constexpr (std::is_same<T, S>::value) {
    const PairJob t = jobs[blockIdx.x];
    const T* vb = static_cast<const T*>(t.v) + (long long)c0 * t.ldv;
    T* ob = static_cast<T*>(t.out) + (long long)c0 * t.ldo;
    wg_gemv_mfma<T, NR>(static_cast<const T*>(t.B), t.ld, t.n, t.m, vb, t.ldv, ob, t.ldo, vs);
  }
}

// Right-hand sides c0 .. c0 + NR - 1 of every job. Workgroups below nbig take
// one workgroup-sized job each; the others take four wave-sized jobs.
template <class T, class S, int NR>
__global__ void __launch_bounds__(256) k_pair_gemv(const PairJob* __restrict__ jobs, long long nbig, long long count,
                                                   int c0) {
  constexpr int VW = vwidth<S>();
  __shared__ __align__(16) T red[lds_elems<T>()];
  __shared__ __align__(16) T vs[lds_elems<T>()];
  if ((long long)blockIdx.x >= nbig) {
    const int wave = wave_id();
    const long long j = nbig + ((long long)blockIdx.x - nbig) * 4 + wave;
    if (j >= count) return;
    const PairJob t = jobs[j];
    T* vw = vs + wave * (lds_elems<T>() / 4);
    if (t.vec)
      wave_gemv<T, S, VW, NR>(t, c0, vw);
    else
      wave_gemv<T, S, 1, NR>(t, c0, vw);
    return;
  }
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
    wg_gemv<T, S, VW, NR>(B, t.ld, t.n, t.m, vf, ef, red, vs);
  else
    wg_gemv<T, S, 1, NR>(B, t.ld, t.n, t.m, vf, ef, red, vs);
}

// sum over p in [p0, p1) of t[toff[p] + off]: four independent partial sums,
// so that four loads are in flight, added in a fixed order
template <class T>
__device__ inline T sum_terms(const T* __restrict__ t, const long long* __restrict__ toff, long long p0, long long p1,
                              long long off) {
  T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
  long long p = p0;
  for (; p + 3 < p1; p += 4) {
    a0 += t[toff[p] + off];
    a1 += t[toff[p + 1] + off];
    a2 += t[toff[p + 2] + off];
    a3 += t[toff[p + 3] + off];
  }
  for (; p < p1; p++) a0 += t[toff[p] + off];
  return (a0 + a1) + (a2 + a3);
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
  __shared__ __align__(16) T red[lds_elems<T>()];
  __shared__ __align__(16) T vs[lds_elems<T>()];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* w = reinterpret_cast<T*>(dsm);  // kmax x NR
  const int i = blockIdx.x;
  const int M = dims[i], k = ranks[i];
  for (int e = threadIdx.x; e < k * NR; e += 256) {
    const int r = e % k, q = e / k;
    T s = w_in[offZ[i] + r + q * ld.a];
    s += sum_terms(tF, toffF, frp[i], frp[i + 1], r + q * ld.b);
    w[r + q * kmax] = s;
  }
  __syncthreads();
  T* yi = y + offX[i];
  const long long n0 = tN ? nrp[i] : 0, n1 = tN ? nrp[i + 1] : 0;
  auto emit = [&](int r, int q, T s) {
    if (n1 > n0) s += sum_terms(tN, toffN, n0, n1, r + q * ld.b);
    yi[r + q * ld.d] = s;
  };
  if (k == 0) {
    for (int q = 0; q < NR; q++)
      for (int r = threadIdx.x; r < M; r += 256) emit(r, q, T(0));
    return;
  }
  wg_gemv<T, T, V, NR>(Q + offQ[i], M, M, k, [&](int c, int q) { return w[c + q * kmax]; }, emit, red, vs);
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
    acc -= sum_terms(t, toff, nrp[i], nrp[i + 1], r + q * ld.b);
    z[offZ[i] + r + q * ld.c] = acc;
  }
}

// Backward substitution, the vector only: v_i = [w_i ; x_i[s:] - sum_near t_p]
// into v (offsets offX), for a separate launch of the products U_i v_i; the
// same arithmetic as the assembly in k_comb_bwd. Strides: a = x, b = w,
// c = t, d = v.
template <class T>
__global__ void __launch_bounds__(256) k_bwd_gather(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                    const long long* __restrict__ offX, const long long* __restrict__ offZ,
                                                    const long long* __restrict__ nrp, const long long* __restrict__ toff,
                                                    const T* __restrict__ x, const T* __restrict__ w,
                                                    const T* __restrict__ t, T* __restrict__ v, Strides ld, int nr) {
  const int i = blockIdx.x;
  const int M = dims[i], s = ranks[i];
  for (int e = threadIdx.x; e < M * nr; e += 256) {
    const int r = e % M, q = e / M;
    T acc;
    if (r < s) {
      acc = w[offZ[i] + r + q * ld.b];
    } else {
      acc = x[offX[i] + r + q * ld.a];
      acc -= sum_terms(t, toff, nrp[i], nrp[i + 1], (r - s) + q * ld.c);
    }
    v[offX[i] + r + q * ld.d] = acc;
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
                                                  const T* __restrict__ t, T* __restrict__ y, Strides ld, int mmax,
                                                  int rows) {
  __shared__ __align__(16) T red[lds_elems<T>()];
  __shared__ __align__(16) T vs[lds_elems<T>()];
  extern __shared__ __align__(16) unsigned char dsm[];
  T* xv = reinterpret_cast<T*>(dsm);  // mmax x NR
  const int i = blockIdx.x;
  const int r0 = blockIdx.y * rows;  // this workgroup's rows of y_i: r0 .. r0 + rows - 1
  const int M = dims[i], s = ranks[i];
  if (r0 >= M) return;
  const int Mc = min(rows, M - r0);
  for (int e = threadIdx.x; e < M * NR; e += 256) {
    const int r = e % M, q = e / M;
    T acc;
    if (r < s) {
      acc = w[offZ[i] + r + q * ld.b];
    } else {
      acc = x[offX[i] + r + q * ld.a];
      acc -= sum_terms(t, toff, nrp[i], nrp[i + 1], (r - s) + q * ld.c);
    }
    xv[r + q * mmax] = acc;
  }
  __syncthreads();
  T* yi = y + offX[i] + r0;
  wg_gemv<T, S, V, NR>(U + offQ[i] + r0, M, Mc, M, [&](int c, int q) { return xv[c + q * mmax]; },
                       [&](int r, int q, T v) { yi[r + q * ld.d] = v; }, red, vs);
}

__global__ void k_to_bf16(bf16_t* __restrict__ dst, const float* __restrict__ src, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const unsigned b = __float_as_uint(src[e]);
    dst[e] = (bf16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
  }
}

// dynamic LDS of a per-box sum (its box vector for every right-hand side of
// the launch), on top of the 2 RED_BYTES of the block product: 64 KiB in all
constexpr size_t DYN_LDS_BYTES = 65536 - 2 * RED_BYTES;

// blocks up to this size (and with n <= 64 rows per vector lane) go to one
// wave; bfloat16 blocks up to twice this size, so that a 256 x 256 leaf factor
// of the substitution is one wave's job (each lane owns 4 whole rows, and no
// reduction across waves is needed)
constexpr long long WAVE_JOB_BYTES = 65536;

}  // namespace

template <class T, class S>
void pair_jobs_append(std::vector<PairJob>& jobs, const std::vector<PairTask>& tasks, const S* B, long long nB, const T* v,
                      long long nv, long long ldv, T* out, long long nout, long long ldo, bool matrix_cores) {
  MXG_REQUIRE(ldv >= nv && ldo >= nout, "right-hand-side strides shorter than the vectors");
  constexpr long long W = vwidth<S>();
  for (const PairTask& t : tasks) {
    MXG_REQUIRE(t.n >= 0 && t.m >= 0 && t.ld >= std::max(t.n, 1) && t.b >= 0 && t.v >= 0 && t.out >= 0, "bad pair task");
    MXG_REQUIRE(t.out + t.n <= nout && t.v + t.m <= nv, "pair task outside its vectors");
    MXG_REQUIRE(t.n == 0 || t.m == 0 || t.b + (long long)(t.m - 1) * t.ld + t.n <= nB, "pair task outside its matrix");
    const S* b = B + t.b;
    const int vec = ((reinterpret_cast<uintptr_t>(b) % 16) == 0 && t.ld % W == 0 && t.n % W == 0) ? 1 : 0;
    const long long wave_bytes = std::is_same<S, bf16_t>::value ? 2 * WAVE_JOB_BYTES : WAVE_JOB_BYTES;
    const int wave = (t.n <= 64 * (vec ? W : 1) && (long long)t.n * t.m * (long long)sizeof(S) <= wave_bytes) ? 1 : 0;
    // the matrix cores take the aligned blocks stored in the compute type of
    // the lists that ask for them: the matvec's near and coupling lists when
    // its tuning selects them. The upward products and the substitution ran
    // slower on them at N = 2^20 (FP64 substitution 4.32 against 3.69 ms,
    // bfloat16 leaf substitution 1.76 against 0.82 ms) and stay on the
    // vector units.
    const int mx = (matrix_cores && vec && std::is_same<T, S>::value) ? 1 : 0;
    jobs.push_back(PairJob{b, v + t.v, out + t.out, ldv, ldo, t.n, t.m, t.ld, vec, wave, mx});
  }
}
template void pair_jobs_append<float, float>(std::vector<PairJob>&, const std::vector<PairTask>&, const float*, long long,
                                             const float*, long long, long long, float*, long long, long long, bool);
template void pair_jobs_append<double, double>(std::vector<PairJob>&, const std::vector<PairTask>&, const double*,
                                               long long, const double*, long long, long long, double*, long long,
                                               long long, bool);
template void pair_jobs_append<float, bf16_t>(std::vector<PairJob>&, const std::vector<PairTask>&, const bf16_t*,
                                              long long, const float*, long long, long long, float*, long long, long long, bool);

void pair_list_upload(PairList& pl, const std::vector<PairJob>& jobs, hipStream_t s) {
  MXG_REQUIRE(jobs.size() < (1ull << 31), "pair list too long");
  std::vector<PairJob> sorted;
  sorted.reserve(jobs.size());
  for (int mx : {1, 0}) {
    for (const PairJob& j : jobs)
      if (j.mx == mx && !j.wave) sorted.push_back(j);
    if (mx) pl.nbig_mx = (long long)sorted.size();
    else pl.nbig = (long long)sorted.size() - pl.count_mx;
    for (const PairJob& j : jobs)
      if (j.mx == mx && j.wave) sorted.push_back(j);
    if (mx) pl.count_mx = (long long)sorted.size();
  }
  pl.count = (long long)sorted.size();
  if (pl.count) pl.jobs.upload(sorted, s);
}

template <class T, class S>
void pair_gemv(const PairList& pl, int nrhs, hipStream_t s) {
  if (pl.count == 0) return;
  MXG_REQUIRE(nrhs >= 1 && nrhs <= MAX_RHS, "right-hand sides out of range");
  // matrix-core jobs [0, count_mx), vector-unit jobs [count_mx, count)
  const PairJob* jm = pl.jobs.p;
  const PairJob* jv = pl.jobs.p + pl.count_mx;
  const long long nv = pl.count - pl.count_mx;
  const long long gm = pl.nbig_mx + (pl.count_mx - pl.nbig_mx + 3) / 4;
  const long long gv = pl.nbig + (nv - pl.nbig + 3) / 4;
  MXG_REQUIRE(gm < (1ll << 31) && gv < (1ll << 31), "pair list too long");
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(nrhs - c0);
    const dim3 b(256);
    if (gm > 0) {
      const dim3 g((unsigned)gm);
      switch (nr) {
        case 8: hipLaunchKernelGGL((k_pair_gemv_mx<T, S, 8>), g, b, 0, s, jm, pl.nbig_mx, pl.count_mx, c0); break;
        case 4: hipLaunchKernelGGL((k_pair_gemv_mx<T, S, 4>), g, b, 0, s, jm, pl.nbig_mx, pl.count_mx, c0); break;
        case 2: hipLaunchKernelGGL((k_pair_gemv_mx<T, S, 2>), g, b, 0, s, jm, pl.nbig_mx, pl.count_mx, c0); break;
        default: hipLaunchKernelGGL((k_pair_gemv_mx<T, S, 1>), g, b, 0, s, jm, pl.nbig_mx, pl.count_mx, c0); break;
      }
      MXG_HIP(hipGetLastError());
    }
    if (gv > 0) {
      const dim3 g((unsigned)gv);
      switch (nr) {
        case 8: hipLaunchKernelGGL((k_pair_gemv<T, S, 8>), g, b, 0, s, jv, pl.nbig, nv, c0); break;
        case 4: hipLaunchKernelGGL((k_pair_gemv<T, S, 4>), g, b, 0, s, jv, pl.nbig, nv, c0); break;
        case 2: hipLaunchKernelGGL((k_pair_gemv<T, S, 2>), g, b, 0, s, jv, pl.nbig, nv, c0); break;
        default: hipLaunchKernelGGL((k_pair_gemv<T, S, 1>), g, b, 0, s, jv, pl.nbig, nv, c0); break;
      }
      MXG_HIP(hipGetLastError());
    }
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
  const int fit = (int)(DYN_LDS_BYTES / ((size_t)kmax * sizeof(T)));
  MXG_REQUIRE(fit >= 1, "rank too large for the downward pass");
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(std::min(nrhs - c0, fit));
    const size_t shm = (size_t)kmax * nr * sizeof(T);
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

template <class T>
void bwd_gather(const Level& Lv, const long long* d_toff, const T* x, long long ldx, const T* w, long long ldw, const T* t,
                long long ldt, T* v, long long ldv, int nrhs, hipStream_t s) {
  if (Lv.nb == 0) return;
  hipLaunchKernelGGL(k_bwd_gather<T>, dim3((unsigned)Lv.nb), dim3(256), 0, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offX.p,
                     Lv.d_offZ.p, Lv.d_nrp.p, d_toff, x, w, t, v, Strides{ldx, ldw, ldt, ldv}, nrhs);
  MXG_HIP(hipGetLastError());
}
template void bwd_gather<float>(const Level&, const long long*, const float*, long long, const float*, long long,
                                const float*, long long, float*, long long, int, hipStream_t);
template void bwd_gather<double>(const Level&, const long long*, const double*, long long, const double*, long long,
                                 const double*, long long, double*, long long, int, hipStream_t);

template <class T, class S>
void comb_bwd(const Level& Lv, bool vec, const long long* d_toff, const S* U, const T* x, long long ldx, const T* w,
              long long ldw, const T* t, long long ldt, T* y, long long ldy, int nrhs, hipStream_t s) {
  if (Lv.nb == 0) return;
  const int mmax = (int)std::max<long long>(Lv.maxM, 1);
  const int fit = (int)(DYN_LDS_BYTES / ((size_t)mmax * sizeof(T)));
  MXG_REQUIRE(fit >= 1, "box too large for the backward pass");
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(std::min(nrhs - c0, fit));
    const size_t shm = (size_t)mmax * nr * sizeof(T);
    const Strides ld{ldx, ldw, ldt, ldy};
    const T* xi = x + c0 * ldx;
    const T* wi = w + c0 * ldw;
    const T* ti = t + c0 * ldt;
    T* yo = y + c0 * ldy;
    // levels with few boxes split each box's product into chunks of rows
    const int rows = (size_t)Lv.nb < SPLIT_BELOW ? SPLIT_ROWS : mmax;
    const dim3 g((unsigned)Lv.nb, (unsigned)((mmax + rows - 1) / rows)), b(256);
#define MXG_COMB_BWD(VV, NN)                                                                                          \
  hipLaunchKernelGGL((k_comb_bwd<T, S, VV, NN>), g, b, shm, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p, Lv.d_offX.p,  \
                     Lv.d_offZ.p, Lv.d_nrp.p, d_toff, U, xi, wi, ti, yo, ld, mmax, rows)
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

using namespace mxg;

namespace {
template <class T, class S>
void one_block(mxg_ctx_t ctx, long long n, long long m, const void* B, long long ldb, const void* X, long long ldx, void* Y,
               long long ldy, int nrhs, bool matrix_cores) {
  std::vector<PairJob> jobs;
  const std::vector<PairTask> t{PairTask{0, 0, 0, (int)n, (int)m, (int)ldb, 0}};
  pair_jobs_append<T, S>(jobs, t, static_cast<const S*>(B), (m - 1) * ldb + n, static_cast<const T*>(X), m, ldx,
                         static_cast<T*>(Y), n, ldy, matrix_cores);
  PairList pl;
  pair_list_upload(pl, jobs, ctx->stream);
  pair_gemv<T, S>(pl, nrhs, ctx->stream);
  MXG_HIP(hipStreamSynchronize(ctx->stream));
}
}  // namespace

extern "C" int mxg_block_gemv(mxg_ctx_t ctx, int storage, long long n, long long m, const void* d_B, long long ldb,
                              const void* d_X, long long ldx, void* d_Y, long long ldy, long long nrhs, int matrix_cores) {
  return guarded([&] {
    MXG_REQUIRE(ctx, "ctx is null");
    MXG_REQUIRE(n >= 1 && m >= 1 && n < (1 << 30) && m < (1 << 30), "block size out of range");
    MXG_REQUIRE(ldb >= n && ldx >= m && ldy >= n && ldb < (1 << 30), "bad leading dimensions");
    MXG_REQUIRE(nrhs >= 1 && nrhs <= MAX_RHS, "nrhs must be 1..8");
    MXG_REQUIRE(d_B && d_X && d_Y, "null pointer");
    MXG_HIP(hipSetDevice(ctx->device));
    if (storage == MXG_FP64)
      one_block<double, double>(ctx, n, m, d_B, ldb, d_X, ldx, d_Y, ldy, (int)nrhs, matrix_cores != 0);
    else if (storage == MXG_FP32)
      one_block<float, float>(ctx, n, m, d_B, ldb, d_X, ldx, d_Y, ldy, (int)nrhs, matrix_cores != 0);
    else if (storage == MXG_BF16)
      one_block<float, bf16_t>(ctx, n, m, d_B, ldb, d_X, ldx, d_Y, ldy, (int)nrhs, matrix_cores != 0);
    else
      throw Error(MXG_ERR_ARG, "unknown storage");
  });
}
