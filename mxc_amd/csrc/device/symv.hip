// Block products of a symmetric kernel matrix, each stored block read once.
//
// The near blocks of an H^2 matrix built from a kernel on one point set come
// in transposed pairs: A_ji = A_ij^T bit for bit (K depends on the squared
// distance, and (a - b)^2 = (b - a)^2 exactly in floating point); so do the
// couplings of interpolative bases, C_ij = K(S_i, S_j). The terms
// u_(i,j) = A_ij x_j and u_(j,i) = A_ji x_i of a pair are therefore both
// computed from the block with i <= j, which one workgroup streams from HBM
// once:
//  * direct product A x_j on the vector units: lane l owns V consecutive rows
//    (one 16-byte load per column), the four waves take every fourth column
//    in order, x_j's values are broadcast from LDS, and the waves' partial
//    rows are summed in LDS in wave order at the end of a row chunk;
//  * transposed product A^T x_i on the matrix cores: every loaded column is
//    also stored into a row-major LDS panel (32 columns), and
//    out^T (16 x 16) += x_i^T (16 x 4) * A (4 x 16) runs on
//    v_mfma_{f32_16x16x4f32, f64_16x16x4f64} with the right-hand sides as the
//    16 rows (zero past NR), two waves per 16-column tile, one for each half
//    of the chunk's rows; the two halves are added in order.
// Every output is summed in an order that depends on neither the number of
// right-hand sides of the launch nor on the other right-hand sides' values.
#include "h2.hpp"

namespace mxg {
namespace {

constexpr int SP = 32;  // columns of a panel

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d4_t __attribute__((ext_vector_type(4)));
template <class T> struct MfmaAcc;
template <> struct MfmaAcc<float> { typedef f4_t type; };
template <> struct MfmaAcc<double> { typedef d4_t type; };
__device__ inline f4_t mfma4(float a, float b, f4_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline d4_t mfma4(double a, double b, d4_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// Row of the 16 x 16 accumulator held in entry q of lane l (its column is l % 16).
template <class T>
__device__ inline int acc_row(int lane, int q) {
  return sizeof(T) == 4 ? 4 * (lane >> 4) + q : (lane >> 4) + 4 * q;
}

// LDS index of (row r, column c) of the row-major panel. The XOR on the
// column keeps two access patterns free of bank conflicts: the stores of the
// direct product (a lane's rows are V apart across lanes, one column) and the
// matrix-core operand reads (16 consecutive columns of one row per 16 lanes).
template <class T>
__device__ inline int pslot(int r, int c) {
  if constexpr (sizeof(T) == 4)
    return r * SP + (c ^ (((r >> 2) ^ ((r & 1) << 4)) & 31));
  else
    return r * SP + (c ^ ((r >> 1) & 15));
}

__device__ inline int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Right-hand sides c0 .. c0 + NR - 1 of job blockIdx.x.
template <class T, int NR>
__global__ void __launch_bounds__(256) k_sym_gemv(const SymJob* __restrict__ jobs, int c0) {
  constexpr int V = 16 / (int)sizeof(T);  // rows per lane
  constexpr int RC = 64 * V;              // rows per chunk
  constexpr int U = SP / 4;               // columns per wave and panel
  typedef T VT __attribute__((ext_vector_type(V)));
  typedef typename MfmaAcc<T>::type AT;
  static_assert(4 * RC * NR <= RC * SP, "row sums must fit the panel");
  __shared__ __align__(16) T pan[RC * SP];  // the panel; the direct product's row sums at a chunk's end
  __shared__ __align__(16) T xs[RC * 16];   // x_i of the chunk, [r][16], zero past NR and past the rows
  __shared__ __align__(16) T xp[SP * NR];   // x_j of the panel, [c][k]
  __shared__ __align__(16) T tr[2 * 64 * 4];  // transposed partial sums of the lower half of the rows
  const SymJob job = jobs[blockIdx.x];
  const T* __restrict__ B = static_cast<const T*>(job.B);
  const T* __restrict__ xi = static_cast<const T*>(job.xi) + (long long)c0 * job.ldv;
  const T* __restrict__ xj = static_cast<const T*>(job.xj) + (long long)c0 * job.ldv;
  T* out1 = static_cast<T*>(job.out1) + (long long)c0 * job.ldo;
  T* out2 = job.out2 ? static_cast<T*>(job.out2) + (long long)c0 * job.ldo : nullptr;
  const int n = job.n, m = job.m, ld = job.ld;
  const long long ldv = job.ldv, ldo = job.ldo;
  const int t = threadIdx.x, lane = t & 63, wave = wave_index();
  if (n == 0) {  // A^T x_i of an empty box i is zero
    if (out2)
      for (int e = t; e < m * NR; e += 256) out2[e % m + (long long)(e / m) * ldo] = T(0);
    return;
  }
  const int rl = lane * V;  // the lane's first row in the chunk
  const int tile = wave & 1, half = wave >> 1;
  for (int r0 = 0; r0 < n; r0 += RC) {
    const int rows = min(RC, n - r0);
    __syncthreads();  // the previous chunk's reads of xs and pan are done
    if (out2 && t < RC) {
      // row t of x_i: loads coalesced across the threads, one row of 16 written at once
      T row[16];
#pragma unroll
      for (int k = 0; k < 16; k++) row[k] = (k < NR && t < rows) ? xi[r0 + t + (long long)k * ldv] : T(0);
#pragma unroll
      for (int k = 0; k < 16; k += V) {
        VT v;
#pragma unroll
        for (int q = 0; q < V; q++) v[q] = row[k + q];
        *reinterpret_cast<VT*>(xs + t * 16 + k) = v;
      }
    }
    VT acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = VT(0);
    for (int cb = 0; cb < m; cb += SP) {
      const int cn = min(SP, m - cb);
      __syncthreads();  // the previous panel's reads of pan and xp are done
      if (t < cn)  // column t of the panel: loads coalesced across the threads
#pragma unroll
        for (int k = 0; k < NR; k++) xp[t * NR + k] = xj[cb + t + (long long)k * ldv];
      const T* Bp = B + r0 + (long long)cb * ld;
      VT a[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int c = wave + 4 * u;
        a[u] = VT(0);
        if (c < cn && rl < rows) {
          const T* col = Bp + (long long)c * ld + rl;
          if (job.vec) {
            a[u] = *reinterpret_cast<const VT*>(col);
          } else {
#pragma unroll
            for (int q = 0; q < V; q++)
              if (rl + q < rows) a[u][q] = col[q];
          }
#pragma unroll
          for (int q = 0; q < V; q++)
            if (rl + q >= rows) a[u][q] = T(0);  // rows past n within a padded leading dimension
        }
      }
      __syncthreads();  // xp is staged
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int c = wave + 4 * u;
        if (c < cn) {
          const T* w = xp + c * NR;
#pragma unroll
          for (int k = 0; k < NR; k++) acc[k] += a[u] * w[k];
#pragma unroll
          for (int q = 0; q < V; q++) pan[pslot<T>(rl + q, c)] = a[u][q];
        }
      }
      if (out2) {
        __syncthreads();  // the panel is stored
        AT cacc = AT(0);
        const bool live = 16 * tile < cn;  // wave-uniform
        if (live) {
          // four independent accumulator chains (rows r, r + 4, r + 8, r + 12
          // of every 16), added in a fixed order
          AT c1 = AT(0), c2 = AT(0), c3 = AT(0);
          const int rb = half * (RC / 2);
          const int re = min(rb + RC / 2, (rows + 3) & ~3);
          const int cc = 16 * tile + (lane & 15), dr = lane >> 4;
          int r = rb;
          for (; r + 12 < re; r += 16) {
            cacc = mfma4(xs[(r + dr) * 16 + (lane & 15)], pan[pslot<T>(r + dr, cc)], cacc);
            c1 = mfma4(xs[(r + 4 + dr) * 16 + (lane & 15)], pan[pslot<T>(r + 4 + dr, cc)], c1);
            c2 = mfma4(xs[(r + 8 + dr) * 16 + (lane & 15)], pan[pslot<T>(r + 8 + dr, cc)], c2);
            c3 = mfma4(xs[(r + 12 + dr) * 16 + (lane & 15)], pan[pslot<T>(r + 12 + dr, cc)], c3);
          }
          for (; r < re; r += 4) cacc = mfma4(xs[(r + dr) * 16 + (lane & 15)], pan[pslot<T>(r + dr, cc)], cacc);
          cacc = (cacc + c1) + (c2 + c3);
          if (half == 1)
#pragma unroll
            for (int q = 0; q < 4; q++) tr[(tile * 64 + lane) * 4 + q] = cacc[q];
        }
        __syncthreads();
        if (live && half == 0) {
          const int col = 16 * tile + (lane & 15);
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int k = acc_row<T>(lane, q);
            const T v = cacc[q] + tr[(tile * 64 + lane) * 4 + q];
            if (k < NR && col < cn) {
              T* o = out2 + cb + col + (long long)k * ldo;
              *o = (r0 == 0) ? v : *o + v;
            }
          }
        }
      }
    }
    // the direct product's partial rows of the four waves, summed in wave order
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; k++)
#pragma unroll
      for (int q = 0; q < V; q++) pan[(k * 4 + wave) * RC + rl + q] = acc[k][q];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const T* rk = pan + k * 4 * RC;
      for (int r = t; r < rows; r += 256) out1[r0 + r + (long long)k * ldo] = ((rk[r] + rk[RC + r]) + rk[2 * RC + r]) + rk[3 * RC + r];
    }
  }
}

}  // namespace

template <class T>
void sym_jobs_append(std::vector<SymJob>& jobs, const std::vector<SymTask>& tasks, const T* B, long long nB, const T* x,
                     long long nx, long long ldv, T* out, long long nout, long long ldo) {
  MXG_REQUIRE(ldv >= nx && ldo >= nout, "right-hand-side strides shorter than the vectors");
  constexpr long long W = 16 / sizeof(T);
  for (const SymTask& s : tasks) {
    MXG_REQUIRE(s.n >= 0 && s.m >= 0 && s.ld >= std::max(s.n, 1) && s.b >= 0 && s.xi >= 0 && s.xj >= 0 && s.o1 >= 0,
                "bad symmetric task");
    MXG_REQUIRE(s.xi + s.n <= nx && s.xj + s.m <= nx, "symmetric task outside its input vector");
    MXG_REQUIRE(s.o1 + s.n <= nout && (s.o2 < 0 || s.o2 + s.m <= nout), "symmetric task outside its terms");
    MXG_REQUIRE(s.n == 0 || s.m == 0 || s.b + (long long)(s.m - 1) * s.ld + s.n <= nB, "symmetric task outside its matrix");
    const T* b = B + s.b;
    // 16-byte loads of whole vectors, up to n rounded up to a vector, which
    // the leading dimension must cover (padded FP32 coupling blocks do)
    const long long nw = (s.n + W - 1) / W * W;
    const int vec = ((reinterpret_cast<uintptr_t>(b) % 16) == 0 && s.ld % W == 0 && nw <= s.ld &&
                     (s.m == 0 || s.b + (long long)(s.m - 1) * s.ld + nw <= nB)) ? 1 : 0;
    jobs.push_back(SymJob{b, x + s.xi, x + s.xj, out + s.o1, s.o2 >= 0 ? out + s.o2 : nullptr, ldv, ldo, s.n, s.m, s.ld,
                          vec});
  }
}
template void sym_jobs_append<float>(std::vector<SymJob>&, const std::vector<SymTask>&, const float*, long long,
                                     const float*, long long, long long, float*, long long, long long);
template void sym_jobs_append<double>(std::vector<SymJob>&, const std::vector<SymTask>&, const double*, long long,
                                      const double*, long long, long long, double*, long long, long long);

void sym_list_upload(SymList& sl, const std::vector<SymJob>& jobs, hipStream_t s) {
  MXG_REQUIRE(jobs.size() < (1ull << 31), "symmetric job list too long");
  sl.count = (long long)jobs.size();
  if (sl.count) sl.jobs.upload(jobs, s);
}

template <class T>
void sym_gemv(const SymList& sl, int nrhs, hipStream_t s) {
  if (sl.count == 0) return;
  MXG_REQUIRE(nrhs >= 1 && nrhs <= MAX_RHS, "right-hand sides out of range");
  for (int c0 = 0; c0 < nrhs;) {
    const int nr = batch_width(nrhs - c0);
    const dim3 g((unsigned)sl.count), b(256);
    switch (nr) {
      case 8: hipLaunchKernelGGL((k_sym_gemv<T, 8>), g, b, 0, s, sl.jobs.p, c0); break;
      case 4: hipLaunchKernelGGL((k_sym_gemv<T, 4>), g, b, 0, s, sl.jobs.p, c0); break;
      case 2: hipLaunchKernelGGL((k_sym_gemv<T, 2>), g, b, 0, s, sl.jobs.p, c0); break;
      default: hipLaunchKernelGGL((k_sym_gemv<T, 1>), g, b, 0, s, sl.jobs.p, c0); break;
    }
    MXG_HIP(hipGetLastError());
    c0 += nr;
  }
}
template void sym_gemv<float>(const SymList&, int, hipStream_t);
template void sym_gemv<double>(const SymList&, int, hipStream_t);

}  // namespace mxg
