// H^2 matrix on the device: level construction and the matvec passes.
//
// Construction of one level follows H2Matrix::construct (h2matrix.cpp:72-249)
// at a single rank:
//   leaf:  skeleton = the box's bodies, Q_in = I, near blocks generated;
//   upper: skeleton = the children's skeleton prefixes, Q_in = blockdiag of the
//          children's R, and (for a factorable matrix) each parent near block
//          receives the couplings of its far child pairs (h2matrix.cpp:204-233);
//   basis (compute_basis, h2matrix.cpp:13-63) for every box at once:
//          G = K(far sample, skeleton) -> row reduction (a Rademacher sketch
//          Omega G, or the Householder R of the reference) -> Eigen-rule CPQR ->
//          rank, T = R11^{-1} R12, skeleton permutation ->
//          orthogonal: QR of triu(Q_in) P [I; T^T] gives Q (M x M) and R;
//          interpolative: Q = [P [I; T^T], 0], R = I;
//   couplings (h2matrix.cpp:222-242): C = triu(R_y) K(S_y, S_x) triu(R_x)^T
//          (orthogonal) or K(S_y, S_x) (interpolative).
// The matvec (h2matrix.cpp:251-344) runs as lists of block products
// (bgemv.hip): the upward pass level by level, the coupling products of
// every level in one launch and, on the context's second stream beside
// them, the leaf near products in one launch; then the downward per-box
// sums level by level, of which only the leaf level waits for the near
// products.
#include "h2.hpp"

#include <algorithm>
#include <numeric>

namespace mxg {
namespace {

// ---------------- small batched kernels ----------------

template <class T>
__global__ void __launch_bounds__(256) k_copy2d(const Copy2d<T>* __restrict__ tasks, const long long* __restrict__ prefix,
                                                int ntasks) {
  const long long blk = blockIdx.x;
  const int t = find_task(prefix, ntasks, blk);
  const Copy2d<T> c = tasks[t];
  const long long e = (blk - prefix[t]) * 256 + threadIdx.x;
  if (e >= (long long)c.m * c.n) return;
  const int i = (int)(e % c.m), j = (int)(e / c.m);
  const T v = c.src[i + (long long)j * c.lds];
  if (c.trans)
    c.dst[j + (long long)i * c.ldd] = v;
  else
    c.dst[i + (long long)j * c.ldd] = v;
}

template <class Td, class Ts>
__global__ void k_convert(Td* __restrict__ dst, const Ts* __restrict__ src, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (Td)src[i];
}

__global__ void k_leaf_init(int nb, const long long* __restrict__ offX, const long long* __restrict__ offQ,
                            const int* __restrict__ dims, const long long* __restrict__ begin, int* __restrict__ S,
                            double* __restrict__ Q) {
  const int i = blockIdx.x;
  const int M = dims[i];
  for (int t = threadIdx.x; t < M; t += blockDim.x) {
    S[offX[i] + t] = (int)(begin[i] + t);
    Q[offQ[i] + t + (long long)t * M] = 1.;
  }
}

// Interpolative decomposition post-processing of one box (h2matrix.cpp:92-116).
struct PostTask {
  const double* R0;  // CPQR output, K x M (ld ldr)
  const int* perm;   // M
  double* Q;         // M x M (holds Q_in on entry)
  double* R;         // M x M
  int* S;            // M skeleton indices
  int* Stmp;         // M scratch
  double* T;         // r x (M - r) scratch
  double* W;         // M x r: P [I; T^T]
  int M, r, ldr, mode;  // mode 1: rank 0 or M; mode 2: 0 < r < M interpolative; mode 3: same, orthogonal
};

__global__ void __launch_bounds__(256) k_post(const PostTask* __restrict__ tasks) {
  const PostTask t = tasks[blockIdx.x];
  const int M = t.M, r = t.r;
  if (t.mode == 1) {
    // R = triu(Q_in), Q = I
    for (long long e = threadIdx.x; e < (long long)M * M; e += 256) {
      const int i = (int)(e % M), j = (int)(e / M);
      t.R[e] = (i <= j) ? t.Q[e] : 0.;
    }
    __syncthreads();
    for (long long e = threadIdx.x; e < (long long)M * M; e += 256) t.Q[e] = ((e % M) == (e / M)) ? 1. : 0.;
    return;
  }
  // T = R11^{-1} R12, one column per thread (back substitution)
  for (int c = threadIdx.x; c < M - r; c += 256) {
    double* x = t.T + (long long)c * r;
    const double* b = t.R0 + (long long)(r + c) * t.ldr;
    for (int i = r - 1; i >= 0; i--) {
      double s = b[i];
      for (int j = i + 1; j < r; j++) s -= t.R0[i + (long long)j * t.ldr] * x[j];
      x[i] = s / t.R0[i + (long long)i * t.ldr];
    }
  }
  for (int j = threadIdx.x; j < M; j += 256) t.Stmp[j] = t.S[t.perm[j]];
  __syncthreads();
  for (int j = threadIdx.x; j < M; j += 256) t.S[j] = t.Stmp[j];
  // W = P [I; T^T]: row perm[j] of W is row j of [I; T^T]
  for (long long e = threadIdx.x; e < (long long)M * r; e += 256) {
    const int j = (int)(e % M), c = (int)(e / M);
    const double v = (j < r) ? (j == c ? 1. : 0.) : t.T[c + (long long)(j - r) * r];
    t.W[t.perm[j] + (long long)c * M] = v;
  }
  if (t.mode == 2) {
    __syncthreads();
    for (long long e = threadIdx.x; e < (long long)M * M; e += 256) {
      const int i = (int)(e % M), j = (int)(e / M);
      t.Q[e] = (j < r) ? t.W[i + (long long)j * M] : 0.;
      t.R[e] = (i == j && j < r) ? 1. : 0.;
    }
  }
}

// QT_i = Q_i[:, :k]^T (k x M, column major), the row-major basis the upward
// pass streams with the block-sparse product kernel.
__global__ void __launch_bounds__(256) k_basis_transpose(const int* __restrict__ dims, const int* __restrict__ ranks,
                                                         const long long* __restrict__ offQ,
                                                         const long long* __restrict__ offQT, const double* __restrict__ Q,
                                                         double* __restrict__ QT) {
  const int i = blockIdx.x;
  const int M = dims[i], k = ranks[i];
  const double* Qi = Q + offQ[i];
  double* Ti = QT + offQT[i];
  for (long long e = threadIdx.x; e < (long long)M * k; e += 256) {
    const int r = (int)(e % M), c = (int)(e / M);
    Ti[c + (long long)r * k] = Qi[e];
  }
}

long long lookup(const std::vector<long long>& rp, const std::vector<long long>& col, long long i, long long j) {
  if (i < 0 || (long long)rp.size() <= i + 1) return -1;
  auto b = col.begin() + rp[i], e = col.begin() + rp[i + 1];
  auto it = std::lower_bound(b, e, j);
  return (it != e && *it == j) ? (long long)(it - col.begin()) : -1;
}

template <class T>
std::vector<T> to_vec(const long long* p, long long n) {
  std::vector<T> v(n > 0 ? n : 0);
  for (long long i = 0; i < n; i++) v[i] = (T)p[i];
  return v;
}

}  // namespace

template <class T>
void copy2d_run(const std::vector<Copy2d<T>>& all, hipStream_t s, Stage* st) {
  std::vector<Copy2d<T>> tasks;
  std::vector<long long> prefix{0};
  for (const Copy2d<T>& c : all) {
    const long long cnt = (long long)c.m * c.n;
    if (cnt <= 0) continue;
    tasks.push_back(c);
    prefix.push_back(prefix.back() + (cnt + 255) / 256);
  }
  if (tasks.empty()) return;
  MXG_REQUIRE(prefix.back() < (1ll << 31), "copy grid too large");
  DevBuf<Copy2d<T>> dt;
  DevBuf<long long> dp;
  const Copy2d<T>* pt;
  const long long* pp;
  if (st) {
    pt = st->put(tasks, s);
    pp = st->put(prefix, s);
  } else {
    dt.upload(tasks, s);
    dp.upload(prefix, s);
    pt = dt.p;
    pp = dp.p;
  }
  hipLaunchKernelGGL(k_copy2d<T>, dim3((unsigned)prefix.back()), dim3(256), 0, s, pt, pp, (int)tasks.size());
  MXG_HIP(hipGetLastError());
  if (!st) MXG_HIP(hipStreamSynchronize(s));
}

template void copy2d_run<double>(const std::vector<Copy2d<double>>&, hipStream_t, Stage*);
template void copy2d_run<float>(const std::vector<Copy2d<float>>&, hipStream_t, Stage*);
template void copy2d_run<int>(const std::vector<Copy2d<int>>&, hipStream_t, Stage*);

template <class T>
void copy2d_batched(const std::vector<const T*>& src, const std::vector<T*>& dst, const std::vector<int>& m,
                    const std::vector<int>& n, const std::vector<int>& lds, const std::vector<int>& ldd, hipStream_t s) {
  std::vector<Copy2d<T>> tasks;
  for (size_t t = 0; t < src.size(); t++) tasks.push_back(Copy2d<T>{src[t], dst[t], m[t], n[t], lds[t], ldd[t], 0, 0});
  copy2d_run<T>(tasks, s);
}

template void copy2d_batched<double>(const std::vector<const double*>&, const std::vector<double*>&, const std::vector<int>&,
                                     const std::vector<int>&, const std::vector<int>&, const std::vector<int>&, hipStream_t);
template void copy2d_batched<float>(const std::vector<const float*>&, const std::vector<float*>&, const std::vector<int>&,
                                    const std::vector<int>&, const std::vector<int>&, const std::vector<int>&, hipStream_t);
template void copy2d_batched<int>(const std::vector<const int*>&, const std::vector<int*>&, const std::vector<int>&,
                                  const std::vector<int>&, const std::vector<int>&, const std::vector<int>&, hipStream_t);

template <class Tdst, class Tsrc>
void convert(Tdst* dst, const Tsrc* src, size_t n, hipStream_t s) {
  if (n == 0) return;
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL((k_convert<Tdst, Tsrc>), dim3(blocks), dim3(256), 0, s, dst, src, n);
  MXG_HIP(hipGetLastError());
}
template void convert<float, double>(float*, const double*, size_t, hipStream_t);
template void convert<double, float>(double*, const float*, size_t, hipStream_t);
template void convert<double, double>(double*, const double*, size_t, hipStream_t);

namespace {

// Budget of doubles (8 GiB) for the far-sample matrices G and the sketch's
// slice partials of one construction chunk: the boxes of a chunk share one
// sketch launch per tile height, so larger chunks mean fewer, fuller launches.
constexpr size_t G_BUDGET = size_t(1) << 30;

// The far-body rows of a level's boxes on the device: box i's rows are
// d[off[i] .. off[i] + nf[i]).
struct FarRows {
  std::vector<long long> nf, off;
  DevBuf<int> d;
};

// From explicit index lists: far_off[nb + 1], far_idx[far_off[nb]].
void far_rows_explicit(mxg_h2_s* h, long long nb, const long long* far_off, const long long* far_idx, FarRows& fr) {
  fr.nf.assign(nb, 0);
  fr.off.assign(nb + 1, 0);
  if (!far_off) return;
  for (long long i = 0; i < nb; i++) {
    fr.nf[i] = far_off[i + 1] - far_off[i];
    MXG_REQUIRE(fr.nf[i] >= 0, "far offsets must not decrease");
    fr.off[i + 1] = far_off[i + 1] - far_off[0];
  }
  const long long nfar = fr.off[nb];
  if (nfar == 0) return;
  std::vector<int> fi(nfar);
  const long long* src = far_idx + far_off[0];
  for (long long t = 0; t < nfar; t++) {
    MXG_REQUIRE(src[t] >= 0 && src[t] < h->n, "far body index out of range");
    fi[t] = (int)src[t];
  }
  fr.d.upload(fi, h->ctx->stream);
  MXG_HIP(hipStreamSynchronize(h->ctx->stream));
}

// Box b's far rows from its runs of consecutive body indices: runs are taken
// 256 at a time, their counts scanned in LDS, and the workgroup writes the
// rows in order, each thread finding its row's run by binary search.
__global__ void __launch_bounds__(256) k_expand_runs(const int2* __restrict__ runs, const long long* __restrict__ run_off,
                                                     const long long* __restrict__ row_off, int* __restrict__ rows) {
  __shared__ long long scan[256];
  __shared__ int first[256];
  const long long b = blockIdx.x;
  const long long r0 = run_off[b], r1 = run_off[b + 1];
  const int tid = threadIdx.x;
  long long base = row_off[b];
  for (long long rb = r0; rb < r1; rb += 256) {
    const long long r = rb + tid;
    const int2 run = r < r1 ? runs[r] : make_int2(0, 0);
    first[tid] = run.x;
    scan[tid] = run.y;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const long long v = tid >= o ? scan[tid - o] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const long long tot = scan[255];
    for (long long e = tid; e < tot; e += 256) {
      int lo = 0, hi = 255;  // first run whose inclusive count exceeds e
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (scan[mid] > e) hi = mid;
        else lo = mid + 1;
      }
      rows[base + e] = first[lo] + (int)(e - (lo ? scan[lo - 1] : 0));
    }
    base += tot;
    __syncthreads();  // the batch's scan and firsts are read before the next batch overwrites them
  }
}

// From runs: box i's rows are the runs run_off[i] .. run_off[i + 1] - 1 in
// order, run k covering bodies runs[2k] .. runs[2k] + runs[2k + 1] - 1.
void far_rows_runs(mxg_h2_s* h, long long nb, const long long* run_off, const long long* runs, FarRows& fr) {
  hipStream_t s = h->ctx->stream;
  fr.nf.assign(nb, 0);
  fr.off.assign(nb + 1, 0);
  if (!run_off) return;
  const long long r_base = run_off[0], nruns = run_off[nb] - r_base;
  std::vector<int2> packed(std::max<long long>(nruns, 1));
  std::vector<long long> roff(nb + 1);
  for (long long i = 0; i < nb; i++) {
    MXG_REQUIRE(run_off[i + 1] >= run_off[i], "run offsets must not decrease");
    long long rows = 0;
    for (long long k = run_off[i]; k < run_off[i + 1]; k++) {
      const long long f = runs[2 * k], c = runs[2 * k + 1];
      MXG_REQUIRE(c >= 0 && f >= 0 && f + c <= h->n, "far run out of range");
      packed[k - r_base] = make_int2((int)f, (int)c);
      rows += c;
    }
    fr.nf[i] = rows;
    fr.off[i + 1] = fr.off[i] + rows;
    roff[i] = run_off[i] - r_base;
  }
  roff[nb] = nruns;
  const long long nfar = fr.off[nb];
  if (nfar == 0) return;
  MXG_REQUIRE(nfar < (1ll << 40), "far lists too long");
  DevBuf<int2> d_runs;
  DevBuf<long long> d_roff, d_off;
  d_runs.upload(packed, s);
  d_roff.upload(roff, s);
  d_off.upload(fr.off, s);
  fr.d.alloc(nfar);
  hipLaunchKernelGGL(k_expand_runs, dim3((unsigned)nb), dim3(256), 0, s, d_runs.p, d_roff.p, d_off.p, fr.d.p);
  MXG_HIP(hipGetLastError());
  MXG_HIP(hipStreamSynchronize(s));
}

void construct_basis(mxg_h2_s* h, Level& Lv, long long l, const FarRows& fr, double tol, long long maxrank, int orth) {
  hipStream_t s = h->ctx->stream;
  const long long nb = Lv.nb;
  std::vector<long long> nf(nb, 0);
  if (l > 0)
    for (long long i = 0; i < nb; i++) nf[i] = fr.nf[i];
  const int* d_far = fr.d.p;
  const std::vector<long long>& far_off = fr.off;
  DevBuf<int> d_perm, d_rank, d_stmp;
  d_perm.alloc(std::max<long long>(Lv.lenX, 1));
  d_stmp.alloc(std::max<long long>(Lv.lenX, 1));
  d_rank.alloc_zero(nb, s);
  std::fill(Lv.ranks.begin(), Lv.ranks.end(), 0);

  // far rows are reduced by the Rademacher sketch unless the Householder R
  // factor of the reference is requested
  const bool sketch = !(h->flags & MXG_H2_HOUSEHOLDER_ROWS);
  std::vector<long long> ldg(nb, 0);
  for (long long i = 0; i < nb; i++) {
    const long long M = Lv.dims[i];
    ldg[i] = (sketch && nf[i] > M) ? M : nf[i];
  }

  long long b0 = 0;
  while (b0 < nb) {
    // chunk [b0, b1) of boxes whose row-reduction buffers fit the budget
    long long b1 = b0;
    size_t gsz = 0, psz = 0;
    std::vector<size_t> goff;
    while (b1 < nb) {
      const long long M = Lv.dims[b1];
      const long long K = std::min(M, nf[b1]);
      const size_t need = K > 0 ? (size_t)ldg[b1] * M : 0;
      const size_t pneed = (K > 0 && sketch && nf[b1] > M) ? sketch_partial_size(M, nf[b1]) : 0;
      if (b1 > b0 && gsz + need + psz + pneed > G_BUDGET) break;
      goff.push_back(gsz);
      gsz += need;
      psz += pneed;
      b1++;
    }
    DevBuf<double> G, tau, work, part;
    G.alloc(std::max<size_t>(gsz, 1));
    part.alloc(std::max<size_t>(psz, 1));
    size_t tsz = 0, wsz = 0;
    for (long long i = b0; i < b1; i++) {
      tsz += (size_t)std::min(Lv.dims[i], nf[i]);
      wsz += 3 * (size_t)Lv.dims[i];
    }
    tau.alloc(std::max<size_t>(tsz, 1));
    work.alloc(std::max<size_t>(wsz, 1));

    std::vector<GenTask> gens;
    std::vector<SketchBox> sks;
    std::vector<QrTask> qrs;
    std::vector<CpqrTask> cps;
    size_t toff = 0, woff = 0;
    for (long long i = b0; i < b1; i++) {
      const long long M = Lv.dims[i], K = std::min(M, nf[i]);
      if (K <= 0) continue;
      double* Gi = G.p + goff[i - b0];
      if (sketch && nf[i] > M) {
        sks.push_back(SketchBox{d_far + far_off[i], Lv.S.p + Lv.offX[i], Gi, nf[i], (int)M, 0});
      } else {
        gens.push_back(GenTask{d_far + far_off[i], Lv.S.p + Lv.offX[i], 0, 0, Gi, (int)nf[i], (int)M, (int)nf[i], 0});
        if (K < nf[i]) qrs.push_back(QrTask{Gi, tau.p + toff, (int)nf[i], (int)M, (int)nf[i], 1});
      }
      toff += K;
      long long steps = maxrank > 0 ? std::min(K, maxrank) : K;
      cps.push_back(CpqrTask{Gi, d_perm.p + Lv.offX[i], d_rank.p + i, work.p + woff, (int)K, (int)M, (int)ldg[i], (int)steps,
                             tol, (int)maxrank, 0});
      woff += 3 * M;
    }
    gen_tasks<double>(gens, h->bodies.p, h->k, s);
    sketch_rows(h->ctx, sks, h->bodies.p, h->k, part.p, s);
    qr_batched(qrs, s);
    cpqr_batched(cps, s);
    std::vector<int> rk(nb);
    MXG_HIP(hipMemcpyAsync(rk.data(), d_rank.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    MXG_HIP(hipStreamSynchronize(s));

    // post-processing (h2matrix.cpp:92-116)
    std::vector<PostTask> posts;
    std::vector<long long> orth_boxes;
    size_t tneed = 0, wneed = 0;
    for (long long i = b0; i < b1; i++) {
      const long long M = Lv.dims[i], K = std::min(M, nf[i]);
      if (K <= 0) continue;
      const long long r = rk[i];
      Lv.ranks[i] = r;
      if (0 < r && r < M) {
        tneed += (size_t)r * (M - r);
        wneed += (size_t)M * r;
      }
    }
    DevBuf<double> Tm, Wm, P, ptau;
    Tm.alloc(std::max<size_t>(tneed, 1));
    Wm.alloc_zero(std::max<size_t>(wneed, 1), s);
    std::vector<size_t> wof(nb, 0);
    size_t to = 0, wo = 0;
    for (long long i = b0; i < b1; i++) {
      const long long M = Lv.dims[i], K = std::min(M, nf[i]);
      if (K <= 0) continue;
      const long long r = Lv.ranks[i];
      PostTask pt{};
      pt.Q = Lv.Q.p + Lv.offQ[i];
      pt.R = Lv.R.p + Lv.offQ[i];
      pt.M = (int)M;
      pt.r = (int)r;
      if (0 < r && r < M) {
        pt.R0 = G.p + goff[i - b0];
        pt.ldr = (int)ldg[i];
        pt.perm = d_perm.p + Lv.offX[i];
        pt.S = Lv.S.p + Lv.offX[i];
        pt.Stmp = d_stmp.p + Lv.offX[i];
        pt.T = Tm.p + to;
        pt.W = Wm.p + wo;
        pt.mode = orth ? 3 : 2;
        wof[i] = wo;
        to += (size_t)r * (M - r);
        wo += (size_t)M * r;
        if (orth) orth_boxes.push_back(i);
      } else {
        pt.mode = 1;
      }
      posts.push_back(pt);
    }
    if (!posts.empty()) {
      DevBuf<PostTask> dt;
      dt.upload(posts, s);
      hipLaunchKernelGGL(k_post, dim3((unsigned)posts.size()), dim3(256), 0, s, dt.p);
      MXG_HIP(hipGetLastError());
      MXG_HIP(hipStreamSynchronize(s));
    }
    if (!orth_boxes.empty()) {
      // P = triu(Q_in) W, QR(P), Q = full Q, R = triu (h2matrix.cpp:100-105)
      size_t psz = 0, tz = 0;
      for (long long i : orth_boxes) { psz += (size_t)Lv.dims[i] * Lv.ranks[i]; tz += Lv.ranks[i]; }
      P.alloc(psz);
      ptau.alloc(tz);
      std::vector<GemmTask<double>> gt;
      std::vector<GemmSeg<double>> gs;
      std::vector<QrTask> pq;
      std::vector<FormQTask> fq;
      size_t po = 0, pt2 = 0;
      for (long long i : orth_boxes) {
        const int M = (int)Lv.dims[i], r = (int)Lv.ranks[i];
        double* Pi = P.p + po;
        gs.push_back(GemmSeg<double>{Lv.Q.p + Lv.offQ[i], Wm.p + wof[i], M, M, M, 0, 0, 1, 0});
        gt.push_back(GemmTask<double>{Pi, M, M, r, (int)gs.size() - 1, 1, 1., 0.});
        pq.push_back(QrTask{Pi, ptau.p + pt2, M, r, M, 0});
        fq.push_back(FormQTask{Pi, ptau.p + pt2, Lv.Q.p + Lv.offQ[i], Lv.R.p + Lv.offQ[i], M, r, M, 0});
        po += (size_t)M * r;
        pt2 += r;
      }
      gemm_batched<double>(gt, gs, s);
      qr_batched(pq, s);
      formq_batched(fq, s);
    }
    b0 = b1;
  }
}

void construct_coupling(mxg_h2_s* h, Level& Lv, int orth) {
  hipStream_t s = h->ctx->stream;
  const long long nb = Lv.nb;
  Lv.sym_far = !orth;  // C = K(S_y, S_x) as generated, without the R factors
  if (Lv.C.n == 0) return;
  DevBuf<double> K1, K2;
  std::vector<GenTask> gens;
  if (orth) {
    K1.alloc(Lv.C.n);
    K2.alloc(Lv.C.n);
  }
  for (long long y = 0; y < nb; y++)
    for (long long ij = Lv.frp[y]; ij < Lv.frp[y + 1]; ij++) {
      const long long x = Lv.fcol[ij];
      const int ky = (int)Lv.ranks[y], kx = (int)Lv.ranks[x];
      if (ky == 0 || kx == 0) continue;
      double* out = (orth ? K1.p : Lv.C.p) + Lv.offC[ij];
      gens.push_back(GenTask{Lv.S.p + Lv.offX[y], Lv.S.p + Lv.offX[x], 0, 0, out, ky, kx, ky, 0});
    }
  gen_tasks<double>(gens, h->bodies.p, h->k, s);
  if (!orth) return;
  std::vector<GemmTask<double>> t1, t2;
  std::vector<GemmSeg<double>> s1, s2;
  for (long long y = 0; y < nb; y++)
    for (long long ij = Lv.frp[y]; ij < Lv.frp[y + 1]; ij++) {
      const long long x = Lv.fcol[ij];
      const int ky = (int)Lv.ranks[y], kx = (int)Lv.ranks[x];
      if (ky == 0 || kx == 0) continue;
      const double* Ry = Lv.R.p + Lv.offQ[y];
      const double* Rx = Lv.R.p + Lv.offQ[x];
      s1.push_back(GemmSeg<double>{Ry, K1.p + Lv.offC[ij], (int)Lv.dims[y], ky, ky, 0, 0, 1, 0});
      t1.push_back(GemmTask<double>{K2.p + Lv.offC[ij], ky, ky, kx, (int)s1.size() - 1, 1, 1., 0.});
      s2.push_back(GemmSeg<double>{K2.p + Lv.offC[ij], Rx, ky, (int)Lv.dims[x], kx, 0, 1, 0, 1});
      t2.push_back(GemmTask<double>{Lv.C.p + Lv.offC[ij], ky, ky, kx, (int)s2.size() - 1, 1, 1., 0.});
    }
  gemm_batched<double>(t1, s1, s);
  gemm_batched<double>(t2, s2, s);
}

void upload_indices(mxg_h2_s* h, Level& Lv) {
  hipStream_t s = h->ctx->stream;
  Lv.d_offX.upload(Lv.offX, s);
  Lv.d_offZ.upload(Lv.offZ, s);
  Lv.d_offQ.upload(Lv.offQ, s);
  Lv.d_offA.upload(Lv.offA, s);
  Lv.d_offC.upload(Lv.offC, s);
  Lv.d_nrp.upload(Lv.nrp, s);
  Lv.d_frp.upload(Lv.frp, s);
  // host staging must outlive the asynchronous copies: synchronize before it goes
  std::vector<int> dims(Lv.dims.begin(), Lv.dims.end()), ranks(Lv.ranks.begin(), Lv.ranks.end());
  std::vector<int> ncol(Lv.ncol.begin(), Lv.ncol.end()), fcol(Lv.fcol.begin(), Lv.fcol.end());
  Lv.d_dims.upload(dims, s);
  Lv.d_ranks.upload(ranks, s);
  Lv.d_ncol.upload(ncol, s);
  Lv.d_fcol.upload(fcol, s);
  MXG_HIP(hipStreamSynchronize(s));
}

}  // namespace

// Matvec tasks of a constructed level and the term offsets of its downward
// pass (the device jobs are made per precision on first use).
// QT_i = Q_i[:, :k]^T of every box (the upward pass streams its columns).
void basis_transpose(mxg_h2_s* h, Level& Lv) {
  hipStream_t s = h->ctx->stream;
  const long long nb = Lv.nb;
  Lv.offQT.assign(nb + 1, 0);
  for (long long i = 0; i < nb; i++) Lv.offQT[i + 1] = Lv.offQT[i] + Lv.dims[i] * Lv.ranks[i];
  Lv.QT.alloc(std::max<long long>(Lv.offQT[nb], 1));
  DevBuf<long long> d_offQT;
  d_offQT.upload(Lv.offQT, s);
  hipLaunchKernelGGL(k_basis_transpose, dim3((unsigned)nb), dim3(256), 0, s, Lv.d_dims.p, Lv.d_ranks.p, Lv.d_offQ.p, d_offQT.p,
                     Lv.Q.p, Lv.QT.p);
  MXG_HIP(hipGetLastError());
  MXG_HIP(hipStreamSynchronize(s));
}

// Position of the pair (j, i) in the level-local CSR (rp, col), whose rows
// hold sorted columns; the lists of a matrix on one point set are symmetric.
long long transposed_pos(const std::vector<long long>& rp, const std::vector<long long>& col, long long i,
                         long long j) {
  const long long* b = col.data() + rp[j];
  const long long* e = col.data() + rp[j + 1];
  const long long* f = std::lower_bound(b, e, i);
  MXG_REQUIRE(f != e && *f == i, "block lists are not symmetric");
  return rp[j] + (f - b);
}

// Far pairs i < j of a level with symmetric couplings as symmetric tasks over
// the coupling blocks (offsets offC, leading dimension ld(i)) and the terms
// (offsets toff).
template <class LD>
std::vector<SymTask> far_sym_tasks(const Level& Lv, const std::vector<long long>& offC, const std::vector<long long>& toff,
                                   LD ld) {
  std::vector<SymTask> out;
  for (long long i = 0; i < Lv.nb; i++)
    for (long long p = Lv.frp[i]; p < Lv.frp[i + 1]; p++) {
      const long long j = Lv.fcol[p];
      if (j <= i) continue;
      const long long pt = transposed_pos(Lv.frp, Lv.fcol, i, j);
      out.push_back(SymTask{offC[p], Lv.offZ[i], Lv.offZ[j], toff[p], toff[pt], (int)Lv.ranks[i], (int)Lv.ranks[j],
                            (int)std::max<long long>(ld(i), 1), 0});
    }
  return out;
}

void build_matvec_lists(mxg_h2_s* h, Level& Lv, bool leaf) {
  hipStream_t s = h->ctx->stream;
  const long long nb = Lv.nb;
  basis_transpose(h, Lv);
  Lv.up_t.clear();
  Lv.far_t.clear();
  Lv.near_t.clear();
  for (long long i = 0; i < nb; i++)
    if (Lv.ranks[i] > 0)
      Lv.up_t.push_back(PairTask{Lv.offQT[i], Lv.offX[i], Lv.offZ[i], (int)Lv.ranks[i], (int)Lv.dims[i], (int)Lv.ranks[i], 0});
  Lv.toffF.assign(Lv.frp[nb] + 1, 0);
  for (long long i = 0; i < nb; i++)
    for (long long p = Lv.frp[i]; p < Lv.frp[i + 1]; p++) {
      const long long j = Lv.fcol[p];
      Lv.toffF[p + 1] = Lv.toffF[p] + Lv.ranks[i];
      if (Lv.ranks[i] > 0)
        Lv.far_t.push_back(PairTask{Lv.offC[p], Lv.offZ[j], Lv.toffF[p], (int)Lv.ranks[i], (int)Lv.ranks[j], (int)Lv.ranks[i], 0});
    }
  Lv.lenTF = Lv.toffF.back();
  Lv.far_sym_t.clear();
  if (Lv.sym_far) Lv.far_sym_t = far_sym_tasks(Lv, Lv.offC, Lv.toffF, [&](long long i) { return Lv.ranks[i]; });
  Lv.toffN.assign(Lv.nrp[nb] + 1, 0);
  Lv.lenTN = 0;
  if (leaf) {
    for (long long i = 0; i < nb; i++)
      for (long long p = Lv.nrp[i]; p < Lv.nrp[i + 1]; p++) {
        const long long j = Lv.ncol[p];
        Lv.toffN[p + 1] = Lv.toffN[p] + Lv.dims[i];
        if (Lv.dims[i] > 0)
          Lv.near_t.push_back(PairTask{Lv.offA[p], Lv.offX[j], Lv.toffN[p], (int)Lv.dims[i], (int)Lv.dims[j], (int)Lv.dims[i], 0});
      }
    Lv.lenTN = Lv.toffN.back();
    // each near pair once (i <= j), with the term position of its transpose
    Lv.near_sym_t.clear();
    for (long long i = 0; i < nb; i++)
      for (long long p = Lv.nrp[i]; p < Lv.nrp[i + 1]; p++) {
        const long long j = Lv.ncol[p];
        if (j < i) continue;
        const long long pt = (j != i) ? Lv.toffN[transposed_pos(Lv.nrp, Lv.ncol, i, j)] : -1;
        Lv.near_sym_t.push_back(SymTask{Lv.offA[p], Lv.offX[i], Lv.offX[j], Lv.toffN[p], pt, (int)Lv.dims[i],
                                        (int)Lv.dims[j], (int)Lv.dims[i], 0});
      }
  }
  Lv.d_toffF.upload(Lv.toffF, s);
  Lv.d_toffN.upload(Lv.toffN, s);
  Lv.vecQ[0] = Lv.vecQ[1] = true;
  for (long long i = 0; i < nb; i++) {
    if (Lv.dims[i] % 2) Lv.vecQ[0] = false;
    if (Lv.dims[i] % 4) Lv.vecQ[1] = false;
  }
  MXG_HIP(hipStreamSynchronize(s));
}

namespace {

template <class T>
struct Ops;
template <>
struct Ops<double> {
  static const double* Q(Level& L) { return L.Q.p; }
  static const double* QT(Level& L) { return L.QT.p; }
  static const double* C(Level& L) { return L.C.p; }
  static const double* A(Level& L) { return L.A.p; }
  static double* Z(mxg_h2_s* h, long long l) { return h->Z[l].p; }
  static double* W(mxg_h2_s* h, long long l) { return h->W[l].p; }
  static double* x(mxg_h2_s* h) { return h->x64.p; }
  static DevBuf<double>& t(mxg_h2_s* h) { return h->t64; }
};
template <>
struct Ops<float> {
  static const float* Q(Level& L) { return L.Q32.p; }
  static const float* QT(Level& L) { return L.QT32.p; }
  static const float* C(Level& L) { return L.C32.p; }
  static const float* A(Level& L) { return L.A32.p; }
  static float* Z(mxg_h2_s* h, long long l) { return h->Z32[l].p; }
  static float* W(mxg_h2_s* h, long long l) { return h->W32[l].p; }
  static float* x(mxg_h2_s* h) { return h->x32.p; }
  static DevBuf<float>& t(mxg_h2_s* h) { return h->t32; }
};

// Device jobs of one precision: the upward pass per level, one list with the
// far terms of every level and one with the near terms of the leaf level. The
// vectors of a level hold MAX_RHS right-hand sides, lenZ (or n, or the term
// count) apart.
template <class T>
void ensure_matvec_jobs(mxg_h2_s* h) {
  constexpr int pi = sizeof(T) == 8 ? 0 : 1;
  if (h->jobs_ready[pi]) return;
  hipStream_t s = h->ctx->stream;
  const long long L = h->L;
  long long tlen = 0;
  for (long long l = 0; l <= L; l++) {
    h->lv[l].tbaseF[pi] = tlen;
    tlen += pi ? h->lv[l].lenTF32 : h->lv[l].lenTF;
  }
  h->lv[L].tbaseN[pi] = tlen;
  tlen += h->lv[L].lenTN;
  h->tlen[pi] = std::max<long long>(tlen, 1);
  DevBuf<T>& tb = Ops<T>::t(h);
  tb.alloc((size_t)h->tlen[pi] * MAX_RHS);
  if constexpr (pi == 0) h->x64.reserve((size_t)h->n * MAX_RHS);
  const MatvecTuning& tn = h->tune[pi];
  std::vector<PairJob> prod;
  std::vector<SymJob> prod_sym;
  for (long long l = 0; l <= L; l++) {
    Level& Lv = h->lv[l];
    std::vector<PairJob> up;
    const T* xin = (l == L) ? Ops<T>::x(h) : Ops<T>::Z(h, l + 1);
    const long long lzin = (l == L) ? h->n : h->lv[l + 1].lenZ;
    if (Lv.lenZ > 0) {
      pair_jobs_append<T, T>(up, spread(Lv.up_t, SPLIT_BELOW, tn.split_rows), Ops<T>::QT(Lv), (long long)Lv.offQT[Lv.nb], xin, Lv.lenX, lzin,
                             Ops<T>::Z(h, l), Lv.lenZ, Lv.lenZ);
      if (Lv.sym_far && tn.far_sym)
        sym_jobs_append<T>(prod_sym, pi ? Lv.far_sym_t32 : Lv.far_sym_t, Ops<T>::C(Lv),
                           pi ? (long long)Lv.C32.n : (long long)Lv.C.n, Ops<T>::Z(h, l), Lv.lenZ, Lv.lenZ,
                           tb.p + Lv.tbaseF[pi], pi ? Lv.lenTF32 : Lv.lenTF, h->tlen[pi]);
      else
        pair_jobs_append<T, T>(prod, pi ? Lv.far_t32 : Lv.far_t, Ops<T>::C(Lv),
                               pi ? (long long)Lv.C32.n : (long long)Lv.C.n, Ops<T>::Z(h, l), Lv.lenZ, Lv.lenZ,
                               tb.p + Lv.tbaseF[pi], pi ? Lv.lenTF32 : Lv.lenTF, h->tlen[pi], tn.matrix_cores != 0);
    }
    pair_list_upload(Lv.up_pl[pi], up, s);
  }
  pair_list_upload(h->prod_pl[pi], prod, s);
  sym_list_upload(h->prod_sym[pi], prod_sym, s);
  Level& Lf = h->lv[L];
  std::vector<PairJob> near;
  const bool near_sym = h->sym_near && tn.near_sym;
  if (!near_sym)
    pair_jobs_append<T, T>(near, Lf.near_t, Ops<T>::A(Lf), (long long)Lf.A.n, Ops<T>::x(h), h->n, h->n,
                           tb.p + Lf.tbaseN[pi], Lf.lenTN, h->tlen[pi], tn.matrix_cores != 0);
  pair_list_upload(h->near_pl[pi], near, s);
  std::vector<SymJob> sym;
  if (near_sym)
    sym_jobs_append<T>(sym, Lf.near_sym_t, Ops<T>::A(Lf), (long long)Lf.A.n, Ops<T>::x(h), h->n, h->n,
                       tb.p + Lf.tbaseN[pi], Lf.lenTN, h->tlen[pi]);
  sym_list_upload(h->near_sym[pi], sym, s);
  MXG_HIP(hipStreamSynchronize(s));
  h->jobs_ready[pi] = true;
}

template <class Td, class Ts>
__global__ void k_convert2d(Td* __restrict__ dst, long long ldd, const Ts* __restrict__ src, long long lds, long long n,
                            int ncols) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int c = 0; c < ncols; c++) dst[r + c * ldd] = (Td)src[r + c * lds];
}

}  // namespace

template <class Td, class Ts>
void convert2d(Td* dst, long long ldd, const Ts* src, long long lds, long long n, int ncols, hipStream_t s) {
  if (n == 0 || ncols == 0) return;
  hipLaunchKernelGGL((k_convert2d<Td, Ts>), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, dst, ldd, src, lds, n, ncols);
  MXG_HIP(hipGetLastError());
}
template void convert2d<double, double>(double*, long long, const double*, long long, long long, int, hipStream_t);
template void convert2d<float, double>(float*, long long, const double*, long long, long long, int, hipStream_t);
template void convert2d<double, float>(double*, long long, const float*, long long, long long, int, hipStream_t);

// Y = A X for nrhs <= MAX_RHS right-hand sides staged in the precision's
// input buffer (n apart); Y has leading dimension ldy.
template <class T>
void h2_matvec_levels(mxg_h2_s* h, T* y, long long ldy, int nrhs) {
  hipStream_t s = h->ctx->stream;
  const long long L = h->L;
  constexpr bool f64 = sizeof(T) == 8;
  constexpr int pi = f64 ? 0 : 1;
  ensure_matvec_jobs<T>(h);
  T* tb = Ops<T>::t(h).p;
  const long long lt = h->tlen[pi];
  // the leaf near terms depend on x alone: they run on the second stream,
  // beside the upward pass and the far terms, which leave most of the GPU idle
  hipStream_t sa = h->ctx->aux;
  ProfScope whole(h->ctx, f64 ? MXG_PROF_MATVEC_FP64 : MXG_PROF_MATVEC_FP32);
  MXG_HIP(hipEventRecord(h->ctx->fork, s));
  MXG_HIP(hipStreamWaitEvent(sa, h->ctx->fork, 0));
  {
    ProfScope prof(h->ctx, f64 ? MXG_PROF_MATVEC_LEAF_FP64 : MXG_PROF_MATVEC_LEAF_FP32, sa);
    // one of the two lists is empty
    sym_gemv<T>(h->near_sym[pi], nrhs, sa);
    pair_gemv<T, T>(h->near_pl[pi], nrhs, sa);
  }
  MXG_HIP(hipEventRecord(h->ctx->join, sa));
  // upward pass: z_i = Q_i[:, :k]^T x_i
  for (long long l = L; l >= 0; l--) {
    if (h->lv[l].lenZ == 0) continue;
    ProfScope prof(h->ctx, MXG_PROF_MATVEC_UPPER);
    pair_gemv<T, T>(h->lv[l].up_pl[pi], nrhs, s);
  }
  // every far term of every level, one launch
  {
    ProfScope prof(h->ctx, MXG_PROF_MATVEC_FAR);
    pair_gemv<T, T>(h->prod_pl[pi], nrhs, s);
    sym_gemv<T>(h->prod_sym[pi], nrhs, s);
  }
  // downward pass: w = w_in + far terms, y = Q w (+ near terms at the leaf);
  // only the leaf level waits for the near terms
  for (long long l = 0; l <= L; l++) {
    Level& Lv = h->lv[l];
    const bool leaf = (l == L);
    if (leaf) MXG_HIP(hipStreamWaitEvent(s, h->ctx->join, 0));
    T* out = leaf ? y : Ops<T>::W(h, l + 1);
    const long long ldo = leaf ? ldy : h->lv[l + 1].lenZ;
    if (!leaf && ldo == 0) continue;
    ProfScope prof(h->ctx, MXG_PROF_MATVEC_UPPER);
    comb_down<T>(Lv, Lv.vecQ[pi], pi ? Lv.d_toffF32.p : Lv.d_toffF.p, Lv.d_toffN.p, Ops<T>::Q(Lv), Ops<T>::W(h, l),
                 std::max<long long>(Lv.lenZ, 1), tb + Lv.tbaseF[pi], leaf ? tb + Lv.tbaseN[pi] : nullptr, lt, out, ldo,
                 nrhs, s);
  }
}

template void h2_matvec_levels<double>(mxg_h2_s*, double*, long long, int);
template void h2_matvec_levels<float>(mxg_h2_s*, float*, long long, int);

void make_shadow32(mxg_h2_s* h);

void h2_apply(mxg_h2_s* h, const double* x, long long ldx, double* y, long long ldy, long long nrhs, int precision) {
  for (long long l = 0; l <= h->L; l++) MXG_REQUIRE(h->lv[l].built, "H2 matrix not fully constructed");
  MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32, "unknown precision");
  MXG_REQUIRE(nrhs >= 1 && ldx >= h->n && ldy >= h->n, "bad right-hand sides");
  hipStream_t s = h->ctx->stream;
  if (precision == MXG_FP32) make_shadow32(h);
  const long long n = h->n;
  for (long long c0 = 0; c0 < nrhs; c0 += MAX_RHS) {
    const int nr = (int)std::min<long long>(MAX_RHS, nrhs - c0);
    if (precision == MXG_FP64) {
      ensure_matvec_jobs<double>(h);
      convert2d(h->x64.p, n, x + c0 * ldx, ldx, n, nr, s);
      h2_matvec_levels<double>(h, y + c0 * ldy, ldy, nr);
    } else {
      ensure_matvec_jobs<float>(h);
      convert2d(h->x32.p, n, x + c0 * ldx, ldx, n, nr, s);
      h2_matvec_levels<float>(h, h->y32.p, n, nr);
      convert2d(y + c0 * ldy, ldy, h->y32.p, n, n, nr, s);
    }
  }
}

}  // namespace mxg

using namespace mxg;

namespace mxg {

namespace {
// One coupling block copied to FP32 with its rows padded to prow (zeros).
struct PadTask {
  long long src, dst;
  int rows, prow, cols, pad;
};

__global__ void __launch_bounds__(256) k_pad_rows(const PadTask* __restrict__ tasks, const double* __restrict__ src,
                                                  float* __restrict__ dst) {
  const PadTask t = tasks[blockIdx.x];
  const long long ne = (long long)t.prow * t.cols;
  for (long long e = threadIdx.x; e < ne; e += blockDim.x) {
    const int r = (int)(e % t.prow), c = (int)(e / t.prow);
    dst[t.dst + e] = r < t.rows ? (float)src[t.src + r + (long long)c * t.rows] : 0.f;
  }
}
}  // namespace

// The FP32 coupling blocks with rows padded to a multiple of 4, their far
// tasks and term offsets (each term holds the padded rows; the padding rows
// of a product are zero and never read).
void pad_couplings32(Level& Lv, hipStream_t s) {
  const long long nb = Lv.nb, np = Lv.frp[nb];
  Lv.offC32.assign(np + 1, 0);
  Lv.toffF32.assign(np + 1, 0);
  Lv.far_t32.clear();
  std::vector<PadTask> pt;
  for (long long i = 0; i < nb; i++) {
    const long long k = Lv.ranks[i], kp = (k + 3) / 4 * 4;
    for (long long p = Lv.frp[i]; p < Lv.frp[i + 1]; p++) {
      const long long j = Lv.fcol[p];
      Lv.offC32[p + 1] = Lv.offC32[p] + kp * Lv.ranks[j];
      Lv.toffF32[p + 1] = Lv.toffF32[p] + kp;
      if (k == 0) continue;
      pt.push_back(PadTask{Lv.offC[p], Lv.offC32[p], (int)k, (int)kp, (int)Lv.ranks[j], 0});
      Lv.far_t32.push_back(PairTask{Lv.offC32[p], Lv.offZ[j], Lv.toffF32[p], (int)kp, (int)Lv.ranks[j], (int)kp, 0});
    }
  }
  Lv.lenTF32 = Lv.toffF32[np];
  Lv.far_sym_t32.clear();
  if (Lv.sym_far)
    Lv.far_sym_t32 = far_sym_tasks(Lv, Lv.offC32, Lv.toffF32, [&](long long i) { return (Lv.ranks[i] + 3) / 4 * 4; });
  Lv.C32.alloc(std::max<long long>(Lv.offC32[np], 1));
  Lv.d_toffF32.upload(Lv.toffF32, s);
  if (pt.empty()) return;
  MXG_REQUIRE(pt.size() < (1ull << 31), "too many coupling blocks");
  DevBuf<PadTask> d;
  d.upload(pt, s);
  hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)pt.size()), dim3(256), 0, s, d.p, Lv.C.p, Lv.C32.p);
  MXG_HIP(hipGetLastError());
  MXG_HIP(hipStreamSynchronize(s));
}

void make_shadow32(mxg_h2_s* h) {
  if (h->shadow32) return;
  hipStream_t s = h->ctx->stream;
  h->Z32.resize(h->L + 1);
  h->W32.resize(h->L + 1);
  for (long long l = 0; l <= h->L; l++) {
    Level& Lv = h->lv[l];
    Lv.Q32.alloc(Lv.Q.n);
    convert(Lv.Q32.p, Lv.Q.p, Lv.Q.n, s);
    pad_couplings32(Lv, s);
    Lv.QT32.alloc(Lv.QT.n);
    convert(Lv.QT32.p, Lv.QT.p, Lv.QT.n, s);
    if (l == h->L) {
      Lv.A32.alloc(Lv.A.n);
      convert(Lv.A32.p, Lv.A.p, Lv.A.n, s);
    }
    h->Z32[l].alloc(std::max<long long>(Lv.lenZ, 1) * MAX_RHS);
    h->W32[l].alloc_zero(std::max<long long>(Lv.lenZ, 1) * MAX_RHS, s);
  }
  h->x32.alloc(h->n * MAX_RHS);
  h->y32.alloc(h->n * MAX_RHS);
  h->jobs_ready[1] = false;
  MXG_HIP(hipStreamSynchronize(s));
  h->shadow32 = true;
}

namespace {
// Box sizes, lists, offsets and zeroed arrays of level l (levels are set up
// from the leaves upward: an upper level's box sizes must be the sums of its
// children's ranks, and the leaves must cover every body).
void level_structure(mxg_h2_s* h, long long l, long long nb, const long long* dims, const long long* near_rp,
                     const long long* near_col, const long long* far_rp, const long long* far_col) {
  MXG_REQUIRE(nb == (1ll << l), "nboxes must be 2^level");
  MXG_REQUIRE(l == h->L || h->lv[l + 1].built, "set up levels from the leaves upward");
  hipStream_t s = h->ctx->stream;
  Level& Lv = h->lv[l];
  Level* low = (l < h->L) ? &h->lv[l + 1] : nullptr;
  Lv = Level();
  Lv.nb = nb;
  Lv.dims.assign(dims, dims + nb);
  Lv.nrp.assign(near_rp, near_rp + nb + 1);
  Lv.frp.assign(far_rp, far_rp + nb + 1);
  MXG_REQUIRE(Lv.nrp[0] == 0 && Lv.frp[0] == 0, "CSR row pointers must start at 0");
  for (long long i = 0; i < nb; i++)
    MXG_REQUIRE(Lv.nrp[i + 1] >= Lv.nrp[i] && Lv.frp[i + 1] >= Lv.frp[i], "CSR row pointers must not decrease");
  MXG_REQUIRE(Lv.nrp[nb] == 0 || near_col, "null near columns");
  MXG_REQUIRE(Lv.frp[nb] == 0 || far_col, "null far columns");
  Lv.ncol.assign(near_col, near_col + Lv.nrp[nb]);
  Lv.fcol.assign(far_col, far_col + Lv.frp[nb]);
  for (long long t = 0; t < Lv.nrp[nb]; t++) MXG_REQUIRE(Lv.ncol[t] >= 0 && Lv.ncol[t] < nb, "near column out of range");
  for (long long t = 0; t < Lv.frp[nb]; t++) MXG_REQUIRE(Lv.fcol[t] >= 0 && Lv.fcol[t] < nb, "far column out of range");
  for (long long i = 0; i < nb; i++) {
    MXG_REQUIRE(dims[i] >= 0 && dims[i] <= 4096, "box dimension out of range (max 4096)");
    if (low) MXG_REQUIRE(dims[i] == low->ranks[2 * i] + low->ranks[2 * i + 1], "upper dims must be the sum of the child ranks");
  }
  if (!low) {
    long long total = 0;
    for (long long i = 0; i < nb; i++) total += dims[i];
    MXG_REQUIRE(total == h->n, "leaf dims must sum to nbodies");
  }
  Lv.ranks.assign(nb, 0);
  Lv.offX.assign(nb + 1, 0);
  Lv.offQ.assign(nb + 1, 0);
  Lv.diag.assign(nb, -1);
  for (long long i = 0; i < nb; i++) {
    Lv.offX[i + 1] = Lv.offX[i] + dims[i];
    Lv.offQ[i + 1] = Lv.offQ[i] + dims[i] * dims[i];
    Lv.maxM = std::max(Lv.maxM, dims[i]);
    Lv.diag[i] = lookup(Lv.nrp, Lv.ncol, i, i);
    MXG_REQUIRE(Lv.diag[i] >= 0, "every box needs its diagonal near block");
  }
  Lv.lenX = Lv.offX[nb];
  const long long nnzA = Lv.nrp[nb];
  Lv.offA.assign(nnzA + 1, 0);
  for (long long i = 0; i < nb; i++)
    for (long long t = Lv.nrp[i]; t < Lv.nrp[i + 1]; t++) Lv.offA[t + 1] = Lv.offA[t] + dims[i] * dims[Lv.ncol[t]];
  Lv.up_off.assign(nnzA, -1);
  Lv.up_ld.assign(nnzA, 0);
  Lv.S.alloc(std::max<long long>(Lv.lenX, 1));
  Lv.Q.alloc_zero(Lv.offQ[nb], s);
  Lv.R.alloc_zero(Lv.offQ[nb], s);
  if (!low || (h->flags & MXG_H2_FACTORABLE)) Lv.A.alloc_zero(Lv.offA[nnzA], s);
}

// For every near pair (i, j) of level l and child pair (y, x) of it: a near
// child pair hands its skeleton block to the parent block at up_off
// (factorizeCopyNext, h2matrix.cpp:427-446; LowerIndA of comm-mpi.cpp:117-149);
// with copy_far, a far child pair's coupling block is copied into the parent
// near block (h2matrix.cpp:172-173), when the parent keeps its near blocks.
void link_children(mxg_h2_s* h, Level& Lv, Level& low, bool copy_far) {
  std::vector<const double*> cs;
  std::vector<double*> cd;
  std::vector<int> cm, cn, cls, cld;
  const bool keepA = Lv.A.n > 0;
  for (long long i = 0; i < Lv.nb; i++) {
    const long long M = Lv.dims[i];
    for (int cy = 0; cy < 2; cy++) {
      const long long y = 2 * i + cy;
      const long long offy = cy ? low.ranks[2 * i] : 0, ny = low.ranks[y];
      if (ny == 0) continue;
      for (long long ij = Lv.nrp[i]; ij < Lv.nrp[i + 1]; ij++) {
        const long long j = Lv.ncol[ij];
        for (int cx = 0; cx < 2; cx++) {
          const long long x = 2 * j + cx;
          const long long offx = cx ? low.ranks[2 * j] : 0, nx = low.ranks[x];
          if (nx == 0) continue;
          const long long lowN = lookup(low.nrp, low.ncol, y, x);
          const long long lowC = lookup(low.frp, low.fcol, y, x);
          const long long dp = Lv.offA[ij] + offy + offx * M;
          if (lowN >= 0) {
            low.up_off[lowN] = dp;
            low.up_ld[lowN] = (int)M;
          } else if (lowC >= 0 && keepA && copy_far) {
            cs.push_back(low.C.p + low.offC[lowC]);
            cd.push_back(Lv.A.p + dp);
            cm.push_back((int)ny); cn.push_back((int)nx); cls.push_back((int)ny); cld.push_back((int)M);
          }
        }
      }
    }
  }
  copy2d_batched<double>(cs, cd, cm, cn, cls, cld, h->ctx->stream);
}

// Layout that depends on the ranks: skeleton vectors and zeroed couplings.
void level_ranks(mxg_h2_s* h, Level& Lv) {
  const long long nb = Lv.nb;
  Lv.offZ.assign(nb + 1, 0);
  for (long long i = 0; i < nb; i++) {
    Lv.offZ[i + 1] = Lv.offZ[i] + Lv.ranks[i];
    Lv.maxK = std::max(Lv.maxK, Lv.ranks[i]);
  }
  Lv.lenZ = Lv.offZ[nb];
  const long long nnzC = Lv.frp[nb];
  Lv.offC.assign(nnzC + 1, 0);
  for (long long i = 0; i < nb; i++)
    for (long long t = Lv.frp[i]; t < Lv.frp[i + 1]; t++) Lv.offC[t + 1] = Lv.offC[t] + Lv.ranks[i] * Lv.ranks[Lv.fcol[t]];
  Lv.C.alloc_zero(Lv.offC[nnzC], h->ctx->stream);
}

// Index arrays, matvec lists and vector buffers of a level whose data is set.
void level_finish(mxg_h2_s* h, Level& Lv, bool leaf) {
  hipStream_t s = h->ctx->stream;
  const long long l = (long long)(&Lv - h->lv.data());
  upload_indices(h, Lv);
  build_matvec_lists(h, Lv, leaf);
  h->jobs_ready[0] = h->jobs_ready[1] = false;
  h->Z[l].alloc_zero(std::max<long long>(Lv.lenZ, 1) * MAX_RHS, s);
  h->W[l].alloc_zero(std::max<long long>(Lv.lenZ, 1) * MAX_RHS, s);
  MXG_HIP(hipStreamSynchronize(s));
  Lv.built = true;
  h->shadow32 = false;
}
}  // namespace

}  // namespace mxg

namespace mxg {
namespace {
// Times every candidate launch shape of one precision's matvec for nrhs
// right-hand sides (a warm-up, then reps timed products with HIP events) and
// keeps the fastest; returns its time per product (ms).
template <class T>
double autotune(mxg_h2_s* h, int nrhs, int reps) {
  constexpr int pi = sizeof(T) == 8 ? 0 : 1;
  hipStream_t s = h->ctx->stream;
  const long long n = h->n;
  DevBuf<T> y;
  y.alloc((size_t)n * nrhs);
  bool any_far = false;
  for (long long l = 0; l <= h->L; l++) any_far = any_far || h->lv[l].sym_far;
  std::vector<MatvecTuning> cand;
  for (int ns : {1, 0})
    for (int fs : {1, 0})
      for (int mc : {1, 0})
        for (int rows : {16, 32, 64}) {
          if ((ns && !h->sym_near) || (fs && !any_far) || (mc && ns && fs)) continue;
          MatvecTuning t;
          t.near_sym = ns;
          t.far_sym = fs;
          t.split_rows = rows;
          t.matrix_cores = mc;
          cand.push_back(t);
        }
  hipEvent_t e0, e1;
  MXG_HIP(hipEventCreate(&e0));
  MXG_HIP(hipEventCreate(&e1));
  double best = -1.;
  MatvecTuning pick = h->tune[pi];
  try {
    for (const MatvecTuning& t : cand) {
      h->tune[pi] = t;
      h->jobs_ready[pi] = false;
      ensure_matvec_jobs<T>(h);
      MXG_HIP(hipMemsetAsync(Ops<T>::x(h), 0, (size_t)n * MAX_RHS * sizeof(T), s));
      h2_matvec_levels<T>(h, y.p, n, nrhs);
      MXG_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) h2_matvec_levels<T>(h, y.p, n, nrhs);
      MXG_HIP(hipEventRecord(e1, s));
      MXG_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MXG_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double per = (double)ms / reps;
      if (best < 0. || per < best) {
        best = per;
        pick = t;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->tune[pi] = pick;
  h->jobs_ready[pi] = false;
  ensure_matvec_jobs<T>(h);
  return best;
}
}  // namespace
}  // namespace mxg

extern "C" {

int mxg_h2_create(mxg_ctx_t ctx, mxg_kernel_t k, const double* bodies, long long nbodies, long long levels, int flags,
                  mxg_h2_t* out) {
  return guarded([&] {
    MXG_REQUIRE(ctx && out && bodies, "null argument");
    MXG_REQUIRE(nbodies > 0 && nbodies < (1ll << 31), "nbodies out of range");
    MXG_REQUIRE(levels >= 0 && levels < 31 && (1ll << levels) <= nbodies, "levels out of range");
    MXG_REQUIRE(k.id >= MXG_KERNEL_LAPLACE && k.id <= MXG_KERNEL_GAUSSIAN, "unknown kernel id");
    MXG_HIP(hipSetDevice(ctx->device));
    auto* h = new mxg_h2_s();
    try {
      h->ctx = ctx;
      h->k = k;
      h->n = nbodies;
      h->L = levels;
      h->flags = flags;
      h->bodies.upload(bodies, 3 * nbodies, ctx->stream);
      h->lv.resize(levels + 1);
      h->Z.resize(levels + 1);
      h->W.resize(levels + 1);
      MXG_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int mxg_h2_destroy(mxg_h2_t h) {
  return guarded([&] {
    if (!h) return;
    (void)hipStreamSynchronize(h->ctx->stream);
    delete h;
  });
}

}  // extern "C"

namespace mxg {
namespace {
void construct_level(mxg_h2_t h, long long l, long long nb, const long long* dims, const long long* leaf_begin,
                     const long long* near_rp, const long long* near_col, const long long* far_rp, const long long* far_col,
                     const FarRows& fr, double tol, long long maxrank, int orth, long long* ranks_out) {
  {
    MXG_REQUIRE(h && dims && near_rp && far_rp && ranks_out, "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L, "level out of range");
    MXG_REQUIRE(l < h->L || leaf_begin, "leaf level needs leaf_begin");
    MXG_HIP(hipSetDevice(h->ctx->device));
    hipStream_t s = h->ctx->stream;
    level_structure(h, l, nb, dims, near_rp, near_col, far_rp, far_col);
    Level& Lv = h->lv[l];
    Level* low = (l < h->L) ? &h->lv[l + 1] : nullptr;
    if (!low) {
      for (long long i = 0; i < nb; i++)
        MXG_REQUIRE(leaf_begin[i] >= 0 && leaf_begin[i] + dims[i] <= h->n, "leaf body range out of bounds");
      DevBuf<long long> d_begin;
      d_begin.upload(leaf_begin, nb, s);
      upload_indices(h, Lv);
      hipLaunchKernelGGL(k_leaf_init, dim3((unsigned)nb), dim3(256), 0, s, (int)nb, Lv.d_offX.p, Lv.d_offQ.p, Lv.d_dims.p,
                         d_begin.p, Lv.S.p, Lv.Q.p);
      MXG_HIP(hipGetLastError());
      std::vector<GenTask> gens;
      for (long long i = 0; i < nb; i++)
        for (long long t = Lv.nrp[i]; t < Lv.nrp[i + 1]; t++) {
          const long long j = Lv.ncol[t];
          gens.push_back(GenTask{nullptr, nullptr, leaf_begin[i], leaf_begin[j], Lv.A.p + Lv.offA[t], (int)dims[i], (int)dims[j],
                                 (int)std::max<long long>(dims[i], 1), 0});
        }
      gen_tasks<double>(gens, h->bodies.p, h->k, s);
    } else {
      // skeleton = the children's skeleton prefixes, Q_in = blockdiag of the children's R
      std::vector<const int*> ss;
      std::vector<int*> sd;
      std::vector<const double*> qs;
      std::vector<double*> qd;
      std::vector<int> sm, sn, sls, sld, qm, qn, qls, qld;
      for (long long i = 0; i < nb; i++) {
        const long long M = dims[i];
        for (int cy = 0; cy < 2; cy++) {
          const long long y = 2 * i + cy;
          const long long offy = cy ? low->ranks[2 * i] : 0, ny = low->ranks[y];
          if (ny == 0) continue;
          ss.push_back(low->S.p + low->offX[y]);
          sd.push_back(Lv.S.p + Lv.offX[i] + offy);
          sm.push_back((int)ny); sn.push_back(1); sls.push_back((int)ny); sld.push_back((int)ny);
          qs.push_back(low->R.p + low->offQ[y]);
          qd.push_back(Lv.Q.p + Lv.offQ[i] + offy + offy * M);
          qm.push_back((int)ny); qn.push_back((int)ny); qls.push_back((int)low->dims[y]); qld.push_back((int)M);
        }
      }
      copy2d_batched<int>(ss, sd, sm, sn, sls, sld, s);
      copy2d_batched<double>(qs, qd, qm, qn, qls, qld, s);
      link_children(h, Lv, *low, true);
    }
    construct_basis(h, Lv, l, fr, tol, maxrank, orth);
    for (long long i = 0; i < nb; i++) ranks_out[i] = Lv.ranks[i];
    level_ranks(h, Lv);
    construct_coupling(h, Lv, orth);
    level_finish(h, Lv, !low);
  }
}
}  // namespace
}  // namespace mxg

extern "C" {

int mxg_h2_construct_level(mxg_h2_t h, long long l, long long nb, const long long* dims, const long long* leaf_begin,
                           const long long* near_rp, const long long* near_col, const long long* far_rp,
                           const long long* far_col, const long long* far_off, const long long* far_idx, double tol,
                           long long maxrank, int orth, long long* ranks_out) {
  return guarded([&] {
    MXG_REQUIRE(h && dims && ranks_out, "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L, "level out of range");
    MXG_HIP(hipSetDevice(h->ctx->device));
    FarRows fr;
    if (l > 0) far_rows_explicit(h, nb, far_off, far_idx, fr);
    else far_rows_explicit(h, nb, nullptr, nullptr, fr);
    construct_level(h, l, nb, dims, leaf_begin, near_rp, near_col, far_rp, far_col, fr, tol, maxrank, orth, ranks_out);
  });
}

int mxg_h2_construct_level_runs(mxg_h2_t h, long long l, long long nb, const long long* dims, const long long* leaf_begin,
                                const long long* near_rp, const long long* near_col, const long long* far_rp,
                                const long long* far_col, const long long* far_run_off, const long long* far_runs,
                                double tol, long long maxrank, int orth, long long* ranks_out) {
  return guarded([&] {
    MXG_REQUIRE(h && dims && ranks_out, "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L, "level out of range");
    MXG_REQUIRE(!far_run_off || far_runs || far_run_off[nb] == far_run_off[0], "null far runs");
    MXG_HIP(hipSetDevice(h->ctx->device));
    FarRows fr;
    far_rows_runs(h, nb, l > 0 ? far_run_off : nullptr, far_runs, fr);
    construct_level(h, l, nb, dims, leaf_begin, near_rp, near_col, far_rp, far_col, fr, tol, maxrank, orth, ranks_out);
  });
}

int mxg_h2_define_level(mxg_h2_t h, long long l, long long nb, const long long* dims, const long long* ranks,
                        const long long* near_rp, const long long* near_col, const long long* far_rp,
                        const long long* far_col) {
  return guarded([&] {
    MXG_REQUIRE(h && dims && ranks && near_rp && far_rp, "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L, "level out of range");
    MXG_HIP(hipSetDevice(h->ctx->device));
    hipStream_t s = h->ctx->stream;
    for (long long i = 0; i < nb; i++) MXG_REQUIRE(ranks[i] >= 0 && ranks[i] <= dims[i], "rank out of range (0..dim)");
    if (l == 0) for (long long i = 0; i < nb; i++) MXG_REQUIRE(ranks[i] == 0, "the root has rank 0");
    level_structure(h, l, nb, dims, near_rp, near_col, far_rp, far_col);
    Level& Lv = h->lv[l];
    Level* low = (l < h->L) ? &h->lv[l + 1] : nullptr;
    if (!low) {
      // the caller's near blocks need not be symmetric
      h->sym_near = false;
      h->jobs_ready[0] = h->jobs_ready[1] = false;
      // the leaf's skeleton is its bodies in tree order, Q_in = I (h2matrix.cpp:179-188)
      upload_indices(h, Lv);
      DevBuf<long long> d_begin;
      d_begin.upload(Lv.offX.data(), nb, s);
      hipLaunchKernelGGL(k_leaf_init, dim3((unsigned)nb), dim3(256), 0, s, (int)nb, Lv.d_offX.p, Lv.d_offQ.p, Lv.d_dims.p,
                         d_begin.p, Lv.S.p, Lv.Q.p);
      MXG_HIP(hipGetLastError());
      MXG_HIP(hipStreamSynchronize(s));
    } else {
      link_children(h, Lv, *low, false);
    }
    Lv.ranks.assign(ranks, ranks + nb);
    level_ranks(h, Lv);
    level_finish(h, Lv, !low);
  });
}

long long mxg_h2_level_size(mxg_h2_t h, long long l, int what) {
  if (!h || l < 0 || l > h->L) return -1;
  const Level& Lv = h->lv[l];
  switch (what) {
    case 0: return (long long)Lv.A.n;
    case 1: return (long long)Lv.Q.n;
    case 2: return (long long)Lv.R.n;
    case 3: return (long long)Lv.C.n;
    case 4: return Lv.lenX;
    default: return -1;
  }
}

int mxg_h2_download(mxg_h2_t h, long long l, int what, double* dst, long long count) {
  return guarded([&] {
    MXG_REQUIRE(h && dst, "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L && h->lv[l].built, "level not constructed");
    long long n = mxg_h2_level_size(h, l, what);
    MXG_REQUIRE(n >= 0 && count == n, "count does not match the level size");
    if (n == 0) return;
    Level& Lv = h->lv[l];
    hipStream_t s = h->ctx->stream;
    if (what == 4) {
      std::vector<int> tmp(n);
      MXG_HIP(hipMemcpyAsync(tmp.data(), Lv.S.p, n * sizeof(int), hipMemcpyDeviceToHost, s));
      MXG_HIP(hipStreamSynchronize(s));
      for (long long i = 0; i < n; i++) dst[i] = tmp[i];
      return;
    }
    const double* src = what == 0 ? Lv.A.p : what == 1 ? Lv.Q.p : what == 2 ? Lv.R.p : Lv.C.p;
    MXG_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    MXG_HIP(hipStreamSynchronize(s));
  });
}

int mxg_h2_upload(mxg_h2_t h, long long l, int what, const double* src, long long count) {
  return guarded([&] {
    MXG_REQUIRE(h && (src || count == 0), "null argument");
    MXG_REQUIRE(l >= 0 && l <= h->L && h->lv[l].built, "level not set up");
    const long long n = mxg_h2_level_size(h, l, what);
    MXG_REQUIRE(n >= 0 && count == n, "count does not match the level size");
    MXG_HIP(hipSetDevice(h->ctx->device));
    if (n == 0) return;
    Level& Lv = h->lv[l];
    hipStream_t s = h->ctx->stream;
    if (what == 4) {
      std::vector<int> tmp(n);
      for (long long i = 0; i < n; i++) {
        MXG_REQUIRE(src[i] >= 0 && src[i] < (double)h->n && src[i] == std::floor(src[i]), "body index out of range");
        tmp[i] = (int)src[i];
      }
      MXG_HIP(hipMemcpyAsync(Lv.S.p, tmp.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
      MXG_HIP(hipStreamSynchronize(s));
      return;
    }
    if (what == 3 && Lv.sym_far) {
      // the caller's couplings need not be symmetric
      Lv.sym_far = false;
      Lv.far_sym_t.clear();
      Lv.far_sym_t32.clear();
      h->jobs_ready[0] = h->jobs_ready[1] = false;
    }
    if (what == 0 && l == h->L && h->sym_near) {
      // the caller's near blocks need not be symmetric: general products from now on
      h->sym_near = false;
      h->jobs_ready[0] = h->jobs_ready[1] = false;
    }
    double* dst = what == 0 ? Lv.A.p : what == 1 ? Lv.Q.p : what == 2 ? Lv.R.p : Lv.C.p;
    MXG_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, s));
    MXG_HIP(hipStreamSynchronize(s));
    if (what == 1) basis_transpose(h, Lv);
    h->shadow32 = false;  // FP32 copies are made again on their next use
  });
}


int mxg_h2_autotune(mxg_h2_t h, int precision, long long nrhs, int reps, double* best_ms) {
  return guarded([&] {
    MXG_REQUIRE(h && best_ms, "null argument");
    MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32, "unknown precision");
    MXG_REQUIRE(nrhs >= 1 && nrhs <= MAX_RHS && reps >= 1 && reps <= 1000, "nrhs must be 1..8 and reps 1..1000");
    for (long long l = 0; l <= h->L; l++) MXG_REQUIRE(h->lv[l].built, "H2 matrix not fully constructed");
    MXG_HIP(hipSetDevice(h->ctx->device));
    if (precision == MXG_FP32) {
      make_shadow32(h);
      *best_ms = autotune<float>(h, (int)nrhs, reps);
    } else {
      *best_ms = autotune<double>(h, (int)nrhs, reps);
    }
    MXG_HIP(hipStreamSynchronize(h->ctx->stream));
  });
}

int mxg_h2_get_tuning(mxg_h2_t h, int precision, int* near_sym, int* far_sym, int* split_rows, int* matrix_cores) {
  return guarded([&] {
    MXG_REQUIRE(h && near_sym && far_sym && split_rows && matrix_cores, "null argument");
    MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32, "unknown precision");
    const MatvecTuning& t = h->tune[precision == MXG_FP64 ? 0 : 1];
    *near_sym = t.near_sym && h->sym_near;
    bool any_far = false;
    for (long long l = 0; l <= h->L; l++) any_far = any_far || h->lv[l].sym_far;
    *far_sym = t.far_sym && any_far;
    *split_rows = t.split_rows;
    *matrix_cores = t.matrix_cores;
  });
}

int mxg_h2_set_tuning(mxg_h2_t h, int precision, int near_sym, int far_sym, int split_rows, int matrix_cores) {
  return guarded([&] {
    MXG_REQUIRE(h, "null argument");
    MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32, "unknown precision");
    MXG_REQUIRE(split_rows >= 8 && split_rows <= 256 && split_rows % 4 == 0, "split_rows must be a multiple of 4 in 8..256");
    const int pi = precision == MXG_FP64 ? 0 : 1;
    h->tune[pi].near_sym = near_sym != 0;
    h->tune[pi].far_sym = far_sym != 0;
    h->tune[pi].split_rows = split_rows;
    h->tune[pi].matrix_cores = matrix_cores != 0;
    h->jobs_ready[pi] = false;
  });
}

int mxg_h2_matvec(mxg_h2_t h, const double* d_x, double* d_y, long long nrhs, long long ld, int precision) {
  return guarded([&] {
    MXG_REQUIRE(h && d_x && d_y, "null argument");
    MXG_REQUIRE(nrhs >= 1 && ld >= h->n, "bad nrhs/ld");
    MXG_HIP(hipSetDevice(h->ctx->device));
    h2_apply(h, d_x, ld, d_y, ld, nrhs, precision);
    MXG_HIP(hipStreamSynchronize(h->ctx->stream));
  });
}

}  // extern "C"
