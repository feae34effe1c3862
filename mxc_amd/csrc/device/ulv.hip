// ULV factorization and substitution of an H^2 matrix on the device, in FP64
// or FP32 (templated on the factor precision T); an FP32 factorization may
// keep bfloat16 copies of its factors (S) for the substitution, which still
// accumulates in FP32.
//
// Factorization of one level (H2Matrix::factorize, h2matrix.cpp:346-425, and
// its batched formulation device_factorize.cu:56-142), for boxes i with
// M = dims, s = rank, r = M - s, and near pairs (i, j):
//   1  A_ii <- U_i^T A_ii U_i                          (two batched GEMMs)
//   2  rr_i^{-1} = (A_ii[s:, s:])^{-1}                 (batched LU + inverse)
//   3  V_i = [U_i[:, :s]^T ; rr_i^{-1} U_i[:, s:]^T]   (stored for the solves)
//   4  A_ii[s:, :s] <- rr^{-1} A_ii[s:, :s];  A_ii[:s, :s] -= A_ii[:s, s:] A_ii[s:, :s]
//   5  A_ij <- V_i A_ij U_j for j != i, only the ss, sr and rs blocks (the rr
//      block of an off-diagonal pair is never read again)
//   6  A_ij[:s, :s] -= A_ii.sr A_ij.rs + A_ij.sr A_jj.rs   (j != i)
//      A_ii[:s, :s] -= sum_j (A_ij.sr rr_j^{-1}) A_ij.sr^T  (single Schur
//      complement onto the skeleton, no fill-in between non-neighbours)
//   7  A_ij[:s, :s] -> the parent's near block (factorizeCopyNext,
//      h2matrix.cpp:427-446)
// Every step is one batched launch over the whole level; mxg_ulv_create_batch
// shares each launch among the same level steps of several matrices. The
// launches go out with no wait on the host between them: their task lists
// are staged through pinned memory kept on the context, and every level's
// scratch comes from an arena kept on the context, valid until the next
// factorization there (which reuses it instead of allocating).
//
// Substitution (h2matrix.cpp:448-537): forward L -> 0, x_i = V_i y_i,
// z_i = x_i[:s] - sum_j A_ij.sr x_j[s_j:]; backward 0 -> L,
// x_i = [w_i ; x_i[s:] - sum_j A_ij.rs w_j], y_i = U_i x_i. Vectors hand over
// between levels in place (a level's z is its parent's y, the parent's y is
// the child level's w), as in the matvec. Each level's products run as lists
// of block products for up to 8 right-hand sides at once (bgemv.hip), the
// sums of the terms per box in one more launch. At the leaf, the backward
// sums write x_i to a buffer and the products U_i x_i run as their own list,
// where a 256 x 256 bfloat16 factor is one wave's job.
#include <memory>

#include "h2.hpp"

namespace mxg {
namespace {

// Scratch budget (elements) of one off-diagonal chunk of step 5.
constexpr size_t PAIR_BUDGET = size_t(1) << 28;

}  // namespace

// GEMM tasks of one launch; segment indices are local to the list.
template <class T>
struct GemmList {
  std::vector<GemmTask<T>> t;
  std::vector<GemmSeg<T>> g;
  int seg(const GemmSeg<T>& s) {
    g.push_back(s);
    return (int)g.size() - 1;
  }
  // a task over the segments [seg0, end of the list)
  void task(T* C, int ldc, int m, int n, int seg0, double alpha, double beta) {
    t.push_back(GemmTask<T>{C, ldc, m, n, seg0, (int)g.size() - seg0, alpha, beta});
  }
  void append(const GemmList& o) {
    const int base = (int)g.size();
    for (GemmTask<T> x : o.t) {
      x.seg0 += base;
      t.push_back(x);
    }
    g.insert(g.end(), o.g.begin(), o.g.end());
  }
  bool bf16 = false;  // FP32 products with bfloat16 inputs
  void run(hipStream_t s, Stage* st) const { gemm_batched<T>(t, g, s, bf16, st); }
};

// The launches of one level's factorization in their order. Plans of the
// levels of several matrices merge phase by phase, so that one launch of each
// phase serves all of them.
template <class T>
struct LevelPlan {
  GemmList<T> rot1, rot2;                 // 1
  std::vector<LuTask<T>> lu;              // 2
  std::vector<Copy2d<T>> vtop;            // 3
  GemmList<T> vbot;                       // 3
  GemmList<T> rs, ss;                     // 4
  std::vector<Copy2d<T>> rscopy;          // 4
  std::vector<GemmList<T>> offa, offb;    // 5, per chunk of pairs
  GemmList<T> sch1, sch2;                 // 6
  Arena* arena = nullptr;  // the scratch: valid until the arena is rewound
  T* alloc(size_t n) {
    MXG_REQUIRE(arena, "level plan without scratch arena");
    return static_cast<T*>(arena->take(std::max<size_t>(n, 1) * sizeof(T)));
  }
  void merge(LevelPlan&& o) {
    rot1.append(o.rot1);
    rot2.append(o.rot2);
    lu.insert(lu.end(), o.lu.begin(), o.lu.end());
    vtop.insert(vtop.end(), o.vtop.begin(), o.vtop.end());
    vbot.append(o.vbot);
    rs.append(o.rs);
    ss.append(o.ss);
    rscopy.insert(rscopy.end(), o.rscopy.begin(), o.rscopy.end());
    if (offa.size() < o.offa.size()) {
      offa.resize(o.offa.size());
      offb.resize(o.offb.size());
    }
    for (size_t c = 0; c < o.offa.size(); c++) {
      offa[c].append(o.offa[c]);
      offb[c].append(o.offb[c]);
    }
    sch1.append(o.sch1);
    sch2.append(o.sch2);
    sch1.bf16 = sch1.bf16 || o.sch1.bf16;
    sch2.bf16 = sch2.bf16 || o.sch2.bf16;
  }
  // the launches, in order, with their descriptors staged in st: they have
  // not necessarily run when this returns
  void run(hipStream_t s, Stage* st) const {
    rot1.run(s, st);
    rot2.run(s, st);
    lu_inverse_batched<T>(lu, s, st);
    copy2d_run<T>(vtop, s, st);
    vbot.run(s, st);
    rs.run(s, st);
    ss.run(s, st);
    copy2d_run<T>(rscopy, s, st);
    for (size_t c = 0; c < offa.size(); c++) {
      offa[c].run(s, st);
      offb[c].run(s, st);
    }
    sch1.run(s, st);
    sch2.run(s, st);
  }
};

template <class T, class S>
struct UlvImpl;

// Factorizes several H^2 matrices (same precision, same stream) level by
// level from their leaves: sweep t factors level L_p - t of every matrix p
// that has it, each phase of the sweep one launch over all of them, then
// hands their skeleton blocks to the parent levels. The trees may differ in
// depth, box sizes, ranks and near lists; every matrix gets the same results
// as when factorized alone (the tasks are the same, only their launch is
// shared).
//
// No launch waits for the previous one on the host: the descriptor arrays go
// through the context's stage, every level's scratch comes from the
// context's arena and stays valid until the end, and the plans are kept, so
// the host plans level t + 1 while the GPU runs level t.
template <class T, class S>
void factorize_batch(const std::vector<UlvImpl<T, S>*>& fs, hipStream_t s) {
  long long depth = 0;
  for (auto* f : fs) depth = std::max(depth, f->h->L);
  Stage& st = fs.front()->h->ctx->stage;
  st.rewind(s);
  Arena& arena = fs.front()->h->ctx->fscratch;
  arena.rewind(s);
  std::vector<LevelPlan<T>> plans(depth + 1);
  try {
    for (long long t = 0; t <= depth; t++) {
      LevelPlan<T>& P = plans[t];
      std::vector<Copy2d<T>> next;
      for (auto* f : fs) {
        const long long l = f->h->L - t;
        if (l < 0) continue;
        LevelPlan<T> q;
        q.arena = &arena;
        f->plan_level(l, q);
        P.merge(std::move(q));
        if (l > 0) f->plan_copy_next(l, next);
      }
      P.run(s, &st);
      copy2d_run<T>(next, s, &st);
    }
    for (auto* f : fs) f->finish();
  } catch (...) {
    // the launches already issued may still be running on the factors, which
    // the caller frees on an error: wait for them first
    (void)hipStreamSynchronize(s);
    throw;
  }
  MXG_HIP(hipStreamSynchronize(s));
}

struct UlvBase {
  virtual ~UlvBase() = default;
  virtual void factorize() = 0;
  virtual void solve(double* x, long long ldx, long long nrhs) = 0;
  virtual void info(int* out) = 0;
};

// T: factorization and substitution arithmetic; S: storage of the factors the
// substitution streams (T, or bfloat16 with T = float).
template <class T, class S = T>
struct UlvImpl : UlvBase {
  static constexpr bool kCopy = !std::is_same<T, S>::value;
  mxg_h2_s* h;
  bool schur_bf16 = false;  // Schur-complement products with bfloat16 inputs
  struct Lv {
    DevBuf<T> A, U, V, RRinv, X, Z;
    DevBuf<S> As, Us, Vs;  // the factors in storage precision (kCopy only)
    std::vector<long long> offRR;
    DevBuf<int> info;
    int info_sum = 0;
    // substitution task lists: x = V y per box; sr terms A_ij[:s_i, s_j:] x_j[s_j:]
    // (forward) and rs terms A_ij[s_i:, :s_j] w_j (backward) per near pair
    PairList v_pl, fwd_pl, bwd_pl;
    // leaf only: y_i = U_i v_i over the gathered backward vectors (ubuf)
    PairList u_pl;
    DevBuf<long long> d_toffU, d_toffB;
    bool vecM = false;
  };
  std::vector<Lv> lv;
  DevBuf<T> vin, tbuf;  // n and tlen entries per right-hand side
  DevBuf<T> ubuf;       // the leaf's backward vectors [w_i ; x_i[s:] - sum t], n per right-hand side
  long long tlen = 1;

  explicit UlvImpl(mxg_h2_s* hh, bool sbf16 = false) : h(hh), schur_bf16(sbf16) {
    hipStream_t s = h->ctx->stream;
    lv.resize(h->L + 1);
    for (long long l = 0; l <= h->L; l++) {
      Level& H = h->lv[l];
      Lv& F = lv[l];
      MXG_REQUIRE(H.A.n == (size_t)H.offA[H.nrp.back()], "H2 matrix was built without MXG_H2_FACTORABLE");
      F.A.alloc(H.A.n);
      convert(F.A.p, H.A.p, H.A.n, s);
      F.U.alloc(H.Q.n);
      convert(F.U.p, H.Q.p, H.Q.n, s);
      F.V.alloc_zero(H.Q.n, s);
      F.offRR.assign(H.nb + 1, 0);
      for (long long i = 0; i < H.nb; i++) {
        const long long r = H.dims[i] - H.ranks[i];
        F.offRR[i + 1] = F.offRR[i] + r * r;
      }
      F.RRinv.alloc(std::max<long long>(F.offRR.back(), 1));
      F.X.alloc(std::max<long long>(H.lenX, 1) * MAX_RHS);
      F.Z.alloc(std::max<long long>(H.lenZ, 1) * MAX_RHS);
      F.info.alloc_zero(std::max<long long>(H.nb, 1), s);
      if constexpr (kCopy) {
        F.As.alloc(std::max<size_t>(F.A.n, 1));
        F.Us.alloc(std::max<size_t>(F.U.n, 1));
        F.Vs.alloc(std::max<size_t>(F.V.n, 1));
      }
    }
    vin.alloc(h->n * MAX_RHS);
    ubuf.alloc(h->n * MAX_RHS);
    long long tmax = 1;
    for (long long l = 0; l <= h->L; l++) {
      const Level& H = h->lv[l];
      long long u = 0, b = 0;
      for (long long i = 0; i < H.nb; i++) {
        const long long cnt = H.nrp[i + 1] - H.nrp[i];
        u += cnt * H.ranks[i];
        b += cnt * (H.dims[i] - H.ranks[i]);
      }
      tmax = std::max(tmax, std::max(u, b));
    }
    tlen = tmax;
    tbuf.alloc((size_t)tmax * MAX_RHS);
    for (long long l = 0; l <= h->L; l++) build_solve_lists(l, tmax);
    MXG_HIP(hipStreamSynchronize(s));
  }

  static const S* Astore(Lv& F) {
    if constexpr (kCopy) return F.As.p; else return F.A.p;
  }
  static const S* Ustore(Lv& F) {
    if constexpr (kCopy) return F.Us.p; else return F.U.p;
  }
  static const S* Vstore(Lv& F) {
    if constexpr (kCopy) return F.Vs.p; else return F.V.p;
  }

  // Jobs of the substitution at level l (terms in tbuf, tlen elements).
  void build_solve_lists(long long l, long long tlen) {
    hipStream_t s = h->ctx->stream;
    Level& H = h->lv[l];
    Lv& F = lv[l];
    const long long nb = H.nb;
    std::vector<PairTask> vt, ft, bt;
    std::vector<long long> toffU(H.nrp[nb] + 1, 0), toffB(H.nrp[nb] + 1, 0);
    F.vecM = true;
    for (long long i = 0; i < nb; i++) {
      const long long M = H.dims[i];
      if (M % (16 / (long long)sizeof(S))) F.vecM = false;
      if (M > 0) vt.push_back(PairTask{H.offQ[i], H.offX[i], H.offX[i], (int)M, (int)M, (int)M, 0});
      const long long si = H.ranks[i];
      for (long long p = H.nrp[i]; p < H.nrp[i + 1]; p++) {
        const long long j = H.ncol[p];
        const long long Mj = H.dims[j], sj = H.ranks[j];
        toffU[p + 1] = toffU[p] + si;
        toffB[p + 1] = toffB[p] + (M - si);
        if (si > 0) ft.push_back(PairTask{H.offA[p] + sj * M, H.offX[j] + sj, toffU[p], (int)si, (int)(Mj - sj), (int)M, 0});
        if (M - si > 0) bt.push_back(PairTask{H.offA[p] + si, H.offZ[j], toffB[p], (int)(M - si), (int)sj, (int)M, 0});
      }
    }
    const T* y = (l == h->L) ? vin.p : lv[l + 1].Z.p;
    const long long ylen = (l == h->L) ? h->n : std::max<long long>(h->lv[l + 1].lenZ, 1);
    const long long lx = std::max<long long>(H.lenX, 1), lz = std::max<long long>(H.lenZ, 1);
    std::vector<PairJob> jv, jf, jb;
    pair_jobs_append<T, S>(jv, spread(vt), Vstore(F), H.offQ[nb], y, ylen, ylen, F.X.p, lx, lx);
    pair_jobs_append<T, S>(jf, spread(ft), Astore(F), (long long)F.A.n, F.X.p, lx, lx, tbuf.p, tlen, tlen);
    pair_jobs_append<T, S>(jb, spread(bt), Astore(F), (long long)F.A.n, F.Z.p, lz, lz, tbuf.p, tlen, tlen);
    pair_list_upload(F.v_pl, jv, s);
    pair_list_upload(F.fwd_pl, jf, s);
    pair_list_upload(F.bwd_pl, jb, s);
    if (l == h->L) {
      // the leaf's U products as their own launch: a 256 x 256 bfloat16 box
      // factor is then one wave's job rather than one workgroup's
      std::vector<PairJob> ju;
      pair_jobs_append<T, S>(ju, spread(vt), Ustore(F), H.offQ[nb], ubuf.p, lx, lx, vin.p, h->n, h->n);
      pair_list_upload(F.u_pl, ju, s);
    }
    F.d_toffU.upload(toffU, s);
    F.d_toffB.upload(toffB, s);
    MXG_HIP(hipStreamSynchronize(s));
  }

  void factorize() override { factorize_batch<T, S>({this}, h->ctx->stream); }

  // After every level is factored: the bfloat16 copies of the factors and the
  // singular-block counts per level.
  void finish() {
    hipStream_t s = h->ctx->stream;
    if constexpr (kCopy) {
      for (long long l = 0; l <= h->L; l++) {
        Lv& F = lv[l];
        to_bf16(F.As.p, F.A.p, F.A.n, s);
        to_bf16(F.Us.p, F.U.p, F.U.n, s);
        to_bf16(F.Vs.p, F.V.p, F.V.n, s);
      }
    }
    for (long long l = 0; l <= h->L; l++) {
      std::vector<int> inf(h->lv[l].nb);
      MXG_HIP(hipMemcpyAsync(inf.data(), lv[l].info.p, inf.size() * sizeof(int), hipMemcpyDeviceToHost, s));
      MXG_HIP(hipStreamSynchronize(s));
      lv[l].info_sum = 0;
      for (int v : inf) lv[l].info_sum += v;
    }
  }

  // The launches of level l's factorization, appended to P (steps 1-6 of the
  // header; their scratch comes from P's arena).
  void plan_level(long long l, LevelPlan<T>& P) {
    Level& H = h->lv[l];
    Lv& F = lv[l];
    P.sch1.bf16 = P.sch2.bf16 = schur_bf16;
    const long long nb = H.nb;
    auto M_ = [&](long long i) { return (int)H.dims[i]; };
    auto S_ = [&](long long i) { return (int)H.ranks[i]; };
    T* A = F.A.p;
    T* U = F.U.p;
    T* V = F.V.p;
    auto Aij = [&](long long ij) { return A + H.offA[ij]; };
    auto Ui = [&](long long i) { return U + H.offQ[i]; };
    auto Vi = [&](long long i) { return V + H.offQ[i]; };
    auto RRi = [&](long long i) { return F.RRinv.p + F.offRR[i]; };

    // 1: two-sided rotation of the diagonal blocks
    {
      T* T1 = P.alloc(H.offQ[nb]);
      for (long long i = 0; i < nb; i++) {
        const int M = M_(i);
        if (M == 0) continue;
        int g = P.rot1.seg(GemmSeg<T>{Aij(H.diag[i]), Ui(i), M, M, M, 0, 0, 0, 0});
        P.rot1.task(T1 + H.offQ[i], M, M, M, g, 1., 0.);
        g = P.rot2.seg(GemmSeg<T>{Ui(i), T1 + H.offQ[i], M, M, M, 1, 0, 0, 0});
        P.rot2.task(Aij(H.diag[i]), M, M, M, g, 1., 0.);
      }
    }
    // 2: rr inverses
    {
      size_t wsz = 0;
      for (long long i = 0; i < nb; i++) {
        const long long r = M_(i) - S_(i);
        wsz += (size_t)r * r + 2 * r;
      }
      T* work = P.alloc(wsz);
      size_t wo = 0;
      for (long long i = 0; i < nb; i++) {
        const int M = M_(i), sI = S_(i), r = M - sI;
        if (r == 0) continue;
        P.lu.push_back(LuTask<T>{Aij(H.diag[i]) + sI + (long long)sI * M, RRi(i), work + wo, F.info.p + i, r, M});
        wo += (size_t)r * r + 2 * r;
      }
    }
    // 3: V = [U[:, :s]^T ; rr^{-1} U[:, s:]^T]
    for (long long i = 0; i < nb; i++) {
      const int M = M_(i), sI = S_(i), r = M - sI;
      if (sI > 0) P.vtop.push_back(Copy2d<T>{Ui(i), Vi(i), M, sI, M, M, 1, 0});
      if (r == 0) continue;
      const int g = P.vbot.seg(GemmSeg<T>{RRi(i), Ui(i) + (long long)sI * M, r, M, r, 0, 1, 0, 0});
      P.vbot.task(Vi(i) + sI, M, r, M, g, 1., 0.);
    }
    // 4: rs <- rr^{-1} rs; ss -= sr rs
    {
      size_t tsz = 0;
      for (long long i = 0; i < nb; i++) tsz += (size_t)(M_(i) - S_(i)) * S_(i);
      T* T2 = P.alloc(tsz);
      size_t to = 0;
      for (long long i = 0; i < nb; i++) {
        const int M = M_(i), sI = S_(i), r = M - sI;
        if (r == 0 || sI == 0) continue;
        T* Aii = Aij(H.diag[i]);
        int g = P.rs.seg(GemmSeg<T>{RRi(i), Aii + sI, r, M, r, 0, 0, 0, 0});
        P.rs.task(T2 + to, r, r, sI, g, 1., 0.);
        g = P.ss.seg(GemmSeg<T>{Aii + (long long)sI * M, T2 + to, M, r, r, 0, 0, 0, 0});
        P.ss.task(Aii, M, sI, sI, g, -1., 1.);
        P.rscopy.push_back(Copy2d<T>{T2 + to, Aii + sI, r, sI, r, M, 0, 0});
        to += (size_t)r * sI;
      }
    }
    // 5: off-diagonal rotations, in chunks of pairs
    {
      std::vector<long long> pairs;
      for (long long i = 0; i < nb; i++)
        for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++)
          if (ij != H.diag[i]) pairs.push_back(ij);
      std::vector<long long> rowof(H.nrp.back());
      for (long long i = 0; i < nb; i++)
        for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++) rowof[ij] = i;
      size_t p0 = 0, chunk = 0;
      while (p0 < pairs.size()) {
        size_t p1 = p0, need = 0;
        std::vector<size_t> off;
        while (p1 < pairs.size()) {
          const long long ij = pairs[p1], i = rowof[ij], j = H.ncol[ij];
          const size_t e = (size_t)S_(i) * M_(j) + (size_t)M_(i) * S_(j);
          if (p1 > p0 && need + e > PAIR_BUDGET) break;
          off.push_back(need);
          need += e;
          p1++;
        }
        T* Pbuf = P.alloc(need);
        if (P.offa.size() <= chunk) {
          P.offa.resize(chunk + 1);
          P.offb.resize(chunk + 1);
        }
        GemmList<T>& ga = P.offa[chunk];
        GemmList<T>& gb = P.offb[chunk];
        for (size_t q = p0; q < p1; q++) {
          const long long ij = pairs[q], i = rowof[ij], j = H.ncol[ij];
          const int M = M_(i), N = M_(j), si = S_(i), sj = S_(j), ri = M - si;
          T* Pq = Pbuf + off[q - p0];            // s_i x N
          T* Qq = Pq + (size_t)si * N;           // M x s_j
          if (si > 0) {
            int g = ga.seg(GemmSeg<T>{Ui(i), Aij(ij), M, M, M, 1, 0, 0, 0});
            ga.task(Pq, si, si, N, g, 1., 0.);
            g = gb.seg(GemmSeg<T>{Pq, Ui(j), si, N, N, 0, 0, 0, 0});
            gb.task(Aij(ij), M, si, N, g, 1., 0.);
          }
          if (sj > 0 && ri > 0) {
            int g = ga.seg(GemmSeg<T>{Aij(ij), Ui(j), M, N, N, 0, 0, 0, 0});
            ga.task(Qq, M, M, sj, g, 1., 0.);
            g = gb.seg(GemmSeg<T>{Vi(i) + si, Qq, M, M, M, 0, 0, 0, 0});
            gb.task(Aij(ij) + si, M, ri, sj, g, 1., 0.);
          }
        }
        p0 = p1;
        chunk++;
      }
    }
    // 6: Schur complements onto the skeleton
    {
      size_t t3 = 0;
      for (long long i = 0; i < nb; i++)
        for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++)
          if (ij != H.diag[i]) t3 += (size_t)S_(i) * (M_(H.ncol[ij]) - S_(H.ncol[ij]));
      T* T3 = P.alloc(t3);
      std::vector<size_t> t3off(H.nrp.back(), 0);
      size_t to = 0;
      for (long long i = 0; i < nb; i++) {
        const int M = M_(i), si = S_(i), ri = M - si;
        T* Aii = Aij(H.diag[i]);
        for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++) {
          if (ij == H.diag[i]) continue;
          const long long j = H.ncol[ij];
          const int N = M_(j), sj = S_(j), rj = N - sj;
          T* Aq = Aij(ij);
          if (si > 0 && sj > 0) {
            const int g = (int)P.sch1.g.size();
            if (ri > 0) P.sch1.seg(GemmSeg<T>{Aii + (long long)si * M, Aq + si, M, M, ri, 0, 0, 0, 0});
            if (rj > 0) P.sch1.seg(GemmSeg<T>{Aq + (long long)sj * M, Aij(H.diag[j]) + sj, M, N, rj, 0, 0, 0, 0});
            P.sch1.task(Aq, M, si, sj, g, -1., 1.);
          }
          if (si > 0 && rj > 0) {
            t3off[ij] = to;
            const int g = P.sch1.seg(GemmSeg<T>{Aq + (long long)sj * M, RRi(j), M, rj, rj, 0, 0, 0, 0});
            P.sch1.task(T3 + to, si, si, rj, g, 1., 0.);
            to += (size_t)si * rj;
          }
        }
      }
      for (long long i = 0; i < nb; i++) {
        const int M = M_(i), si = S_(i);
        if (si == 0) continue;
        const int g = (int)P.sch2.g.size();
        for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++) {
          if (ij == H.diag[i]) continue;
          const long long j = H.ncol[ij];
          const int sj = S_(j), rj = M_(j) - sj;
          if (rj == 0) continue;
          P.sch2.seg(GemmSeg<T>{T3 + t3off[ij], Aij(ij) + (long long)sj * M, si, M, rj, 0, 1, 0, 0});
        }
        if ((int)P.sch2.g.size() > g) P.sch2.task(Aij(H.diag[i]), M, si, si, g, -1., 1.);
      }
    }
  }

  // 7: the skeleton blocks of level l's near pairs into the parent level's
  // near blocks (factorizeCopyNext)
  void plan_copy_next(long long l, std::vector<Copy2d<T>>& out) {
    Level& H = h->lv[l];
    Lv& F = lv[l];
    Lv& Pl = lv[l - 1];
    for (long long i = 0; i < H.nb; i++)
      for (long long ij = H.nrp[i]; ij < H.nrp[i + 1]; ij++) {
        const long long j = H.ncol[ij];
        const int si = (int)H.ranks[i], sj = (int)H.ranks[j];
        if (si == 0 || sj == 0 || H.up_off[ij] < 0) continue;
        out.push_back(Copy2d<T>{F.A.p + H.offA[ij], Pl.A.p + H.up_off[ij], si, sj, (int)H.dims[i], H.up_ld[ij], 0, 0});
      }
  }

  void solve(double* x, long long ldx, long long nrhs) override {
    hipStream_t s = h->ctx->stream;
    const long long L = h->L, n = h->n;
    T* v = vin.p;
    for (long long c0 = 0; c0 < nrhs; c0 += MAX_RHS) {
      const int nr = (int)std::min<long long>(MAX_RHS, nrhs - c0);
      convert2d(v, n, x + c0 * ldx, ldx, n, nr, s);
      ProfScope whole(h->ctx, MXG_PROF_ULV_SOLVE);
      for (long long l = L; l >= 0; l--) {
        Level& H = h->lv[l];
        Lv& F = lv[l];
        const long long lx = std::max<long long>(H.lenX, 1), lz = std::max<long long>(H.lenZ, 1);
        ProfScope prof(h->ctx, l == L ? MXG_PROF_ULV_FWD_LEAF : -1);
        pair_gemv<T, S>(F.v_pl, nr, s);
        if (H.lenZ > 0) {
          pair_gemv<T, S>(F.fwd_pl, nr, s);
          comb_fwd<T>(H, F.d_toffU.p, F.X.p, lx, tbuf.p, tlen, F.Z.p, lz, nr, s);
        }
      }
      for (long long l = 0; l <= L; l++) {
        Level& H = h->lv[l];
        Lv& F = lv[l];
        const long long lx = std::max<long long>(H.lenX, 1), lz = std::max<long long>(H.lenZ, 1);
        // w of level l is the output of level l - 1 (stored in this level's Z buffer)
        T* out = (l == L) ? v : lv[l + 1].Z.p;
        const long long lo = (l == L) ? n : std::max<long long>(h->lv[l + 1].lenZ, 1);
        ProfScope prof(h->ctx, l == L ? MXG_PROF_ULV_BWD_LEAF : -1);
        pair_gemv<T, S>(F.bwd_pl, nr, s);
        if (l == L) {
          bwd_gather<T>(H, F.d_toffB.p, F.X.p, lx, F.Z.p, lz, tbuf.p, tlen, ubuf.p, lx, nr, s);
          pair_gemv<T, S>(F.u_pl, nr, s);
        } else {
          comb_bwd<T, S>(H, F.vecM, F.d_toffB.p, Ustore(F), F.X.p, lx, F.Z.p, lz, tbuf.p, tlen, out, lo, nr, s);
        }
      }
      whole.end();
      convert2d(x + c0 * ldx, ldx, v, n, n, nr, s);
    }
  }

  void info(int* out) override {
    for (long long l = 0; l <= h->L; l++) out[l] = lv[l].info_sum;
  }
};

}  // namespace mxg

struct mxg_ulv_s {
  mxg_h2_s* h = nullptr;
  int precision = MXG_FP64;
  double seconds = 0.;
  std::unique_ptr<mxg::UlvBase> impl;
};

namespace mxg {
void ulv_apply(mxg_ulv_s* f, double* x, long long ldx, long long nrhs) { f->impl->solve(x, ldx, nrhs); }
}  // namespace mxg

using namespace mxg;

extern "C" {

int mxg_ulv_create(mxg_h2_t h, int precision, mxg_ulv_t* out) {
  return guarded([&] {
    MXG_REQUIRE(h && out, "null argument");
    for (long long l = 0; l <= h->L; l++) MXG_REQUIRE(h->lv[l].built, "H2 matrix not fully constructed");
    MXG_REQUIRE(h->flags & MXG_H2_FACTORABLE, "H2 matrix was created without MXG_H2_FACTORABLE");
    MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32 || precision == MXG_BF16 || precision == MXG_BF16_SCHUR,
                "unknown precision");
    MXG_HIP(hipSetDevice(h->ctx->device));
    auto f = std::make_unique<mxg_ulv_s>();
    f->h = h;
    f->precision = precision;
    if (precision == MXG_FP64)
      f->impl = std::make_unique<UlvImpl<double>>(h);
    else if (precision == MXG_FP32)
      f->impl = std::make_unique<UlvImpl<float>>(h);
    else
      f->impl = std::make_unique<UlvImpl<float, bf16_t>>(h, precision == MXG_BF16_SCHUR);
    // device time of the factorization proper (every level's launches and the
    // host work that issues them), bracketed by events on the context stream
    hipEvent_t e0, e1;
    MXG_HIP(hipEventCreate(&e0));
    MXG_HIP(hipEventCreate(&e1));
    MXG_HIP(hipEventRecord(e0, h->ctx->stream));
    f->impl->factorize();
    MXG_HIP(hipEventRecord(e1, h->ctx->stream));
    MXG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MXG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    f->seconds = 1e-3 * ms;
    *out = f.release();
  });
}

int mxg_ulv_create_batch(const mxg_h2_t* hs, long long count, int precision, mxg_ulv_t* out) {
  return guarded([&] {
    MXG_REQUIRE(hs && out && count >= 1, "null argument or empty batch");
    MXG_REQUIRE(precision == MXG_FP64 || precision == MXG_FP32 || precision == MXG_BF16 || precision == MXG_BF16_SCHUR,
                "unknown precision");
    mxg_ctx_t ctx = hs[0] ? hs[0]->ctx : nullptr;
    for (long long p = 0; p < count; p++) {
      MXG_REQUIRE(hs[p], "null matrix in the batch");
      MXG_REQUIRE(hs[p]->ctx == ctx, "every matrix of a batch must live on one context");
      MXG_REQUIRE(hs[p]->flags & MXG_H2_FACTORABLE, "H2 matrix was created without MXG_H2_FACTORABLE");
      for (long long l = 0; l <= hs[p]->L; l++) MXG_REQUIRE(hs[p]->lv[l].built, "H2 matrix not fully constructed");
      for (long long q = 0; q < p; q++) MXG_REQUIRE(hs[q] != hs[p], "a matrix appears twice in the batch");
    }
    MXG_HIP(hipSetDevice(ctx->device));
    std::vector<std::unique_ptr<mxg_ulv_s>> fs(count);
    auto run = [&](auto* tag) {
      using Impl = std::remove_pointer_t<decltype(tag)>;
      std::vector<Impl*> impls;
      for (long long p = 0; p < count; p++) {
        fs[p] = std::make_unique<mxg_ulv_s>();
        fs[p]->h = hs[p];
        fs[p]->precision = precision;
        auto impl = std::make_unique<Impl>(hs[p], precision == MXG_BF16_SCHUR);
        impls.push_back(impl.get());
        fs[p]->impl = std::move(impl);
      }
      hipEvent_t e0, e1;
      MXG_HIP(hipEventCreate(&e0));
      MXG_HIP(hipEventCreate(&e1));
      MXG_HIP(hipEventRecord(e0, ctx->stream));
      factorize_batch(impls, ctx->stream);
      MXG_HIP(hipEventRecord(e1, ctx->stream));
      MXG_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MXG_HIP(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      for (auto& f : fs) f->seconds = 1e-3 * ms;
    };
    if (precision == MXG_FP64)
      run((UlvImpl<double>*)nullptr);
    else if (precision == MXG_FP32)
      run((UlvImpl<float>*)nullptr);
    else
      run((UlvImpl<float, bf16_t>*)nullptr);
    for (long long p = 0; p < count; p++) out[p] = fs[p].release();
  });
}

int mxg_ulv_destroy(mxg_ulv_t f) {
  return guarded([&] {
    if (!f) return;
    (void)hipStreamSynchronize(f->h->ctx->stream);
    delete f;
  });
}

int mxg_ulv_info(mxg_ulv_t f, int* info_per_level) {
  return guarded([&] {
    MXG_REQUIRE(f && info_per_level, "null argument");
    f->impl->info(info_per_level);
  });
}

int mxg_ulv_solve(mxg_ulv_t f, double* d_x, long long nrhs, long long ld) {
  return guarded([&] {
    MXG_REQUIRE(f && d_x, "null argument");
    MXG_REQUIRE(nrhs >= 1 && ld >= f->h->n, "bad nrhs/ld");
    MXG_HIP(hipSetDevice(f->h->ctx->device));
    f->impl->solve(d_x, ld, nrhs);
    MXG_HIP(hipStreamSynchronize(f->h->ctx->stream));
  });
}

double mxg_ulv_factor_seconds(mxg_ulv_t f) { return f ? f->seconds : -1.; }

}  // extern "C"
