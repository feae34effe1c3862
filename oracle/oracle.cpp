// CPU oracle: FP64 restatement of the reference H^2 construct / matvec / ULV
// factorization / substitution / GMRES path (delta-leader/mxc). Test
// infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg are its only callers, always as the checker.
//
// File:line citations refer to the reference tree. Complex arithmetic of the
// reference is restated in real FP64 (conj = identity, adjoint = transpose),
// which is exact for the real-valued kernels on this path.
//
// Builder-defined semantics (not in the reference):
//   * orc_sphere_rnd: points uniform on the unit sphere (no reference generator).
//   * orc_far_sample: the `sampling` cap on far bodies per box (the reference
//     HEAD uses every far body; sampling <= 0 restores that).
#include "oracle.h"
#include "dense.hpp"

#include <array>
#include <numeric>
#include <random>
#include <utility>
#include <omp.h>
#include <cstdlib>

using orc::gemm;

namespace {

struct Kern {
  int kid;
  double diag, alpha;
  // kernel.hpp:28-61 (Laplace3D, Yukawa3D, Gaussian) with d == 0 mapped to `diag`.
  double operator()(double d) const {
    if (kid == ORC_GAUSSIAN) return d == 0. ? 1.0001 : std::exp(-alpha * d * d);
    if (d == 0.) return diag;
    if (kid == ORC_YUKAWA) return std::exp(-alpha * d) / d;
    return 1. / d;
  }
};

// kernel.cpp:71-85 gen_matrix: A[iy + ix*m] = K(hypot(bi[iy] - bj[ix])).
void gen_block(const Kern& k, long m, long n, const double* bi, const double* bj, double* A, long lda) {
  for (long ix = 0; ix < n; ix++)
    for (long iy = 0; iy < m; iy++) {
      double d = std::hypot(bi[3 * iy] - bj[3 * ix], bi[3 * iy + 1] - bj[3 * ix + 1], bi[3 * iy + 2] - bj[3 * ix + 2]);
      A[iy + ix * lda] = k(d);
    }
}

struct Cell {
  long child[2] = {-1, -1};
  long body[2] = {-1, -1};
  double R[3] = {0, 0, 0}, C[3] = {0, 0, 0};
};

// build_tree.cpp:8-22
void get_bounds(const double* b, long n, double* R, double* C) {
  for (int d = 0; d < 3; d++) {
    double lo = b[d], hi = b[d];
    for (long i = 1; i < n; i++) {
      lo = std::min(lo, b[3 * i + d]);
      hi = std::max(hi, b[3 * i + d]);
    }
    C[d] = (lo + hi) * 0.5;
    R[d] = (lo == hi && lo == 0.) ? 0. : ((hi - lo) * 0.5 + 1.e-8);
  }
}

// build_tree.cpp:67-99
void build_tree(std::vector<Cell>& cells, double* bodies, long n, long levels) {
  cells[0].body[0] = 0;
  cells[0].body[1] = n;
  get_bounds(bodies, n, cells[0].R, cells[0].C);
  long nleaf = 1L << levels;
  for (long i = 0; i < nleaf - 1; i++) {
    Cell& ci = cells[i];
    int sdim = 0;
    for (int d = 1; d < 3; d++)
      if (ci.R[d] > ci.R[sdim]) sdim = d;
    auto* b3 = reinterpret_cast<std::array<double, 3>*>(bodies + 3 * ci.body[0]);
    auto* e3 = reinterpret_cast<std::array<double, 3>*>(bodies + 3 * ci.body[1]);
    std::sort(b3, e3, [=](const std::array<double, 3>& x, const std::array<double, 3>& y) { return x[sdim] < y[sdim]; });
    long c0 = 2 * i + 1;
    ci.child[0] = c0;
    ci.child[1] = c0 + 2;
    long loc = ci.body[0] + (ci.body[1] - ci.body[0]) / 2;
    cells[c0].body[0] = ci.body[0];
    cells[c0].body[1] = loc;
    cells[c0 + 1].body[0] = loc;
    cells[c0 + 1].body[1] = ci.body[1];
    get_bounds(bodies + 3 * ci.body[0], loc - ci.body[0], cells[c0].R, cells[c0].C);
    get_bounds(bodies + 3 * loc, ci.body[1] - loc, cells[c0 + 1].R, cells[c0 + 1].C);
  }
}

struct CSR {
  std::vector<long> rp, col;
};

// build_tree.cpp:24-39 (squared-distance admissibility, recursion into inadmissible pairs)
void get_list(bool far, std::vector<std::pair<long, long>>& rels, const std::vector<Cell>& c, long i, long j, double theta) {
  double dC = 0., r1 = 0., r2 = 0.;
  for (int d = 0; d < 3; d++) {
    double t = c[i].C[d] - c[j].C[d];
    dC += t * t;
    r1 += c[i].R[d] * c[i].R[d];
    r2 += c[j].R[d] * c[j].R[d];
  }
  bool admis = dC > theta * (r1 + r2);
  if (admis ? far : !far) rels.emplace_back(i, j);
  if (!admis && c[i].child[0] >= 0 && c[j].child[0] >= 0)
    for (long a = c[i].child[0]; a < c[i].child[1]; a++)
      for (long b = c[j].child[0]; b < c[j].child[1]; b++) get_list(far, rels, c, a, b, theta);
}

// build_tree.cpp:41-58
CSR make_csr(bool far, const std::vector<Cell>& cells, double theta) {
  std::vector<std::pair<long, long>> lil;
  get_list(far, lil, cells, 0, 0, theta);
  std::sort(lil.begin(), lil.end());
  CSR r;
  long nc = cells.size();
  r.rp.assign(nc + 1, 0);
  r.col.resize(lil.size());
  for (size_t t = 0; t < lil.size(); t++) {
    r.col[t] = lil[t].second;
    r.rp[lil[t].first + 1]++;
  }
  for (long i = 0; i < nc; i++) r.rp[i + 1] += r.rp[i];
  return r;
}

// build_tree.cpp:60-65 / h2matrix.cpp:122-127
long lookup(const std::vector<long>& rp, const std::vector<long>& col, long i, long j) {
  if (i < 0 || (long)rp.size() <= i + 1) return -1;
  for (long k = rp[i]; k < rp[i + 1]; k++)
    if (col[k] == j) return k;
  return -1;
}

// Far-body list of cell `cell` (h-matrix.cpp:121-127,160-165: the parent's list
// followed by the bodies of the cell's own basis-far cells, in CSR order), as
// global body indices, with the `sampling` cap.
// Sampling rule (builder-defined): when the list is longer than `sampling`,
// its bodies are grouped into distance shells from the matrix's theta lists:
// shell 0 = the cell's near cells except itself (only when `with_near`, i.e.
// the basis list holds every body outside the cell), then the far cells of
// the cell, of its parent, ..., of its level-1 ancestor.
//  * far lists (the accurate operator): the R = max(3, ceil(L / 2)) closest
//    shells are kept whole (L = tree levels), every farther shell keeps
//    min(n, sampling) bodies;
//  * lists of every outside body (the rank-capped preconditioner): a total
//    of `sampling` bodies, quotas halving from the closest shell outward.
// A shell of n bodies with quota q keeps positions floor(t*n/q),
// t = 0..q-1; shells are emitted farthest first. sampling <= 0 or a list no
// longer than sampling keeps the full list in reference order.
long resolved_shells(long levels) { return std::max(3L, (levels + 1) / 2); }
std::vector<long> far_sample(const std::vector<Cell>& cells, const CSR& bfar, const CSR& near, const CSR& far, bool with_near,
                             long cell, long sampling) {
  std::vector<long> chain;
  for (long c = cell; c > 0; c = (c - 1) / 2) chain.push_back(c);
  std::reverse(chain.begin(), chain.end());  // level 1 .. own level
  long total = 0;
  for (long c : chain)
    for (long k = bfar.rp[c]; k < bfar.rp[c + 1]; k++) total += cells[bfar.col[k]].body[1] - cells[bfar.col[k]].body[0];
  std::vector<long> out;
  if (sampling <= 0 || total <= sampling) {
    for (long c : chain)
      for (long k = bfar.rp[c]; k < bfar.rp[c + 1]; k++)
        for (long b = cells[bfar.col[k]].body[0]; b < cells[bfar.col[k]].body[1]; b++) out.push_back(b);
    return out;
  }
  // shells, closest first: lists of far cells
  std::vector<std::vector<long>> shells;
  if (with_near) {
    shells.emplace_back();
    for (long k = near.rp[cell]; k < near.rp[cell + 1]; k++)
      if (near.col[k] != cell) shells.back().push_back(near.col[k]);
  }
  for (long a = (long)chain.size() - 1; a >= 0; a--) {
    shells.emplace_back();
    for (long k = far.rp[chain[a]]; k < far.rp[chain[a] + 1]; k++) shells.back().push_back(far.col[k]);
  }
  long nsh = shells.size();
  std::vector<long> segn(nsh, 0), q(nsh, 0);
  for (long a = 0; a < nsh; a++)
    for (long f : shells[a]) segn[a] += cells[f].body[1] - cells[f].body[0];
  if (with_near) {
    // halving quotas from the closest shell outward, the last shell takes the
    // rest; unused quota is handed back from the closest shell outward
    long rem = sampling;
    for (long a = 0; a < nsh; a++) {
      long share = (a == nsh - 1) ? rem : (rem + 1) / 2;
      q[a] = std::min(segn[a], share);
      rem -= q[a];
    }
    for (long a = 0; a < nsh && rem > 0; a++) {
      long extra = std::min(segn[a] - q[a], rem);
      q[a] += extra;
      rem -= extra;
    }
  } else {
    long levels = 0;
    while ((2L << levels) - 1 < (long)cells.size()) levels++;
    const long resolved = resolved_shells(levels);
    for (long a = 0; a < nsh; a++) q[a] = (a < resolved) ? segn[a] : std::min(segn[a], sampling);
  }
  for (long a = nsh - 1; a >= 0; a--) {
    if (q[a] == 0) continue;
    std::vector<std::pair<long, long>> ranges;
    for (long f : shells[a]) ranges.emplace_back(cells[f].body[0], cells[f].body[1]);
    long n = segn[a], qa = q[a];
    long r = 0, base = 0;
    for (long t = 0; t < qa; t++) {
      long pos = (long)(((__int128)t * n) / qa);
      while (pos >= base + (ranges[r].second - ranges[r].first)) {
        base += ranges[r].second - ranges[r].first;
        r++;
      }
      out.push_back(ranges[r].first + (pos - base));
    }
  }
  return out;
}

struct Level {
  long nb = 0, gbegin = 0;
  std::vector<long> dims, ranks;
  std::vector<long> arp, acol, crp, ccol;
  std::vector<long> offA, offC, offQ, offX, offZ;
  std::vector<double> A, C, Q, R, X, Y, Z, W, S;
  std::vector<long> NA, upperStride;
  long lenX = 0, lenZ = 0;
  int info = 0;
};

}  // namespace

struct orc_problem {
  long n = 0, leaf = 0, levels = 0;
  std::vector<double> bodies;
  std::vector<Cell> cells;
};

struct orc_h2 {
  const orc_problem* p = nullptr;
  Kern k{};
  long L = 0;
  double theta = 0., tol = 0.;
  long rank = 0, leveled_rank = 0, sampling = 0;
  bool orth = false;
  bool hss = false;  // basis far lists are the sibling lists (every body outside the box)
  int nthreads = 1;
  CSR near, far, basisfar;
  std::vector<Level> lv;
  long maxrank(long l) const { return rank > 0 ? (L - l) * leveled_rank + rank : 0; }  // solver.cpp:29
};

namespace {

// Row reduction of the far-field block (builder-defined): when a box has more
// far rows N than columns M, the reference's R factor of the N x M block is
// replaced by the M x M sketch Omega G with Rademacher entries
// Omega(i, t) from a splitmix64-style mix of (t, i / 64). CPQR of R0 equals
// CPQR of G in exact arithmetic; the sketch is a subspace embedding of G's
// columns, so CPQR of Omega G is a randomized interpolative decomposition of
// G (Halko, Martinsson, Tropp 2011). g_sketch = false restores the QR.
bool g_sketch = true;

// 64 sketch signs per word: Omega(i, t) = -1 when bit (i mod 64) of
// sketch_word(t, i / 64) is set.
inline unsigned long long sketch_word(unsigned long long t, unsigned long long b) {
  unsigned long long z = t * 0x9E3779B97F4A7C15ull + b * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double rademacher(unsigned long long t, unsigned long long i) {
  return ((sketch_word(t, i >> 6) >> (i & 63)) & 1) ? -1. : 1.;
}

// h2matrix.cpp:13-63 compute_basis, generalized to (tol, maxrank, orth):
//   reference tolerance mode: tol = eps, maxrank = 0, orth = false;
//   reference fixed-rank mode: tol = 0, maxrank = rank_func(l), orth = true.
long compute_basis(const Kern& k, double tol, long maxrank, bool orth, long M, long N, double* X, const double* F,
                   double* Qo, double* Ro) {
  long K = std::min(M, N), rank = 0;
  if (K <= 0) return 0;
  std::vector<double> RX((size_t)K * M, 0.);
  if (K < N && g_sketch) {
    // Y = Omega K(F, X) (M x M): row reduction by a Rademacher sketch, in
    // chunks of far rows (replaces the R factor of h2matrix.cpp:78-79)
    const long CH = 2048;
    std::vector<double> XF((size_t)CH * M), Om((size_t)M * CH);
    for (long t0 = 0; t0 < N; t0 += CH) {
      long c = std::min(CH, N - t0);
      gen_block(k, c, M, F + 3 * t0, X, XF.data(), c);
      for (long t = 0; t < c; t++)
        for (long i = 0; i < M; i++) Om[i + t * M] = rademacher(t0 + t, i);
      gemm('N', 'N', M, M, c, 1., Om.data(), M, XF.data(), c, t0 == 0 ? 0. : 1., RX.data(), K);
    }
  } else if (K < N) {
    std::vector<double> XF((size_t)N * M);
    gen_block(k, N, M, F, X, XF.data(), N);
    std::vector<double> tau(std::min(N, M));
    orc::householder_qr(N, M, XF.data(), N, tau.data());
    for (long j = 0; j < M; j++)
      for (long i = 0; i <= std::min(j, K - 1); i++) RX[i + j * K] = XF[i + j * N];
  } else {
    gen_block(k, N, M, F, X, RX.data(), K);
  }
  orc::CPQRInfo info = orc::cpqr(K, M, RX.data(), K);
  rank = (tol > 0.) ? orc::cpqr_rank(info, RX.data(), K, tol) : K;
  if (maxrank > 0) rank = std::min(rank, maxrank);

  if (0 < rank && rank < M) {
    // C.topRows(rank) = R.topRows(rank); solve R11 T = R12 (h2matrix.cpp:93-95)
    std::vector<double> T((size_t)rank * (M - rank));
    for (long j = 0; j < M - rank; j++)
      for (long i = 0; i < rank; i++) T[i + j * rank] = RX[i + (rank + j) * K];
    orc::trsm_upper(rank, M - rank, RX.data(), K, T.data(), rank);
    // permute the body coordinates (h2matrix.cpp:97-98)
    std::vector<double> Xp(3 * M);
    for (long j = 0; j < M; j++)
      for (int d = 0; d < 3; d++) Xp[3 * j + d] = X[3 * info.perm[j] + d];
    std::copy(Xp.begin(), Xp.end(), X);
    // W = P [I; T^T]  (M x rank), row perm[j] holds row j of [I; T^T]
    std::vector<double> Wm((size_t)M * rank, 0.);
    for (long j = 0; j < M; j++)
      for (long c = 0; c < rank; c++) {
        double v = (j < rank) ? (j == c ? 1. : 0.) : T[c + (j - rank) * rank];
        Wm[info.perm[j] + c * M] = v;
      }
    if (orth) {
      // RX = triu(Q_in) * W; QR; Q = full Q, R[:rank,:rank] = triu (h2matrix.cpp:100-105)
      std::vector<double> Qin(Qo, Qo + M * M), P((size_t)M * rank, 0.);
      for (long j = 0; j < M; j++)
        for (long i = j + 1; i < M; i++) Qin[i + j * M] = 0.;
      gemm('N', 'N', M, rank, M, 1., Qin.data(), M, Wm.data(), M, 0., P.data(), M);
      std::vector<double> tau(rank);
      orc::householder_qr(M, rank, P.data(), M, tau.data());
      orc::form_q(M, rank, P.data(), M, tau.data(), Qo, M);
      std::fill(Ro, Ro + M * M, 0.);
      for (long j = 0; j < rank; j++)
        for (long i = 0; i <= j; i++) Ro[i + j * M] = P[i + j * M];
    } else {
      // Q = [W, 0], R = I_rank (h2matrix.cpp:107-111)
      std::fill(Qo, Qo + M * M, 0.);
      std::copy(Wm.begin(), Wm.end(), Qo);
      std::fill(Ro, Ro + M * M, 0.);
      for (long i = 0; i < rank; i++) Ro[i + i * M] = 1.;
    }
  } else {
    // C = triu(Q_in); Q = I (h2matrix.cpp:114-116)
    for (long j = 0; j < M; j++)
      for (long i = 0; i < M; i++) {
        Ro[i + j * M] = (i <= j) ? Qo[i + j * M] : 0.;
        Qo[i + j * M] = (i == j) ? 1. : 0.;
      }
  }
  return rank;
}

void local_csr(const CSR& g, long gbegin, long nb, std::vector<long>& rp, std::vector<long>& col) {
  rp.resize(nb + 1);
  long base = g.rp[gbegin];
  for (long i = 0; i <= nb; i++) rp[i] = g.rp[gbegin + i] - base;
  col.resize(rp[nb]);
  for (long t = 0; t < rp[nb]; t++) col[t] = g.col[base + t] - gbegin;
}

// h2matrix.cpp:72-249 H2Matrix::construct at one rank (comm-mpi.cpp:49-186 with P = 1).
void construct_level(orc_h2& h, long l) {
  const orc_problem& p = *h.p;
  Level& Lv = h.lv[l];
  Level* low = (l < h.L) ? &h.lv[l + 1] : nullptr;
  Lv.nb = 1L << l;
  Lv.gbegin = Lv.nb - 1;
  long nb = Lv.nb;
  local_csr(h.near, Lv.gbegin, nb, Lv.arp, Lv.acol);
  local_csr(h.far, Lv.gbegin, nb, Lv.crp, Lv.ccol);
  Lv.dims.assign(nb, 0);
  Lv.ranks.assign(nb, 0);
  Lv.upperStride.assign(nb, 0);
  for (long i = 0; i < nb; i++) {
    const Cell& c = p.cells[Lv.gbegin + i];
    Lv.dims[i] = low ? (low->ranks[2 * i] + low->ranks[2 * i + 1]) : (c.body[1] - c.body[0]);
  }
  Lv.offX.assign(nb + 1, 0);
  Lv.offQ.assign(nb + 1, 0);
  for (long i = 0; i < nb; i++) {
    Lv.offX[i + 1] = Lv.offX[i] + Lv.dims[i];
    Lv.offQ[i + 1] = Lv.offQ[i] + Lv.dims[i] * Lv.dims[i];
  }
  Lv.lenX = Lv.offX[nb];
  long nnzA = Lv.arp[nb];
  Lv.offA.assign(nnzA + 1, 0);
  for (long i = 0; i < nb; i++)
    for (long t = Lv.arp[i]; t < Lv.arp[i + 1]; t++) Lv.offA[t + 1] = Lv.offA[t] + Lv.dims[i] * Lv.dims[Lv.acol[t]];
  Lv.NA.assign(nnzA, -1);
  Lv.X.assign(Lv.lenX, 0.);
  Lv.Y.assign(Lv.lenX, 0.);
  Lv.Q.assign(Lv.offQ[nb], 0.);
  Lv.R.assign(Lv.offQ[nb], 0.);
  Lv.S.assign(3 * Lv.lenX, 0.);
  Lv.A.assign(Lv.offA[nnzA], 0.);

  if (Lv.lenX > 0) {
    for (long i = 0; i < nb; i++) {
      long M = Lv.dims[i];
      double* Qi = Lv.Q.data() + Lv.offQ[i];
      if (low) {
        for (long cy = 0; cy < 2; cy++) {
          long y = 2 * i + cy;
          long offy = cy ? low->ranks[2 * i] : 0, ny = low->ranks[y];
          std::copy(low->S.begin() + 3 * low->offX[y], low->S.begin() + 3 * low->offX[y] + 3 * ny, Lv.S.begin() + 3 * (Lv.offX[i] + offy));
          const double* Ry = low->R.data() + low->offQ[y];
          for (long c = 0; c < ny; c++)
            for (long r = 0; r < ny; r++) Qi[(offy + r) + (offy + c) * M] = Ry[r + c * low->dims[y]];
          if (M > 0) {
            low->upperStride[y] = M;
            for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
              long j = Lv.acol[ij];
              for (long cx = 0; cx < 2; cx++) {
                long x = 2 * j + cx;
                long offx = cx ? low->ranks[2 * j] : 0, nx = low->ranks[x];
                long lowN = lookup(low->arp, low->acol, y, x);
                long lowC = lookup(low->crp, low->ccol, y, x);
                long dp = Lv.offA[ij] + offy + offx * M;
                if (lowN >= 0)
                  low->NA[lowN] = dp;
                else if (lowC >= 0) {
                  const double* Cyx = low->C.data() + low->offC[lowC];
                  for (long c = 0; c < nx; c++)
                    for (long r = 0; r < ny; r++) Lv.A[dp + r + c * M] = Cyx[r + c * ny];
                }
              }
            }
          }
        }
      } else {
        const Cell& c = p.cells[Lv.gbegin + i];
        std::copy(p.bodies.begin() + 3 * c.body[0], p.bodies.begin() + 3 * c.body[1], Lv.S.begin() + 3 * Lv.offX[i]);
        for (long r = 0; r < M; r++) Qi[r + r * M] = 1.;
      }
    }
    if (!low) {
#pragma omp parallel for schedule(dynamic) num_threads(h.nthreads)
      for (long i = 0; i < nb; i++) {
        const Cell& c = p.cells[Lv.gbegin + i];
        for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
          const Cell& cj = p.cells[Lv.gbegin + Lv.acol[ij]];
          gen_block(h.k, Lv.dims[i], Lv.dims[Lv.acol[ij]], p.bodies.data() + 3 * c.body[0], p.bodies.data() + 3 * cj.body[0],
                    Lv.A.data() + Lv.offA[ij], Lv.dims[i]);
        }
      }
    }
    // basis (h2matrix.cpp:195-200); the root has no far list (solver.cpp:31)
#pragma omp parallel for schedule(dynamic) num_threads(h.nthreads)
    for (long i = 0; i < nb; i++) {
      long M = Lv.dims[i];
      std::vector<double> F;
      if (l > 0) {
        std::vector<long> idx = far_sample(p.cells, h.basisfar, h.near, h.far, h.hss, Lv.gbegin + i, h.sampling);
        F.resize(3 * idx.size());
        for (size_t t = 0; t < idx.size(); t++)
          for (int d = 0; d < 3; d++) F[3 * t + d] = p.bodies[3 * idx[t] + d];
      }
      long nf = F.size() / 3;
      Lv.ranks[i] = compute_basis(h.k, h.tol, h.maxrank(l), h.orth, M, nf, Lv.S.data() + 3 * Lv.offX[i], F.data(),
                                  Lv.Q.data() + Lv.offQ[i], Lv.R.data() + Lv.offQ[i]);
    }
  }

  Lv.offZ.assign(nb + 1, 0);
  for (long i = 0; i < nb; i++) Lv.offZ[i + 1] = Lv.offZ[i] + Lv.ranks[i];
  Lv.lenZ = Lv.offZ[nb];
  long nnzC = Lv.crp[nb];
  Lv.offC.assign(nnzC + 1, 0);
  for (long i = 0; i < nb; i++)
    for (long t = Lv.crp[i]; t < Lv.crp[i + 1]; t++) Lv.offC[t + 1] = Lv.offC[t] + Lv.ranks[i] * Lv.ranks[Lv.ccol[t]];
  Lv.C.assign(Lv.offC[nnzC], 0.);
  Lv.Z.assign(Lv.lenZ, 0.);
  Lv.W.assign(Lv.lenZ, 0.);
  if (Lv.lenZ > 0) {
    // coupling blocks (h2matrix.cpp:222-242)
#pragma omp parallel for schedule(dynamic) num_threads(h.nthreads)
    for (long y = 0; y < nb; y++) {
      long M = Lv.ranks[y];
      const double* Ry = Lv.R.data() + Lv.offQ[y];
      for (long ij = Lv.crp[y]; ij < Lv.crp[y + 1]; ij++) {
        long x = Lv.ccol[ij], N = Lv.ranks[x];
        double* Cyx = Lv.C.data() + Lv.offC[ij];
        if (M == 0 || N == 0) continue;
        if (h.orth) {
          std::vector<double> Ayx((size_t)M * N), T((size_t)M * N);
          gen_block(h.k, M, N, Lv.S.data() + 3 * Lv.offX[y], Lv.S.data() + 3 * Lv.offX[x], Ayx.data(), M);
          std::vector<double> Ru((size_t)M * M, 0.), Rv((size_t)N * N, 0.);
          const double* Rx = Lv.R.data() + Lv.offQ[x];
          for (long c = 0; c < M; c++)
            for (long r = 0; r <= c; r++) Ru[r + c * M] = Ry[r + c * Lv.dims[y]];
          for (long c = 0; c < N; c++)
            for (long r = 0; r <= c; r++) Rv[r + c * N] = Rx[r + c * Lv.dims[x]];
          gemm('N', 'N', M, N, M, 1., Ru.data(), M, Ayx.data(), M, 0., T.data(), M);
          gemm('N', 'T', M, N, N, 1., T.data(), M, Rv.data(), N, 0., Cyx, M);
        } else {
          gen_block(h.k, M, N, Lv.S.data() + 3 * Lv.offX[y], Lv.S.data() + 3 * Lv.offX[x], Cyx, M);
        }
      }
    }
  }
}

// h2matrix.cpp:251-271
// Threads of the per-box loops of the solve passes (each box is independent,
// so the result does not depend on the count).
int g_threads = 1;

void up_pass(Level& Lv, const double* xin) {
  std::copy(xin, xin + Lv.lenX, Lv.X.begin());
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i], N = Lv.ranks[i];
    if (N > 0) orc::gemv('T', M, N, 1., Lv.Q.data() + Lv.offQ[i], M, Lv.X.data() + Lv.offX[i], 0., Lv.Z.data() + Lv.offZ[i]);
  }
}

// h2matrix.cpp:273-301
void down_pass(Level& Lv, double* yout) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i], K = Lv.ranks[i];
    double* y = Lv.Y.data() + Lv.offX[i];
    if (K > 0) {
      double* w = Lv.W.data() + Lv.offZ[i];
      for (long ij = Lv.crp[i]; ij < Lv.crp[i + 1]; ij++) {
        long j = Lv.ccol[ij];
        orc::gemv('N', K, Lv.ranks[j], 1., Lv.C.data() + Lv.offC[ij], K, Lv.Z.data() + Lv.offZ[j], 1., w);
      }
      orc::gemv('N', M, K, 1., Lv.Q.data() + Lv.offQ[i], M, w, 0., y);
    } else {
      std::fill(y, y + M, 0.);
    }
  }
  std::copy(Lv.Y.begin(), Lv.Y.end(), yout);
}

// h2matrix.cpp:303-344
void leaf_pass(Level& Lv, double* x) {
  std::copy(x, x + Lv.lenX, Lv.X.begin());
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i], K = Lv.ranks[i];
    double* y = Lv.Y.data() + Lv.offX[i];
    std::fill(y, y + M, 0.);
    if (K > 0) {
      double* w = Lv.W.data() + Lv.offZ[i];
      for (long ij = Lv.crp[i]; ij < Lv.crp[i + 1]; ij++) {
        long j = Lv.ccol[ij];
        orc::gemv('N', K, Lv.ranks[j], 1., Lv.C.data() + Lv.offC[ij], K, Lv.Z.data() + Lv.offZ[j], 1., w);
      }
      orc::gemv('N', M, K, 1., Lv.Q.data() + Lv.offQ[i], M, w, 1., y);
    }
    for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
      long j = Lv.acol[ij];
      orc::gemv('N', M, Lv.dims[j], 1., Lv.A.data() + Lv.offA[ij], M, Lv.X.data() + Lv.offX[j], 1., y);
    }
  }
  std::copy(Lv.Y.begin(), Lv.Y.end(), x);
}

// h2matrix.cpp:346-425
void factorize_level(orc_h2& h, Level& Lv) {
  long nb = Lv.nb;
  long dmax = 0;
  for (long i = 0; i < nb; i++) dmax = std::max(dmax, Lv.dims[i]);
  std::vector<double> B((size_t)nb * dmax * dmax, 0.);
  Lv.info = 0;
  int info = 0;
#pragma omp parallel for schedule(dynamic) num_threads(h.nthreads) reduction(+ : info)
  for (long i = 0; i < nb; i++) {
    long diag = lookup(Lv.arp, Lv.acol, i, i);
    long M = Lv.dims[i], Ms = Lv.ranks[i], Mr = M - Ms;
    const double* Ui = Lv.Q.data() + Lv.offQ[i];
    double* V = Lv.R.data() + Lv.offQ[i];
    double* Aii = Lv.A.data() + Lv.offA[diag];
    double* b = B.data() + (size_t)i * dmax * dmax;
    if (M == 0) continue;
    gemm('T', 'T', M, M, M, 1., Ui, M, Aii, M, 0., b, dmax);   // b = Ui^T Aii^T
    gemm('T', 'T', M, M, M, 1., Ui, M, b, dmax, 0., Aii, M);   // Aii = Ui^T b^T
    for (long c = 0; c < M; c++)
      for (long r = 0; r < Ms; r++) V[r + c * M] = Ui[c + r * M];
    if (Mr > 0) {
      std::vector<double> rr((size_t)Mr * Mr), rhs((size_t)Mr * M);
      for (long c = 0; c < Mr; c++)
        for (long r = 0; r < Mr; r++) rr[r + c * Mr] = Aii[(Ms + r) + (Ms + c) * M];
      std::vector<long> piv(Mr);
      double logdet = orc::lu_factor(Mr, rr.data(), Mr, piv.data());
      for (long c = 0; c < M; c++)
        for (long r = 0; r < Mr; r++) rhs[r + c * Mr] = Ui[c + (Ms + r) * M];
      orc::lu_solve(Mr, M, rr.data(), Mr, piv.data(), rhs.data(), Mr);
      for (long c = 0; c < M; c++)
        for (long r = 0; r < Mr; r++) V[(Ms + r) + c * M] = rhs[r + c * Mr];
      if (Ms > 0) {
        gemm('N', 'T', Mr, Ms, M, 1., V + Ms, M, b, dmax, 0., Aii + Ms, M);                  // A_rs
        gemm('N', 'N', Ms, Ms, Mr, -1., Aii + Ms * M, M, Aii + Ms, M, 1., Aii, M);           // A_ss -= A_sr A_rs
      }
      info += (logdet <= std::log(std::numeric_limits<double>::min())) ? 1 : 0;
    }
    for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
      if (ij == diag) continue;
      long j = Lv.acol[ij], N = Lv.dims[j];
      const double* Uj = Lv.Q.data() + Lv.offQ[j];
      double* Aij = Lv.A.data() + Lv.offA[ij];
      gemm('T', 'T', N, M, N, 1., Uj, N, Aij, M, 0., b, dmax);   // b = Uj^T Aij^T (N x M)
      gemm('N', 'T', M, N, M, 1., V, M, b, dmax, 0., Aij, M);    // Aij = V b^T
    }
    // b = [A_rs | rr^{-1}] (h2matrix.cpp:399-400)
    for (long c = 0; c < Ms; c++)
      for (long r = 0; r < Mr; r++) b[r + c * dmax] = Aii[(Ms + r) + c * M];
    if (Mr > 0) gemm('N', 'N', Mr, Mr, M, 1., V + Ms, M, Ui + Ms * M, M, 0., b + Ms * dmax, dmax);
  }
  Lv.info = info;

#pragma omp parallel for schedule(dynamic) num_threads(h.nthreads)
  for (long i = 0; i < nb; i++) {
    long diag = lookup(Lv.arp, Lv.acol, i, i);
    long M = Lv.dims[i], Ms = Lv.ranks[i], Mr = M - Ms;
    double* Aii = Lv.A.data() + Lv.offA[diag];
    for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
      if (ij == diag) continue;
      long j = Lv.acol[ij], N = Lv.dims[j], Ns = Lv.ranks[j], Nr = N - Ns;
      double* Aij = Lv.A.data() + Lv.offA[ij];
      const double* Bj = B.data() + (size_t)j * dmax * dmax;
      if (Ms > 0 && Ns > 0) {
        gemm('N', 'N', Ms, Ns, Mr, -1., Aii + Ms * M, M, Aij + Ms, M, 1., Aij, M);
        gemm('N', 'N', Ms, Ns, Nr, -1., Aij + Ns * M, M, Bj, dmax, 1., Aij, M);
      }
      if (Ms > 0 && Nr > 0) {
        std::vector<double> T((size_t)Ms * Nr);
        gemm('N', 'N', Ms, Nr, Nr, 1., Aij + Ns * M, M, Bj + Ns * dmax, dmax, 0., T.data(), Ms);
        gemm('N', 'T', Ms, Ms, Nr, -1., T.data(), Ms, Aij + Ns * M, M, 1., Aii, M);
      }
    }
  }
}

// h2matrix.cpp:427-446
void copy_next(Level& up, const Level& Lv) {
  for (long i = 0; i < Lv.nb; i++)
    for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
      long j = Lv.acol[ij], M = Lv.dims[i], Ms = Lv.ranks[i], Ns = Lv.ranks[j];
      if (Ms > 0 && Ns > 0) {
        const double* Aij = Lv.A.data() + Lv.offA[ij];
        long ld = Lv.upperStride[i];
        for (long c = 0; c < Ns; c++)
          for (long r = 0; r < Ms; r++) up.A[Lv.NA[ij] + r + c * ld] = Aij[r + c * M];
      }
    }
}

// h2matrix.cpp:448-493
void forward(Level& Lv, const double* xin) {
  std::copy(xin, xin + Lv.lenX, Lv.Y.begin());
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i];
    if (M > 0) orc::gemv('N', M, M, 1., Lv.R.data() + Lv.offQ[i], M, Lv.Y.data() + Lv.offX[i], 0., Lv.X.data() + Lv.offX[i]);
  }
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i], Ms = Lv.ranks[i];
    if (Ms <= 0) continue;
    double* z = Lv.Z.data() + Lv.offZ[i];
    std::copy(Lv.X.begin() + Lv.offX[i], Lv.X.begin() + Lv.offX[i] + Ms, z);
    for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
      long j = Lv.acol[ij], N = Lv.dims[j], Ns = Lv.ranks[j], Nr = N - Ns;
      if (Nr > 0)
        orc::gemv('N', Ms, Nr, -1., Lv.A.data() + Lv.offA[ij] + Ns * M, M, Lv.X.data() + Lv.offX[j] + Ns, 1., z);
    }
  }
}

// h2matrix.cpp:495-537
void backward(Level& Lv, double* yout) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i], Ms = Lv.ranks[i], Mr = M - Ms;
    double* x = Lv.X.data() + Lv.offX[i];
    std::copy(Lv.W.begin() + Lv.offZ[i], Lv.W.begin() + Lv.offZ[i] + Ms, x);
    if (Mr > 0)
      for (long ij = Lv.arp[i]; ij < Lv.arp[i + 1]; ij++) {
        long j = Lv.acol[ij], Ns = Lv.ranks[j];
        if (Ns > 0) orc::gemv('N', Mr, Ns, -1., Lv.A.data() + Lv.offA[ij] + Ms, M, Lv.W.data() + Lv.offZ[j], 1., x + Ms);
      }
  }
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads)
  for (long i = 0; i < Lv.nb; i++) {
    long M = Lv.dims[i];
    if (M > 0) orc::gemv('N', M, M, 1., Lv.Q.data() + Lv.offQ[i], M, Lv.X.data() + Lv.offX[i], 0., Lv.Y.data() + Lv.offX[i]);
  }
  std::copy(Lv.Y.begin(), Lv.Y.end(), yout);
}

}  // namespace

extern "C" {

// test_funcs.hpp:31-41
void orc_uniform_cube_rnd(double* bodies, long long n, double diameter, long long dim, unsigned int seed) {
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> u(0., diameter);
  for (long long b = 0; b < n; b++)
    for (int i = 0; i < 3; i++) bodies[3 * b + i] = i < dim ? u(gen) : 0.;
}

// Builder-defined (no reference generator): uniform on the unit sphere as the
// golden-angle points (test_funcs.hpp:43-61) at radius 1 under a uniform
// random rotation; the unit quaternion comes from three U(0, 1) draws of
// std::mt19937(seed) (Shoemake 1992: q = (sqrt(1-u1) sin 2pi u2,
// sqrt(1-u1) cos 2pi u2, sqrt(u1) sin 2pi u3, sqrt(u1) cos 2pi u3)).
void orc_sphere_rnd(double* bodies, long long n, unsigned int seed) {
  orc_mesh_sphere(bodies, n, 1.);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> u(0., 1.);
  const double two_pi = 6.283185307179586476925286766559;
  double u1 = u(gen), u2 = u(gen), u3 = u(gen);
  double x = std::sqrt(1. - u1) * std::sin(two_pi * u2), y = std::sqrt(1. - u1) * std::cos(two_pi * u2);
  double z = std::sqrt(u1) * std::sin(two_pi * u3), w = std::sqrt(u1) * std::cos(two_pi * u3);
  double R[9] = {1. - 2. * (y * y + z * z), 2. * (x * y - z * w), 2. * (x * z + y * w),
                 2. * (x * y + z * w), 1. - 2. * (x * x + z * z), 2. * (y * z - x * w),
                 2. * (x * z - y * w), 2. * (y * z + x * w), 1. - 2. * (x * x + y * y)};
  for (long long i = 0; i < n; i++) {
    double p[3] = {bodies[3 * i], bodies[3 * i + 1], bodies[3 * i + 2]};
    for (int d = 0; d < 3; d++) bodies[3 * i + d] = (R[3 * d] * p[0] + R[3 * d + 1] * p[1]) + R[3 * d + 2] * p[2];
  }
}

// test_funcs.hpp:43-61
void orc_mesh_sphere(double* bodies, long long n, double r) {
  const double phi = 3.14159265358979323846 * (3. - std::sqrt(5.));
  const double d = r + r, r2 = r * r;
  for (long long i = 0; i < n; ++i) {
    const double y = r - ((double)i / (double)(n - 1)) * d;
    const double radius = std::sqrt(r2 - y * y);
    const double theta = (double)i * phi;
    bodies[i * 3] = radius * std::cos(theta);
    bodies[i * 3 + 1] = y;
    bodies[i * 3 + 2] = radius * std::sin(theta);
  }
}

// main.cpp:73-76 with a seeded std::mt19937_64 and bounds [lo, hi)
void orc_uniform_vec(double* x, long long n, double lo, double hi, unsigned long long seed) {
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> u(lo, hi);
  for (long long i = 0; i < n; i++) x[i] = u(gen);
}

// main.cpp:53-55
long long orc_levels(long long n, long long leaf) {
  leaf = n < leaf ? n : leaf;
  return (long long)std::log2((double)n / leaf);
}

void orc_build_tree(double* bodies, long long n, long long levels, long long* child, long long* body, double* R, double* C) {
  long nc = (2L << levels) - 1;
  std::vector<Cell> cells(nc);
  build_tree(cells, bodies, n, levels);
  for (long i = 0; i < nc; i++) {
    child[2 * i] = cells[i].child[0];
    child[2 * i + 1] = cells[i].child[1];
    body[2 * i] = cells[i].body[0];
    body[2 * i + 1] = cells[i].body[1];
    for (int d = 0; d < 3; d++) {
      R[3 * i + d] = cells[i].R[d];
      C[3 * i + d] = cells[i].C[d];
    }
  }
}

static std::vector<Cell> cells_from(long long levels, const long long* child, const long long* body, const double* R,
                                    const double* C) {
  long nc = (2L << levels) - 1;
  std::vector<Cell> cells(nc);
  for (long i = 0; i < nc; i++) {
    cells[i].child[0] = child ? child[2 * i] : -1;
    cells[i].child[1] = child ? child[2 * i + 1] : -1;
    cells[i].body[0] = body ? body[2 * i] : -1;
    cells[i].body[1] = body ? body[2 * i + 1] : -1;
    for (int d = 0; d < 3; d++) {
      cells[i].R[d] = R ? R[3 * i + d] : 0.;
      cells[i].C[d] = C ? C[3 * i + d] : 0.;
    }
  }
  return cells;
}

long long orc_csr(int far, double theta, long long levels, const long long* child, const double* R, const double* C,
                  long long* rowptr, long long* col, long long cap) {
  std::vector<Cell> cells = cells_from(levels, child, nullptr, R, C);
  CSR c = make_csr(far != 0, cells, theta);
  long long nnz = c.col.size();
  if (cap >= nnz) {
    for (size_t i = 0; i < c.rp.size(); i++) rowptr[i] = c.rp[i];
    for (long long t = 0; t < nnz; t++) col[t] = c.col[t];
  }
  return nnz;
}

long long orc_far_sample(long long levels, const long long* body, const long long* bfar_rp, const long long* bfar_col,
                         const long long* near_rp, const long long* near_col, const long long* far_rp,
                         const long long* far_col, int with_near, long long cell, long long sampling, long long* out,
                         long long cap) {
  long nc = (2L << levels) - 1;
  std::vector<Cell> cells = cells_from(levels, nullptr, body, nullptr, nullptr);
  auto mk = [&](const long long* rp, const long long* col) {
    CSR c;
    c.rp.assign(rp, rp + nc + 1);
    c.col.assign(col, col + rp[nc]);
    return c;
  };
  CSR bf = mk(bfar_rp, bfar_col), nr = mk(near_rp, near_col), fr = mk(far_rp, far_col);
  std::vector<long> idx = far_sample(cells, bf, nr, fr, with_near != 0, cell, sampling);
  if (cap >= (long long)idx.size())
    for (size_t t = 0; t < idx.size(); t++) out[t] = idx[t];
  return idx.size();
}

double orc_kernel_eval(int kid, double diag, double alpha, double d) { return Kern{kid, diag, alpha}(d); }

void orc_set_sketch(int on) { g_sketch = on != 0; }
double orc_rademacher(unsigned long long t, unsigned long long i) { return rademacher(t, i); }

void orc_gen_matrix(int kid, double diag, double alpha, long long m, long long n, const double* bi, const double* bj, double* A) {
  gen_block(Kern{kid, diag, alpha}, m, n, bi, bj, A, m);
}

// Dense check y = A x (kernel.cpp:25-35 + main.cpp:83-84,114), matrix-free.
void orc_dense_matvec(int kid, double diag, double alpha, const double* bodies, long long n, const double* x, double* y, int nthreads) {
  Kern k{kid, diag, alpha};
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (long long i = 0; i < n; i++) {
    double s = 0.;
    const double* p = bodies + 3 * i;
    for (long long j = 0; j < n; j++) {
      const double* q = bodies + 3 * j;
      s += k(std::hypot(p[0] - q[0], p[1] - q[1], p[2] - q[2])) * x[j];
    }
    y[i] = s;
  }
}

void orc_qr_r(long long m, long long n, double* A, long long lda) {
  std::vector<double> tau(std::min(m, n));
  orc::householder_qr(m, n, A, lda, tau.data());
  for (long long j = 0; j < n; j++)
    for (long long i = j + 1; i < m; i++) A[i + j * lda] = 0.;
}

long long orc_cpqr(long long m, long long n, double* A, long long lda, long long* perm, double tol, double* maxpivot, long long* nonzero) {
  orc::CPQRInfo info = orc::cpqr(m, n, A, lda);
  for (long long j = 0; j < n; j++) perm[j] = info.perm[j];
  *maxpivot = info.maxpivot;
  *nonzero = info.nonzero_pivots;
  for (long long j = 0; j < n; j++)
    for (long long i = j + 1; i < m; i++) A[i + j * lda] = 0.;
  return tol > 0. ? orc::cpqr_rank(info, A, lda, tol) : std::min(m, n);
}

double orc_lu_solve(long long n, long long nrhs, double* A, double* B) {
  std::vector<long> piv(n);
  double ld = orc::lu_factor(n, A, n, piv.data());
  orc::lu_solve(n, nrhs, A, n, piv.data(), B, n);
  return ld;
}

orc_problem* orc_problem_create(const double* points, long long n, long long leaf) {
  orc_problem* p = new orc_problem();
  p->n = n;
  p->leaf = leaf;
  p->levels = orc_levels(n, leaf);
  p->bodies.assign(points, points + 3 * n);
  p->cells.resize((2L << p->levels) - 1);
  build_tree(p->cells, p->bodies.data(), n, p->levels);
  return p;
}

void orc_problem_destroy(orc_problem* p) { delete p; }
long long orc_problem_levels(const orc_problem* p) { return p->levels; }
void orc_problem_bodies(const orc_problem* p, double* out) { std::copy(p->bodies.begin(), p->bodies.end(), out); }

// solver.cpp:12-41 H2MatrixSolver constructor at one rank.
orc_h2* orc_h2_create(const orc_problem* p, int kid, double diag, double alpha, double theta, double tol, long long rank,
                      long long leveled_rank, int orth, int hss_far, long long sampling, int nthreads) {
  orc_h2* h = new orc_h2();
  h->p = p;
  h->k = Kern{kid, diag, alpha};
  h->L = p->levels;
  h->theta = theta;
  h->tol = tol;
  h->rank = rank;
  h->leveled_rank = leveled_rank;
  h->orth = orth != 0;
  h->sampling = sampling;
  h->nthreads = nthreads > 0 ? nthreads : 1;
  h->near = make_csr(false, p->cells, theta);
  h->far = make_csr(true, p->cells, theta);
  h->basisfar = hss_far ? make_csr(true, p->cells, 0.) : h->far;
  h->hss = hss_far != 0;
  h->lv.resize(h->L + 1);
  for (long l = h->L; l >= 0; l--) construct_level(*h, l);
  return h;
}

void orc_h2_destroy(orc_h2* h) { delete h; }
// threads of the matrix's later passes (matvec, factorization, substitution, GMRES)
void orc_h2_set_threads(orc_h2* h, int nthreads) { h->nthreads = nthreads > 0 ? nthreads : 1; }
long long orc_h2_nboxes(const orc_h2* h, long long level) { return h->lv[level].nb; }
void orc_h2_dims(const orc_h2* h, long long level, long long* dims, long long* ranks) {
  const Level& Lv = h->lv[level];
  for (long i = 0; i < Lv.nb; i++) {
    if (dims) dims[i] = Lv.dims[i];
    if (ranks) ranks[i] = Lv.ranks[i];
  }
}
// Level data in the layout of matrix_container.cpp:6-36 at one rank; returns
// the count. what: 0 = near blocks A (M_i x M_j column major, in the order of
// the level's near list; in place, so after orc_h2_factorize they hold the
// factored blocks), 1 = bases Q, 2 = R (M*M per box; V after the
// factorization), 3 = couplings C (k_i x k_j, in the order of the far list).
long long orc_h2_download(const orc_h2* h, long long level, int what, double* out) {
  const Level& Lv = h->lv[level];
  const std::vector<double>& v = what == 0 ? Lv.A : what == 1 ? Lv.Q : what == 2 ? Lv.R : Lv.C;
  if (out) std::copy(v.begin(), v.end(), out);
  return (long long)v.size();
}
// Level-local near (far = 0) or far (far = 1) list (comm-mpi.cpp:105-115 at
// P = 1): rowptr nb + 1 entries, columns box indices within the level. Fills
// rp/col when given; returns the number of entries.
long long orc_h2_level_list(const orc_h2* h, long long level, int far, long long* rp, long long* col) {
  const Level& Lv = h->lv[level];
  const std::vector<long>& r = far ? Lv.crp : Lv.arp;
  const std::vector<long>& c = far ? Lv.ccol : Lv.acol;
  if (rp) std::copy(r.begin(), r.end(), rp);
  if (col) std::copy(c.begin(), c.end(), col);
  return (long long)c.size();
}
long long orc_h2_near_count(const orc_h2* h, long long level) { return h->lv[level].arp.back(); }
long long orc_h2_far_count(const orc_h2* h, long long level) { return h->lv[level].crp.back(); }

// solver.cpp:76-88
void orc_h2_matvec(orc_h2* h, double* x) {
  long L = h->L;
  g_threads = h->nthreads;
  up_pass(h->lv[L], x);
  for (long l = L - 1; l >= 0; l--) up_pass(h->lv[l], h->lv[l + 1].Z.data());
  for (long l = 0; l < L; l++) down_pass(h->lv[l], h->lv[l + 1].W.data());
  leaf_pass(h->lv[L], x);
}

// Far field of one level alone: y_approx = the H^2 matvec restricted to the
// couplings of `level` (upward pass, couplings of that level, downward pass
// without couplings, no near field); y_exact = the same interactions summed
// directly, sum over far pairs (i, j) of the level of K(X_i, X_j) x_j.
// A diagnostic that attributes the construction error to a level.
void orc_h2_far_level(orc_h2* h, long long level, const double* x, double* y_approx, double* y_exact) {
  const long L = h->L;
  g_threads = h->nthreads;
  up_pass(h->lv[L], x);
  for (long l = L - 1; l >= level; l--) up_pass(h->lv[l], h->lv[l + 1].Z.data());
  Level& T = h->lv[level];
  std::vector<double> y(T.lenX, 0.);
  for (long i = 0; i < T.nb; i++) {
    long M = T.dims[i], K = T.ranks[i];
    if (K == 0) continue;
    std::vector<double> w(K, 0.);
    for (long ij = T.crp[i]; ij < T.crp[i + 1]; ij++) {
      long j = T.ccol[ij];
      orc::gemv('N', K, T.ranks[j], 1., T.C.data() + T.offC[ij], K, T.Z.data() + T.offZ[j], 1., w.data());
    }
    orc::gemv('N', M, K, 1., T.Q.data() + T.offQ[i], M, w.data(), 0., y.data() + T.offX[i]);
  }
  for (long l = level + 1; l <= L; l++) {
    Level& D = h->lv[l];
    std::vector<double> yl(D.lenX, 0.);
    for (long i = 0; i < D.nb; i++) {
      long M = D.dims[i], K = D.ranks[i];
      if (K > 0) orc::gemv('N', M, K, 1., D.Q.data() + D.offQ[i], M, y.data() + D.offZ[i], 0., yl.data() + D.offX[i]);
    }
    y.swap(yl);
  }
  std::copy(y.begin(), y.end(), y_approx);
  // direct sums
  const std::vector<Cell>& cells = h->p->cells;
  const double* B = h->p->bodies.data();
  const long first = (1L << level) - 1;
  std::fill(y_exact, y_exact + h->lv[L].lenX, 0.);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
  for (long i = 0; i < T.nb; i++) {
    const Cell& ci = cells[first + i];
    for (long ij = T.crp[i]; ij < T.crp[i + 1]; ij++) {
      const Cell& cj = cells[first + T.ccol[ij]];
      for (long a = ci.body[0]; a < ci.body[1]; a++) {
        double s = 0.;
        for (long b = cj.body[0]; b < cj.body[1]; b++) {
          double dx = B[3 * a] - B[3 * b], dy = B[3 * a + 1] - B[3 * b + 1], dz = B[3 * a + 2] - B[3 * b + 2];
          s += h->k(std::sqrt(dx * dx + dy * dy + dz * dz)) * x[b];
        }
        y_exact[a] += s;
      }
    }
  }
}

// solver.cpp:90-100
int orc_h2_factorize(orc_h2* h) {
  int info = 0;
  for (long l = h->L; l >= 0; l--) {
    factorize_level(*h, h->lv[l]);
    if (l > 0) copy_next(h->lv[l - 1], h->lv[l]);
    info += h->lv[l].info;
  }
  return info;
}

// solver.cpp:121-132
void orc_h2_precondition(orc_h2* h, double* x) {
  long L = h->L;
  g_threads = h->nthreads;
  forward(h->lv[L], x);
  for (long l = L - 1; l >= 0; l--) forward(h->lv[l], h->lv[l + 1].Z.data());
  for (long l = 0; l < L; l++) backward(h->lv[l], h->lv[l + 1].W.data());
  backward(h->lv[L], x);
}

// solver.cpp:144-201 (left-preconditioned restarted GMRES, CGS1, Householder least squares)
long long orc_h2_gmres(orc_h2* A, orc_h2* M, double tol, double* x, const double* b, long long inner, long long outer, double* resid) {
  long N = A->lv[A->L].lenX;
  long ld = inner + 1;
  double nb2 = 0.;
  for (long i = 0; i < N; i++) nb2 += b[i] * b[i];
  double normb = std::sqrt(nb2);
  if (normb == 0.) normb = 1.;
  std::vector<double> R(b, b + N);
  resid[0] = 1.;
  long long iters = 0;
  while (iters < outer && tol <= resid[iters]) {
    orc_h2_precondition(M, R.data());
    double beta = 0.;
    for (long i = 0; i < N; i++) beta += R[i] * R[i];
    beta = std::sqrt(beta);
    std::vector<double> H((size_t)ld * inner, 0.), v((size_t)N * ld, 0.);
    for (long i = 0; i < N; i++) v[i] = R[i] * (1. / beta);
    for (long i = 0; i < inner; i++) {
      std::copy(v.begin() + i * N, v.begin() + (i + 1) * N, R.begin());
      orc_h2_matvec(A, R.data());
      orc_h2_precondition(M, R.data());
      orc::gemv('T', N, i + 1, 1., v.data(), N, R.data(), 0., H.data() + i * ld);
      orc::gemv('N', N, i + 1, -1., v.data(), N, H.data() + i * ld, 1., R.data());
      double s = 0.;
      for (long t = 0; t < N; t++) s += R[t] * R[t];
      H[(i + 1) + i * ld] = std::sqrt(s);
      double inv = 1. / H[(i + 1) + i * ld];
      for (long t = 0; t < N; t++) v[t + (i + 1) * N] = R[t] * inv;
    }
    // least squares min || beta e1 - H y || by Householder QR (solver.cpp:187-191)
    std::vector<double> s(ld, 0.), tau(inner);
    s[0] = beta;
    orc::householder_qr(ld, inner, H.data(), ld, tau.data());
    for (long k = 0; k < inner; k++) orc::apply_householder_left(ld - k, 1, H.data() + (k + 1) + k * ld, tau[k], s.data() + k, ld);
    orc::trsm_upper(inner, 1, H.data(), ld, s.data(), ld);
    orc::gemv('N', N, inner, 1., v.data(), N, s.data(), 1., x);
    for (long t = 0; t < N; t++) R[t] = -x[t];
    orc_h2_matvec(A, R.data());
    double r2 = 0.;
    for (long t = 0; t < N; t++) {
      R[t] += b[t];
      r2 += R[t] * R[t];
    }
    resid[++iters] = std::sqrt(r2) / normb;
  }
  return iters;
}

// solver.cpp:229-238
double orc_rel_err(long long n, const double* x, const double* ref) {
  double e0 = 0., e1 = 0.;
  for (long long i = 0; i < n; i++) {
    double d = x[i] - ref[i];
    e0 += d * d;
    e1 += ref[i] * ref[i];
  }
  return std::sqrt(e0 / e1);
}

}  // extern "C"
