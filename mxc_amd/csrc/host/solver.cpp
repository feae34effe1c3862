#include "solver.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace mxc {

namespace {
void check(int rc) {
  if (rc != MXG_OK) throw std::runtime_error(std::string("device layer: ") + mxg_last_error());
}
}  // namespace

H2MatrixSolver::H2MatrixSolver(mxg_ctx_t c, const MatrixAccessor& eval, double epi, long long rank, long long leveled_rank,
                               const std::vector<Cell>& cells, double theta, const double bodies[], long long lv,
                               long long sampling, bool factorable)
    : ctx(c) {
  const bool fix_rank = (epi == 0.);  // solver.cpp:28
  opts.tol = fix_rank ? 0. : epi;
  opts.rank = fix_rank ? rank : 0;
  opts.leveled_rank = fix_rank ? leveled_rank : 0;
  opts.orth = fix_rank;               // compute_basis(..., 1. <= epi) with epi = rank_func(l)
  opts.hss_far = fix_rank;            // solver.cpp:32
  opts.sampling = sampling;
  build(eval, cells, theta, bodies, lv, factorable);
}

H2MatrixSolver::H2MatrixSolver(mxg_ctx_t c, const MatrixAccessor& eval, const BasisOptions& opt, const std::vector<Cell>& cells,
                               double theta, const double bodies[], long long lv, bool factorable)
    : ctx(c), opts(opt) {
  build(eval, cells, theta, bodies, lv, factorable);
}

void H2MatrixSolver::build(const MatrixAccessor& eval, const std::vector<Cell>& cells, double theta, const double bodies[],
                           long long lv, bool factorable) {
  levels = lv;
  const long long nbodies = cells[0].Body[1] - cells[0].Body[0];
  CSR Near('N', cells, cells, theta);
  CSR Far('F', cells, cells, theta);
  CSR BasisFar = opts.hss_far ? CSR('F', cells, cells, 0.) : Far;
  const int flags = (factorable ? MXG_H2_FACTORABLE : 0) | (opts.householder_rows ? MXG_H2_HOUSEHOLDER_ROWS : 0);
  check(mxg_h2_create(ctx, eval.desc(), bodies, nbodies, levels, flags, &dev));
  A.assign(levels + 1, H2Matrix());
  // MXC_PROGRESS set: one line per constructed level on stderr
  const bool progress = std::getenv("MXC_PROGRESS") != nullptr;
  for (long long l = levels; l >= 0; l--) {
    const auto t0 = std::chrono::steady_clock::now();
    // rank_func(l) = (levels - l) * leveled_rank + rank (solver.cpp:29)
    long long maxrank = opts.rank > 0 ? (levels - l) * opts.leveled_rank + opts.rank : 0;
    A[l].construct(dev, opts, maxrank, cells, Near, Far, BasisFar, l, l < levels ? &A[l + 1] : nullptr);
    if (progress) {
      long long rsum = 0, rmax = 0;
      for (long long r : A[l].DimsLr) {
        rsum += r;
        rmax = r > rmax ? r : rmax;
      }
      std::fprintf(stderr, "[mxc] level %lld built in %.1f s: %lld boxes, ranks sum %lld max %lld\n", l,
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), A[l].nboxes, rsum, rmax);
      std::fflush(stderr);
    }
  }
  local_bodies = {0, nbodies};
}

H2MatrixSolver::H2MatrixSolver(H2MatrixSolver&& o) noexcept { *this = std::move(o); }

H2MatrixSolver& H2MatrixSolver::operator=(H2MatrixSolver&& o) noexcept {
  if (this != &o) {
    release();
    levels = o.levels;
    A = std::move(o.A);
    ctx = o.ctx;
    dev = o.dev;
    ulv = o.ulv;
    local_bodies = o.local_bodies;
    resid = std::move(o.resid);
    iters = o.iters;
    inner_iters = o.inner_iters;
    opts = o.opts;
    o.dev = nullptr;
    o.ulv = nullptr;
    o.levels = -1;
  }
  return *this;
}

H2MatrixSolver::~H2MatrixSolver() { release(); }

void H2MatrixSolver::release() {
  if (ulv) mxg_ulv_destroy(ulv);
  if (dev) mxg_h2_destroy(dev);
  ulv = nullptr;
  dev = nullptr;
}

void H2MatrixSolver::matVecMul(const double* X_dev, double* Y_dev, int precision) {
  if (levels < 0) return;
  check(mxg_h2_matvec(dev, X_dev, Y_dev, 1, local_bodies.second, precision));
}

void H2MatrixSolver::matVecMul(double* X_dev) {
  if (levels < 0) return;
  const long long n = local_bodies.second;
  void* tmp = nullptr;
  check(mxg_malloc(ctx, n * sizeof(double), &tmp));
  int rc = mxg_h2_matvec(dev, X_dev, static_cast<double*>(tmp), 1, n, MXG_FP64);
  if (rc == MXG_OK) rc = mxg_memcpy_d2d(ctx, X_dev, tmp, n * sizeof(double));
  mxg_free(ctx, tmp);
  check(rc);
}

int H2MatrixSolver::factorizeM(int precision) {
  if (ulv) check(mxg_ulv_destroy(ulv));
  ulv = nullptr;
  check(mxg_ulv_create(dev, precision, &ulv));
  std::vector<int> info(levels + 1, 0);
  check(mxg_ulv_info(ulv, info.data()));
  int singular = 0;
  for (long long l = levels; l >= 0; l--)
    if (info[l]) {
      std::printf("singularity detected at level %lld.\n", l);
      singular++;
    }
  return singular;
}

int H2MatrixSolver::factorizeBatch(const std::vector<H2MatrixSolver*>& ms, int precision) {
  std::vector<mxg_h2_t> hs;
  for (H2MatrixSolver* m : ms) {
    if (!m || m->levels < 0) throw std::runtime_error("factorizeBatch: matrix not constructed");
    if (m->ulv) check(mxg_ulv_destroy(m->ulv));
    m->ulv = nullptr;
    hs.push_back(m->dev);
  }
  std::vector<mxg_ulv_t> fs(ms.size(), nullptr);
  check(mxg_ulv_create_batch(hs.data(), (long long)hs.size(), precision, fs.data()));
  int singular = 0;
  for (size_t p = 0; p < ms.size(); p++) {
    ms[p]->ulv = fs[p];
    std::vector<int> info(ms[p]->levels + 1, 0);
    check(mxg_ulv_info(fs[p], info.data()));
    for (int v : info) singular += v ? 1 : 0;
  }
  return singular;
}

void H2MatrixSolver::solvePrecondition(double* X_dev) {
  if (levels < 0) return;
  if (!ulv) throw std::runtime_error("solvePrecondition before factorizeM");
  check(mxg_ulv_solve(ulv, X_dev, 1, local_bodies.second));
}

void H2MatrixSolver::solveGMRES(double tol, H2MatrixSolver& M, double* X_dev, const double* B_dev, long long inner,
                                long long outer, int mode, int inner_precision, mxg_h2_t inner_operator) {
  if (!M.ulv) throw std::runtime_error("solveGMRES: preconditioner not factorized");
  mxg_gmres_opts_t o{tol, (int)inner, (int)outer, mode, inner_precision, inner_operator};
  resid.assign(outer + 1, 0.);
  check(mxg_gmres(dev, M.ulv, &o, B_dev, X_dev, resid.data(), &iters, &inner_iters));
}

double H2MatrixSolver::solveRelErr(long long lenX, const double X[], const double ref[]) {
  double e0 = 0., e1 = 0.;
  for (long long i = 0; i < lenX; i++) {
    const double d = X[i] - ref[i];
    e0 += d * d;
    e1 += ref[i] * ref[i];
  }
  return std::sqrt(e0 / e1);
}

}  // namespace mxc
