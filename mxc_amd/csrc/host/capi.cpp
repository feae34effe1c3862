// C ABI of the host layer (include/mxc_h2.h).
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

#include "build_tree.hpp"
#include "far_bodies.hpp"
#include "kernel.hpp"
#include "mxc_h2.h"
#include "solver.hpp"
#include "test_funcs.hpp"

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

void require(bool c, const char* msg) {
  if (!c) throw std::invalid_argument(msg);
}

std::vector<mxc::Cell> cells_from(long long levels, const long long* child, const long long* body, const double* R,
                                  const double* C) {
  const long long nc = (2ll << levels) - 1;
  std::vector<mxc::Cell> cells(nc);
  for (long long i = 0; i < nc; i++) {
    if (child) cells[i].Child = {child[2 * i], child[2 * i + 1]};
    if (body) cells[i].Body = {body[2 * i], body[2 * i + 1]};
    for (int d = 0; d < 3; d++) {
      if (R) cells[i].R[d] = R[3 * i + d];
      if (C) cells[i].C[d] = C[3 * i + d];
    }
  }
  return cells;
}

std::unique_ptr<mxc::MatrixAccessor> accessor(int kernel, double diag, double alpha) {
  switch (kernel) {
    case MXG_KERNEL_LAPLACE: return std::make_unique<mxc::Laplace3D>(1. / diag);
    case MXG_KERNEL_YUKAWA: return std::make_unique<mxc::Yukawa3D>(1. / diag, alpha);
    case MXG_KERNEL_GAUSSIAN: return std::make_unique<mxc::Gaussian>(alpha);
    default: throw std::invalid_argument("unknown kernel");
  }
}

}  // namespace

struct mxc_problem_s {
  mxg_ctx_t ctx = nullptr;
  long long n = 0, leaf = 0, levels = 0;
  std::vector<double> bodies;
  std::vector<mxc::Cell> cells;
  void* d_bodies = nullptr;
};

struct mxc_solver_s {
  mxc_problem_s* p = nullptr;
  mxc::H2MatrixSolver s;
};

extern "C" {

const char* mxc_last_error(void) { return g_err.c_str(); }

int mxc_generate_points(int geometry, long long n, unsigned int seed, double* out) {
  return guard([&] {
    require(out && n > 0, "bad arguments");
    if (geometry == 0) mxc::uniform_unit_cube_rnd(out, n, 1., 3, seed);
    else if (geometry == 1) mxc::sphere_surface_rnd(out, n, seed);
    else if (geometry == 2) mxc::mesh_sphere(out, n, std::sqrt(n / (4. * M_PI)));
    else throw std::invalid_argument("unknown geometry");
  });
}

int mxc_uniform_vector(double* x, long long n, double lo, double hi, unsigned long long seed) {
  return guard([&] {
    require(x && n >= 0, "bad arguments");
    mxc::uniform_vector(x, n, lo, hi, seed);
  });
}

long long mxc_tree_levels(long long n, long long leaf) { return mxc::treeLevels(n, leaf); }

int mxc_build_tree(double* bodies, long long n, long long levels, long long* child, long long* body, double* R, double* C) {
  return guard([&] {
    require(bodies && n > 0 && levels >= 0 && (1ll << levels) <= n, "bad arguments");
    const long long nc = (2ll << levels) - 1;
    std::vector<mxc::Cell> cells(nc);
    mxc::buildBinaryTree(cells.data(), bodies, n, levels);
    for (long long i = 0; i < nc; i++) {
      child[2 * i] = cells[i].Child[0];
      child[2 * i + 1] = cells[i].Child[1];
      body[2 * i] = cells[i].Body[0];
      body[2 * i + 1] = cells[i].Body[1];
      for (int d = 0; d < 3; d++) {
        R[3 * i + d] = cells[i].R[d];
        C[3 * i + d] = cells[i].C[d];
      }
    }
  });
}

long long mxc_csr(int far, double theta, long long levels, const long long* child, const long long* body, const double* R,
                  const double* C, long long* rowptr, long long* col, long long cap) {
  try {
    std::vector<mxc::Cell> cells = cells_from(levels, child, body, R, C);
    mxc::CSR c(far ? 'F' : 'N', cells, cells, theta);
    const long long nnz = (long long)c.ColIndex.size();
    if (cap >= nnz) {
      std::copy(c.RowIndex.begin(), c.RowIndex.end(), rowptr);
      std::copy(c.ColIndex.begin(), c.ColIndex.end(), col);
    }
    return nnz;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

long long mxc_far_sample(long long levels, const long long* body, const long long* bfar_rp, const long long* bfar_col,
                         const long long* near_rp, const long long* near_col, const long long* far_rp,
                         const long long* far_col, int with_near, long long cell, long long sampling, long long* out,
                         long long cap) {
  try {
    std::vector<mxc::Cell> cells = cells_from(levels, nullptr, body, nullptr, nullptr);
    const long long nc = (2ll << levels) - 1;
    auto mk = [&](const long long* rp, const long long* col) {
      mxc::CSR c;
      c.RowIndex.assign(rp, rp + nc + 1);
      c.ColIndex.assign(col, col + rp[nc]);
      return c;
    };
    mxc::CSR bf = mk(bfar_rp, bfar_col), nr = mk(near_rp, near_col), fr = mk(far_rp, far_col);
    std::vector<long long> idx = mxc::sampleFarBodies(cells, bf, nr, fr, with_near != 0, cell, sampling);
    if (cap >= (long long)idx.size()) std::copy(idx.begin(), idx.end(), out);
    return (long long)idx.size();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int mxc_problem_create(mxg_ctx_t ctx, const double* points, long long n, long long leaf, mxc_problem_t* out) {
  return guard([&] {
    require(ctx && points && out && n > 0 && leaf > 0, "bad arguments");
    auto p = std::make_unique<mxc_problem_s>();
    p->ctx = ctx;
    p->n = n;
    p->leaf = leaf;
    p->levels = mxc::treeLevels(n, leaf);
    p->bodies.assign(points, points + 3 * n);
    p->cells.resize((2ll << p->levels) - 1);
    mxc::buildBinaryTree(p->cells.data(), p->bodies.data(), n, p->levels);
    if (mxg_malloc(ctx, 3 * n * sizeof(double), &p->d_bodies) != MXG_OK) throw std::runtime_error(mxg_last_error());
    if (mxg_memcpy_h2d(ctx, p->d_bodies, p->bodies.data(), 3 * n * sizeof(double)) != MXG_OK)
      throw std::runtime_error(mxg_last_error());
    *out = p.release();
  });
}

int mxc_problem_destroy(mxc_problem_t p) {
  return guard([&] {
    if (!p) return;
    if (p->d_bodies) mxg_free(p->ctx, p->d_bodies);
    delete p;
  });
}

long long mxc_problem_levels(mxc_problem_t p) { return p ? p->levels : -1; }

int mxc_problem_bodies(mxc_problem_t p, double* out) {
  return guard([&] {
    require(p && out, "bad arguments");
    std::copy(p->bodies.begin(), p->bodies.end(), out);
  });
}

const double* mxc_problem_device_bodies(mxc_problem_t p) { return p ? static_cast<const double*>(p->d_bodies) : nullptr; }
mxg_ctx_t mxc_problem_ctx(mxc_problem_t p) { return p ? p->ctx : nullptr; }

int mxc_solver_create(mxc_problem_t p, const mxc_h2_opts_t* o, mxc_solver_t* out) {
  return guard([&] {
    require(p && o && out, "bad arguments");
    auto acc = accessor(o->kernel, o->diag, o->alpha);
    mxc::BasisOptions b;
    b.tol = o->tol;
    b.rank = o->rank;
    b.leveled_rank = o->leveled_rank;
    b.orth = o->orth != 0;
    b.hss_far = o->hss_far != 0;
    b.sampling = o->sampling;
    b.householder_rows = o->householder_rows != 0;
    b.explicit_far = o->explicit_far != 0;
    auto s = std::make_unique<mxc_solver_s>();
    s->p = p;
    s->s = mxc::H2MatrixSolver(p->ctx, *acc, b, p->cells, o->theta, p->bodies.data(), p->levels, o->factorable != 0);
    *out = s.release();
  });
}

int mxc_solver_destroy(mxc_solver_t s) {
  return guard([&] { delete s; });
}

mxg_h2_t mxc_solver_device(mxc_solver_t s) { return s ? s->s.dev : nullptr; }
mxg_ulv_t mxc_solver_ulv(mxc_solver_t s) { return s ? s->s.ulv : nullptr; }

int mxc_solver_dims(mxc_solver_t s, long long level, long long* dims, long long* ranks) {
  return guard([&] {
    require(s && level >= 0 && level <= s->s.levels, "bad level");
    const mxc::H2Matrix& A = s->s.A[level];
    for (long long i = 0; i < A.nboxes; i++) {
      if (dims) dims[i] = A.Dims[i];
      if (ranks) ranks[i] = A.DimsLr[i];
    }
  });
}

int mxc_solver_counts(mxc_solver_t s, long long level, long long* counts) {
  return guard([&] {
    require(s && counts && level >= 0 && level <= s->s.levels, "bad level");
    counts[0] = s->s.A[level].ARows.back();
    counts[1] = s->s.A[level].CRows.back();
  });
}

int mxc_solver_matvec(mxc_solver_t s, const double* d_x, double* d_y, int precision) {
  return guard([&] {
    require(s, "null solver");
    s->s.matVecMul(d_x, d_y, precision);
  });
}

int mxc_solver_factorize(mxc_solver_t s, int precision, int* info) {
  return guard([&] {
    require(s, "null solver");
    s->s.factorizeM(precision);
    if (info && mxg_ulv_info(s->s.ulv, info) != MXG_OK) throw std::runtime_error(mxg_last_error());
  });
}

int mxc_solver_factorize_batch(mxc_solver_t* s, long long count, int precision, int* info) {
  return guard([&] {
    require(s && count >= 1, "null or empty batch");
    std::vector<mxc::H2MatrixSolver*> ms;
    for (long long p = 0; p < count; p++) {
      require(s[p], "null solver in the batch");
      ms.push_back(&s[p]->s);
    }
    mxc::H2MatrixSolver::factorizeBatch(ms, precision);
    if (!info) return;
    for (auto* m : ms) {
      if (mxg_ulv_info(m->ulv, info) != MXG_OK) throw std::runtime_error(mxg_last_error());
      info += m->levels + 1;
    }
  });
}

int mxc_solver_precondition(mxc_solver_t s, double* d_x) {
  return guard([&] {
    require(s, "null solver");
    s->s.solvePrecondition(d_x);
  });
}

int mxc_solver_gmres(mxc_solver_t A, mxc_solver_t M, const mxg_gmres_opts_t* o, const double* d_b, double* d_x,
                     double* resid, long long* iters, long long* inner) {
  return guard([&] {
    require(A && M && o && resid, "bad arguments");
    A->s.solveGMRES(o->tol, M->s, d_x, d_b, o->restart, o->max_cycles, o->mode, o->inner_precision, o->inner_operator);
    std::copy(A->s.resid.begin(), A->s.resid.begin() + (A->s.iters + 1), resid);
    if (iters) *iters = A->s.iters;
    if (inner) *inner = A->s.inner_iters;
  });
}

int mxc_solver_gmres_batch(mxc_solver_t A, mxc_solver_t M, const mxg_gmres_opts_t* o, long long nrhs, const double* d_B,
                           long long ldb, double* d_X, long long ldx, double* resid, long long* iters, long long* inner) {
  return guard([&] {
    require(A && M && o && resid, "bad arguments");
    if (!M->s.ulv) throw std::runtime_error("preconditioner not factorized");
    if (mxg_gmres_batch(A->s.dev, M->s.ulv, o, nrhs, d_B, ldb, d_X, ldx, resid, iters, inner) != MXG_OK)
      throw std::runtime_error(std::string("device layer: ") + mxg_last_error());
  });
}

int mxc_dense_matvec(mxc_problem_t p, mxg_kernel_t k, const double* d_x, double* d_y) {
  return guard([&] {
    require(p, "null problem");
    if (mxg_dense_matvec(p->ctx, k, static_cast<const double*>(p->d_bodies), p->n, d_x, d_y) != MXG_OK)
      throw std::runtime_error(mxg_last_error());
  });
}

}  // extern "C"
