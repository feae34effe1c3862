// H2MatrixSolver: the reference's host API (solver.hpp:11-47) at one rank,
// with every numeric step on the GPU through the mxg_* C ABI.
//
// Vectors passed to the methods are device pointers to tree-ordered FP64
// arrays of length nbodies (the reference's host-pointer twins copied with
// blocking cudaMemcpy; callers here keep their data in HBM).
#pragma once

#include <utility>
#include <vector>

#include "h2matrix.hpp"

namespace mxc {

class H2MatrixSolver {
 public:
  long long levels = -1;
  std::vector<H2Matrix> A;
  mxg_ctx_t ctx = nullptr;
  mxg_h2_t dev = nullptr;
  mxg_ulv_t ulv = nullptr;
  std::pair<long long, long long> local_bodies{0, 0};
  std::vector<double> resid;
  long long iters = 0;
  long long inner_iters = 0;
  BasisOptions opts;

  H2MatrixSolver() = default;
  // solver.cpp:12-41: epi == 0 selects the fixed-rank orthogonal preconditioner
  // (rank + leveled_rank per level, sibling far lists), otherwise the
  // interpolative operator at tolerance epi. `sampling` caps the far bodies
  // per box; `factorable` keeps the upper-level near blocks for factorizeM().
  H2MatrixSolver(mxg_ctx_t ctx, const MatrixAccessor& eval, double epi, long long rank, long long leveled_rank,
                 const std::vector<Cell>& cells, double theta, const double bodies[], long long levels,
                 long long sampling, bool factorable);
  // Same with an explicit basis policy (e.g. tolerance-truncated orthogonal
  // bases with a rank cap).
  H2MatrixSolver(mxg_ctx_t ctx, const MatrixAccessor& eval, const BasisOptions& opt, const std::vector<Cell>& cells,
                 double theta, const double bodies[], long long levels, bool factorable);
  H2MatrixSolver(const H2MatrixSolver&) = delete;
  H2MatrixSolver& operator=(const H2MatrixSolver&) = delete;
  H2MatrixSolver(H2MatrixSolver&& o) noexcept;
  H2MatrixSolver& operator=(H2MatrixSolver&& o) noexcept;
  ~H2MatrixSolver();

  // y = A x (solver.cpp:76-88); precision MXG_FP32 uses the FP32 operator copy.
  void matVecMul(const double* X_dev, double* Y_dev, int precision = MXG_FP64);
  // x = A x in place, the reference's calling convention.
  void matVecMul(double* X_dev);
  // ULV factorization (solver.cpp:90-100); returns the number of singular levels.
  int factorizeM(int precision = MXG_FP64);
  // ULV factorizations of several independent matrices on one context, each
  // phase of a level sweep one launch shared by all of them
  // (mxg_ulv_create_batch); returns the number of singular levels in total.
  static int factorizeBatch(const std::vector<H2MatrixSolver*>& ms, int precision = MXG_FP32);
  // x = M^{-1} x (solver.cpp:121-132).
  void solvePrecondition(double* X_dev);
  // GMRES (solver.cpp:144-201) with this matrix as the operator and M as the
  // left preconditioner. X holds the initial guess.
  // inner_operator (optional): the device matrix applied in the inner
  // iterations (e.g. the `dev` of this operator built at a looser tolerance);
  // the residual of every cycle uses this matrix.
  void solveGMRES(double tol, H2MatrixSolver& M, double* X_dev, const double* B_dev, long long inner_iters_per_cycle,
                  long long outer_iters, int mode = MXG_GMRES_REFERENCE, int inner_precision = MXG_FP64,
                  mxg_h2_t inner_operator = nullptr);
  // solver.cpp:229-238 (host arrays).
  static double solveRelErr(long long lenX, const double X[], const double ref[]);

 private:
  void build(const MatrixAccessor& eval, const std::vector<Cell>& cells, double theta, const double bodies[],
             long long lv, bool factorable);
  void release();
};

}  // namespace mxc
