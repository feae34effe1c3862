/* mxc_h2.h - C ABI of the C++ host layer (tree, block lists, far-body
 * sampling, H2MatrixSolver) for callers such as Python/ctypes.
 *
 * The host layer mirrors the reference's C++ API: buildBinaryTree and CSR
 * (build_tree.hpp:7-26), the generators of test_funcs.hpp:31-61, and
 * H2MatrixSolver (solver.hpp:11-47); the lorasp CLI uses the C++ classes
 * directly. All numeric work runs on the GPU through mxc_gpu.h. Vectors are
 * device pointers to tree-ordered FP64 arrays.
 */
#ifndef MXC_H2_H
#define MXC_H2_H

#include "mxc_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* mxc_last_error(void);

/* ---- generators (test_funcs.hpp:31-61) ----
 * geometry 0: uniform_unit_cube_rnd (diameter 1, dim 3, std::mt19937(seed));
 *          1: uniform on the unit sphere surface (std::mt19937(seed));
 *          2: mesh_sphere of radius sqrt(n / 4 pi) (seed unused). */
int mxc_generate_points(int geometry, long long n, unsigned int seed, double* out);
/* U(lo, hi) from std::mt19937_64(seed) */
int mxc_uniform_vector(double* x, long long n, double lo, double hi, unsigned long long seed);

/* ---- tree and lists (build_tree.cpp) ---- */
long long mxc_tree_levels(long long n, long long leaf);
/* Sorts bodies into tree order in place; per cell: child[2], body[2], R[3], C[3]. */
int mxc_build_tree(double* bodies, long long n, long long levels, long long* child, long long* body, double* R,
                   double* C);
/* Near (far = 0) or far (far = 1) CSR over all cells; returns nnz, fills rowptr
 * (ncells + 1) and col when cap >= nnz. */
long long mxc_csr(int far, double theta, long long levels, const long long* child, const long long* body,
                  const double* R, const double* C, long long* rowptr, long long* col, long long cap);
/* Sampled far-body list of one cell: bfar = the basis far lists (Far, or the
 * theta = 0 sibling lists), near/far = the matrix's theta lists defining the
 * distance shells, with_near != 0 when bfar covers every body outside the
 * cell. Returns the length and fills out when cap >= length. */
long long mxc_far_sample(long long levels, const long long* body, const long long* bfar_rp, const long long* bfar_col,
                         const long long* near_rp, const long long* near_col, const long long* far_rp,
                         const long long* far_col, int with_near, long long cell, long long sampling, long long* out,
                         long long cap);

/* ---- problem: points + tree on one device ---- */
typedef struct mxc_problem_s* mxc_problem_t;
int mxc_problem_create(mxg_ctx_t ctx, const double* points, long long n, long long leaf, mxc_problem_t* p);
int mxc_problem_destroy(mxc_problem_t p);
long long mxc_problem_levels(mxc_problem_t p);
int mxc_problem_bodies(mxc_problem_t p, double* out); /* tree-ordered bodies, 3n doubles */
const double* mxc_problem_device_bodies(mxc_problem_t p);
mxg_ctx_t mxc_problem_ctx(mxc_problem_t p);

/* ---- H^2 matrix + solver (H2MatrixSolver) ---- */
typedef struct {
  int kernel;   /* MXG_KERNEL_* */
  double diag;  /* K(0) */
  double alpha; /* kernel parameter */
  double theta; /* admissibility */
  double tol;   /* > 0: CPQR threshold rank (reference eps) */
  long long rank;         /* > 0: rank cap at the leaves (reference fixed rank) */
  long long leveled_rank; /* rank cap growth per level */
  int orth;     /* orthogonal (1) or interpolative (0) bases */
  int hss_far;  /* basis far lists from the sibling lists */
  long long sampling;  /* far bodies per box, <= 0 for all */
  int factorable;      /* keep what mxc_solver_factorize needs */
  int householder_rows; /* reduce far rows by the Householder R factor (MXG_H2_HOUSEHOLDER_ROWS) */
  int explicit_far;     /* hand the far lists over as one index per body (mxg_h2_construct_level)
                           instead of runs of consecutive bodies (mxg_h2_construct_level_runs) */
} mxc_h2_opts_t;
typedef struct mxc_solver_s* mxc_solver_t;
int mxc_solver_create(mxc_problem_t p, const mxc_h2_opts_t* o, mxc_solver_t* s);
int mxc_solver_destroy(mxc_solver_t s);
mxg_h2_t mxc_solver_device(mxc_solver_t s);
mxg_ulv_t mxc_solver_ulv(mxc_solver_t s);
/* dims/ranks of the boxes of one level (2^level entries each) */
int mxc_solver_dims(mxc_solver_t s, long long level, long long* dims, long long* ranks);
/* counts[0] near pairs, counts[1] far pairs of one level */
int mxc_solver_counts(mxc_solver_t s, long long level, long long* counts);
int mxc_solver_matvec(mxc_solver_t s, const double* d_x, double* d_y, int precision);
/* info: levels + 1 ints (singular boxes per level) */
int mxc_solver_factorize(mxc_solver_t s, int precision, int* info);
/* ULV factorizations of count solvers on one context in shared launches
 * (mxg_ulv_create_batch); info: levels + 1 ints per solver, concatenated */
int mxc_solver_factorize_batch(mxc_solver_t* s, long long count, int precision, int* info);
int mxc_solver_precondition(mxc_solver_t s, double* d_x);
/* o->inner_operator may name the device matrix of another solver
 * (mxc_solver_device) used for the inner iterations. */
int mxc_solver_gmres(mxc_solver_t A, mxc_solver_t M, const mxg_gmres_opts_t* o, const double* d_b, double* d_x,
                     double* resid, long long* iters, long long* inner);
/* nrhs right-hand sides in lockstep (mxg_gmres_batch): B, X n x nrhs with
 * leading dimensions ldb / ldx; resid nrhs x (max_cycles + 1); iters, inner
 * nrhs entries */
int mxc_solver_gmres_batch(mxc_solver_t A, mxc_solver_t M, const mxg_gmres_opts_t* o, long long nrhs, const double* d_B,
                           long long ldb, double* d_X, long long ldx, double* resid, long long* iters, long long* inner);
/* dense FP64 check y = K x over the problem's bodies */
int mxc_dense_matvec(mxc_problem_t p, mxg_kernel_t k, const double* d_x, double* d_y);

#ifdef __cplusplus
}
#endif
#endif /* MXC_H2_H */
