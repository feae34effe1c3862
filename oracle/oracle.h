/* CPU oracle for the H^2 solve path: an FP64 restatement of the reference
 * algorithm (delta-leader/mxc) used ONLY by tests, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, as the checker. The product path never
 * links, imports or calls anything here.
 *
 * Every routine follows a reference file:line, cited at its definition in
 * oracle.cpp. Arithmetic is real FP64 (the reference's complex<double> with
 * imaginary parts identically zero for the Laplace/Yukawa/Gaussian kernels).
 */
#ifndef MXC_ORACLE_H
#define MXC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids */
enum { ORC_LAPLACE = 0, ORC_YUKAWA = 1, ORC_GAUSSIAN = 2 };

/* point and right-hand-side generators */
void orc_uniform_cube_rnd(double* bodies, long long n, double diameter, long long dim, unsigned int seed);
void orc_sphere_rnd(double* bodies, long long n, unsigned int seed);
void orc_mesh_sphere(double* bodies, long long n, double r);
void orc_uniform_vec(double* x, long long n, double lo, double hi, unsigned long long seed);

/* tree and block lists */
long long orc_levels(long long n, long long leaf);
void orc_build_tree(double* bodies, long long n, long long levels, long long* child, long long* body,
                    double* R, double* C);
long long orc_csr(int far, double theta, long long levels, const long long* child, const double* R,
                  const double* C, long long* rowptr, long long* col, long long cap);
long long orc_far_sample(long long levels, const long long* body, const long long* bfar_rp, const long long* bfar_col,
                         const long long* near_rp, const long long* near_col, const long long* far_rp,
                         const long long* far_col, int with_near, long long cell, long long sampling, long long* out,
                         long long cap);

/* far-row reduction: 1 = Rademacher sketch (default), 0 = the reference's Householder R factor */
void orc_set_sketch(int on);
double orc_rademacher(unsigned long long t, unsigned long long i);

/* kernel entries */
double orc_kernel_eval(int kid, double diag, double alpha, double d);
void orc_gen_matrix(int kid, double diag, double alpha, long long m, long long n, const double* bi,
                    const double* bj, double* A);
void orc_dense_matvec(int kid, double diag, double alpha, const double* bodies, long long n, const double* x,
                      double* y, int nthreads);

/* dense linear algebra entry points (pinned against LAPACK known answers) */
void orc_qr_r(long long m, long long n, double* A, long long lda);
long long orc_cpqr(long long m, long long n, double* A, long long lda, long long* perm, double tol,
                   double* maxpivot, long long* nonzero);
double orc_lu_solve(long long n, long long nrhs, double* A, double* B);

/* H^2 solver (H2MatrixSolver semantics) */
typedef struct orc_problem orc_problem;
typedef struct orc_h2 orc_h2;
orc_problem* orc_problem_create(const double* points, long long n, long long leaf);
void orc_problem_destroy(orc_problem* p);
long long orc_problem_levels(const orc_problem* p);
void orc_problem_bodies(const orc_problem* p, double* out);

/* tol > 0: Eigen CPQR threshold rank; maxrank > 0 caps the rank (fixed rank when tol == 0).
 * far_theta < 0 selects the sibling (HSS) far lists for the basis, else Far(theta).
 * orth != 0 selects the orthogonal basis form. sampling <= 0 keeps every far body. */
orc_h2* orc_h2_create(const orc_problem* p, int kid, double diag, double alpha, double theta, double tol,
                      long long rank, long long leveled_rank, int orth, int hss_far, long long sampling,
                      int nthreads);
void orc_h2_destroy(orc_h2* h);
void orc_h2_set_threads(orc_h2* h, int nthreads);
long long orc_h2_nboxes(const orc_h2* h, long long level);
void orc_h2_dims(const orc_h2* h, long long level, long long* dims, long long* ranks);
/* what: 0 = near blocks A, 1 = bases Q, 2 = R (V after orc_h2_factorize), 3 = couplings C */
long long orc_h2_download(const orc_h2* h, long long level, int what, double* out);
/* level-local near (far = 0) or far (far = 1) list; returns the entry count */
long long orc_h2_level_list(const orc_h2* h, long long level, int far, long long* rp, long long* col);
long long orc_h2_near_count(const orc_h2* h, long long level);
long long orc_h2_far_count(const orc_h2* h, long long level);
void orc_h2_matvec(orc_h2* h, double* x);
int orc_h2_factorize(orc_h2* h);
/* diagnostic: the far field of one level, through the H^2 bases and directly */
void orc_h2_far_level(orc_h2* h, long long level, const double* x, double* y_approx, double* y_exact);
void orc_h2_precondition(orc_h2* h, double* x);
long long orc_h2_gmres(orc_h2* A, orc_h2* M, double tol, double* x, const double* b, long long inner,
                       long long outer, double* resid);
double orc_rel_err(long long n, const double* x, const double* ref);

#ifdef __cplusplus
}
#endif
#endif
