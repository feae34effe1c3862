/* mxc_gpu.h - C ABI of the MI355X (gfx950) device layer of the H^2 solve path.
 *
 * This is the boundary that replaces the reference's CUDA device layer
 * (delta-leader/mxc): gpu_handles.cuh:20-25, device_factorize.cuh:67-78,
 * device_csr_matrix.cuh:65-77 and the device GMRES. Plain C: opaque handles,
 * pointers and sizes, int status codes (0 = success), no C++/STL/MPI/complex
 * types. Arithmetic is real (the reference's complex<double> has zero
 * imaginary parts for these kernels). Device pointers are marked d_.
 *
 * One process drives one GPU; the reference's MPI/NCCL distribution collapses
 * to a single rank holding the whole tree, so there are no communicators.
 */
#ifndef MXC_GPU_H
#define MXC_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
  MXG_OK = 0,
  MXG_ERR_ARG = 1,      /* invalid argument or size beyond a documented limit */
  MXG_ERR_HIP = 2,      /* a HIP runtime call failed */
  MXG_ERR_NOMEM = 3,    /* device allocation failed */
  MXG_ERR_STATE = 4,    /* call out of order (e.g. level not constructed) */
  MXG_ERR_SINGULAR = 5  /* factorization flagged a singular diagonal block */
};

/* kernel functions K(d) with K(0) = diag (kernel.hpp:28-61) */
enum { MXG_KERNEL_LAPLACE = 0, MXG_KERNEL_YUKAWA = 1, MXG_KERNEL_GAUSSIAN = 2 };
typedef struct {
  int id;        /* MXG_KERNEL_* */
  double diag;   /* K(0) */
  double alpha;  /* Yukawa decay / Gaussian width */
} mxg_kernel_t;

/* precision of a factorization or of an operator application; MXG_BF16 (ULV
 * only): factorization in FP32, factors stored in bfloat16 for the
 * substitution, which accumulates in FP32; MXG_BF16_SCHUR (ULV only): the same,
 * with the Schur-complement products of the factorization
 * (h2matrix.cpp:405-424) taking bfloat16 inputs on the bfloat16 matrix cores
 * (FP32 accumulation). mxg_gemm takes MXG_BF16 for FP32 matrices multiplied
 * with bfloat16-rounded inputs. */
enum { MXG_FP64 = 0, MXG_FP32 = 1, MXG_BF16 = 2, MXG_BF16_SCHUR = 3 };

typedef struct mxg_ctx_s* mxg_ctx_t; /* device + stream (replaces deviceHandle_t) */
typedef struct mxg_h2_s* mxg_h2_t;   /* one H^2 matrix resident on the device   */
typedef struct mxg_ulv_s* mxg_ulv_t; /* one ULV factorization of an H^2 matrix  */

/* ---- context: replaces cudaSetDevice/initGpuEnvs/finalizeGpuEnvs
 *      (gpu_handles.cu:5-12, 54-71) ---- */
int mxg_init(int device, mxg_ctx_t* ctx);
int mxg_finalize(mxg_ctx_t ctx);
int mxg_device_count(int* count);
int mxg_synchronize(mxg_ctx_t ctx);
const char* mxg_last_error(void);
/* HIP stream of the context (a hipStream_t) */
void* mxg_stream(mxg_ctx_t ctx);
/* free and total device memory of the context's GPU (hipMemGetInfo), for
 * callers that admit work by its memory footprint */
int mxg_mem_info(mxg_ctx_t ctx, size_t* free_bytes, size_t* total_bytes);

/* device memory for callers without an allocator of their own */
int mxg_malloc(mxg_ctx_t ctx, size_t bytes, void** d_ptr);
int mxg_free(mxg_ctx_t ctx, void* d_ptr);
int mxg_memcpy_h2d(mxg_ctx_t ctx, void* d_dst, const void* src, size_t bytes);
int mxg_memcpy_d2h(mxg_ctx_t ctx, void* dst, const void* d_src, size_t bytes);
int mxg_memcpy_d2d(mxg_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes);
int mxg_memset(mxg_ctx_t ctx, void* d_dst, int value, size_t bytes);

/* C = alpha op(A) op(B) + beta C on device matrices (column major; ta/tb != 0
 * transpose), on the matrix cores; precision MXG_FP64 (double*) or MXG_FP32 (float*). */
int mxg_gemm(mxg_ctx_t ctx, int precision, int ta, int tb, long long m, long long n, long long k, double alpha,
             const void* d_A, long long lda, const void* d_B, long long ldb, double beta, void* d_C, long long ldc);

/* The solve phase's block product on one block, for checks of its own:
 * Y[:, q] = B X[:, q] (q < nrhs, 1..8 right-hand sides) for an n x m column-major
 * B (leading dimension ldb), with the lane layout the matvec and the
 * substitution use for a block of this size (one wave, or a workgroup).
 * storage MXG_FP64: B, X, Y double; MXG_FP32: float; MXG_BF16: B is bfloat16
 * (uint16 holding the upper half of the FP32 pattern), X and Y float, FP32
 * arithmetic. matrix_cores != 0: an FP64 or FP32 block whose columns admit
 * 16-byte loads runs on the matrix cores (v_mfma_f64_16x16x4f64 /
 * v_mfma_f32_16x16x4f32), as the matvec's near and coupling products do when
 * its tuning selects them; otherwise on the vector units. No counterpart in
 * the reference (cuSPARSE SpMV, device_csr_matrix.cu). */
int mxg_block_gemv(mxg_ctx_t ctx, int storage, long long n, long long m, const void* d_B, long long ldb,
                   const void* d_X, long long ldx, void* d_Y, long long ldy, long long nrhs, int matrix_cores);

/* ---- kernel blocks: replaces gen_matrix (kernel.cpp:71-85) on the device.
 * A[i + j*lda] = K(|bi_i - bj_j|), column major, points 3 doubles each. ---- */
int mxg_gen_block(mxg_ctx_t ctx, mxg_kernel_t k, long long m, long long n, const double* d_bi,
                  const double* d_bj, double* d_A, long long lda);

/* Dense FP64 check y = A x, matrix free (main.cpp:83-84,114 with kernel.cpp:25-35). */
int mxg_dense_matvec(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n, const double* d_x,
                     double* d_y);
/* Y = K X for nrhs columns (leading dimensions ldx / ldy): every kernel entry
 * is evaluated once per group of up to 8 columns. */
int mxg_dense_matvec_batch(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n, const double* d_X,
                           long long ldx, double* d_Y, long long ldy, long long nrhs);
/* mxg_dense_matvec_batch without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). The partial sums go to a
 * 64 MiB scratch array that the library holds per GPU for the life of the
 * process, in row slabs, so results equal mxg_dense_matvec_batch's bit for
 * bit. n is limited to 2^28 points (MXG_ERR_ARG beyond). Calls that overlap
 * on one GPU (from several threads or contexts) share the array and race:
 * serialise them. */
int mxg_dense_matvec_batch_noalloc(mxg_ctx_t ctx, mxg_kernel_t k, const double* d_bodies, long long n,
                                   const double* d_X, long long ldx, double* d_Y, long long ldy, long long nrhs);

/* ---- H^2 matrix: replaces createSpMatrixDesc/createDeviceCsr/createMatrixDesc and
 *      the numeric work of H2Matrix::construct (h2matrix.cpp:72-249). ---- */
enum {
  MXG_H2_FACTORABLE = 1, /* keep the assembled upper-level near blocks for mxg_ulv_create */
  /* reduce a box's far rows (nf > M) by the Householder R factor of the far
   * block, as the reference does (h2matrix.cpp:78-79), instead of the default
   * M x M Rademacher sketch Omega * G (Omega(i, t) = -1 when bit i mod 64 of
   * splitmix64-style word (t, i / 64) is set) */
  MXG_H2_HOUSEHOLDER_ROWS = 2
};
/* bodies: host array, tree order (3 doubles per body). */
int mxg_h2_create(mxg_ctx_t ctx, mxg_kernel_t k, const double* bodies, long long nbodies, long long levels,
                  int flags, mxg_h2_t* h);
int mxg_h2_destroy(mxg_h2_t h);

/* Construct one level, leaves first (level = levels) then upward to 0.
 * dims: box sizes (leaf: body counts; upper: sum of the two children's ranks,
 *   which must equal what the lower level returned).
 * leaf_begin: first body of each leaf box (leaf level only, else NULL).
 * near_rp/near_col, far_rp/far_col: level-local CSR lists (columns are box
 *   indices within the level, rows sorted).
 * far_off[nboxes+1], far_idx: sampled far-body indices per box (may be empty).
 * Basis rule (h2matrix.cpp:13-63): rank = Eigen CPQR threshold rank when
 *   tol > 0 (else min(M, #far)), capped at maxrank when maxrank > 0; orth != 0
 *   builds the orthogonal basis, orth == 0 the interpolative one.
 * ranks_out[nboxes]: receives the skeleton ranks. */
int mxg_h2_construct_level(mxg_h2_t h, long long level, long long nboxes, const long long* dims,
                           const long long* leaf_begin, const long long* near_rp, const long long* near_col,
                           const long long* far_rp, const long long* far_col, const long long* far_off,
                           const long long* far_idx, double tol, long long maxrank, int orth,
                           long long* ranks_out);
/* The same with each box's far-body list given as runs of consecutive body
 * indices, in list order: run k covers bodies far_runs[2k] ..
 * far_runs[2k] + far_runs[2k + 1] - 1, box i owns runs far_run_off[i] ..
 * far_run_off[i + 1] - 1. A far cell kept whole is one run, so the list of
 * the reference (h-matrix.cpp:121-127) crosses the bus in a few runs per cell
 * and is expanded on the device. */
int mxg_h2_construct_level_runs(mxg_h2_t h, long long level, long long nboxes, const long long* dims,
                                const long long* leaf_begin, const long long* near_rp, const long long* near_col,
                                const long long* far_rp, const long long* far_col, const long long* far_run_off,
                                const long long* far_runs, double tol, long long maxrank, int orth,
                                long long* ranks_out);

/* Copy of level data to the host (copyDataOutMatrixDesc, device_precondition.cu:204-208).
 * what: 0 = near blocks A, 1 = bases Q (M*M per box), 2 = R (M*M per box), 3 = couplings C,
 *       4 = skeleton body indices (as doubles). count = number of doubles. */
int mxg_h2_download(mxg_h2_t h, long long level, int what, double* dst, long long count);
long long mxg_h2_level_size(mxg_h2_t h, long long level, int what);

/* Level structure of an H^2 matrix built on the host, for callers that hand
 * over their own level data instead of constructing on the device: the
 * reference's createMatrixDesc + copyDataInMatrixDesc
 * (device_precondition.cu:39-154,198-202) as used by
 * H2MatrixSolver::factorizeDeviceM (solver.cpp:102-119), and the
 * createSpMatrixDesc of allocSparseMV (device_csr_matrix.cu:164-225,
 * solver.cpp:59-64). Levels are defined leaves first, as in construction:
 * dims as in mxg_h2_construct_level, ranks[nboxes] the skeleton ranks
 * (0 at the root), near/far the level-local CSR lists. The leaf's skeleton
 * indices are its bodies in tree order and its Q is the identity; every other
 * datum starts at zero and is given with mxg_h2_upload. */
int mxg_h2_define_level(mxg_h2_t h, long long level, long long nboxes, const long long* dims, const long long* ranks,
                        const long long* near_rp, const long long* near_col, const long long* far_rp,
                        const long long* far_col);
/* Host data into a level that is constructed or defined (copyDataInMatrixDesc,
 * device_precondition.cu:198-202). what and layout as in mxg_h2_download; an
 * upper level's near blocks A hold the couplings of the far child pairs
 * (h2matrix.cpp:172-173), the blocks of near child pairs are filled by the
 * factorization. count = mxg_h2_level_size. Factorizations made before the
 * upload keep their own copies and are not affected. */
int mxg_h2_upload(mxg_h2_t h, long long level, int what, const double* src, long long count);

/* Launch-shape autotuner of the matvec (no counterpart in the reference): times
 * every candidate shape of one precision's matvec for nrhs (1..8) right-hand
 * sides, reps products each after a warm-up, and keeps the fastest for later
 * matvecs. Candidates: the read-once symmetric products or the general ones
 * for the leaf near blocks (matrices built from a kernel) and for the
 * couplings (interpolative bases), the matrix cores or the vector units for
 * the general near and coupling products, and 16-, 32- or 64-row chunks for
 * the upward products of levels with fewer than 1024 blocks. Every shape
 * gives the same matvec up to rounding. best_ms: time per product of the
 * kept shape. */
int mxg_h2_autotune(mxg_h2_t h, int precision, long long nrhs, int reps, double* best_ms);
/* The kept shape (near_sym / far_sym: 1 when the symmetric products run;
 * matrix_cores: 1 when the general near and coupling products of aligned
 * blocks run on the matrix cores) and setting one by hand (split_rows: a
 * multiple of 4 in 8..256). */
int mxg_h2_get_tuning(mxg_h2_t h, int precision, int* near_sym, int* far_sym, int* split_rows, int* matrix_cores);
int mxg_h2_set_tuning(mxg_h2_t h, int precision, int near_sym, int far_sym, int split_rows, int matrix_cores);
/* H^2 matvec y = A x on the device (matVecDeviceH2, device_csr_matrix.cu:328-338).
 * x, y: tree-ordered FP64 vectors of length nbodies, nrhs columns with leading
 * dimension ld. precision MXG_FP32 applies an FP32 copy of the operator data
 * (built on first use) with FP32 inputs rounded from x and FP64 outputs. */
int mxg_h2_matvec(mxg_h2_t h, const double* d_x, double* d_y, long long nrhs, long long ld, int precision);

/* ---- ULV factorization: replaces compute_factorize / factorizeCopyNext / check_info
 *      (device_factorize.cu:56-142, h2matrix.cpp:346-446, device_precondition.cu:210-215) ---- */
int mxg_ulv_create(mxg_h2_t h, int precision, mxg_ulv_t* f);
/* Factorizations of count independent H^2 matrices on one context, whose
 * trees may differ in depth, box sizes, ranks and near lists: sweep t factors
 * level L_p - t of every matrix p that has it, and each phase of a sweep (the
 * basis rotations, the rr inverses, V, the off-diagonal rotations, the Schur
 * complements, the hand-over to the parent level) is one launch shared by all
 * of them. out[count] receives one factorization per matrix, each equal to
 * what mxg_ulv_create gives that matrix alone; mxg_ulv_factor_seconds of
 * each is the time of the whole batch. The batched form of
 * compute_factorize (device_factorize.cu:56-142), generalized from the
 * reference's uniform padded blocks of one matrix to variable sizes across
 * matrices. */
int mxg_ulv_create_batch(const mxg_h2_t* h, long long count, int precision, mxg_ulv_t* out);
int mxg_ulv_destroy(mxg_ulv_t f);
/* info_per_level[levels+1]: number of boxes whose rr block had |det| <= DBL_MIN. */
int mxg_ulv_info(mxg_ulv_t f, int* info_per_level);
/* x <- M^{-1} x in place (matSolvePreconditionDeviceH2, device_factorize.cu:244-254). */
int mxg_ulv_solve(mxg_ulv_t f, double* d_x, long long nrhs, long long ld);
/* seconds of the factorization (from the first to the last of its launches
 * on the context stream, HIP events; the allocation and list building of
 * mxg_ulv_create before it are not counted) */
double mxg_ulv_factor_seconds(mxg_ulv_t f);

/* ---- GMRES: replaces solveDeviceGMRES (device_gmres.cu:26-115) ---- */
enum {
  MXG_GMRES_REFERENCE = 0, /* CGS1 Arnoldi, fixed inner iterations (solver.cpp:144-201) */
  MXG_GMRES_CGS2 = 1       /* CGS2 Arnoldi, Givens updates, inner exit on the estimate */
};
typedef struct {
  double tol;          /* stop when ||b - A x|| / ||b|| < tol */
  int restart;         /* inner iterations per cycle */
  int max_cycles;      /* outer cycles */
  int mode;            /* MXG_GMRES_* */
  int inner_precision; /* MXG_FP64, or MXG_FP32 for FP32 inner operator applications */
  mxg_h2_t inner_operator; /* NULL: A; else the H^2 matrix applied in the inner (Arnoldi)
                              iterations, e.g. A built at a looser tolerance, while the
                              residual of every cycle uses A (iterative refinement) */
  int graph;           /* 1: with CGS2 and FP32 inner iterations, the inner steps of a call
                          after its first replay one HIP graph of the step's operator and
                          preconditioner launches, captured from that first step (their
                          input is a fixed staging column); results are bit for bit those
                          of plain launches. Ignored otherwise, or while profiling. */
} mxg_gmres_opts_t;
/* Solves A x = b with left preconditioner M. x holds the initial guess on entry
 * (a zero guess skips the first operator application: r = b).
 * resid[0..max_cycles]: relative residual (H^2 operator) after each cycle.
 * iters_out: cycles run; inner_out: total inner iterations. */
int mxg_gmres(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* opts, const double* d_b, double* d_x,
              double* resid, long long* iters_out, long long* inner_out);
/* Batch of nrhs (1..64) independent right-hand sides solved in lockstep: each
 * has its own Krylov basis and stopping test, the operator and
 * preconditioner applications of a step serve all of them (the H^2 and ULV
 * data are streamed once per step for up to 8 of them). B and X are n x nrhs
 * with leading dimensions ldb / ldx. resid: nrhs rows of max_cycles + 1;
 * iters_out / inner_out: nrhs entries. No counterpart in the reference, whose
 * solveGMRES takes one right-hand side. */
int mxg_gmres_batch(mxg_h2_t A, mxg_ulv_t M, const mxg_gmres_opts_t* opts, long long nrhs, const double* d_B,
                    long long ldb, double* d_X, long long ldx, double* resid, long long* iters_out,
                    long long* inner_out);

/* ---- kernel timing (HIP events on the stream each launch runs on) ----
 * While enabled, every launch of the instrumented kernel classes is bracketed
 * by a pair of HIP events; mxg_profile_read synchronizes and returns the
 * number of launches and their summed device time since the last reset. The
 * leaf near products run on a second stream of the context, concurrently
 * with the upward pass and the coupling products. */
enum {
  MXG_PROF_MATVEC_LEAF_FP64 = 0, /* near-field product of the leaf level (one kernel launch) */
  MXG_PROF_MATVEC_LEAF_FP32 = 1,
  MXG_PROF_ULV_FWD_LEAF = 2,     /* forward substitution, leaf level (x = V y and z) */
  MXG_PROF_ULV_BWD_LEAF = 3,     /* backward substitution, leaf level */
  MXG_PROF_MATVEC_UPPER = 4,     /* upward products and per-box sums of every level */
  MXG_PROF_MATVEC_FAR = 5,       /* coupling products of every level */
  MXG_PROF_MATVEC_FP64 = 6,      /* a whole FP64 matvec, fork to join */
  MXG_PROF_MATVEC_FP32 = 7,      /* a whole FP32 matvec */
  MXG_PROF_ULV_SOLVE = 8,        /* a whole substitution (every level, both sweeps) */
  MXG_PROF_SKETCH = 9,           /* the far-field sketches of a level's basis construction */
  MXG_PROF_CLASSES = 10
};
int mxg_profile_enable(mxg_ctx_t ctx, int enable);
int mxg_profile_reset(mxg_ctx_t ctx);
int mxg_profile_read(mxg_ctx_t ctx, int kernel_class, long long* launches, double* total_ms);
/* Algorithmic work of the timed launches of a class since the last reset, for
 * classes that count it (MXG_PROF_SKETCH: 2 M^2 nf flops per box of M columns
 * and nf sketched far rows), else 0. */
int mxg_profile_work(mxg_ctx_t ctx, int kernel_class, double* work);

#ifdef __cplusplus
}
#endif
#endif /* MXC_GPU_H */
