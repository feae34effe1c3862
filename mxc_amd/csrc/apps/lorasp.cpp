// lorasp: solve a dense kernel system with an H^2 operator, a ULV-factorized
// H^2 preconditioner and GMRES on one MI355X, checked by a dense FP64 matvec.
// Flow and report follow the reference driver (main.cpp:36-217); the argument
// order is the README's: lorasp N theta leaf eps rank sampling [options].
//
// options:
//   --kernel laplace|yukawa|gaussian   (default laplace)   --diag D (K(0), default 1e3)
//   --alpha a (Yukawa decay / Gaussian width, default 1)
//   --geometry cube|sphere|mesh  (default cube)   --seed s (points, default 0)
//   --rhs reference|uniform  (b = A x_true with x_true ~ U(0,1), or b ~ U(-1,1); default reference)
//   --rhs-seed s (default 5489, the seed of a default-constructed std::mt19937_64 as in
//     main.cpp:73-76, so x_true is the reference's)  --leveled-rank r (default 0)  --mode h2|hss (default h2)
//   --prec-eps e (preconditioner truncation tolerance, default 0 = fixed rank)
//   --factor fp64|fp32  --gmres reference|cgs2  --inner fp64|fp32
//   --restart m (default 10)  --outer k (default 50)  --tol t (default eps)
//   --csv file   --device d
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "build_tree.hpp"
#include "kernel.hpp"
#include "solver.hpp"
#include "test_funcs.hpp"

using namespace mxc;

namespace {

struct Args {
  long long N = 2048;
  double theta = 1.;
  long long leaf = 256;
  double eps = 1e-10;
  long long rank = 50;
  long long sampling = 2000;
  std::string kernel = "laplace", geometry = "cube", rhs = "reference", mode = "h2", factor = "fp64", gmres = "reference",
              inner = "fp64";
  double diag = 1e3, alpha = 1., prec_eps = 0., tol = -1.;
  unsigned seed = 0;
  unsigned long long rhs_seed = 5489;
  long long leveled_rank = 0, restart = 10, outer = 50;
  const char* csv = nullptr;
  int device = 0;
};

void die(const std::string& m) {
  std::fprintf(stderr, "lorasp: %s\n", m.c_str());
  std::exit(2);
}

Args parse(int argc, char** argv) {
  Args a;
  int pos = 0;
  for (int i = 1; i < argc; i++) {
    std::string s = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + s);
      return argv[++i];
    };
    if (s.rfind("--", 0) == 0) {
      if (s == "--kernel") a.kernel = val();
      else if (s == "--diag") a.diag = std::atof(val().c_str());
      else if (s == "--alpha") a.alpha = std::atof(val().c_str());
      else if (s == "--geometry") a.geometry = val();
      else if (s == "--seed") a.seed = (unsigned)std::atoll(val().c_str());
      else if (s == "--rhs") a.rhs = val();
      else if (s == "--rhs-seed") a.rhs_seed = std::strtoull(val().c_str(), nullptr, 10);
      else if (s == "--leveled-rank") a.leveled_rank = std::atoll(val().c_str());
      else if (s == "--mode") a.mode = val();
      else if (s == "--prec-eps") a.prec_eps = std::atof(val().c_str());
      else if (s == "--factor") a.factor = val();
      else if (s == "--gmres") a.gmres = val();
      else if (s == "--inner") a.inner = val();
      else if (s == "--restart") a.restart = std::atoll(val().c_str());
      else if (s == "--outer") a.outer = std::atoll(val().c_str());
      else if (s == "--tol") a.tol = std::atof(val().c_str());
      else if (s == "--csv") a.csv = argv[++i];
      else if (s == "--device") a.device = std::atoi(val().c_str());
      else die("unknown option " + s);
      continue;
    }
    switch (pos++) {
      case 0: a.N = std::atoll(s.c_str()); break;
      case 1: a.theta = std::atof(s.c_str()); break;
      case 2: a.leaf = std::atoll(s.c_str()); break;
      case 3: a.eps = std::atof(s.c_str()); break;
      case 4: a.rank = std::atoll(s.c_str()); break;
      case 5: a.sampling = std::atoll(s.c_str()); break;
      default: die("too many positional arguments");
    }
  }
  if (a.N <= 0 || a.leaf <= 0) die("N and leaf must be positive");
  if (a.tol < 0.) a.tol = a.eps;
  return a;
}

void check(int rc) {
  if (rc != MXG_OK) die(std::string("device layer: ") + mxg_last_error());
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Dev {
  mxg_ctx_t ctx;
  double* p = nullptr;
  long long n;
  Dev(mxg_ctx_t c, long long len) : ctx(c), n(len) {
    void* q = nullptr;
    check(mxg_malloc(c, n * sizeof(double), &q));
    p = static_cast<double*>(q);
  }
  ~Dev() { mxg_free(ctx, p); }
  void put(const std::vector<double>& v) { check(mxg_memcpy_h2d(ctx, p, v.data(), n * sizeof(double))); }
  std::vector<double> get() const {
    std::vector<double> v(n);
    check(mxg_memcpy_d2h(ctx, v.data(), p, n * sizeof(double)));
    return v;
  }
};

}  // namespace

int main(int argc, char** argv) {
  Args a = parse(argc, argv);
  mxg_ctx_t ctx = nullptr;
  check(mxg_init(a.device, &ctx));

  const long long leaf = a.N < a.leaf ? a.N : a.leaf;
  const long long levels = treeLevels(a.N, leaf);
  std::vector<double> body(3 * a.N);
  if (a.geometry == "cube") uniform_unit_cube_rnd(body.data(), a.N, 1., 3, a.seed);
  else if (a.geometry == "sphere") sphere_surface_rnd(body.data(), a.N, a.seed);
  else if (a.geometry == "mesh") mesh_sphere(body.data(), a.N, std::sqrt(a.N / (4 * M_PI)));
  else die("unknown geometry");
  std::vector<Cell> cell((2ll << levels) - 1);
  buildBinaryTree(cell.data(), body.data(), a.N, levels);

  std::unique_ptr<MatrixAccessor> eval;
  mxg_kernel_t kd{};
  if (a.kernel == "laplace") eval = std::make_unique<Laplace3D>(1. / a.diag);
  else if (a.kernel == "yukawa") eval = std::make_unique<Yukawa3D>(1. / a.diag, a.alpha);
  else if (a.kernel == "gaussian") eval = std::make_unique<Gaussian>(a.alpha);
  else die("unknown kernel");
  kd = eval->desc();

  Dev dbody(ctx, 3 * a.N);
  dbody.put(body);
  // x_true ~ U(0, 1) (main.cpp:73-76, seeded), b_ref = A_dense x_true
  std::vector<double> xtrue(a.N);
  uniform_vector(xtrue.data(), a.N, 0., 1., a.rhs_seed);
  Dev dxt(ctx, a.N), dbref(ctx, a.N), dy(ctx, a.N);
  dxt.put(xtrue);
  double t = now();
  check(mxg_dense_matvec(ctx, kd, dbody.p, a.N, dxt.p, dbref.p));
  const double dense_time = now() - t;
  std::vector<double> bref = dbref.get();

  t = now();
  H2MatrixSolver matA(ctx, *eval, a.eps, a.rank, a.leveled_rank, cell, a.theta, body.data(), levels, a.sampling, false);
  check(mxg_synchronize(ctx));
  const double construct_time = now() - t;
  t = now();
  matA.matVecMul(dxt.p, dy.p);
  check(mxg_synchronize(ctx));
  const double matvec_time = now() - t;
  const double cerr = H2MatrixSolver::solveRelErr(a.N, dy.get().data(), bref.data());
  std::printf("Construct Err: %g\n", cerr);
  std::printf("H^2-Matrix Construct Time: %g, %g\n", construct_time, 0.);
  std::printf("H^2-Matvec Time: %g, %g\n", matvec_time, 0.);
  std::printf("Dense Matvec Time: %g\n", dense_time);

  t = now();
  H2MatrixSolver matM;
  const double mtheta = (a.mode == "hss") ? 0. : a.theta;
  if (a.mode != "h2" && a.mode != "hss") die("mode must be h2 or hss");
  if (a.prec_eps > 0.) {
    BasisOptions b;
    b.tol = a.prec_eps;
    b.rank = a.rank;
    b.leveled_rank = a.leveled_rank;
    b.orth = true;
    b.hss_far = true;
    b.sampling = a.sampling;
    matM = H2MatrixSolver(ctx, *eval, b, cell, mtheta, body.data(), levels, true);
  } else {
    matM = H2MatrixSolver(ctx, *eval, 0., a.rank, a.leveled_rank, cell, mtheta, body.data(), levels, a.sampling, true);
  }
  check(mxg_synchronize(ctx));
  const double mctime = now() - t;
  matM.matVecMul(dxt.p, dy.p);
  const double cerr_m = H2MatrixSolver::solveRelErr(a.N, dy.get().data(), bref.data());

  t = now();
  const int prec = (a.factor == "fp32") ? MXG_FP32 : MXG_FP64;
  matM.factorizeM(prec);
  check(mxg_synchronize(ctx));
  const double factor_time = now() - t;

  Dev dx(ctx, a.N);
  dx.put(bref);
  t = now();
  matM.solvePrecondition(dx.p);
  check(mxg_synchronize(ctx));
  const double sub_time = now() - t;
  const double serr = H2MatrixSolver::solveRelErr(a.N, dx.get().data(), xtrue.data());
  std::printf("H^2-Preconditioner Construct Time: %g, %g\n", mctime, 0.);
  std::printf("H^2-Preconditioner Construct Err: %g\n", cerr_m);
  std::printf("H^2-Matrix Factorization Time: %g, %g\n", factor_time, 0.);
  std::printf("H^2-Matrix Substitution Time: %g, %g\n", sub_time, 0.);
  std::printf("H^2-Matrix Substitution Err: %g\n", serr);

  std::vector<double> b = bref;
  if (a.rhs == "uniform") uniform_vector(b.data(), a.N, -1., 1., a.rhs_seed);
  else if (a.rhs != "reference") die("rhs must be reference or uniform");
  Dev db(ctx, a.N);
  db.put(b);
  dx.put(std::vector<double>(a.N, 0.));
  const int gmode = (a.gmres == "cgs2") ? MXG_GMRES_CGS2 : MXG_GMRES_REFERENCE;
  const int iprec = (a.inner == "fp32") ? MXG_FP32 : MXG_FP64;
  t = now();
  matA.solveGMRES(a.tol, matM, dx.p, db.p, a.restart, a.outer, gmode, iprec);
  check(mxg_synchronize(ctx));
  const double gmres_time = now() - t;
  std::printf("GMRES Residual: %g, Iters: %lld\n", matA.resid[matA.iters], matA.iters);
  std::printf("GMRES Time: %g, Comm: %g\n", gmres_time, 0.);
  for (long long i = 0; i <= matA.iters; i++) std::printf("iter %lld: %g\n", i, matA.resid[i]);

  // dense FP64 check of the solution
  check(mxg_dense_matvec(ctx, kd, dbody.p, a.N, dx.p, dy.p));
  std::vector<double> ax = dy.get();
  double rn = 0., bn = 0.;
  for (long long i = 0; i < a.N; i++) {
    rn += (b[i] - ax[i]) * (b[i] - ax[i]);
    bn += b[i] * b[i];
  }
  std::printf("Dense Check Residual: %g\n", std::sqrt(rn / bn));
  std::printf("GMRES Inner Iters: %lld\n", matA.inner_iters);

  if (a.csv)
    write_to_csv(a.csv, 1, a.N, a.theta, leaf, a.rank, a.eps, a.mode.c_str(), cerr, construct_time, 0., matvec_time, 0.,
                 dense_time, mctime, 0., cerr_m, factor_time, 0., sub_time, 0., serr, matA.resid[matA.iters],
                 (double)matA.iters, gmres_time, 0., matA.resid.data());
  matA = H2MatrixSolver();
  matM = H2MatrixSolver();
  mxg_finalize(ctx);
  return 0;
}
