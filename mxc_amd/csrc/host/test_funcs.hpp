// Point and right-hand-side generators (drop-in for the reference's
// test_funcs.hpp:31-61 and main.cpp:73-76) plus the CSV report writer
// (test_funcs.hpp:94-111).
#pragma once

#include <cstdio>

namespace mxc {

// test_funcs.hpp:31-41: std::mt19937(seed), U(0, diameter) drawn x, y, z per body.
void uniform_unit_cube_rnd(double* bodies, long long nbodies, double diameter, long long dim, unsigned int seed);

// test_funcs.hpp:43-61: golden-angle points on a sphere of radius r.
void mesh_sphere(double* bodies, long long nbodies, double r);

// Points uniform on the unit sphere (the reference has no such generator):
// the golden-angle points of mesh_sphere at radius 1 under a uniform random
// rotation, the unit quaternion drawn from three U(0, 1) values of
// std::mt19937(seed) (Shoemake 1992). Quasi-uniform points keep every pair
// at least ~sqrt(4 pi / N) apart; i.i.d. uniform points produce
// near-coincident pairs whose 1/r entries swamp a fixed diagonal and make
// the kernel matrix indefinite.
void sphere_surface_rnd(double* bodies, long long nbodies, unsigned int seed);

// Entries U(lo, hi) from std::mt19937_64(seed) (main.cpp:73-76 uses the default seed, U(0, 1)).
void uniform_vector(double* x, long long n, double lo, double hi, unsigned long long seed);

// One CSV row with the reference's 25 columns and the residual history
// (test_funcs.hpp:94-111).
void write_to_csv(const char* fname, int nproc, long long N, double theta, long long leaf_size, long long rank,
                  double epi, const char* mode, double h2cerr, double h2ctime, double h2ctime_comm, double h2mvtime,
                  double h2mvtime_comm, double dense_mvtime, double mctime, double mctime_comm, double mcerr,
                  double factor_time, double factor_time_comm, double sub_time, double sub_time_comm, double sub_err,
                  double gmres_err, double gmres_iters, double gmres_time, double gmres_time_comm,
                  const double* iter_err);

}  // namespace mxc
