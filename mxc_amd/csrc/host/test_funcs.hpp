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

// Points uniform on the unit sphere (no reference generator exists):
// z = 2u - 1, phi = 2 pi v with u then v from std::mt19937(seed) through U(0, 1).
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
