#include "test_funcs.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <random>

namespace mxc {

void uniform_unit_cube_rnd(double* bodies, long long nbodies, double diameter, long long dim, unsigned int seed) {
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(0., diameter);
  for (long long b = 0; b < nbodies; b++)
    for (int d = 0; d < 3; d++) bodies[3 * b + d] = d < dim ? dist(gen) : 0.;
}

void mesh_sphere(double* bodies, long long nbodies, double r) {
  const double golden = M_PI * (3. - std::sqrt(5.));
  for (long long i = 0; i < nbodies; i++) {
    const double y = r - ((double)i / (double)(nbodies - 1)) * (r + r);
    const double rad = std::sqrt(r * r - y * y);
    const double t = (double)i * golden;
    bodies[3 * i] = rad * std::cos(t);
    bodies[3 * i + 1] = y;
    bodies[3 * i + 2] = rad * std::sin(t);
  }
}

void sphere_surface_rnd(double* bodies, long long nbodies, unsigned int seed) {
  // golden-angle points on the unit sphere, then a uniform random rotation:
  // unit quaternion from three U(0, 1) draws (Shoemake 1992)
  mesh_sphere(bodies, nbodies, 1.);
  std::mt19937 gen(seed);
  std::uniform_real_distribution<double> dist(0., 1.);
  const double two_pi = 6.283185307179586476925286766559;
  const double u1 = dist(gen), u2 = dist(gen), u3 = dist(gen);
  const double a = std::sqrt(1. - u1), b = std::sqrt(u1);
  const double qx = a * std::sin(two_pi * u2), qy = a * std::cos(two_pi * u2);
  const double qz = b * std::sin(two_pi * u3), qw = b * std::cos(two_pi * u3);
  const double R[3][3] = {{1. - 2. * (qy * qy + qz * qz), 2. * (qx * qy - qz * qw), 2. * (qx * qz + qy * qw)},
                          {2. * (qx * qy + qz * qw), 1. - 2. * (qx * qx + qz * qz), 2. * (qy * qz - qx * qw)},
                          {2. * (qx * qz - qy * qw), 2. * (qy * qz + qx * qw), 1. - 2. * (qx * qx + qy * qy)}};
  for (long long i = 0; i < nbodies; i++) {
    const double p0 = bodies[3 * i], p1 = bodies[3 * i + 1], p2 = bodies[3 * i + 2];
    for (int d = 0; d < 3; d++) bodies[3 * i + d] = (R[d][0] * p0 + R[d][1] * p1) + R[d][2] * p2;
  }
}

void uniform_vector(double* x, long long n, double lo, double hi, unsigned long long seed) {
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> dist(lo, hi);
  for (long long i = 0; i < n; i++) x[i] = dist(gen);
}

void write_to_csv(const char* fname, int nproc, long long N, double theta, long long leaf_size, long long rank,
                  double epi, const char* mode, double h2cerr, double h2ctime, double h2ctime_comm, double h2mvtime,
                  double h2mvtime_comm, double dense_mvtime, double mctime, double mctime_comm, double mcerr,
                  double factor_time, double factor_time_comm, double sub_time, double sub_time_comm, double sub_err,
                  double gmres_err, double gmres_iters, double gmres_time, double gmres_time_comm,
                  const double* iter_err) {
  std::ofstream f(fname, std::ios_base::app);
  if (f.bad()) return;
  f << nproc << ',' << N << ',' << theta << ',' << leaf_size << ',' << rank << ',' << epi << ',' << mode << ',';
  f << h2cerr << ',' << h2ctime << ',' << h2ctime_comm << ',' << h2mvtime << ',' << h2mvtime_comm << ',' << dense_mvtime << ',';
  f << mctime << ',' << mctime_comm << ',' << mcerr << ',' << factor_time << ',' << factor_time_comm << ',' << sub_time << ','
    << sub_time_comm << ',' << sub_err << ',';
  f << gmres_err << ',' << gmres_iters << ',' << gmres_time << ',' << gmres_time_comm;
  for (long long i = 0; i <= (long long)gmres_iters; i++) f << ',' << iter_err[i];
  f << '\n';
}

}  // namespace mxc
