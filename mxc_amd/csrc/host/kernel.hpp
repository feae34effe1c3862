// Kernel functors (drop-in for the reference's kernel.hpp:23-61).
//
// Each functor evaluates K(d) on the host and describes itself to the device
// layer (kernel id + parameters), where the HIP kernels evaluate the same
// formula. K(0) is the fixed diagonal value `singularity` = 1/s, as in the
// reference constructors Laplace3D(s) and Yukawa3D(s, a).
#pragma once

#include <cmath>

#include "mxc_gpu.h"

namespace mxc {

class MatrixAccessor {
 public:
  virtual ~MatrixAccessor() = default;
  virtual double operator()(double d) const = 0;
  virtual mxg_kernel_t desc() const = 0;
};

// kernel.hpp:28-38: 1/d, K(0) = 1/s.
class Laplace3D : public MatrixAccessor {
 public:
  double singularity;
  explicit Laplace3D(double s) : singularity(1. / s) {}
  double operator()(double d) const override { return d == 0. ? singularity : 1. / d; }
  mxg_kernel_t desc() const override { return mxg_kernel_t{MXG_KERNEL_LAPLACE, singularity, 0.}; }
};

// kernel.hpp:40-50: exp(-alpha d)/d, K(0) = 1/s.
class Yukawa3D : public MatrixAccessor {
 public:
  double singularity, alpha;
  Yukawa3D(double s, double a) : singularity(1. / s), alpha(a) {}
  double operator()(double d) const override { return d == 0. ? singularity : std::exp(-alpha * d) / d; }
  mxg_kernel_t desc() const override { return mxg_kernel_t{MXG_KERNEL_YUKAWA, singularity, alpha}; }
};

// kernel.hpp:52-61: exp(-alpha d^2), K(0) = 1.0001.
class Gaussian : public MatrixAccessor {
 public:
  double alpha;
  explicit Gaussian(double a) : alpha(a) {}
  double operator()(double d) const override { return d == 0. ? 1.0001 : std::exp(-alpha * d * d); }
  mxg_kernel_t desc() const override { return mxg_kernel_t{MXG_KERNEL_GAUSSIAN, 1.0001, alpha}; }
};

}  // namespace mxc
