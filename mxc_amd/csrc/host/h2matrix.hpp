// One level of an H^2 matrix, host side (drop-in for the reference's
// h2matrix.hpp:11-56 at a single rank). The numeric data lives on the GPU in
// an mxg_h2_t; this class owns the level's layout and drives the device
// construction (H2Matrix::construct, h2matrix.cpp:72-249).
#pragma once

#include <vector>

#include "build_tree.hpp"
#include "far_bodies.hpp"
#include "kernel.hpp"
#include "mxc_gpu.h"

namespace mxc {

// Basis policy of one H^2 matrix (compute_basis, h2matrix.cpp:13-63).
struct BasisOptions {
  double tol = 0.;            // > 0: Eigen CPQR threshold rank
  long long rank = 0;         // > 0: rank cap at the leaf level
  long long leveled_rank = 0; // cap grows by this much per level above the leaf (solver.cpp:29)
  bool orth = false;          // orthogonal (true) or interpolative (false) basis
  bool hss_far = false;       // basis far lists from the sibling (theta = 0) lists (solver.cpp:32)
  long long sampling = 0;     // far bodies per box (<= 0: all)
  bool householder_rows = false;  // far rows reduced by the Householder R factor, not the sketch
  bool explicit_far = false;      // far lists as one index per body instead of runs
};

class H2Matrix {
 public:
  long long level = 0;
  long long nboxes = 0;
  long long lenX = 0;
  long long LowerZ = 0;  // always 0 at one rank (comm-mpi.cpp:172)
  std::vector<long long> Dims, DimsLr;
  std::vector<long long> ARows, ACols, CRows, CCols;  // level-local near/far CSR

  // Builds the layout of `level` from the global lists and constructs the
  // level on the device. `lower` is the already constructed child level
  // (nullptr at the leaf level).
  void construct(mxg_h2_t dev, const BasisOptions& opt, long long maxrank, const std::vector<Cell>& cells,
                 const CSR& Near, const CSR& Far, const CSR& BasisFar, long long lvl, const H2Matrix* lower);
};

}  // namespace mxc
