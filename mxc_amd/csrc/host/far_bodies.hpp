// Far-body lists for the shared-basis construction, with the sampling cap.
//
// Replaces the only part of the reference's Hmatrix that reaches the results:
// the per-box far-body list m (h-matrix.cpp:121-127,160-165,169-175), which is
// the parent's list followed by the bodies of the box's own far-list cells in
// CSR order. The reference's randomized SVDs (h-matrix.cpp:103-120) never
// reach a result and are not built.
//
// Sampling (the `sampling` argument of lorasp; the reference HEAD has no cap):
// the list is a sequence of per-level segments, coarsest ancestor first. With
// s = sampling < list length, quotas go from the box's own level upward, each
// level taking at most half of the remaining budget (the coarsest takes the
// rest); quota a segment cannot use is then handed back from the finest level
// upward. A segment of n bodies with quota q keeps positions floor(t n / q),
// t = 0..q-1. sampling <= 0, or a list no longer than s, keeps every body in
// reference order.
#pragma once

#include <vector>

#include "build_tree.hpp"

namespace mxc {

class FarBodies {
 public:
  std::vector<long long> offsets;  // nboxes + 1
  std::vector<long long> indices;  // tree-order body indices

  FarBodies() = default;
  // level >= 1; boxes are the cells [2^level - 1, 2^(level+1) - 1).
  FarBodies(const std::vector<Cell>& cells, const CSR& far, long long level, long long sampling);

  long long fbodies_size_at_i(long long i) const { return offsets[i + 1] - offsets[i]; }
  const long long* fbodies_at_i(long long i) const { return indices.data() + offsets[i]; }
};

// Sampled far-body list of one cell (exposed for tests).
std::vector<long long> sampleFarBodies(const std::vector<Cell>& cells, const CSR& far, long long cell, long long sampling);

}  // namespace mxc
