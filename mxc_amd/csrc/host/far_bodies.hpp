// Far-body lists for the shared-basis construction, with the sampling cap.
//
// Replaces the only part of the reference's Hmatrix that reaches the results:
// the per-box far-body list m (h-matrix.cpp:121-127,160-165,169-175), which is
// the parent's list followed by the bodies of the box's own far-list cells in
// CSR order, built from the basis far lists (Far, or the sibling lists of
// HSS_Far for the preconditioner, solver.cpp:32). The reference's randomized
// SVDs (h-matrix.cpp:103-120) never reach a result and are not built.
//
// Sampling (the `sampling` argument of lorasp; the reference HEAD has no cap).
// When the list is longer than s = sampling, its bodies are grouped into
// distance shells from the matrix's own near/far structure (theta): shell 0 =
// the box's near cells other than itself (only when the list holds every
// body outside the box, as with the sibling lists), shell 1 = the box's own
// far cells, shell 2 = its parent's far cells, and so on up to level 1. For a
// box at level l these shells partition exactly the same set of bodies as the
// reference list. Quotas depend on what the list is for:
//  * far lists of the accurate operator (no shell 0): the R closest shells
//    are kept whole, R = max(3, ceil(L / 2)) for a tree of L levels, and
//    every farther shell of n bodies keeps min(n, s) of them. A basis built
//    to tol ~1e-12 must see the nearer far cells in full: with R = 3 the
//    operator error of the unit-sphere Yukawa problem grows from 1.4e-13
//    (L = 8) to 4.5e-10 (L = 10) and 1.1e-9 (L = 12); with R = ceil(L / 2)
//    it is 1.7e-13 (L = 10) and 2.3e-13 (L = 12);
//  * lists of every outside body (the rank-capped preconditioner): s bodies
//    in total, quotas going from the closest shell outward, each shell taking
//    at most half of the remaining budget (the last takes the rest); quota a
//    shell cannot use is then handed back from the closest shell outward.
// A shell of n bodies (in cell CSR order, tree order within a cell) with
// quota q keeps positions floor(t n / q), t = 0..q-1. The sample is emitted
// farthest shell first. sampling <= 0, or a list no longer than s, keeps
// every body in reference order.
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
  // basis_far: the lists the reference gathers from; near/far: the matrix's
  // theta lists that define the shells; with_near: basis_far covers every
  // body outside the box (shell 0 present).
  FarBodies(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
            long long level, long long sampling);

  long long fbodies_size_at_i(long long i) const { return offsets[i + 1] - offsets[i]; }
  const long long* fbodies_at_i(long long i) const { return indices.data() + offsets[i]; }
};

// The same lists as runs of consecutive body indices, in list order: run k
// covers bodies runs[2k] .. runs[2k] + runs[2k + 1] - 1. A shell kept whole is
// one run per cell, so a level's lists take a few runs per far cell instead
// of one index per body; the device expands them
// (mxg_h2_construct_level_runs).
class FarRuns {
 public:
  std::vector<long long> offsets;  // nboxes + 1, into the runs
  std::vector<long long> runs;     // 2 per run: first body, count

  FarRuns() = default;
  FarRuns(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
          long long level, long long sampling);
};

// Sampled far-body list of one cell as runs (first body, count) appended to
// out, and the list itself (its runs expanded).
void sampleFarRuns(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
                   long long cell, long long sampling, std::vector<long long>& out);
std::vector<long long> sampleFarBodies(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near,
                                       const CSR& far, bool with_near, long long cell, long long sampling);

}  // namespace mxc
