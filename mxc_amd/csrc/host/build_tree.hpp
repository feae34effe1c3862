// Cluster tree and block lists (drop-in for the reference's build_tree.hpp:7-26).
#pragma once

#include <array>
#include <vector>

namespace mxc {

// One cell of the heap-ordered binary cluster tree (build_tree.hpp:7-15):
// children of cell i are 2i+1 and 2i+2; Body = [begin, end) into the
// tree-ordered body array; R = half extents (+1e-8), C = centre.
class Cell {
 public:
  std::array<long long, 2> Child{-1, -1};
  std::array<long long, 2> Body{-1, -1};
  std::array<double, 3> R{0., 0., 0.};
  std::array<double, 3> C{0., 0., 0.};
};

// Near ('N') or far ('F') block list over all cells (build_tree.hpp:17-24).
// A pair (i, j) is admissible iff |C_i - C_j|^2 > theta (|R_i|^2 + |R_j|^2)
// (build_tree.cpp:25-29). Near holds the inadmissible pairs of every level,
// far the admissible pairs whose parents are inadmissible.
class CSR {
 public:
  std::vector<long long> RowIndex;
  std::vector<long long> ColIndex;

  CSR() = default;
  CSR(char NoF, const std::vector<Cell>& ci, const std::vector<Cell>& cj, double theta);
  long long lookupIJ(long long i, long long j) const;
};

// Median split on the longest axis, levels deep (build_tree.cpp:67-99).
// Permutes `bodies` (3 doubles per body) into tree order in place.
void buildBinaryTree(Cell* cells, double* bodies, long long nbodies, long long levels);

// Number of tree levels for N bodies and a leaf size (main.cpp:53-55).
long long treeLevels(long long nbodies, long long leaf_size);

}  // namespace mxc
