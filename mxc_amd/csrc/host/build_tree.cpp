// Cluster tree and near/far block lists, single-rank host side.
#include "build_tree.hpp"

#include <algorithm>
#include <cmath>

namespace mxc {

namespace {

// Bounding box of a body range: centre and half extent + 1e-8 (build_tree.cpp:8-22).
void bounds(const double* b, long long n, Cell& c) {
  std::array<double, 3> lo{b[0], b[1], b[2]}, hi = lo;
  for (long long i = 1; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double v = b[3 * i + d];
      lo[d] = v < lo[d] ? v : lo[d];
      hi[d] = v > hi[d] ? v : hi[d];
    }
  for (int d = 0; d < 3; d++) {
    c.C[d] = (lo[d] + hi[d]) * 0.5;
    c.R[d] = (lo[d] == hi[d] && lo[d] == 0.) ? 0. : ((hi[d] - lo[d]) * 0.5 + 1.e-8);
  }
}

bool admissible(const Cell& a, const Cell& b, double theta) {
  double dc = 0., ra = 0., rb = 0.;
  for (int d = 0; d < 3; d++) {
    double t = a.C[d] - b.C[d];
    dc += t * t;
    ra += a.R[d] * a.R[d];
    rb += b.R[d] * b.R[d];
  }
  return dc > theta * (ra + rb);
}

}  // namespace

long long treeLevels(long long nbodies, long long leaf_size) {
  long long leaf = nbodies < leaf_size ? nbodies : leaf_size;
  return (long long)std::log2((double)nbodies / (double)leaf);
}

void buildBinaryTree(Cell* cells, double* bodies, long long nbodies, long long levels) {
  cells[0].Body = {0, nbodies};
  bounds(bodies, nbodies, cells[0]);
  const long long inner = (1ll << levels) - 1;
  for (long long i = 0; i < inner; i++) {
    Cell& c = cells[i];
    int axis = (int)(std::max_element(c.R.begin(), c.R.end()) - c.R.begin());
    auto* first = reinterpret_cast<std::array<double, 3>*>(bodies + 3 * c.Body[0]);
    auto* last = reinterpret_cast<std::array<double, 3>*>(bodies + 3 * c.Body[1]);
    std::sort(first, last, [axis](const std::array<double, 3>& p, const std::array<double, 3>& q) { return p[axis] < q[axis]; });
    long long mid = c.Body[0] + (c.Body[1] - c.Body[0]) / 2;
    long long l = 2 * i + 1;
    c.Child = {l, l + 2};
    cells[l].Body = {c.Body[0], mid};
    cells[l + 1].Body = {mid, c.Body[1]};
    bounds(bodies + 3 * c.Body[0], mid - c.Body[0], cells[l]);
    bounds(bodies + 3 * mid, c.Body[1] - mid, cells[l + 1]);
  }
}

// Breadth-first dual traversal: the inadmissible pairs of one level spawn the
// child pairs of the next. Produces the same sorted lists as the reference's
// depth-first getList + sort (build_tree.cpp:24-58).
CSR::CSR(char NoF, const std::vector<Cell>& ci, const std::vector<Cell>& cj, double theta) {
  const bool want_far = (NoF == 'F' || NoF == 'f');
  const long long ncells = (long long)ci.size();
  std::vector<std::vector<long long>> rows(ncells);
  std::vector<std::pair<long long, long long>> frontier{{0, 0}}, next;
  while (!frontier.empty()) {
    next.clear();
    for (const auto& p : frontier) {
      bool adm = admissible(ci[p.first], cj[p.second], theta);
      if (adm == want_far) rows[p.first].push_back(p.second);
      if (!adm && ci[p.first].Child[0] >= 0 && cj[p.second].Child[0] >= 0)
        for (long long a = ci[p.first].Child[0]; a < ci[p.first].Child[1]; a++)
          for (long long b = cj[p.second].Child[0]; b < cj[p.second].Child[1]; b++) next.emplace_back(a, b);
    }
    frontier.swap(next);
  }
  RowIndex.assign(ncells + 1, 0);
  for (long long i = 0; i < ncells; i++) {
    std::sort(rows[i].begin(), rows[i].end());
    RowIndex[i + 1] = RowIndex[i] + (long long)rows[i].size();
  }
  ColIndex.reserve(RowIndex[ncells]);
  for (auto& r : rows) ColIndex.insert(ColIndex.end(), r.begin(), r.end());
}

long long CSR::lookupIJ(long long i, long long j) const {
  if (i < 0 || (long long)RowIndex.size() <= i + 1) return -1;
  auto b = ColIndex.begin() + RowIndex[i], e = ColIndex.begin() + RowIndex[i + 1];
  auto it = std::lower_bound(b, e, j);
  return (it != e && *it == j) ? (long long)(it - ColIndex.begin()) : -1;
}

}  // namespace mxc
