#include "far_bodies.hpp"

#include <algorithm>

namespace mxc {

namespace {

// Closest distance shells kept whole in a far list of a tree with `levels`
// levels: half the depth, at least three.
long long resolvedShells(long long levels) { return std::max(3ll, (levels + 1) / 2); }

long long treeLevels(const std::vector<Cell>& cells) {
  long long l = 0;
  while ((2ll << l) - 1 < (long long)cells.size()) l++;
  return l;
}

struct Shell {
  std::vector<std::pair<long long, long long>> ranges;  // body ranges of the shell's cells
  long long count = 0;
  long long quota = 0;
  void add(const Cell& c) {
    ranges.emplace_back(c.Body[0], c.Body[1]);
    count += c.Body[1] - c.Body[0];
  }
};

// ancestors of `cell` from level 1 down to the cell itself
std::vector<long long> ancestry(long long cell) {
  std::vector<long long> path;
  for (long long c = cell; c > 0; c = (c - 1) >> 1) path.push_back(c);
  std::reverse(path.begin(), path.end());
  return path;
}

}  // namespace

namespace {
// appends bodies [first, first + count) to a run list, extending its last run
// when the two are adjacent
void add_run(std::vector<long long>& out, long long first, long long count) {
  if (count <= 0) return;
  const size_t n = out.size();
  if (n >= 2 && out[n - 2] + out[n - 1] == first) out[n - 1] += count;
  else {
    out.push_back(first);
    out.push_back(count);
  }
}
}  // namespace

void sampleFarRuns(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
                   long long cell, long long sampling, std::vector<long long>& out) {
  const std::vector<long long> path = ancestry(cell);
  long long total = 0;
  for (long long a : path)
    for (long long k = basis_far.RowIndex[a]; k < basis_far.RowIndex[a + 1]; k++) {
      const Cell& f = cells[basis_far.ColIndex[k]];
      total += f.Body[1] - f.Body[0];
    }
  if (sampling <= 0 || total <= sampling) {
    for (long long a : path)
      for (long long k = basis_far.RowIndex[a]; k < basis_far.RowIndex[a + 1]; k++) {
        const Cell& f = cells[basis_far.ColIndex[k]];
        add_run(out, f.Body[0], f.Body[1] - f.Body[0]);
      }
    return;
  }

  // shells, closest first
  std::vector<Shell> sh;
  if (with_near) {
    sh.emplace_back();
    for (long long k = near.RowIndex[cell]; k < near.RowIndex[cell + 1]; k++)
      if (near.ColIndex[k] != cell) sh.back().add(cells[near.ColIndex[k]]);
  }
  for (size_t a = path.size(); a-- > 0;) {
    sh.emplace_back();
    for (long long k = far.RowIndex[path[a]]; k < far.RowIndex[path[a] + 1]; k++) sh.back().add(cells[far.ColIndex[k]]);
  }

  if (with_near) {
    long long left = sampling;
    for (size_t k = 0; k < sh.size(); k++) {
      long long cap = (k + 1 == sh.size()) ? left : (left + 1) / 2;
      sh[k].quota = std::min(sh[k].count, cap);
      left -= sh[k].quota;
    }
    for (size_t k = 0; k < sh.size() && left > 0; k++) {
      long long more = std::min(sh[k].count - sh[k].quota, left);
      sh[k].quota += more;
      left -= more;
    }
  } else {
    const size_t resolved = (size_t)resolvedShells(treeLevels(cells));
    for (size_t k = 0; k < sh.size(); k++) sh[k].quota = (k < resolved) ? sh[k].count : std::min(sh[k].count, sampling);
  }

  for (size_t k = sh.size(); k-- > 0;) {
    const Shell& s = sh[k];
    if (s.quota == s.count) {  // positions floor(t n / n) = t: every range whole
      for (const auto& r : s.ranges) add_run(out, r.first, r.second - r.first);
      continue;
    }
    size_t r = 0;
    long long before = 0;  // bodies in ranges[0..r)
    for (long long t = 0; t < s.quota; t++) {
      long long pos = (long long)(((__int128)t * s.count) / s.quota);
      while (pos - before >= s.ranges[r].second - s.ranges[r].first) {
        before += s.ranges[r].second - s.ranges[r].first;
        r++;
      }
      add_run(out, s.ranges[r].first + (pos - before), 1);
    }
  }
}

std::vector<long long> sampleFarBodies(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near,
                                       const CSR& far, bool with_near, long long cell, long long sampling) {
  std::vector<long long> runs;
  sampleFarRuns(cells, basis_far, near, far, with_near, cell, sampling, runs);
  long long total = 0;
  for (size_t k = 1; k < runs.size(); k += 2) total += runs[k];
  std::vector<long long> out;
  out.reserve(total);
  for (size_t k = 0; k < runs.size(); k += 2)
    for (long long b = runs[k]; b < runs[k] + runs[k + 1]; b++) out.push_back(b);
  return out;
}

FarBodies::FarBodies(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
                     long long level, long long sampling) {
  const long long nb = 1ll << level, first = nb - 1;
  std::vector<std::vector<long long>> lists(nb);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < nb; i++) lists[i] = sampleFarBodies(cells, basis_far, near, far, with_near, first + i, sampling);
  offsets.assign(nb + 1, 0);
  for (long long i = 0; i < nb; i++) offsets[i + 1] = offsets[i] + (long long)lists[i].size();
  indices.resize(offsets[nb]);
  for (long long i = 0; i < nb; i++) std::copy(lists[i].begin(), lists[i].end(), indices.begin() + offsets[i]);
}

FarRuns::FarRuns(const std::vector<Cell>& cells, const CSR& basis_far, const CSR& near, const CSR& far, bool with_near,
                 long long level, long long sampling) {
  const long long nb = 1ll << level, first = nb - 1;
  std::vector<std::vector<long long>> lists(nb);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < nb; i++) sampleFarRuns(cells, basis_far, near, far, with_near, first + i, sampling, lists[i]);
  offsets.assign(nb + 1, 0);
  std::vector<size_t> at(nb + 1, 0);
  for (long long i = 0; i < nb; i++) {
    offsets[i + 1] = offsets[i] + (long long)lists[i].size() / 2;
    at[i + 1] = at[i] + lists[i].size();
  }
  runs.resize(at[nb]);
#pragma omp parallel for schedule(dynamic, 64)
  for (long long i = 0; i < nb; i++) std::copy(lists[i].begin(), lists[i].end(), runs.begin() + at[i]);
}

}  // namespace mxc
