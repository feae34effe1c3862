#include "far_bodies.hpp"

#include <algorithm>

namespace mxc {

namespace {

struct Segment {
  std::vector<std::pair<long long, long long>> ranges;  // body ranges of the far cells
  long long count = 0;
  long long quota = 0;
};

}  // namespace

std::vector<long long> sampleFarBodies(const std::vector<Cell>& cells, const CSR& far, long long cell, long long sampling) {
  // ancestors from level 1 down to the cell itself
  std::vector<long long> path;
  for (long long c = cell; c > 0; c = (c - 1) >> 1) path.push_back(c);
  std::reverse(path.begin(), path.end());

  std::vector<Segment> seg(path.size());
  long long total = 0;
  for (size_t a = 0; a < path.size(); a++) {
    for (long long k = far.RowIndex[path[a]]; k < far.RowIndex[path[a] + 1]; k++) {
      const Cell& f = cells[far.ColIndex[k]];
      seg[a].ranges.emplace_back(f.Body[0], f.Body[1]);
      seg[a].count += f.Body[1] - f.Body[0];
    }
    seg[a].quota = seg[a].count;
    total += seg[a].count;
  }

  if (sampling > 0 && total > sampling) {
    long long left = sampling;
    for (size_t a = seg.size(); a-- > 0;) {
      long long cap = (a == 0) ? left : (left + 1) / 2;
      seg[a].quota = std::min(seg[a].count, cap);
      left -= seg[a].quota;
    }
    for (size_t a = seg.size(); a-- > 0 && left > 0;) {
      long long more = std::min(seg[a].count - seg[a].quota, left);
      seg[a].quota += more;
      left -= more;
    }
  }

  std::vector<long long> out;
  out.reserve(std::min(total, sampling > 0 ? sampling : total));
  for (const Segment& s : seg) {
    if (s.quota == 0) continue;
    size_t r = 0;
    long long before = 0;  // bodies in ranges[0..r)
    for (long long t = 0; t < s.quota; t++) {
      long long pos = (long long)(((__int128)t * s.count) / s.quota);
      while (pos - before >= s.ranges[r].second - s.ranges[r].first) {
        before += s.ranges[r].second - s.ranges[r].first;
        r++;
      }
      out.push_back(s.ranges[r].first + (pos - before));
    }
  }
  return out;
}

FarBodies::FarBodies(const std::vector<Cell>& cells, const CSR& far, long long level, long long sampling) {
  const long long nb = 1ll << level, first = nb - 1;
  std::vector<std::vector<long long>> lists(nb);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long i = 0; i < nb; i++) lists[i] = sampleFarBodies(cells, far, first + i, sampling);
  offsets.assign(nb + 1, 0);
  for (long long i = 0; i < nb; i++) offsets[i + 1] = offsets[i] + (long long)lists[i].size();
  indices.resize(offsets[nb]);
  for (long long i = 0; i < nb; i++) std::copy(lists[i].begin(), lists[i].end(), indices.begin() + offsets[i]);
}

}  // namespace mxc
