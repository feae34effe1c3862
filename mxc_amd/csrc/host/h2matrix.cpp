#include "h2matrix.hpp"

#include <stdexcept>
#include <string>

namespace mxc {

namespace {

// Level-local CSR of the rows [first, first + nb) of a global list, columns
// rebased to the level (comm-mpi.cpp:105-115 at one rank).
void levelCSR(const CSR& g, long long first, long long nb, std::vector<long long>& rows, std::vector<long long>& cols) {
  rows.resize(nb + 1);
  const long long base = g.RowIndex[first];
  for (long long i = 0; i <= nb; i++) rows[i] = g.RowIndex[first + i] - base;
  cols.resize(rows[nb]);
  for (long long t = 0; t < rows[nb]; t++) cols[t] = g.ColIndex[base + t] - first;
}

void check(int rc) {
  if (rc != MXG_OK) throw std::runtime_error(std::string("device layer: ") + mxg_last_error());
}

}  // namespace

void H2Matrix::construct(mxg_h2_t dev, const BasisOptions& opt, long long maxrank, const std::vector<Cell>& cells,
                         const CSR& Near, const CSR& Far, const CSR& BasisFar, long long lvl, const H2Matrix* lower) {
  level = lvl;
  nboxes = 1ll << lvl;
  const long long first = nboxes - 1;
  levelCSR(Near, first, nboxes, ARows, ACols);
  levelCSR(Far, first, nboxes, CRows, CCols);
  Dims.assign(nboxes, 0);
  DimsLr.assign(nboxes, 0);
  std::vector<long long> leaf_begin;
  if (lower) {
    for (long long i = 0; i < nboxes; i++) Dims[i] = lower->DimsLr[2 * i] + lower->DimsLr[2 * i + 1];
  } else {
    leaf_begin.resize(nboxes);
    for (long long i = 0; i < nboxes; i++) {
      Dims[i] = cells[first + i].Body[1] - cells[first + i].Body[0];
      leaf_begin[i] = cells[first + i].Body[0];
    }
  }
  lenX = 0;
  for (long long d : Dims) lenX += d;
  LowerZ = 0;

  const long long* lb = lower ? nullptr : leaf_begin.data();
  if (opt.explicit_far) {
    FarBodies fb;
    if (lvl > 0) fb = FarBodies(cells, BasisFar, Near, Far, opt.hss_far, lvl, opt.sampling);
    check(mxg_h2_construct_level(dev, lvl, nboxes, Dims.data(), lb, ARows.data(), ACols.data(), CRows.data(),
                                 CCols.data(), lvl > 0 ? fb.offsets.data() : nullptr,
                                 lvl > 0 ? fb.indices.data() : nullptr, opt.tol, maxrank, opt.orth ? 1 : 0,
                                 DimsLr.data()));
  } else {
    // runs of consecutive bodies: a few per far cell instead of one index per
    // body (2.9e9 far rows over the levels of the N = 2^20 sphere operator)
    FarRuns fr;
    if (lvl > 0) fr = FarRuns(cells, BasisFar, Near, Far, opt.hss_far, lvl, opt.sampling);
    check(mxg_h2_construct_level_runs(dev, lvl, nboxes, Dims.data(), lb, ARows.data(), ACols.data(), CRows.data(),
                                      CCols.data(), lvl > 0 ? fr.offsets.data() : nullptr,
                                      lvl > 0 ? fr.runs.data() : nullptr, opt.tol, maxrank, opt.orth ? 1 : 0,
                                      DimsLr.data()));
  }
}

}  // namespace mxc
