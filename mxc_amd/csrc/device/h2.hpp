// Internal layout of an H^2 matrix resident on the device.
//
// Per level l (boxes i = 0..nb-1, tree order):
//   dims[i]  = M_i   box size (leaf: bodies; upper: sum of the children's ranks)
//   ranks[i] = k_i   skeleton rank
//   offX/offZ        prefix sums of M and k: vector layouts. The upper level's
//                    X is the lower level's Z and the lower level's W is the
//                    upper level's Y, so the passes hand vectors over in place.
//   offQ             prefix sum of M^2: Q (basis, M x M) and R (M x M) per box
//   near CSR + offA  near blocks A_ij (M_i x M_j, column major)
//   far CSR + offC   coupling blocks C_ij (k_i x k_j, column major)
//   S                skeleton body indices (int32, M per box, permuted so the
//                    k_i skeleton points come first)
#pragma once

#include "internal.hpp"

namespace mxg {

struct Level {
  long long nb = 0;
  bool built = false;
  std::vector<long long> dims, ranks, nrp, ncol, frp, fcol;
  std::vector<long long> offX, offZ, offQ, offA, offC, diag;
  long long lenX = 0, lenZ = 0, maxM = 0, maxK = 0;
  DevBuf<double> A, Q, R, C;
  DevBuf<int> S;
  DevBuf<long long> d_offX, d_offZ, d_offQ, d_offA, d_offC, d_nrp, d_frp;
  DevBuf<int> d_dims, d_ranks, d_ncol, d_fcol;
  // Schur hand-off to the parent: destination offset in the parent's near
  // blocks and its leading dimension, per near pair of this level (-1: none).
  std::vector<long long> up_off;
  std::vector<int> up_ld;
  // FP32 copies of the operator data (inner matvec of mixed precision GMRES)
  DevBuf<float> A32, Q32, C32;
};

}  // namespace mxg

struct mxg_h2_s {
  mxg_ctx_t ctx = nullptr;
  mxg_kernel_t k{};
  long long n = 0, L = 0;
  int flags = 0;
  mxg::DevBuf<double> bodies;
  std::vector<mxg::Level> lv;
  std::vector<mxg::DevBuf<double>> Z, W;
  std::vector<mxg::DevBuf<float>> Z32, W32;
  mxg::DevBuf<float> x32, y32;
  bool shadow32 = false;
};

namespace mxg {
// launch helpers used by ulv.hip / gmres.hip
template <class T>
void copy2d_batched(const std::vector<const T*>& src, const std::vector<T*>& dst, const std::vector<int>& m,
                    const std::vector<int>& n, const std::vector<int>& lds, const std::vector<int>& ldd, hipStream_t s);
template <class T>
void h2_matvec_levels(mxg_h2_s* h, const T* x, T* y);
}  // namespace mxg
