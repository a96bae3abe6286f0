// Compiled-in network / jet shapes and the per-dimension launch hooks.
#pragma once

#include "fbpx_common.hpp"

// X(XD, H, NH, D1M, D2M) for every compiled shape of one input dimension.
#define FBPX_JETS_XD1(X, H, NH) X(1, H, NH, 0, 0) X(1, H, NH, 1, 0) X(1, H, NH, 1, 1)
#define FBPX_JETS_XD2(X, H, NH) \
  X(2, H, NH, 0, 0) X(2, H, NH, 3, 0) X(2, H, NH, 3, 1) X(2, H, NH, 3, 2) X(2, H, NH, 3, 3)
#define FBPX_JETS_XD3(X, H, NH) X(3, H, NH, 0, 0) X(3, H, NH, 7, 0) X(3, H, NH, 7, 3) X(3, H, NH, 7, 7)

#define FBPX_NETS(X, JETS) \
  JETS(X, 16, 1) JETS(X, 32, 1) JETS(X, 64, 1) JETS(X, 16, 2) JETS(X, 32, 2) JETS(X, 64, 2)

// Schemas (XD, D1M, D2M) used by the point-level kernels.
#define FBPX_SCHEMAS(X)                                       \
  X(1, 0, 0) X(1, 1, 0) X(1, 1, 1)                            \
  X(2, 0, 0) X(2, 3, 0) X(2, 3, 1) X(2, 3, 2) X(2, 3, 3)      \
  X(3, 0, 0) X(3, 7, 0) X(3, 7, 3) X(3, 7, 7)

namespace fbpx {

struct NetArgs {
  const float* theta;
  const float* dom;
  const float* x;
  fbpx_worklist wl;
  const uint8_t* chunk_train;
  const float* gseg;
  float* out;  // pairbuf (forward) or partial (backward)
  const float* g_scale = nullptr;  // backward: NULL or d L / d loss for unit-seed gseg
  AdamFold adam{};                 // backward: the optimizer step folded in (adam.mu != NULL)
};

// Return FBPX_EUNSUPPORTED if the shape is not compiled, else launch status.
int launch_fwd_xd1(const fbpx_net& n, const NetArgs& a, hipStream_t s);
int launch_fwd_xd2(const fbpx_net& n, const NetArgs& a, hipStream_t s);
int launch_fwd_xd3(const fbpx_net& n, const NetArgs& a, hipStream_t s);
int launch_bwd_xd1(const fbpx_net& n, const NetArgs& a, hipStream_t s);
int launch_bwd_xd2(const fbpx_net& n, const NetArgs& a, hipStream_t s);
int launch_bwd_xd3(const fbpx_net& n, const NetArgs& a, hipStream_t s);
// Resident 256-thread workgroups per CU of the forward and backward kernels.
int occupancy_xd1(const fbpx_net& n, int* fwd, int* bwd);
int occupancy_xd2(const fbpx_net& n, int* fwd, int* bwd);
int occupancy_xd3(const fbpx_net& n, int* fwd, int* bwd);
bool schema_supported(int xd, int d1, int d2);

}  // namespace fbpx
