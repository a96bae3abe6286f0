// fbpx_residual_loss_noalloc: the fused residual loss with its work space in
// a file-scope device array, for callers that cannot allocate device memory
// (code under hipGraph capture, tight serving loops).
//
// The residual launch writes one partial sum of r^2 per workgroup and has at
// most FBPX_RESIDUAL_BLOCKS workgroups (its grid is capped there and strides
// over the points), so the array below holds every partial for any n.  Each
// call overwrites the partials before a second one-workgroup launch sums them
// in a fixed order: nothing in the array has to start at zero.  Every call on
// one GPU shares that GPU's copy of the array, so calls that overlap on one
// GPU (other streams, other threads) race; calls on one stream do not.
#include <mutex>

#include "fbpx_common.hpp"

namespace {

__device__ __align__(256) float scratch[FBPX_RESIDUAL_BLOCKS];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch` (the array's name is a host
// stand-in in host code), looked up once per device with that device current.
int scratch_ptr(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fbpx::hip_check(e, "residual_loss_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    fbpx::set_error("residual_loss_noalloc: device %d outside 0..%d", dev, kMaxDevices - 1);
    return FBPX_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fbpx::hip_check(e, "residual_loss_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) {
      fbpx::set_error("residual_loss_noalloc: no device address for the work space on device %d", dev);
      return FBPX_EHIP;
    }
    g_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_ptr[dev];
  return FBPX_OK;
}

}  // namespace

extern "C" {

int fbpx_residual_loss_noalloc(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                               const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                               const float* coef, float weight, float* r, float* loss, int32_t accumulate) {
  if (n < 0 || n_terms < 1 || n_terms > FBPX_MAX_TERMS) {
    fbpx::set_error("residual_loss_noalloc: n=%lld n_terms=%d (work space: %d partial sums, at most %d terms)",
                    (long long)n, n_terms, FBPX_RESIDUAL_BLOCKS, FBPX_MAX_TERMS);
    return FBPX_EINVAL;
  }
  float* ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc) return rc;
  // done = NULL: the partials are summed by a second launch, so no counter
  // has to be zero on entry
  return fbpx_residual_loss(stream, n, n_cols, col, col_idx, n_terms, fa, fb, coef, weight, r, ws, loss, accumulate,
                            nullptr);
}

}  // extern "C"
