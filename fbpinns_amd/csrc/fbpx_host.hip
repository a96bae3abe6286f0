// C-ABI entry points for the pair kernels, error reporting and shape queries.
#include <stdarg.h>
#include <stdio.h>

#include "fbpx_shapes.hpp"

namespace fbpx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return FBPX_OK;
  set_error("%s: %s", what, hipGetErrorString(e));
  return FBPX_EHIP;
}

static int check_net(const fbpx_net* net) {
  if (!net) {
    set_error("null net");
    return FBPX_EINVAL;
  }
  if (net->xd < 1 || net->xd > FBPX_MAX_XD || (net->d2mask & ~net->d1mask) || net->d1mask >= (1 << net->xd) ||
      net->d1mask < 0 || net->d2mask < 0) {
    set_error("bad net xd=%d d1mask=%d d2mask=%d", net->xd, net->d1mask, net->d2mask);
    return FBPX_EINVAL;
  }
  if (!(net->h == 16 || net->h == 32 || net->h == 64) || !(net->nh == 1 || net->nh == 2)) {
    set_error("network [%d, %d x %d, 1] not compiled (hidden width 16/32/64, 1 or 2 hidden layers)", net->xd,
              net->h, net->nh);
    return FBPX_EUNSUPPORTED;
  }
  if (!schema_supported(net->xd, net->d1mask, net->d2mask)) {
    set_error("jet schema xd=%d d1mask=%d d2mask=%d not compiled", net->xd, net->d1mask, net->d2mask);
    return FBPX_EUNSUPPORTED;
  }
  return FBPX_OK;
}

static int check_pair_args(const fbpx_net* net, const float* theta, const float* dom, const float* x,
                           const fbpx_worklist* wl, const void* out) {
  int rc = check_net(net);
  if (rc) return rc;
  if (!wl || wl->n_pairs < 0 || wl->n_chunks < 0 || wl->n_fbins < 0 || wl->n_bbins < 0) {
    set_error("bad work list");
    return FBPX_EINVAL;
  }
  if (wl->n_chunks > 0 && (!theta || !dom || !x || !out || !wl->chunk_model || !wl->chunk_start ||
                           !wl->pair_point)) {
    set_error("null pointer in pair kernel arguments");
    return FBPX_EINVAL;
  }
  if (wl->n_chunks > 0 && (wl->n_fbins < 1 || wl->n_bbins < 1 || !wl->fbin_chunk_start || !wl->bbin_chunk_start)) {
    set_error("work list has chunks but no bins (n_fbins=%d n_bbins=%d)", wl->n_fbins, wl->n_bbins);
    return FBPX_EINVAL;
  }
  if ((wl->fwd_wave_bins != 0 && wl->fwd_wave_bins != 1) || (wl->bwd_wave_bins != 0 && wl->bwd_wave_bins != 1)) {
    set_error("fwd_wave_bins / bwd_wave_bins must be 0 or 1 (got %d, %d)", wl->fwd_wave_bins, wl->bwd_wave_bins);
    return FBPX_EINVAL;
  }
  if ((wl->fwd_wave_bins || wl->bwd_wave_bins) && net->nh != 1) {
    set_error("one wave per bin runs one-hidden-layer networks only (nh=%d)", net->nh);
    return FBPX_EINVAL;
  }
  return FBPX_OK;
}

typedef int (*launch_fn)(const fbpx_net&, const NetArgs&, hipStream_t);

static int dispatch(const launch_fn* fns, const fbpx_net* net, const NetArgs& a, hipStream_t s) {
  const int rc = fns[net->xd - 1](*net, a, s);
  if (rc == FBPX_EUNSUPPORTED)
    set_error("no kernel for xd=%d h=%d nh=%d d1mask=%d d2mask=%d", net->xd, net->h, net->nh, net->d1mask,
              net->d2mask);
  return rc;
}

}  // namespace fbpx

using namespace fbpx;

extern "C" {

int fbpx_forward(void* stream, const fbpx_net* net, const float* theta, const float* dom, const float* x,
                 const fbpx_worklist* wl, float* pairbuf) {
  int rc = check_pair_args(net, theta, dom, x, wl, pairbuf);
  if (rc) return rc;
  static const launch_fn fns[3] = {launch_fwd_xd1, launch_fwd_xd2, launch_fwd_xd3};
  NetArgs a{theta, dom, x, *wl, nullptr, nullptr, pairbuf};
  return dispatch(fns, net, a, (hipStream_t)stream);
}

int fbpx_backward(void* stream, const fbpx_net* net, const float* theta, const float* dom, const float* x,
                  const fbpx_worklist* wl, const uint8_t* chunk_train, const float* gseg, const float* g_scale,
                  float* partial) {
  int rc = check_pair_args(net, theta, dom, x, wl, partial);
  if (rc) return rc;
  if (wl->n_chunks > 0 && (!gseg || !wl->pair_seg)) {
    set_error("backward: null gseg / pair_seg");
    return FBPX_EINVAL;
  }
  static const launch_fn fns[3] = {launch_bwd_xd1, launch_bwd_xd2, launch_bwd_xd3};
  NetArgs a{theta, dom, x, *wl, chunk_train, gseg, partial, g_scale};
  return dispatch(fns, net, a, (hipStream_t)stream);
}

int fbpx_backward_adam(void* stream, const fbpx_net* net, float* theta, const float* dom, const float* x,
                       const fbpx_worklist* wl, const uint8_t* chunk_train, const float* gseg, const float* g_scale,
                       float* mu, float* nu, int32_t* count, int32_t* count_done, double lr, double b1, double b2,
                       double eps) {
  int rc = check_pair_args(net, theta, dom, x, wl, theta);
  if (rc) return rc;
  if (net->nh != 1) {
    set_error("backward_adam: one-hidden-layer networks only (nh=%d)", net->nh);
    return FBPX_EINVAL;
  }
  if (!mu || !nu || !count || !count_done || (wl->n_chunks > 0 && (!gseg || !wl->pair_seg))) {
    set_error("backward_adam: null mu / nu / count / count_done / gseg / pair_seg");
    return FBPX_EINVAL;
  }
  if (wl->n_chunks == 0) return fbpx_count_increment(stream, count);
  static const launch_fn fns[3] = {launch_bwd_xd1, launch_bwd_xd2, launch_bwd_xd3};
  NetArgs a{theta, dom, x, *wl, chunk_train, gseg, nullptr, g_scale,
            AdamFold{theta, mu, nu, count, count_done, adam_const(lr, b1, b2, eps)}};
  return dispatch(fns, net, a, (hipStream_t)stream);
}

int fbpx_occupancy(const fbpx_net* net, int32_t* fwd_per_cu, int32_t* bwd_per_cu) {
  int rc = check_net(net);
  if (rc) return rc;
  if (!fwd_per_cu || !bwd_per_cu) {
    set_error("occupancy: null output");
    return FBPX_EINVAL;
  }
  typedef int (*occ_fn)(const fbpx_net&, int*, int*);
  static const occ_fn fns[3] = {occupancy_xd1, occupancy_xd2, occupancy_xd3};
  int f = 0, b = 0;
  rc = fns[net->xd - 1](*net, &f, &b);
  if (rc == FBPX_EUNSUPPORTED)
    set_error("no kernel for xd=%d h=%d nh=%d d1mask=%d d2mask=%d", net->xd, net->h, net->nh, net->d1mask,
              net->d2mask);
  *fwd_per_cu = f;
  *bwd_per_cu = b;
  return rc;
}

int32_t fbpx_param_count(const fbpx_net* net) {
  if (!net || net->xd < 1 || net->h < 1 || net->nh < 1 || net->nh > 2) return 0;
  return param_count(net->xd, net->h, net->nh);
}

int32_t fbpx_supported(const fbpx_net* net) { return check_net(net) == FBPX_OK ? 1 : 0; }

const char* fbpx_last_error(void) { return g_err; }

const char* fbpx_version(void) { return "fbpx 0.1 gfx950"; }

}  // extern "C"
