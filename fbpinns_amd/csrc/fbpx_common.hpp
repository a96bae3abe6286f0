// Shared device/host helpers for libfbpx (gfx950).
//
// Jet layout (per point or per pair): J floats
//   [value, d/dx_i for i in D1 (ascending), d2/dx_i^2 for i in D2 (ascending)]
// This is the set of derivatives that the reference's nested jvp tree
// (trainers.py:63-107, 213-247) produces for required_ujs of order <= 2 with
// repeated indices; it is carried forward through every layer at once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fbpx.h"

typedef float f4 __attribute__((ext_vector_type(4)));

namespace fbpx {

constexpr float kPi = 3.14159265358979323846f;

__host__ __device__ constexpr int popc(int v) { return v == 0 ? 0 : (v & 1) + popc(v >> 1); }

template <int XD_, int D1M_, int D2M_>
struct Schema {
  static constexpr int XD = XD_;
  static constexpr int D1M = D1M_;
  static constexpr int D2M = D2M_;
  static constexpr int N1 = popc(D1M);
  static constexpr int N2 = popc(D2M);
  static constexpr int J = 1 + N1 + N2;
  static_assert((D2M & ~D1M) == 0, "second-order dims must carry first order");
  static_assert(D1M < (1 << XD), "mask exceeds dimension");
  __host__ __device__ static constexpr bool has1(int i) { return (D1M >> i) & 1; }
  __host__ __device__ static constexpr bool has2(int i) { return (D2M >> i) & 1; }
  __host__ __device__ static constexpr int col1(int i) { return 1 + popc(D1M & ((1 << i) - 1)); }
  __host__ __device__ static constexpr int col2(int i) { return 1 + N1 + popc(D2M & ((1 << i) - 1)); }
};

// Offsets of one subdomain's parameter row; layer sizes [XD, H x NH, 1],
// W stored (out, in) row-major as in networks.py:50-58.
template <int XD, int H, int NH>
struct Layout {
  static constexpr int W1 = 0;
  static constexpr int B1 = W1 + H * XD;
  static constexpr int W2 = B1 + H;
  static constexpr int B2 = W2 + (NH == 2 ? H * H : 0);
  static constexpr int WL = B2 + (NH == 2 ? H : 0);
  static constexpr int BL = WL + H;
  static constexpr int P = BL + 1;
};

__host__ __device__ inline int param_count(int xd, int h, int nh) {
  return h * xd + h + (nh == 2 ? h * h + h : 0) + h + 1;
}

// Static row of one subdomain (decompositions.py:127-128,181).
struct Dom {
  float xmin[FBPX_MAX_XD], xmax[FBPX_MAX_XD];
  float umu, usd, flag;
};

__device__ inline void load_dom(const float* dom, int model, Dom& d) {
  const float* r = dom + (size_t)model * FBPX_DOM_STRIDE;
#pragma unroll
  for (int i = 0; i < FBPX_MAX_XD; ++i) {
    d.xmin[i] = r[i];
    d.xmax[i] = r[4 + i];
  }
  d.umu = r[8];
  d.usd = r[9];
  d.flag = r[10];
}

// tanh(z) = 1 - 2 / (1 + e^{2z}) with the hardware base-2 exponential and
// reciprocal: five instructions, branch-free, exact limits (e^{2z} -> inf
// gives 1, -> 0 gives -1).  Absolute error <= 2.5e-7; the rounding of
// 1 - 2r near z = 0 costs relative, not absolute, accuracy.
__device__ inline float tanh_f(float z) {
  constexpr float k2log2e = 2.8853900817779268f;   // 2 / ln 2
  const float e = __builtin_amdgcn_exp2f(z * k2log2e);
  return fmaf(-2.f, __builtin_amdgcn_rcpf(1.f + e), 1.f);
}

__device__ inline f4 mfma16(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// cos and sin of a half angle |phi| <= pi/2 (no range reduction): even and
// odd polynomials of degree 10 and 11, least-squares fits in phi^2 on
// Chebyshev nodes refined toward equal ripple.  With these float32
// coefficients the error is < 2.1e-8 in exact arithmetic and < 1.5e-7
// absolute in float32 evaluation.
__device__ inline void half_angle_cs(float phi, float& c, float& s) {
  const float y = phi * phi;
  float pc = -2.6051492e-07f;
  pc = fmaf(pc, y, 2.4760162e-05f);
  pc = fmaf(pc, y, -1.3888362e-03f);
  pc = fmaf(pc, y, 4.1666638e-02f);
  pc = fmaf(pc, y, -0.5f);
  c = fmaf(pc, y, 1.f);
  float ps = -2.3888729e-08f;
  ps = fmaf(ps, y, 2.7525236e-06f);
  ps = fmaf(ps, y, -1.9840860e-04f);
  ps = fmaf(ps, y, 8.3333310e-03f);
  ps = fmaf(ps, y, -1.6666667e-01f);
  s = phi * fmaf(ps, y, 1.f);
}

// Window jets of one pair (windows.py:25-35 with decompositions.py:196-199).
// Pairs only exist for points inside the box (inclusive), where both
// heaviside factors equal 1 and have zero derivative, and the cosine angle
// theta = pi (x - mu) / sd lies in [-pi, pi].  With phi = theta / 2,
// c = cos(phi), s = sin(phi) (half-angle identities, k = pi / sd):
//   ((1 + cos theta) / 2)^2 = c^4,   d/dx = -2 k s c^3,
//   d2/dx2 = k^2 (sin^2 theta - cos theta (1 + cos theta)) / 2 = k^2 c^2 (3 - 4 c^2).
template <class S>
__device__ inline void window_jets(const float* x, const Dom& d, float* w) {
  constexpr int XD = S::XD;
  float c[XD], c1[XD], c2[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    const float mu = (d.xmin[i] + d.xmax[i]) * 0.5f;
    const float sd = (d.xmax[i] - d.xmin[i]) * 0.5f;
    const float k = kPi / sd;
    // clamped so that rounding at the box edge cannot leave the fitted range
    const float phi = __builtin_amdgcn_fmed3f((x[i] - mu) * (0.5f * k), -0.5f * kPi, 0.5f * kPi);
    float co, si;
    half_angle_cs(phi, co, si);
    // 1 + cos(theta) = 2 c^2 rounds to 0 in the reference's float32 form once
    // c^2 <= 2^-26: there the window and its derivatives are exactly 0, so a
    // point on the edges of all its boxes has window sum 0 (u = 0/0 = NaN,
    // trainers.py:166) as in the reference
    const float cc = fabsf(co) <= 0x1p-13f ? 0.f : co * co;
    c[i] = cc * cc;
    c1[i] = (-2.f * k) * (si * (cc * co));
    c2[i] = (k * k) * (cc * fmaf(-4.f, cc, 3.f));
  }
  float prod = 1.f;
#pragma unroll
  for (int i = 0; i < XD; ++i) prod *= c[i];
  w[0] = d.flag * prod + (1.f - d.flag);
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    float others = 1.f;
#pragma unroll
    for (int j = 0; j < XD; ++j)
      if (j != i) others *= c[j];
    if (S::has1(i)) w[S::col1(i)] = d.flag * c1[i] * others;
    if (S::has2(i)) w[S::col2(i)] = d.flag * c2[i] * others;
  }
}

// Leibniz product of two jets (uw = u * w), diagonal second order.
template <class S>
__device__ inline void jet_mul(const float* u, const float* w, float* out) {
  out[0] = u[0] * w[0];
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (S::has1(i)) {
      const int a = S::col1(i);
      out[a] = u[a] * w[0] + u[0] * w[a];
      if (S::has2(i)) {
        const int b = S::col2(i);
        out[b] = u[b] * w[0] + 2.f * u[a] * w[a] + u[0] * w[b];
      }
    }
  }
}

// Transpose of jet_mul w.r.t. u: given g = dL/d(uw), returns dL/du.
template <class S>
__device__ inline void jet_mul_T(const float* g, const float* w, float* gu) {
  float g0 = g[0] * w[0];
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (S::has1(i)) {
      const int a = S::col1(i);
      g0 += g[a] * w[a];
      float ga = g[a] * w[0];
      if (S::has2(i)) {
        const int b = S::col2(i);
        g0 += g[b] * w[b];
        ga += 2.f * g[b] * w[a];
        gu[b] = g[b] * w[0];
      }
      gu[a] = ga;
    }
  }
  gu[0] = g0;
}

// optax.adam (scale_by_adam, eps_root = 0, then scale_by_learning_rate and
// apply_updates), element-wise in FP32 with the same operation order:
//   mu = omb1*g + b1*mu ; nu = omb2*g^2 + b2*nu ; c = count + 1
//   p  = p + nlr * ((mu / (1 - b1^c)) / (sqrt(nu / (1 - b2^c)) + eps))
// omb1 = f32(1 - b1) and nlr = f32(-lr) are formed in double on the host, as
// Python forms them before they meet the FP32 arrays.
struct AdamConst {
  float b1, b2, omb1, omb2, nlr, eps;
};

inline AdamConst adam_const(double lr, double b1, double b2, double eps) {
  AdamConst k;
  k.b1 = (float)b1;
  k.b2 = (float)b2;
  k.omb1 = (float)(1.0 - b1);
  k.omb2 = (float)(1.0 - b2);
  k.nlr = (float)(-lr);
  k.eps = (float)eps;
  return k;
}

// The optimizer step folded into the backward (mu == NULL: off): each
// chunk's gradient row is applied to its subdomain row of theta in place.
struct AdamFold {
  float* theta;
  float* mu;
  float* nu;
  int32_t* count;
  int32_t* count_done;
  AdamConst k;
};

// bc[0] = 1 - b1^c, bc[1] = 1 - b2^c for c = n + 1, n the step count
__device__ inline void adam_bias_n(int32_t n, const AdamConst& k, float* bc) {
  const float cf = (float)(n + 1);
  bc[0] = 1.f - powf(k.b1, cf);
  bc[1] = 1.f - powf(k.b2, cf);
}

__device__ inline void adam_bias(const int32_t* count, const AdamConst& k, float* bc) {
  adam_bias_n(count[0], k, bc);
}

__device__ inline void adam_elem(float& p, float& m, float& v, float g, float bc1, float bc2, const AdamConst& k) {
#pragma clang fp contract(off)
  m = k.omb1 * g + k.b1 * m;
  v = k.omb2 * (g * g) + k.b2 * v;
  const float mh = m / bc1;
  const float vh = v / bc2;
  const float u = mh / (sqrtf(vh) + k.eps);
  p = p + k.nlr * u;
}

// Thread-local error text; defined in fbpx_host.hip.
void set_error(const char* fmt, ...);
int hip_check(hipError_t e, const char* what);

}  // namespace fbpx
