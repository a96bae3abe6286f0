// Point-level kernels: the inside test / take fill, the sorted POU reduction
// and its transpose, and the masked Adam update.
#include "fbpx_shapes.hpp"

namespace fbpx {

// ------------------------------------------------------------------ takes
//
// One thread per point; the boxes are staged through LDS in tiles.  The test
// is the reference's inclusive FP32 comparison xmin <= x <= xmax on every
// axis (decompositions.py:225-226); fill writes ascending box ids, which is
// the row-major order of jnp.nonzero on the (n, m) inside matrix.

constexpr int kBoxTile = 512;

template <int XD, bool FILL>
__global__ void __launch_bounds__(256) inside_kernel(const float* __restrict__ x, int n, const float* __restrict__ xmin,
                                                     const float* __restrict__ xmax, int m,
                                                     const uint8_t* __restrict__ mask, int32_t* __restrict__ counts,
                                                     const int64_t* __restrict__ offsets, int32_t* __restrict__ n_take,
                                                     int32_t* __restrict__ m_take) {
  __shared__ float slo[kBoxTile * XD], shi[kBoxTile * XD];
  __shared__ uint8_t smask[kBoxTile];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < n;
  float xv[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) xv[i] = live ? x[(size_t)p * XD + i] : 0.f;
  int cnt = 0;
  int64_t pos = (FILL && live) ? offsets[p] : 0;
  for (int t0 = 0; t0 < m; t0 += kBoxTile) {
    const int tn = min(kBoxTile, m - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn * XD; i += blockDim.x) {
      slo[i] = xmin[(size_t)t0 * XD + i];
      shi[i] = xmax[(size_t)t0 * XD + i];
    }
    for (int i = threadIdx.x; i < tn; i += blockDim.x) smask[i] = mask ? mask[t0 + i] : 1;
    __syncthreads();
    if (live) {
      for (int j = 0; j < tn; ++j) {
        bool in = smask[j] != 0;
#pragma unroll
        for (int i = 0; i < XD; ++i) in = in && (xv[i] >= slo[j * XD + i]) && (xv[i] <= shi[j * XD + i]);
        if (in) {
          if (FILL) {
            n_take[pos] = p;
            m_take[pos] = t0 + j;
            ++pos;
          } else {
            ++cnt;
          }
        }
      }
    }
  }
  if (!FILL && live) counts[p] = cnt;
}

template <bool FILL>
static int launch_inside(hipStream_t s, const float* x, int n, int xd, const float* xmin, const float* xmax, int m,
                         const uint8_t* mask, int32_t* counts, const int64_t* offsets, int32_t* n_take,
                         int32_t* m_take) {
  if (n < 0 || m < 0 || xd < 1 || xd > FBPX_MAX_XD) {
    set_error("inside: bad sizes n=%d m=%d xd=%d", n, m, xd);
    return FBPX_EINVAL;
  }
  if (n == 0) return FBPX_OK;
  if (!x || (m > 0 && (!xmin || !xmax)) || (!FILL && !counts) || (FILL && (!offsets || !n_take || !m_take))) {
    set_error("inside: null pointer");
    return FBPX_EINVAL;
  }
  dim3 grid((n + 255) / 256), block(256);
  switch (xd) {
    case 1: hipLaunchKernelGGL((inside_kernel<1, FILL>), grid, block, 0, s, x, n, xmin, xmax, m, mask, counts, offsets, n_take, m_take); break;
    case 2: hipLaunchKernelGGL((inside_kernel<2, FILL>), grid, block, 0, s, x, n, xmin, xmax, m, mask, counts, offsets, n_take, m_take); break;
    default: hipLaunchKernelGGL((inside_kernel<3, FILL>), grid, block, 0, s, x, n, xmin, xmax, m, mask, counts, offsets, n_take, m_take); break;
  }
  return hip_check(hipGetLastError(), FILL ? "inside_fill" : "inside_count");
}

// ------------------------------------------------------------------ POU

template <class S>
__global__ void __launch_bounds__(256) window_pairs_kernel(const float* __restrict__ dom, const float* __restrict__ x,
                                                           fbpx_worklist wl, float* __restrict__ wpair) {
  constexpr int J = S::J, XD = S::XD;
  const int chunk = blockIdx.x;
  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk], end = wl.chunk_start[chunk + 1];
  Dom d;
  load_dom(dom, model, d);
  for (int k = beg + threadIdx.x; k < end; k += blockDim.x) {
    const int p = wl.pair_point[k];
    float xv[XD], wj[J];
#pragma unroll
    for (int i = 0; i < XD; ++i) xv[i] = x[(size_t)p * XD + i];
    window_jets<S>(xv, d, wj);
    const size_t slot = wl.pair_slot[k];
#pragma unroll
    for (int c = 0; c < J; ++c) wpair[slot * J + c] = wj[c];
  }
}

template <int J>
__global__ void __launch_bounds__(256) seg_sum_kernel(fbpx_worklist wl, const float* __restrict__ vals,
                                                      float* __restrict__ out) {
  const int seg = blockIdx.x * blockDim.x + threadIdx.x;
  if (seg >= wl.n_segs) return;
  float acc[J];
#pragma unroll
  for (int c = 0; c < J; ++c) acc[c] = 0.f;
  for (size_t j = wl.seg_pair_start[seg]; j < (size_t)wl.seg_pair_start[seg + 1]; ++j) {
#pragma unroll
    for (int c = 0; c < J; ++c) acc[c] += vals[j * J + c];
  }
#pragma unroll
  for (int c = 0; c < J; ++c) out[(size_t)seg * J + c] = acc[c];
}

// q = U / W on jets (trainers.py:165-166 under jvp):
//   q0 = U0/W0, qi = (Ui - q0 Wi)/W0, qii = (Uii - 2 qi Wi - q0 Wii)/W0
template <class S>
__device__ inline void jet_div(const float* U, const float* W, float* q) {
  const float iw = 1.f / W[0];
  q[0] = U[0] * iw;
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (S::has1(i)) {
      const int a = S::col1(i);
      q[a] = (U[a] - q[0] * W[a]) * iw;
      if (S::has2(i)) {
        const int b = S::col2(i);
        q[b] = (U[b] - 2.f * q[a] * W[a] - q[0] * W[b]) * iw;
      }
    }
  }
}

// Transpose of jet_div w.r.t. U (W is parameter-independent).
template <class S>
__device__ inline void jet_div_T(const float* gq, const float* W, float* gU) {
  const float iw = 1.f / W[0];
  float g0 = gq[0];
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (S::has1(i)) {
      const int a = S::col1(i);
      float ga = gq[a];
      if (S::has2(i)) {
        const int b = S::col2(i);
        const float gb = gq[b] * iw;
        gU[b] = gb;
        ga -= 2.f * W[a] * gb;
        g0 -= W[b] * gb;
      }
      const float gaa = ga * iw;
      gU[a] = gaa;
      g0 -= W[a] * gaa;
    }
  }
  gU[0] = g0 * iw;
}

// The affine constraining function on all J jet columns of one point (the
// Leibniz rule of fbpx_affine_jets, same operation order):
//   C_0 = A_0 + B_0 q_0,  C_a = A_a + B_a q_0 + B_0 q_a,
//   C_b = A_b + B_b q_0 + 2 B_a q_a + B_0 q_b   (a = col1(i), b = col2(i)).
template <class S>
__device__ inline void leibniz(const float* Aj, const float* Bj, const float* q, float* c) {
#pragma clang fp contract(off)
  c[0] = Aj[0] + Bj[0] * q[0];
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (!S::has1(i)) continue;
    const int a = S::col1(i);
    c[a] = Aj[a] + Bj[a] * q[0] + Bj[0] * q[a];
    if (S::has2(i)) {
      const int b = S::col2(i);
      c[b] = Aj[b] + Bj[b] * q[0] + 2.f * Bj[a] * q[a] + Bj[0] * q[b];
    }
  }
}

// Its transpose: gq = (d C / d q)^T gc.
template <class S>
__device__ inline void leibniz_T(const float* Bj, const float* gc, float* gq) {
  gq[0] = Bj[0] * gc[0];
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (!S::has1(i)) continue;
    const int a = S::col1(i);
    gq[0] += Bj[a] * gc[a];
    gq[a] = Bj[0] * gc[a];
    if (S::has2(i)) {
      const int b = S::col2(i);
      gq[0] += Bj[b] * gc[b];
      gq[a] += 2.f * Bj[a] * gc[b];
      gq[b] = Bj[0] * gc[b];
    }
  }
}

// The POU-combined jets of point p (its segments' sums of u*w jets divided by
// their window sums with the quotient rule, averaged over the levels,
// trainers.py:162-170) and, with A != NULL, the affine constraining_fn
// applied to them (trainers.py:174).
template <class S>
__device__ inline void pou_point(const fbpx_worklist& wl, const float* __restrict__ pairbuf,
                                 const float* __restrict__ wsum, const float* __restrict__ A,
                                 const float* __restrict__ B, int p, float* acc) {
  constexpr int J = S::J;
#pragma unroll
  for (int c = 0; c < J; ++c) acc[c] = 0.f;
  for (int seg = wl.point_seg_start[p]; seg < wl.point_seg_start[p + 1]; ++seg) {
    float U[J], W[J], q[J];
#pragma unroll
    for (int c = 0; c < J; ++c) {
      U[c] = 0.f;
      W[c] = wsum[(size_t)seg * J + c];
    }
    for (size_t j = wl.seg_pair_start[seg]; j < (size_t)wl.seg_pair_start[seg + 1]; ++j) {
#pragma unroll
      for (int c = 0; c < J; ++c) U[c] += pairbuf[j * J + c];
    }
    jet_div<S>(U, W, q);
#pragma unroll
    for (int c = 0; c < J; ++c) acc[c] += q[c];
  }
  const float inv = 1.f / (float)wl.npou;
#pragma unroll
  for (int c = 0; c < J; ++c) acc[c] = (wl.npou == 1) ? acc[c] : acc[c] * inv;
  if (A != nullptr) {
    float Aj[J], Bj[J], out[J];
#pragma unroll
    for (int c = 0; c < J; ++c) {
      Aj[c] = A[(size_t)c * wl.n_points + p];
      Bj[c] = B[(size_t)c * wl.n_points + p];
    }
    leibniz<S>(Aj, Bj, acc, out);
#pragma unroll
    for (int c = 0; c < J; ++c) acc[c] = out[c];
  }
}

// Its transpose at point p: from d loss / d (the point's jets) gq to the
// rows of gseg of the point's segments.  B: the affine map's [J][n_points]
// jets, or with b_own the point's own J values.
template <class S>
__device__ inline void pou_point_T(const fbpx_worklist& wl, const float* __restrict__ wsum,
                                   const float* __restrict__ B, const float* g, int p, float* __restrict__ gseg,
                                   bool b_own = false) {
  constexpr int J = S::J;
  const float inv = 1.f / (float)wl.npou;
  float gq[J];
#pragma unroll
  for (int c = 0; c < J; ++c) gq[c] = g[c];
  if (B != nullptr) {
    float Bj[J], gc[J];
#pragma unroll
    for (int c = 0; c < J; ++c) {
      Bj[c] = b_own ? B[c] : B[(size_t)c * wl.n_points + p];
      gc[c] = gq[c];
    }
    leibniz_T<S>(Bj, gc, gq);
  }
#pragma unroll
  for (int c = 0; c < J; ++c) gq[c] = (wl.npou == 1) ? gq[c] : gq[c] * inv;
  for (int seg = wl.point_seg_start[p]; seg < wl.point_seg_start[p + 1]; ++seg) {
    float W[J], gU[J];
#pragma unroll
    for (int c = 0; c < J; ++c) W[c] = wsum[(size_t)seg * J + c];
    jet_div_T<S>(gq, W, gU);
#pragma unroll
    for (int c = 0; c < J; ++c) gseg[(size_t)seg * J + c] = gU[c];
  }
}

template <class S>
__global__ void __launch_bounds__(256) pou_reduce_kernel(fbpx_worklist wl, const float* __restrict__ pairbuf,
                                                         const float* __restrict__ wsum, const float* __restrict__ A,
                                                         const float* __restrict__ B, float* __restrict__ ujet) {
  constexpr int J = S::J;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= wl.n_points) return;
  // the affine jets do not depend on the reduction: loaded ahead of its
  // chain of dependent loads, not after it (pou_point's own affine step,
  // with the same arithmetic)
  float Aj[J], Bj[J];
#pragma unroll
  for (int c = 0; c < J; ++c) {
    Aj[c] = A ? A[(size_t)c * wl.n_points + p] : 0.f;
    Bj[c] = B ? B[(size_t)c * wl.n_points + p] : 0.f;
  }
  float acc[J];
  pou_point<S>(wl, pairbuf, wsum, nullptr, nullptr, p, acc);
  if (A != nullptr) {
    float q[J];
#pragma unroll
    for (int c = 0; c < J; ++c) q[c] = acc[c];
    leibniz<S>(Aj, Bj, q, acc);
  }
#pragma unroll
  for (int c = 0; c < J; ++c) ujet[(size_t)c * wl.n_points + p] = acc[c];
}

// g_scale != NULL: g_ujet is scaled by g_scale[0] as it is read (d L / d
// loss applied to jets computed at unit upstream gradient).
template <class S>
__device__ inline void pou_backward_point(const fbpx_worklist& wl, const float* __restrict__ wsum,
                                          const float* __restrict__ B, const float* __restrict__ g_ujet,
                                          const float* __restrict__ g_scale, float* __restrict__ gseg, int p) {
  constexpr int J = S::J;
  if (p >= wl.n_points) return;
  const float sc = g_scale ? g_scale[0] : 1.f;
  float gq[J];
#pragma unroll
  for (int c = 0; c < J; ++c) gq[c] = g_ujet[(size_t)c * wl.n_points + p] * sc;
  pou_point_T<S>(wl, wsum, B, gq, p, gseg);
}

template <class S>
__global__ void __launch_bounds__(256) pou_backward_kernel(fbpx_worklist wl, const float* __restrict__ wsum,
                                                           const float* __restrict__ B,
                                                           const float* __restrict__ g_ujet,
                                                           const float* __restrict__ g_scale,
                                                           float* __restrict__ gseg) {
  pou_backward_point<S>(wl, wsum, B, g_ujet, g_scale, gseg, blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------ affine constraining function

struct AffineArgs {
  int n_req;
  int kind[FBPX_MAX_REQ];   // 0 value, 1 first, 2 second derivative
  int ca[FBPX_MAX_REQ];     // column of the first derivative
  int cb[FBPX_MAX_REQ];     // column of the second derivative
  float* out[FBPX_MAX_REQ];
  const float* gout[FBPX_MAX_REQ];
};

// One thread per point.  Same operation order as the reference's Leibniz
// expansion evaluated term by term (no contraction into FMAs).
__global__ void __launch_bounds__(256) affine_jets_kernel(int n, int ld, const float* __restrict__ A,
                                                          const float* __restrict__ B,
                                                          const float* __restrict__ q, AffineArgs a) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float A0 = A[k], B0 = B[k], q0 = q[k];
  for (int r = 0; r < a.n_req; ++r) {
    float v;
    if (a.kind[r] == 0) {
      v = A0 + B0 * q0;
    } else {
      const size_t ia = (size_t)a.ca[r] * ld + k;
      if (a.kind[r] == 1) {
        v = A[ia] + B[ia] * q0 + B0 * q[ia];
      } else {
        const size_t ib = (size_t)a.cb[r] * ld + k;
        v = A[ib] + B[ib] * q0 + 2.f * B[ia] * q[ia] + B0 * q[ib];
      }
    }
    a.out[r][k] = v;
  }
}

__global__ void __launch_bounds__(256) affine_jets_backward_kernel(int n, int ld, int J,
                                                                   const float* __restrict__ B, AffineArgs a,
                                                                   float* __restrict__ gq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float g[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) g[c] = 0.f;
  const float B0 = B[k];
  for (int r = 0; r < a.n_req; ++r) {
    const float gr = a.gout[r][k];
    const int kind = a.kind[r], ca = a.ca[r], cb = a.cb[r];
    const float Ba = kind >= 1 ? B[(size_t)ca * ld + k] : 0.f;
    const float Bb = kind == 2 ? B[(size_t)cb * ld + k] : 0.f;
    // d out / d q0 = B0, Ba or Bb;  d / d qa = B0 (first) or 2 Ba (second);  d / d qb = B0
    const float d0 = kind == 0 ? B0 : (kind == 1 ? Ba : Bb);
    const float da = kind == 1 ? B0 : 2.f * Ba;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      float t = c == 0 ? d0 * gr : 0.f;
      if (kind >= 1 && c == ca) t += da * gr;
      if (kind == 2 && c == cb) t += B0 * gr;
      g[c] += t;
    }
  }
#pragma unroll
  for (int c = 0; c < 7; ++c)
    if (c < J) gq[(size_t)c * ld + k] = g[c];
}

bool schema_supported(int xd, int d1, int d2) {
#define X(XD, D1, D2) \
  if (xd == XD && d1 == D1 && d2 == D2) return true;
  FBPX_SCHEMAS(X)
#undef X
  return false;
}

// ------------------------------------------------------------------ Adam
//
// (optax.adam's element update, AdamConst and adam_bias: fbpx_common.hpp)

__global__ void __launch_bounds__(256) reduce_rows_kernel(int P, const float* __restrict__ partial,
                                                          const int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ rcs, int n_rows,
                                                          float* __restrict__ grad) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n_rows * P) return;
  const int row = (int)(idx / P), e = (int)(idx % P);
  float g = 0.f;
  // loads unrolled so several chunk rows are in flight; the sum stays in
  // ascending chunk order
#pragma unroll 4
  for (int c = rcs[row]; c < rcs[row + 1]; ++c) g += partial[(size_t)c * P + e];
  grad[(size_t)rows[row] * P + e] = g;
}

// Each workgroup's thread 0 reads the step count once, at the start.  With
// count_done != NULL that thread, after its read, counts the workgroup on
// count_done; the last workgroup to arrive adds 1 to the count (every other
// workgroup has read it by then) and resets count_done to 0 for the next
// launch.  Nothing else is ordered against the count, so no fence is needed.
// Launches of at most this many workgroups finish their own reduction (or
// the step count) in the last workgroup; larger ones use a second launch,
// which costs less than a fence and an atomic per workgroup.
constexpr int kSingleLaunchBlocks = 256;

__device__ inline void adam_count_step(int32_t* count, int32_t* count_done) {
  if (count_done != nullptr && threadIdx.x == 0 && atomicAdd(count_done, 1) == (int)gridDim.x - 1) {
    atomicAdd(count, 1);
    atomicExch(count_done, 0);
  }
}

// PARTIAL false: the gradient rows grad; true: the per-chunk partial rows,
// summed per listed row as fbpx_reduce_rows sums them (ascending chunks,
// same unrolling), so that both forms give the same update bit for bit.
// One element per thread: the launch covers n_rows * P (n_rows >= 1).  The
// first wave forms the bias corrections while its first loads are in
// flight, so that the count's round trip and the powf do not come before
// them.
template <bool PARTIAL>
__global__ void __launch_bounds__(256) adam_rows_kernel(float* __restrict__ theta, float* __restrict__ mu,
                                                        float* __restrict__ nu, int P,
                                                        const float* __restrict__ grad,
                                                        const float* __restrict__ partial,
                                                        const int32_t* __restrict__ rcs,
                                                        const int32_t* __restrict__ rows, int n_rows,
                                                        int32_t* count, int32_t* count_done, AdamConst k) {
  __shared__ float sbc[2];
  const int idx = (int)(blockIdx.x * blockDim.x + threadIdx.x);   // n_rows * P < 2^31 - 255
  const bool on = idx < n_rows * P;
  // two dependent round trips: the row and its chunk range (past the end:
  // row 0's, unused), then the parameter, its moments and the gradient (or
  // the chunk rows) together
  const int row = on ? idx / P : 0, e = on ? idx - row * P : 0;
  const int r = rows[row];
  const int c0 = PARTIAL ? rcs[row] : 0, c1 = PARTIAL ? rcs[row + 1] : 0;
  if (threadIdx.x < 64) {
    float bc[2];
    adam_bias(count, k, bc);
    if (threadIdx.x == 0) {
      sbc[0] = bc[0];
      sbc[1] = bc[1];
    }
  }
  const size_t i = (size_t)r * P + e;
  float p = 0.f, m = 0.f, v = 0.f, gi = 0.f;
  if (on) {
    p = theta[i];
    m = mu[i];
    v = nu[i];
    if constexpr (!PARTIAL) {
      gi = grad[i];
    } else {
#pragma unroll 4
      for (int c = c0; c < c1; ++c) gi += partial[(size_t)c * P + e];
    }
  }
  __syncthreads();
  if (on) {
    adam_elem(p, m, v, gi, sbc[0], sbc[1], k);
    theta[i] = p;
    mu[i] = m;
    nu[i] = v;
  }
  adam_count_step(count, count_done);
}

// As adam_rows_kernel, over a flat vector: one element per thread.
__global__ void __launch_bounds__(256) adam_flat_kernel(float* __restrict__ p, float* __restrict__ mu,
                                                        float* __restrict__ nu, const float* __restrict__ g, int n,
                                                        int32_t* count, int32_t* count_done, AdamConst k) {
  __shared__ float sbc[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float pv = 0.f, m = 0.f, v = 0.f, gv = 0.f;
  if (i < n) {
    pv = p[i];
    m = mu[i];
    v = nu[i];
    gv = g[i];
  }
  if (threadIdx.x == 0) adam_bias(count, k, sbc);
  __syncthreads();
  if (i < n) {
    adam_elem(pv, m, v, gv, sbc[0], sbc[1], k);
    p[i] = pv;
    mu[i] = m;
    nu[i] = v;
  }
  adam_count_step(count, count_done);
}

// Deterministic mean: stage 1 sums a fixed grid of strided ranges into
// scratch[kSumBlocks] with block sums, stage 2 reduces those in one block.
// No atomics and no memset, so it is safe inside a captured graph.
constexpr int kSumBlocks = 256;

// Sum of s over the block in a fixed order: a butterfly over each wave's 64
// lanes, then the waves' partials in wave order.  The total is returned in
// thread 0 only.  sm holds blockDim.x / 64 floats; a caller that reuses sm
// separates the calls by a barrier.
__device__ inline float block_tree_sum(float s, float* sm) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  float t = 0.f;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w];
  return t;
}

__global__ void __launch_bounds__(256) sum_stage1_kernel(const float* __restrict__ x, int64_t n,
                                                         float* __restrict__ scratch) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  float s = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += x[i];
  const float t = block_tree_sum(s, sm);
  if (threadIdx.x == 0) scratch[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256) sum_stage2_kernel(const float* __restrict__ scratch, int nb, int64_t n,
                                                         float* __restrict__ out) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  const float s = (int)threadIdx.x < nb ? scratch[threadIdx.x] : 0.f;
  const float t = block_tree_sum(s, sm);
  if (threadIdx.x == 0) out[0] = t / (float)n;
}

// ------------------------------------------------------------------ fused residual loss

// Column and term tables of one residual (factor index -1: the factor 1).
// Term t's factors are read through its own column pointers, filled on the
// host from the column table (col[fa[t]], col[fb[t]]), so that a workgroup
// loads them all at once instead of a factor index and then the pointer it
// selects.  ia/ib[t] != NULL: that factor's column f is gathered, F(f, k) =
// col[f][col_idx[f][k]].
struct ResidualArgs {
  int n_cols, n_terms;
  const float* pa[FBPX_MAX_TERMS];
  const float* pb[FBPX_MAX_TERMS];
  const int32_t* ia[FBPX_MAX_TERMS];
  const int32_t* ib[FBPX_MAX_TERMS];
  float* gcol[FBPX_MAX_REQ];
  int fa[FBPX_MAX_TERMS], fb[FBPX_MAX_TERMS];
  int gather;   // bit t: ia[t] != NULL; bit 8 + t: ib[t] != NULL
};

// Points per thread and loop trip: independent loads in flight per thread.
constexpr int kResUnroll = 4;

// The factor 1 of a term without a second (or any) column is read from here
// at index 0, so that every factor is an unconditional load and all of a
// point's loads are in flight at once (a load behind a branch on the term
// count would be waited for before the next one is issued).
__device__ float kResOne[1] = {1.f};

// out[j] = scale * sum over the grid's workgroups b of scratch[b * ncol + j]
// (plus out[j]'s value on entry when acc), computed by the last workgroup to
// finish (counted on done, which it re-zeroes for the next launch) in the
// order of sum_columns_kernel, so one launch replaces the partial-sum launch
// and its reduction.  Every thread of every workgroup calls it after thread 0
// has written the workgroup's partials.
__device__ inline void last_block_sum(const float* scratch, int ncol, float scale, float* out, bool acc,
                                      int32_t* done, float* sm) {
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();   // this workgroup's partials are visible before it is counted
    last = atomicAdd(done, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();     // every other workgroup's partials are visible here
  for (int j = 0; j < ncol; ++j) {
    float v = 0.f;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) v += scratch[b * ncol + j];
    const float t = block_tree_sum(v, sm);
    if (threadIdx.x == 0) out[j] = acc ? out[j] + scale * t : scale * t;
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicExch(done, 0);
}

// The index 0 that a factor without a gather index reads, so that the
// index loads of a trip need no branch per factor.
__device__ int32_t kResIdx0[1] = {0};

// Term t's factor columns, index masks (all-ones for a column, 0 for the
// factor 1, which reads kResOne[0]) and gather indices (ia/ib[t] read at
// point k & ja/jb[t]: the column's index, or kResIdx0[0] with mask 0 where
// the factor is not gathered), uniform across the grid.  gather: bit t (8 +
// t) where factor a (b) of term t is gathered.
template <int NT>
struct TermCols {
  const float* pa[NT];
  const float* pb[NT];
  int64_t ma[NT], mb[NT];
  const int32_t* ia[NT];
  const int32_t* ib[NT];
  int64_t ja[NT], jb[NT];
  int gather;
};

// The factors of one loop trip, xa/xb[u][t] at the points k = k0 + u *
// stride (past n: point 0, the result dropped): p[k], p[idx[k]] for a
// gathered column, p[0] for the factor 1 (mask 0).  When any factor is
// gathered, every index of the trip is loaded, unconditionally, before any
// value, so that a thread waits on two round trips, not on two per gathered
// factor (a load behind a branch of its own is waited for inside it).
template <int NT>
__device__ inline void res_factors(const TermCols<NT>& c, int64_t k0, int64_t stride, int64_t n,
                                   float (*xa)[NT], float (*xb)[NT]) {
  int32_t ga[kResUnroll][NT], gb[kResUnroll][NT];
  if (c.gather != 0) {
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u) {
      const int64_t k = k0 + u * stride;
      const int64_t kk = k < n ? k : 0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        ga[u][t] = c.ia[t][kk & c.ja[t]];
        gb[u][t] = c.ib[t][kk & c.jb[t]];
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) ga[u][t] = gb[u][t] = 0;
  }
#pragma unroll
  for (int u = 0; u < kResUnroll; ++u) {
    const int64_t k = k0 + u * stride;
    const int64_t kk = k < n ? k : 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      xa[u][t] = c.pa[t][(((c.gather >> t) & 1) ? (int64_t)ga[u][t] : kk) & c.ma[t]];
      xb[u][t] = c.pb[t][(((c.gather >> (8 + t)) & 1) ? (int64_t)gb[u][t] : kk) & c.mb[t]];
    }
  }
}

static_assert(FBPX_MAX_TERMS == 8, "the residual launches below instantiate 1 .. 8 terms");

template <int NT>
__device__ inline TermCols<NT> term_cols(const ResidualArgs& a) {
  TermCols<NT> c;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const bool ha = a.fa[t] >= 0, hb = a.fb[t] >= 0;
    c.pa[t] = ha ? a.pa[t] : kResOne;
    c.pb[t] = hb ? a.pb[t] : kResOne;
    c.ma[t] = ha ? ~int64_t(0) : 0;
    c.mb[t] = hb ? ~int64_t(0) : 0;
    const bool ga = (a.gather >> t) & 1, gb = (a.gather >> (8 + t)) & 1;
    c.ia[t] = ga ? a.ia[t] : kResIdx0;
    c.ib[t] = gb ? a.ib[t] : kResIdx0;
    c.ja[t] = ga ? ~int64_t(0) : 0;
    c.jb[t] = gb ? ~int64_t(0) : 0;
  }
  c.gather = a.gather;
  return c;
}

// gcol[f][k] = sr[u] * d r / d F(f, k) at the points k = k0 + u * stride < n
// of one loop trip, for every f with gcol[f] != NULL: d r / d F(f, .) is the
// sum over the terms that read column f of coef times the other factor (a
// column read by both factors of a term counts twice).  No loads here, so a
// runtime column count costs nothing.
template <int NT>
__device__ inline void residual_col_grads(const ResidualArgs& a, const float* cf, const float (*xa)[NT],
                                          const float (*xb)[NT], const float* sr, int64_t k0, int64_t stride,
                                          int64_t n) {
  for (int f = 0; f < a.n_cols; ++f) {
    float* gc = a.gcol[f];
    if (gc == nullptr) continue;
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u) {
      const int64_t k = k0 + u * stride;
      float g = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (a.fa[t] == f) g = fmaf(cf[t], xb[u][t], g);
        if (a.fb[t] == f) g = fmaf(cf[t], xa[u][t], g);
      }
      if (k < n) gc[k] = sr[u] * g;
    }
  }
}

// r[k] = sum_t coef[t] term_t(k) for NT = n_terms over the grid-stride
// points of workgroup lb of a grid of nbl workgroups; returns the sum of r^2
// over them in thread 0.  GRAD: also the loss's column gradients at unit
// upstream gradient, gcol[f][k] = w2n r[k] d r / d F(f, k) (the arithmetic of
// residual_backward_kernel with g_loss = 1).
template <int NT, bool GRAD>
__device__ inline float residual_block(int64_t n, const ResidualArgs& a, const float* __restrict__ coef,
                                       float* __restrict__ r, float w2n, int64_t lb, int64_t nbl, float* sm) {
  const TermCols<NT> c = term_cols<NT>(a);
  float cf[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) cf[t] = coef[t];
  const int64_t stride = nbl * blockDim.x;
  float s = 0.f;
  for (int64_t k0 = lb * blockDim.x + threadIdx.x; k0 < n; k0 += kResUnroll * stride) {
    float xa[kResUnroll][NT], xb[kResUnroll][NT], sr[kResUnroll];
    res_factors<NT>(c, k0, stride, n, xa, xb);
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u) {
      const int64_t k = k0 + u * stride;
      float v = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) v = fmaf(cf[t], xa[u][t] * xb[u][t], v);
      sr[u] = w2n * v;
      if (k < n) {
        r[k] = v;
        s = fmaf(v, v, s);
      }
    }
    if constexpr (GRAD) residual_col_grads<NT>(a, cf, xa, xb, sr, k0, stride, n);
  }
  return block_tree_sum(s, sm);
}

// residual_block over the grid; scratch[block] = its sum of r^2, summed into
// loss[0] (scaled, added to it on entry when acc) by the last workgroup when
// done is given, else by sum_columns_kernel.
template <int NT, bool GRAD>
__global__ void __launch_bounds__(256) residual_kernel(int64_t n, ResidualArgs a, const float* __restrict__ coef,
                                                       float* __restrict__ r, float* scratch, float scale,
                                                       float* __restrict__ loss, bool acc, int32_t* done,
                                                       float w2n) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  const float tot = residual_block<NT, GRAD>(n, a, coef, r, w2n, blockIdx.x, gridDim.x, sm);
  if (done && gridDim.x == 1) {
    // the only workgroup: its total is the sum that last_block_sum would
    // form (the other threads add zeros), with no fence, atomic or scratch
    if (threadIdx.x == 0) loss[0] = acc ? loss[0] + scale * tot : scale * tot;
    return;
  }
  if (threadIdx.x == 0) scratch[blockIdx.x] = tot;
  if (done) last_block_sum(scratch, 1, scale, loss, acc, done, sm);
}

// Several residual losses in one grid: residual i takes workgroups
// [b0[i], b0[i + 1]) and runs there exactly as residual_kernel runs it on a
// grid of its own.
struct ResidualSums {
  int n_res;
  int b0[FBPX_POU_LOSS_RES + 1];
  float scale[FBPX_POU_LOSS_RES];
};

struct ResidualSet {
  ResidualSums sums;
  int nt[FBPX_POU_LOSS_RES];
  int64_t n[FBPX_POU_LOSS_RES];
  float w2n[FBPX_POU_LOSS_RES];
  const float* coef[FBPX_POU_LOSS_RES];
  float* r[FBPX_POU_LOSS_RES];
  ResidualArgs a[FBPX_POU_LOSS_RES];
};

// loss[0] = (loss[0] when acc) + scale[0] S_0 + scale[1] S_1 + ..., S_i the
// sum of residual i's workgroup partials in sum_columns_kernel's order, each
// added as a launch of residual_kernel adds it: the loss of the residual
// launches one after another, bit for bit.  Called by every thread of one
// workgroup.
__device__ inline void residual_set_sum(const float* scratch, const ResidualSums& q, float* loss, bool acc,
                                        float* sm) {
  float total = acc ? loss[0] : 0.f;
  for (int i = 0; i < q.n_res; ++i) {
    float v = 0.f;
    for (int b = q.b0[i] + threadIdx.x; b < q.b0[i + 1]; b += blockDim.x) v += scratch[b];
    const float t = block_tree_sum(v, sm);
    if (threadIdx.x == 0) total = (i == 0 && !acc) ? q.scale[i] * t : total + q.scale[i] * t;
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = total;
}

// NTMAX: the largest term count of the set (2, 4 or 8), so that a set of
// short residuals is not compiled with the registers of eight-term ones.
template <int NTMAX, bool GRAD>
__global__ void __launch_bounds__(256) residual_set_kernel(ResidualSet s_arg, float* scratch,
                                                           float* __restrict__ loss, bool acc, int32_t* done) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  __shared__ int last;
  // the set is read where the launch placed it (the first kernel argument,
  // at the start of the argument segment): indexed by the residual, a copy
  // of the argument would be made in per-lane stack memory
  const ResidualSet& s = *(const ResidualSet*)__builtin_amdgcn_kernarg_segment_ptr();
  int i = 0;
  while (i + 1 < s.sums.n_res && (int)blockIdx.x >= s.sums.b0[i + 1]) ++i;
  const int64_t lb = (int64_t)blockIdx.x - s.sums.b0[i], nbl = s.sums.b0[i + 1] - s.sums.b0[i];
  float tot = 0.f;
#define FBPX_SET_CASE(NT)                                                                       \
  case NT:                                                                                      \
    if constexpr (NT <= NTMAX)                                                                  \
      tot = residual_block<NT, GRAD>(s.n[i], s.a[i], s.coef[i], s.r[i], s.w2n[i], lb, nbl, sm); \
    break;
  switch (s.nt[i]) {
    FBPX_SET_CASE(1) FBPX_SET_CASE(2) FBPX_SET_CASE(3) FBPX_SET_CASE(4)
    FBPX_SET_CASE(5) FBPX_SET_CASE(6) FBPX_SET_CASE(7) FBPX_SET_CASE(8)
  }
#undef FBPX_SET_CASE
  if (threadIdx.x == 0) scratch[blockIdx.x] = tot;
  if (!done) return;
  // the last workgroup to finish sums the partials (as last_block_sum)
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();   // this workgroup's partial is visible before it is counted
    last = atomicAdd(done, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();     // every other workgroup's partial is visible here
  residual_set_sum(scratch, s.sums, loss, acc, sm);
  if (threadIdx.x == 0) atomicExch(done, 0);
}

__global__ void __launch_bounds__(256) residual_set_sum_kernel(const float* __restrict__ scratch, ResidualSums q,
                                                               float* __restrict__ loss, bool acc) {
  __shared__ float sm[4];
  residual_set_sum(scratch, q, loss, acc, sm);
}

// gcol[f][k] = s r[k] d r / d col_f, s = g_loss[0] * w2n; scratch[block *
// NT + t] = sum over the block's points of s r term_t.
template <int NT>
__global__ void __launch_bounds__(256) residual_backward_kernel(int64_t n, ResidualArgs a,
                                                                const float* __restrict__ coef,
                                                                const float* __restrict__ r, float w2n,
                                                                const float* __restrict__ g_loss, float* scratch,
                                                                float* __restrict__ gcoef, int32_t* done) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  const TermCols<NT> c = term_cols<NT>(a);
  const float sc = g_loss[0] * w2n;
  float cf[NT], acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    cf[t] = coef[t];
    acc[t] = 0.f;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k0 < n; k0 += kResUnroll * stride) {
    float xa[kResUnroll][NT], xb[kResUnroll][NT], sr[kResUnroll];
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u) {
      const int64_t k = k0 + u * stride;
      const int64_t kk = k < n ? k : 0;
      sr[u] = r[kk];
    }
    res_factors<NT>(c, k0, stride, n, xa, xb);
#pragma unroll
    for (int u = 0; u < kResUnroll; ++u) {
      const int64_t k = k0 + u * stride;
      sr[u] = k < n ? sc * sr[u] : 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = fmaf(sr[u], xa[u][t] * xb[u][t], acc[t]);
    }
    residual_col_grads<NT>(a, cf, xa, xb, sr, k0, stride, n);
  }
  // one workgroup: its totals are the sums (as in residual_kernel)
  const bool only = done && gridDim.x == 1;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const float tot = block_tree_sum(acc[t], sm);
    if (threadIdx.x == 0) {
      if (only)
        gcoef[t] = tot;
      else
        scratch[blockIdx.x * NT + t] = tot;
    }
    __syncthreads();
  }
  if (done && !only) last_block_sum(scratch, NT, 1.f, gcoef, false, done, sm);
}

// out[j] = scale * sum over b < nb of scratch[b * ncol + j] (+ out[j] on
// entry when acc), j < ncol (one block; thread i sums b = i, i + 256, ... in
// order, then a tree).
__global__ void __launch_bounds__(256) sum_columns_kernel(const float* __restrict__ scratch, int nb, int ncol,
                                                          float scale, float* __restrict__ out, bool acc) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  for (int j = 0; j < ncol; ++j) {
    float v = 0.f;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += scratch[b * ncol + j];
    const float t = block_tree_sum(v, sm);
    if (threadIdx.x == 0) out[j] = acc ? out[j] + scale * t : scale * t;
    __syncthreads();
  }
}

// out[s] = sum over j in [seg_start[s], seg_start[s + 1]) of x[perm[j]] (or
// x[j]), times scale[0] when scale != NULL; one block per segment, fixed
// order.
__device__ inline void segment_sum_block(const int32_t* __restrict__ seg_start, const int32_t* __restrict__ perm,
                                         const float* __restrict__ x, const float* __restrict__ scale,
                                         float* __restrict__ out, int s, float* sm) {
  const int j0 = seg_start[s], j1 = seg_start[s + 1];
  float v = 0.f;
  for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) v += x[perm ? perm[j] : j];
  const float t = block_tree_sum(v, sm);
  if (threadIdx.x == 0) out[s] = scale ? t * scale[0] : t;
}

__global__ void __launch_bounds__(256) segment_sum_kernel(const int32_t* __restrict__ seg_start,
                                                          const int32_t* __restrict__ perm,
                                                          const float* __restrict__ x,
                                                          const float* __restrict__ scale, float* __restrict__ out) {
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  segment_sum_block(seg_start, perm, x, scale, out, blockIdx.x, sm);
}

// pou_backward_kernel on workgroups [0, nb_pou) and segment_sum_kernel on
// the n_seg after them, each with its own arithmetic: two independent
// transposes of a fused loss's backward in one launch.
template <class S>
__global__ void __launch_bounds__(256) pou_backward_segsum_kernel(fbpx_worklist wl, const float* __restrict__ wsum,
                                                                  const float* __restrict__ B,
                                                                  const float* __restrict__ g_ujet,
                                                                  const float* __restrict__ g_scale,
                                                                  float* __restrict__ gseg, int nb_pou,
                                                                  const int32_t* __restrict__ seg_start,
                                                                  const int32_t* __restrict__ perm,
                                                                  const float* __restrict__ x,
                                                                  float* __restrict__ out) {
  __shared__ float sm[4];
  if ((int)blockIdx.x < nb_pou)
    pou_backward_point<S>(wl, wsum, B, g_ujet, g_scale, gseg, blockIdx.x * blockDim.x + threadIdx.x);
  else
    segment_sum_block(seg_start, perm, x, g_scale, out, (int)blockIdx.x - nb_pou, sm);
}

// ------------------------------------------------------------------ POU reduction + residual losses + transpose

// v[c] for a runtime c < N (selects; a register array indexed at run time
// would live in scratch memory).
template <int N>
__device__ inline float sel(const float* v, int c) {
  float r = v[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (c == k) ? v[k] : r;
  return r;
}

// The residuals of one fused launch with their host-computed scales, the
// same float expressions as the separate launches use.
struct PouLossArgs {
  int n_res;
  int mem[FBPX_POU_LOSS_RES];       // bit f: column f is a memory column (f < n_cols, jet[f] < 0)
  fbpx_residual res[FBPX_POU_LOSS_RES];
  float scale[FBPX_POU_LOSS_RES];   // weight / n
  float w2n[FBPX_POU_LOSS_RES];     // 2 weight / n
};

// loss = sum over residuals i of scale_i * (sum over the nb workgroups b of
// part[b * n_res + i]) (the first added to loss[0] on entry when acc), by one
// workgroup in a fixed order, as the separate launches accumulate them.
__device__ inline void pou_loss_total(const float* part, int nb, const PouLossArgs& a, float* loss, bool acc,
                                      float* sm) {
  float total = 0.f;   // thread 0's
  for (int i = 0; i < a.n_res; ++i) {
    float v = 0.f;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[b * a.n_res + i];
    const float t = block_tree_sum(v, sm);
    if (threadIdx.x == 0) {
      const float term = a.scale[i] * t;
      total = i > 0 ? total + term : acc ? loss[0] + term : term;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = total;
}

// One thread per point: pou_point (ujet written), every residual whose
// points hold it (r, the loss's squares, d loss / d column at unit upstream
// gradient with the arithmetic of residual_kernel<NT, true>: jet columns
// into the point's jet gradient, memory columns into gcol), then pou_point_T
// of the jet gradient into gseg.  The loads that do not depend on the POU
// reduction (the affine jets, the residuals' memory columns) are issued
// before it, so that a thread's chain of dependent loads is the reduction's
// alone.  part[block * n_res + i]: the workgroup's sum of r^2 of residual i;
// done != NULL: the last workgroup (or the only one, directly) forms the
// loss.
template <class S>
__global__ void __launch_bounds__(256) pou_loss_kernel(fbpx_worklist wl, const float* __restrict__ pairbuf,
                                                       const float* __restrict__ wsum, const float* __restrict__ A,
                                                       const float* __restrict__ B, PouLossArgs a,
                                                       float* __restrict__ ujet, float* __restrict__ gseg,
                                                       float* part, float* loss, bool acc, int32_t* done) {
  constexpr int J = S::J, NC = FBPX_POU_LOSS_COLS, NT = FBPX_MAX_TERMS, NR = FBPX_POU_LOSS_RES;
  __shared__ float sm[4];   // one partial per wave of the 256-thread block
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float sq[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) sq[i] = 0.f;
  if (p < wl.n_points) {
    // independent of the reduction: the affine jets and the memory columns
    float Aj[J], Bj[J];
#pragma unroll
    for (int c = 0; c < J; ++c) {
      Aj[c] = A ? A[(size_t)c * wl.n_points + p] : 0.f;
      Bj[c] = B ? B[(size_t)c * wl.n_points + p] : 0.f;
    }
    // every gathered column's index first, then every value: two loads'
    // latency in a thread's chain, not two per column
    float M[NR][NC];
    int kc[NR][NC];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int mi = a.mem[i];   // 0 past n_res
#pragma unroll
      for (int f = 0; f < NC; ++f) M[i][f] = 0.f;
      if (mi == 0) continue;
      const fbpx_residual& R = a.res[i];
      const int k = p - R.offset;
      const bool on = k >= 0 && k < R.n;
#pragma unroll
      for (int f = 0; f < NC; ++f) {
        kc[i][f] = k;
        if (on && ((mi >> f) & 1) && R.col_idx[f]) kc[i][f] = R.col_idx[f][k];
      }
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int mi = a.mem[i];
      if (mi == 0) continue;
      const fbpx_residual& R = a.res[i];
      const int k = p - R.offset;
      const bool on = k >= 0 && k < R.n;
#pragma unroll
      for (int f = 0; f < NC; ++f)
        if (on && ((mi >> f) & 1)) M[i][f] = R.col[f][kc[i][f]];
    }
    // the POU reduction (pou_point without its affine loads)
    float u[J], g[J];
    pou_point<S>(wl, pairbuf, wsum, nullptr, nullptr, p, u);
    if (A != nullptr) {
      float q[J];
#pragma unroll
      for (int c = 0; c < J; ++c) q[c] = u[c];
      leibniz<S>(Aj, Bj, q, u);
    }
#pragma unroll
    for (int c = 0; c < J; ++c) {
      ujet[(size_t)c * wl.n_points + p] = u[c];
      g[c] = 0.f;
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      if (i >= a.n_res) break;
      const fbpx_residual& R = a.res[i];
      const int k = p - R.offset;
      if (k < 0 || k >= R.n) continue;
      float F[NC];
#pragma unroll
      for (int f = 0; f < NC; ++f) F[f] = (f < R.n_cols && R.jet[f] >= 0) ? sel<J>(u, R.jet[f]) : M[i][f];
      float xa[NT], xb[NT], cf[NT];
      float r = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        xa[t] = xb[t] = cf[t] = 0.f;
        if (t < R.n_terms) {
          cf[t] = R.coef[t];
          xa[t] = R.fa[t] < 0 ? 1.f : sel<NC>(F, R.fa[t]);
          xb[t] = R.fb[t] < 0 ? 1.f : sel<NC>(F, R.fb[t]);
          r = fmaf(cf[t], xa[t] * xb[t], r);
        }
      }
      R.r[k] = r;
      sq[i] = fmaf(r, r, 0.f);
      const float sr = a.w2n[i] * r;
#pragma unroll
      for (int f = 0; f < NC; ++f) {
        if (f >= R.n_cols) break;
        float gf = 0.f;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (t >= R.n_terms) break;
          if (R.fa[t] == f) gf = fmaf(cf[t], xb[t], gf);
          if (R.fb[t] == f) gf = fmaf(cf[t], xa[t], gf);
        }
        const float gv = sr * gf;
        if (R.jet[f] >= 0) {
#pragma unroll
          for (int c = 0; c < J; ++c)
            if (c == R.jet[f]) g[c] = gv;
        } else if (R.gcol[f] != nullptr) {
          R.gcol[f][k] = gv;
        }
      }
    }
    pou_point_T<S>(wl, wsum, B ? Bj : nullptr, g, p, gseg, true);
  }
  // per-residual workgroup sums of r^2
  float tot[NR];
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    tot[i] = 0.f;
    if (i < a.n_res) tot[i] = block_tree_sum(sq[i], sm);
    __syncthreads();
  }
  if (done && gridDim.x == 1) {
    // the only workgroup: its sums are the totals (no fence, counter or scratch)
    if (threadIdx.x == 0) {
      float total = 0.f;
      for (int i = 0; i < a.n_res; ++i) {
        const float term = a.scale[i] * tot[i];
        total = i > 0 ? total + term : acc ? loss[0] + term : term;
      }
      loss[0] = total;
    }
    return;
  }
  if (threadIdx.x == 0)
    for (int i = 0; i < a.n_res; ++i) part[blockIdx.x * a.n_res + i] = tot[i];
  if (done) {
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();   // this workgroup's partials are visible before it is counted
      last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();     // every other workgroup's partials are visible here
    pou_loss_total(part, gridDim.x, a, loss, acc, sm);
    if (threadIdx.x == 0) atomicExch(done, 0);
  }
}

// The second launch of pou_loss_kernel's sum when its grid is too large for
// a last-workgroup sum.
__global__ void __launch_bounds__(256) pou_loss_sum_kernel(const float* part, int nb, PouLossArgs a, float* loss,
                                                           bool acc) {
  __shared__ float sm[4];
  pou_loss_total(part, nb, a, loss, acc, sm);
}

__global__ void count_increment_kernel(int32_t* count) {
  if (threadIdx.x == 0) count[0] = count[0] + 1;
}


}  // namespace fbpx

using namespace fbpx;

static int check_net_schema(const fbpx_net* net, int* J) {
  if (!net) {
    set_error("null net");
    return FBPX_EINVAL;
  }
  if (!schema_supported(net->xd, net->d1mask, net->d2mask)) {
    set_error("jet schema xd=%d d1mask=%d d2mask=%d not compiled", net->xd, net->d1mask, net->d2mask);
    return FBPX_EUNSUPPORTED;
  }
  *J = 1 + popc(net->d1mask) + popc(net->d2mask);
  return FBPX_OK;
}

static int check_wl(const fbpx_worklist* wl) {
  if (!wl || wl->n_pairs < 0 || wl->n_chunks < 0 || wl->n_points < 0 || wl->n_segs < 0 || wl->npou < 1) {
    set_error("bad work list");
    return FBPX_EINVAL;
  }
  return FBPX_OK;
}

extern "C" {

int fbpx_inside_count(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                      int32_t m, const uint8_t* model_mask, int32_t* counts) {
  return launch_inside<false>((hipStream_t)stream, x, n, xd, xmin, xmax, m, model_mask, counts, nullptr, nullptr,
                              nullptr);
}

int fbpx_inside_fill(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                     int32_t m, const uint8_t* model_mask, const int64_t* offsets, int32_t* n_take,
                     int32_t* m_take) {
  return launch_inside<true>((hipStream_t)stream, x, n, xd, xmin, xmax, m, model_mask, nullptr, offsets, n_take,
                             m_take);
}

int fbpx_pou_prepare(void* stream, const fbpx_net* net, const float* dom, const float* x, const fbpx_worklist* wl,
                     float* pair_scratch, float* wsum) {
  int J, rc;
  if ((rc = check_net_schema(net, &J)) || (rc = check_wl(wl))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (wl->n_chunks > 0) {
#define X(XD, D1, D2)                                                                                        \
  if (net->xd == XD && net->d1mask == D1 && net->d2mask == D2)                                               \
    hipLaunchKernelGGL((window_pairs_kernel<Schema<XD, D1, D2>>), dim3(wl->n_chunks), dim3(256), 0, s, dom, x, \
                       *wl, pair_scratch);
    FBPX_SCHEMAS(X)
#undef X
  }
  if ((rc = hip_check(hipGetLastError(), "window_pairs"))) return rc;
  if (wl->n_segs > 0) {
    dim3 grid((wl->n_segs + 255) / 256);
    switch (J) {
#define C(JJ) \
  case JJ: hipLaunchKernelGGL((seg_sum_kernel<JJ>), grid, dim3(256), 0, s, *wl, pair_scratch, wsum); break;
      C(1) C(2) C(3) C(4) C(5) C(6) C(7)
#undef C
      default: set_error("J=%d", J); return FBPX_EUNSUPPORTED;
    }
  }
  return hip_check(hipGetLastError(), "seg_sum");
}

int fbpx_pou_reduce(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* pairbuf,
                    const float* wsum, const float* A, const float* B, float* ujet) {
  int J, rc;
  if ((rc = check_net_schema(net, &J)) || (rc = check_wl(wl))) return rc;
  if ((A == nullptr) != (B == nullptr)) {
    set_error("pou_reduce: A and B are both given or both NULL");
    return FBPX_EINVAL;
  }
  if (wl->n_points == 0) return FBPX_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((wl->n_points + 255) / 256);
#define X(XD, D1, D2)                                                                                      \
  if (net->xd == XD && net->d1mask == D1 && net->d2mask == D2)                                             \
    hipLaunchKernelGGL((pou_reduce_kernel<Schema<XD, D1, D2>>), grid, dim3(256), 0, s, *wl, pairbuf, wsum, A, B, \
                       ujet);
  FBPX_SCHEMAS(X)
#undef X
  return hip_check(hipGetLastError(), "pou_reduce");
}

int fbpx_pou_backward(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* wsum,
                      const float* B, const float* g_ujet, const float* g_scale, float* gseg) {
  int J, rc;
  if ((rc = check_net_schema(net, &J)) || (rc = check_wl(wl))) return rc;
  if (wl->n_points == 0) return FBPX_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((wl->n_points + 255) / 256);
#define X(XD, D1, D2)                                                                                          \
  if (net->xd == XD && net->d1mask == D1 && net->d2mask == D2)                                                 \
    hipLaunchKernelGGL((pou_backward_kernel<Schema<XD, D1, D2>>), grid, dim3(256), 0, s, *wl, wsum, B, g_ujet, \
                       g_scale, gseg);
  FBPX_SCHEMAS(X)
#undef X
  return hip_check(hipGetLastError(), "pou_backward");
}

int fbpx_reduce_rows(void* stream, int32_t P, const float* partial, const int32_t* rows,
                     const int32_t* row_chunk_start, int32_t n_rows, float* grad) {
  if (P <= 0 || n_rows < 0) {
    set_error("reduce_rows: bad sizes P=%d n_rows=%d", P, n_rows);
    return FBPX_EINVAL;
  }
  if (n_rows == 0) return FBPX_OK;
  if (!partial || !rows || !row_chunk_start || !grad) {
    set_error("reduce_rows: null pointer");
    return FBPX_EINVAL;
  }
  const int64_t total = (int64_t)n_rows * P;
  hipLaunchKernelGGL(reduce_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     P, partial, rows, row_chunk_start, n_rows, grad);
  return hip_check(hipGetLastError(), "reduce_rows");
}

static int adam_rows_launch(void* stream, float* theta, float* mu, float* nu, int32_t P, const float* grad,
                            const float* partial, const int32_t* rcs, const int32_t* rows, int32_t n_rows,
                            int32_t* count, int32_t* count_done, double lr, double b1, double b2, double eps) {
  if (P <= 0 || n_rows < 0 || (int64_t)n_rows * P > INT32_MAX - 255) {
    set_error("adam_rows: bad sizes P=%d n_rows=%d (n_rows * P below 2^31 - 255)", P, n_rows);
    return FBPX_EINVAL;
  }
  if (!count) {
    set_error("adam_rows: null count");
    return FBPX_EINVAL;
  }
  if (n_rows == 0) return count_done ? fbpx_count_increment(stream, count) : FBPX_OK;
  if (!theta || !mu || !nu || !rows || (!grad && (!partial || !rcs))) {
    set_error("adam_rows: null pointer");
    return FBPX_EINVAL;
  }
  // the count is advanced in this launch (an atomic per workgroup) for a
  // few workgroups, by a second one-thread launch otherwise
  const int64_t total = (int64_t)n_rows * P, nb = (total + 255) / 256;
  const bool fold = count_done && nb <= kSingleLaunchBlocks;
  if (grad)
    hipLaunchKernelGGL(adam_rows_kernel<false>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, theta, mu,
                       nu, P, grad, partial, rcs, rows, n_rows, count, fold ? count_done : nullptr,
                       adam_const(lr, b1, b2, eps));
  else
    hipLaunchKernelGGL(adam_rows_kernel<true>, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, theta, mu, nu,
                       P, grad, partial, rcs, rows, n_rows, count, fold ? count_done : nullptr,
                       adam_const(lr, b1, b2, eps));
  int rc = hip_check(hipGetLastError(), "adam_rows");
  if (rc == FBPX_OK && count_done && !fold) rc = fbpx_count_increment(stream, count);
  return rc;
}

int fbpx_adam_rows(void* stream, float* theta, float* mu, float* nu, int32_t P, const float* grad,
                   const int32_t* rows, int32_t n_rows, int32_t* count, int32_t* count_done, double lr, double b1,
                   double b2, double eps) {
  if (!grad && n_rows > 0) {
    set_error("adam_rows: null grad");
    return FBPX_EINVAL;
  }
  return adam_rows_launch(stream, theta, mu, nu, P, grad, nullptr, nullptr, rows, n_rows, count, count_done, lr, b1,
                          b2, eps);
}

int fbpx_adam_rows_partial(void* stream, float* theta, float* mu, float* nu, int32_t P, const float* partial,
                           const int32_t* rows, const int32_t* row_chunk_start, int32_t n_rows, int32_t* count,
                           int32_t* count_done, double lr, double b1, double b2, double eps) {
  if ((!partial || !row_chunk_start) && n_rows > 0) {
    set_error("adam_rows_partial: null partial / row_chunk_start");
    return FBPX_EINVAL;
  }
  return adam_rows_launch(stream, theta, mu, nu, P, nullptr, partial, row_chunk_start, rows, n_rows, count,
                          count_done, lr, b1, b2, eps);
}

int fbpx_adam_flat(void* stream, float* p, float* mu, float* nu, const float* g, int32_t n, int32_t* count,
                   int32_t* count_done, double lr, double b1, double b2, double eps) {
  if (n < 0) {
    set_error("adam_flat: n=%d", n);
    return FBPX_EINVAL;
  }
  if (!count) {
    set_error("adam_flat: null count");
    return FBPX_EINVAL;
  }
  if (n == 0) return count_done ? fbpx_count_increment(stream, count) : FBPX_OK;
  if (!p || !mu || !nu || !g) {
    set_error("adam_flat: null pointer");
    return FBPX_EINVAL;
  }
  const int nb = (n + 255) / 256;
  const bool fold = count_done && nb <= kSingleLaunchBlocks;
  hipLaunchKernelGGL(adam_flat_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, p, mu, nu, g, n, count,
                     fold ? count_done : nullptr, adam_const(lr, b1, b2, eps));
  int rc = hip_check(hipGetLastError(), "adam_flat");
  if (rc == FBPX_OK && count_done && !fold) rc = fbpx_count_increment(stream, count);
  return rc;
}

int fbpx_count_increment(void* stream, int32_t* count) {
  if (!count) {
    set_error("count_increment: null pointer");
    return FBPX_EINVAL;
  }
  hipLaunchKernelGGL(count_increment_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, count);
  return hip_check(hipGetLastError(), "count_increment");
}

static int affine_args(const fbpx_net* net, int32_t n, int32_t ld, int32_t n_req, const int32_t* req,
                       AffineArgs* a, int* J, const char* what) {
  int rc;
  if ((rc = check_net_schema(net, J))) return rc;
  if (n < 0 || ld < n || n_req < 1 || n_req > FBPX_MAX_REQ || !req) {
    set_error("%s: bad sizes n=%d ld=%d n_req=%d", what, n, ld, n_req);
    return FBPX_EINVAL;
  }
  a->n_req = n_req;
  for (int r = 0; r < n_req; ++r) {
    const int c = req[r], xd = net->xd;
    a->kind[r] = 0;
    a->ca[r] = a->cb[r] = 0;
    if (c == 0) continue;
    const int i = (c - 1) % xd;
    const bool second = c > xd;
    if (c < 0 || c > 2 * xd || !((net->d1mask >> i) & 1) || (second && !((net->d2mask >> i) & 1))) {
      set_error("%s: request code %d not carried by the schema (d1mask=%d d2mask=%d)", what, c, net->d1mask,
                net->d2mask);
      return FBPX_EINVAL;
    }
    a->kind[r] = second ? 2 : 1;
    a->ca[r] = 1 + popc(net->d1mask & ((1 << i) - 1));
    if (second) a->cb[r] = 1 + popc(net->d1mask) + popc(net->d2mask & ((1 << i) - 1));
  }
  return FBPX_OK;
}

int fbpx_affine_jets(void* stream, const fbpx_net* net, int32_t n, int32_t ld, const float* A, const float* B,
                     const float* q, int32_t n_req, const int32_t* req, float* const* out) {
  AffineArgs a;
  int J, rc;
  if ((rc = affine_args(net, n, ld, n_req, req, &a, &J, "affine_jets"))) return rc;
  if (n == 0) return FBPX_OK;
  if (!A || !B || !q || !out) {
    set_error("affine_jets: null pointer");
    return FBPX_EINVAL;
  }
  for (int r = 0; r < n_req; ++r) {
    if (!out[r]) {
      set_error("affine_jets: null output %d", r);
      return FBPX_EINVAL;
    }
    a.out[r] = out[r];
    a.gout[r] = nullptr;
  }
  hipLaunchKernelGGL(affine_jets_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, ld, A, B,
                     q, a);
  return hip_check(hipGetLastError(), "affine_jets");
}

int fbpx_affine_jets_backward(void* stream, const fbpx_net* net, int32_t n, int32_t ld, const float* B,
                              int32_t n_req, const int32_t* req, const float* const* gout, float* gq) {
  AffineArgs a;
  int J, rc;
  if ((rc = affine_args(net, n, ld, n_req, req, &a, &J, "affine_jets_backward"))) return rc;
  if (n == 0) return FBPX_OK;
  if (!B || !gout || !gq) {
    set_error("affine_jets_backward: null pointer");
    return FBPX_EINVAL;
  }
  for (int r = 0; r < n_req; ++r) {
    if (!gout[r]) {
      set_error("affine_jets_backward: null gradient %d", r);
      return FBPX_EINVAL;
    }
    a.gout[r] = gout[r];
    a.out[r] = nullptr;
  }
  hipLaunchKernelGGL(affine_jets_backward_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                     ld, J, B, a, gq);
  return hip_check(hipGetLastError(), "affine_jets_backward");
}

static int residual_blocks(int64_t n) {
  const int64_t need = (n + 256 * kResUnroll - 1) / (256 * kResUnroll);
  return (int)(need < FBPX_RESIDUAL_BLOCKS ? need : FBPX_RESIDUAL_BLOCKS);
}

static int residual_args(int64_t n, int32_t n_cols, const float* const* col, const int32_t* const* col_idx,
                         int32_t n_terms, const int32_t* fa, const int32_t* fb, float* const* gcol,
                         ResidualArgs* a) {
  if (n <= 0 || n_cols < 1 || n_cols > FBPX_MAX_REQ || n_terms < 1 || n_terms > FBPX_MAX_TERMS || !col || !fa ||
      !fb) {
    set_error("residual loss: bad sizes n=%lld n_cols=%d n_terms=%d (at most %d columns, %d terms)", (long long)n,
              n_cols, n_terms, FBPX_MAX_REQ, FBPX_MAX_TERMS);
    return FBPX_EINVAL;
  }
  *a = ResidualArgs{};
  a->n_cols = n_cols;
  a->n_terms = n_terms;
  for (int f = 0; f < n_cols; ++f) {
    if (!col[f]) {
      set_error("residual loss: null column %d", f);
      return FBPX_EINVAL;
    }
    a->gcol[f] = gcol ? gcol[f] : nullptr;
  }
  for (int t = 0; t < n_terms; ++t) {
    if (fa[t] >= n_cols || fb[t] >= n_cols) {
      set_error("residual loss: term %d factors (%d, %d) outside %d columns", t, fa[t], fb[t], n_cols);
      return FBPX_EINVAL;
    }
    a->fa[t] = fa[t] < 0 ? -1 : fa[t];
    a->fb[t] = fb[t] < 0 ? -1 : fb[t];
    if (a->fa[t] >= 0) {
      a->pa[t] = col[a->fa[t]];
      a->ia[t] = col_idx ? col_idx[a->fa[t]] : nullptr;
    }
    if (a->fb[t] >= 0) {
      a->pb[t] = col[a->fb[t]];
      a->ib[t] = col_idx ? col_idx[a->fb[t]] : nullptr;
    }
    a->gather |= (a->ia[t] != nullptr) << t | (a->ib[t] != nullptr) << (8 + t);
  }
  return FBPX_OK;
}

// The residual loss launch; gcol != NULL: with the column gradients at unit
// upstream gradient (residual_kernel<NT, true>).
static int residual_loss_launch(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                                const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa,
                                const int32_t* fb, const float* coef, float weight, float* r, float* scratch,
                                float* loss, int32_t accumulate, int32_t* done, float* const* gcol) {
  ResidualArgs a;
  int rc = residual_args(n, n_cols, col, col_idx, n_terms, fa, fb, gcol, &a);
  if (rc) return rc;
  if (!coef || !r || !scratch || !loss) {
    set_error("residual loss: null coef / r / scratch / loss");
    return FBPX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nb = residual_blocks(n);
  const float scale = weight / (float)n;
  const float w2n = 2.f * weight / (float)n;   // as in the transpose's launch
  const bool acc = accumulate != 0;
  // the last-workgroup sum (a fence and an atomic per workgroup) pays off
  // against a second launch only for a few workgroups
  int32_t* const one = nb <= kSingleLaunchBlocks ? done : nullptr;
#define FBPX_RES_CASE(NT)                                                                                     \
  case NT:                                                                                                    \
    if (gcol)                                                                                                 \
      hipLaunchKernelGGL((residual_kernel<NT, true>), dim3(nb), dim3(256), 0, s, n, a, coef, r, scratch, scale, \
                         loss, acc, one, w2n);                                                                \
    else                                                                                                      \
      hipLaunchKernelGGL((residual_kernel<NT, false>), dim3(nb), dim3(256), 0, s, n, a, coef, r, scratch,     \
                         scale, loss, acc, one, w2n);                                                         \
    break;
  switch (n_terms) {
    FBPX_RES_CASE(1) FBPX_RES_CASE(2) FBPX_RES_CASE(3) FBPX_RES_CASE(4)
    FBPX_RES_CASE(5) FBPX_RES_CASE(6) FBPX_RES_CASE(7) FBPX_RES_CASE(8)
  }
#undef FBPX_RES_CASE
  if (!one) hipLaunchKernelGGL(sum_columns_kernel, dim3(1), dim3(256), 0, s, scratch, nb, 1, scale, loss, acc);
  return hip_check(hipGetLastError(), "residual_loss");
}

int fbpx_residual_loss(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                       const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                       const float* coef, float weight, float* r, float* scratch, float* loss, int32_t accumulate,
                       int32_t* done) {
  return residual_loss_launch(stream, n, n_cols, col, col_idx, n_terms, fa, fb, coef, weight, r, scratch, loss,
                              accumulate, done, nullptr);
}

int fbpx_residual_loss_grad(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                            const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                            const float* coef, float weight, float* r, float* scratch, float* loss,
                            int32_t accumulate, int32_t* done, float* const* gcol) {
  if (!gcol) {
    set_error("residual_loss_grad: null gcol table");
    return FBPX_EINVAL;
  }
  return residual_loss_launch(stream, n, n_cols, col, col_idx, n_terms, fa, fb, coef, weight, r, scratch, loss,
                              accumulate, done, gcol);
}

int fbpx_residual_losses(void* stream, int32_t n_res, const fbpx_residual* res, float* scratch,
                         int64_t scratch_len, float* loss, int32_t accumulate, int32_t* done) {
  if (n_res < 1 || n_res > FBPX_POU_LOSS_RES || !res || !scratch || !loss) {
    set_error("residual_losses: n_res=%d (1 .. %d) or a null pointer", n_res, FBPX_POU_LOSS_RES);
    return FBPX_EINVAL;
  }
  ResidualSet s{};
  s.sums.n_res = n_res;
  bool grad = false;
  for (int i = 0; i < n_res; ++i) {
    const fbpx_residual& R = res[i];
    if (R.offset != 0 || R.n < 1 || R.n_cols < 1 || R.n_cols > FBPX_POU_LOSS_COLS || !R.coef || !R.r) {
      set_error("residual_losses: residual %d: offset %d (must be 0), %d points, %d columns (1 .. %d) or a null "
                "coef / r", i, R.offset, R.n, R.n_cols, FBPX_POU_LOSS_COLS);
      return FBPX_EINVAL;
    }
    for (int f = 0; f < R.n_cols; ++f) {
      if (R.jet[f] != -1) {
        set_error("residual_losses: residual %d column %d: jet %d (columns here are device memory, jet -1)", i, f,
                  R.jet[f]);
        return FBPX_EINVAL;
      }
      grad = grad || R.gcol[f] != nullptr;
    }
    int rc = residual_args(R.n, R.n_cols, R.col, R.col_idx, R.n_terms, R.fa, R.fb, R.gcol, &s.a[i]);
    if (rc) return rc;
    s.nt[i] = R.n_terms;
    s.n[i] = R.n;
    s.coef[i] = R.coef;
    s.r[i] = R.r;
    s.sums.scale[i] = R.weight / (float)R.n;
    s.w2n[i] = 2.f * R.weight / (float)R.n;   // as in fbpx_residual_loss
    s.sums.b0[i + 1] = s.sums.b0[i] + residual_blocks(R.n);
  }
  const int nb = s.sums.b0[n_res];
  if (scratch_len < nb) {
    set_error("residual_losses: scratch of %lld floats for %d workgroups", (long long)scratch_len, nb);
    return FBPX_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const bool acc = accumulate != 0;
  int32_t* const one = nb <= kSingleLaunchBlocks ? done : nullptr;
  int ntmax = 0;
  for (int i = 0; i < n_res; ++i) ntmax = s.nt[i] > ntmax ? s.nt[i] : ntmax;
#define FBPX_SET_LAUNCH(NTMAX)                                                                                \
  if (grad)                                                                                                   \
    hipLaunchKernelGGL((residual_set_kernel<NTMAX, true>), dim3(nb), dim3(256), 0, st, s, scratch, loss, acc, \
                       one);                                                                                  \
  else                                                                                                        \
    hipLaunchKernelGGL((residual_set_kernel<NTMAX, false>), dim3(nb), dim3(256), 0, st, s, scratch, loss, acc, \
                       one);
  if (ntmax <= 2) {
    FBPX_SET_LAUNCH(2)
  } else if (ntmax <= 4) {
    FBPX_SET_LAUNCH(4)
  } else {
    FBPX_SET_LAUNCH(8)
  }
#undef FBPX_SET_LAUNCH
  if (!one) hipLaunchKernelGGL(residual_set_sum_kernel, dim3(1), dim3(256), 0, st, scratch, s.sums, loss, acc);
  return hip_check(hipGetLastError(), "residual_losses");
}

int fbpx_residual_loss_backward(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                                const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa,
                                const int32_t* fb, const float* coef, const float* r, float weight,
                                const float* g_loss, float* const* gcol, float* scratch, float* gcoef,
                                int32_t* done) {
  ResidualArgs a;
  int rc = residual_args(n, n_cols, col, col_idx, n_terms, fa, fb, gcol, &a);
  if (rc) return rc;
  if (!coef || !r || !g_loss || !scratch) {
    set_error("residual loss backward: null coef / r / g_loss / scratch");
    return FBPX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nb = residual_blocks(n);
  const float w2n = 2.f * weight / (float)n;
#define FBPX_RES_CASE(NT)                                                                                 \
  case NT:                                                                                                \
    hipLaunchKernelGGL(residual_backward_kernel<NT>, dim3(nb), dim3(256), 0, s, n, a, coef, r, w2n, g_loss, \
                       scratch, gcoef, one);                                                              \
    break;
  int32_t* const one = gcoef && nb <= kSingleLaunchBlocks ? done : nullptr;
  switch (n_terms) {
    FBPX_RES_CASE(1) FBPX_RES_CASE(2) FBPX_RES_CASE(3) FBPX_RES_CASE(4)
    FBPX_RES_CASE(5) FBPX_RES_CASE(6) FBPX_RES_CASE(7) FBPX_RES_CASE(8)
  }
#undef FBPX_RES_CASE
  if (gcoef && !one)
    hipLaunchKernelGGL(sum_columns_kernel, dim3(1), dim3(256), 0, s, scratch, nb, n_terms, 1.f, gcoef, false);
  return hip_check(hipGetLastError(), "residual_loss_backward");
}

int fbpx_pou_loss(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* pairbuf,
                  const float* wsum, const float* A, const float* B, int32_t n_res, const fbpx_residual* res,
                  float* ujet, float* gseg, float* scratch, int64_t scratch_len, float* loss, int32_t accumulate,
                  int32_t* done) {
  int J, rc;
  if ((rc = check_net_schema(net, &J)) || (rc = check_wl(wl))) return rc;
  if ((A == nullptr) != (B == nullptr)) {
    set_error("pou_loss: A and B are both given or both NULL");
    return FBPX_EINVAL;
  }
  if (n_res < 1 || n_res > FBPX_POU_LOSS_RES || !res || !pairbuf || !wsum || !ujet || !gseg || !scratch || !loss) {
    set_error("pou_loss: n_res=%d (1 .. %d) or a null pointer", n_res, FBPX_POU_LOSS_RES);
    return FBPX_EINVAL;
  }
  const int64_t nb = ((int64_t)wl->n_points + 255) / 256;
  if (wl->n_points < 1 || scratch_len < nb * n_res) {
    set_error("pou_loss: %d points, scratch of %lld floats (need %lld)", wl->n_points, (long long)scratch_len,
              (long long)(nb * n_res));
    return FBPX_EINVAL;
  }
  PouLossArgs a{};
  a.n_res = n_res;
  for (int i = 0; i < n_res; ++i) {
    const fbpx_residual& R = res[i];
    if (R.n < 1 || R.offset < 0 || (int64_t)R.offset + R.n > wl->n_points || R.n_cols < 1 ||
        R.n_cols > FBPX_POU_LOSS_COLS || R.n_terms < 1 || R.n_terms > FBPX_MAX_TERMS || !R.coef || !R.r) {
      set_error("pou_loss: residual %d: points [%d, %d + %d) of %d, %d columns (at most %d), %d terms (at most %d)",
                i, R.offset, R.offset, R.n, wl->n_points, R.n_cols, FBPX_POU_LOSS_COLS, R.n_terms, FBPX_MAX_TERMS);
      return FBPX_EINVAL;
    }
    for (int f = 0; f < R.n_cols; ++f) {
      if (R.jet[f] >= J || (R.jet[f] < 0 && !R.col[f])) {
        set_error("pou_loss: residual %d column %d: jet column %d of %d, memory column %p", i, f, R.jet[f], J,
                  (const void*)R.col[f]);
        return FBPX_EINVAL;
      }
    }
    for (int t = 0; t < R.n_terms; ++t) {
      if (R.fa[t] >= R.n_cols || R.fb[t] >= R.n_cols) {
        set_error("pou_loss: residual %d term %d factors (%d, %d) outside %d columns", i, t, R.fa[t], R.fb[t],
                  R.n_cols);
        return FBPX_EINVAL;
      }
    }
    a.res[i] = R;
    for (int f = R.n_cols; f < FBPX_POU_LOSS_COLS; ++f) a.res[i].jet[f] = -1;
    for (int f = 0; f < R.n_cols; ++f) a.mem[i] |= (R.jet[f] < 0) << f;
    for (int t = 0; t < R.n_terms; ++t) {
      a.res[i].fa[t] = R.fa[t] < 0 ? -1 : R.fa[t];
      a.res[i].fb[t] = R.fb[t] < 0 ? -1 : R.fb[t];
    }
    a.scale[i] = R.weight / (float)R.n;        // as fbpx_residual_loss
    a.w2n[i] = 2.f * R.weight / (float)R.n;    // as its transpose
  }
  hipStream_t s = (hipStream_t)stream;
  const bool acc = accumulate != 0;
  int32_t* const one = nb <= kSingleLaunchBlocks ? done : nullptr;
#define X(XD, D1, D2)                                                                                        \
  if (net->xd == XD && net->d1mask == D1 && net->d2mask == D2)                                               \
    hipLaunchKernelGGL((pou_loss_kernel<Schema<XD, D1, D2>>), dim3((unsigned)nb), dim3(256), 0, s, *wl, pairbuf, \
                       wsum, A, B, a, ujet, gseg, scratch, loss, acc, one);
  FBPX_SCHEMAS(X)
#undef X
  if (!one) hipLaunchKernelGGL(pou_loss_sum_kernel, dim3(1), dim3(256), 0, s, scratch, (int)nb, a, loss, acc);
  return hip_check(hipGetLastError(), "pou_loss");
}

int fbpx_segment_sum(void* stream, int32_t n_seg, const int32_t* seg_start, const int32_t* perm, const float* x,
                     const float* scale, float* out) {
  if (n_seg < 0 || (n_seg > 0 && (!seg_start || !x || !out))) {
    set_error("segment_sum: n_seg=%d, null seg_start / x / out", n_seg);
    return FBPX_EINVAL;
  }
  if (n_seg == 0) return FBPX_OK;
  hipLaunchKernelGGL(segment_sum_kernel, dim3(n_seg), dim3(256), 0, (hipStream_t)stream, seg_start, perm, x, scale,
                     out);
  return hip_check(hipGetLastError(), "segment_sum");
}

int fbpx_pou_backward_segment_sum(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* wsum,
                                  const float* B, const float* g_ujet, const float* g_scale, float* gseg,
                                  int32_t n_seg, const int32_t* seg_start, const int32_t* perm, const float* x,
                                  float* out) {
  int J, rc;
  if ((rc = check_net_schema(net, &J)) || (rc = check_wl(wl))) return rc;
  if (n_seg < 0 || (n_seg > 0 && (!seg_start || !x || !out))) {
    set_error("pou_backward_segment_sum: n_seg=%d, null seg_start / x / out", n_seg);
    return FBPX_EINVAL;
  }
  const int nb_pou = (wl->n_points + 255) / 256;
  if (nb_pou + n_seg == 0) return FBPX_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid((unsigned)(nb_pou + n_seg));
#define X(XD, D1, D2)                                                                                             \
  if (net->xd == XD && net->d1mask == D1 && net->d2mask == D2)                                                    \
    hipLaunchKernelGGL((pou_backward_segsum_kernel<Schema<XD, D1, D2>>), grid, dim3(256), 0, s, *wl, wsum, B,     \
                       g_ujet, g_scale, gseg, nb_pou, seg_start, perm, x, out);
  FBPX_SCHEMAS(X)
#undef X
  return hip_check(hipGetLastError(), "pou_backward_segment_sum");
}

int fbpx_mean(void* stream, const float* x, int64_t n, float* scratch, float* out) {
  if (n <= 0 || !x || !scratch || !out) {
    set_error("mean: bad arguments (n=%lld)", (long long)n);
    return FBPX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t need = (n + 255) / 256;
  const int nb = (int)(need < kSumBlocks ? need : kSumBlocks);
  hipLaunchKernelGGL(sum_stage1_kernel, dim3(nb), dim3(256), 0, s, x, n, scratch);
  hipLaunchKernelGGL(sum_stage2_kernel, dim3(1), dim3(256), 0, s, scratch, nb, n, out);
  return hip_check(hipGetLastError(), "mean");
}

}  // extern "C"
