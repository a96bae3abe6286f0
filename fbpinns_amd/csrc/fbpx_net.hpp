// Forward and backward jet kernels over (point, subdomain) pairs.
//
// Work decomposition: one workgroup (4 waves) per chunk of <= FBPX_CHUNK pairs
// of a single subdomain; each wave processes 16 pairs at a time ("sub-tile").
// Lane l handles pair (l & 15) of the sub-tile and hidden units
//   u = 16*rb + 4*(l >> 4) + r,   rb < H/16, r < 4,
// which is exactly the row set that lane holds in the D layout of
// v_mfma_f32_16x16x4f32 (D[i][j]: lane = j + 16*(i/4), reg = i%4).  Because
// the contraction index of an MFMA may be permuted freely, K-step (kb, ks)
// feeds the B operand from register [kb][ks] (unit 16kb + 4g + ks for lane
// group g) and the A operand with the matching weight column, so a layer's
// output registers are directly the next layer's input registers.
//
// Reference math (per pair, trainers.py:113-118 with the tangents that the
// nested jvp of trainers.py:213-247 pushes through):
//   xh = (x - mu)/sd, tangent e_i/sd_i          decompositions.py:184-188
//   z = W a + b, a = tanh(z) with jets
//     a0 = h, a_i = d1 z_i, a_ii = d1 z_ii + d2 z_i^2 (d1 = 1-h^2, d2 = -2 h d1)
//   u = (WL a + bL) * sd_u + mu_u               networks.py:60-68, 200-201
//   uw = u * w (Leibniz), w = cosine window     windows.py:25-35
#pragma once

#include "fbpx_common.hpp"

namespace fbpx {

// Layer-1 jets for this lane's units: A[c][rb][r].
template <class S, int H, class L>
__device__ inline void layer1(const float* sp, const float* xh, const float* isd, float (*A)[H / 16][4], int g) {
  constexpr int XD = S::XD;
#pragma unroll
  for (int rb = 0; rb < H / 16; ++rb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * rb + 4 * g + r;
      float z = 0.f;
#pragma unroll
      for (int i = 0; i < XD; ++i) z = fmaf(sp[L::W1 + u * XD + i], xh[i], z);
      z += sp[L::B1 + u];
      const float h = tanh_f(z);
      const float d1 = 1.f - h * h;
      const float d2 = -2.f * h * d1;
      A[0][rb][r] = h;
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        if (S::has1(i)) {
          const float zt = sp[L::W1 + u * XD + i] * isd[i];
          A[S::col1(i)][rb][r] = d1 * zt;
          if (S::has2(i)) A[S::col2(i)][rb][r] = d2 * zt * zt;
        }
      }
    }
  }
}

// Padded LDS copies of W2: wt[k*(H+4) + o] = W2[o][k] (read as the forward A
// operand) and wr[o*(H+4) + k] = W2[o][k] (read as the W2^T A operand).  The
// stride H+4 puts lane groups g = 0, 1 of a half-wave 16 banks apart, so both
// read patterns are bank-conflict-free.
template <int H>
__device__ inline void stage_w2(const float* w2, float* wt, float* wr) {
  for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
    const int o = i / H, k = i % H;
    const float v = w2[i];
    if (wt) wt[k * (H + 4) + o] = v;
    if (wr) wr[o * (H + 4) + k] = v;
  }
}

// Hidden layer 2 pre-activation jets via MFMA: Z[c][ob] (D layout).
template <class S, int H, class L>
__device__ inline void layer2_mfma(const float* sp, const float* wt, const float (*A)[H / 16][4], f4 (*Z)[H / 16],
                                   int pl, int g) {
  constexpr int J = S::J, NRB = H / 16;
#pragma unroll
  for (int ob = 0; ob < NRB; ++ob) {
    const int o = 16 * ob + 4 * g;
    f4 b = {sp[L::B2 + o], sp[L::B2 + o + 1], sp[L::B2 + o + 2], sp[L::B2 + o + 3]};
    Z[0][ob] = b;
#pragma unroll
    for (int c = 1; c < J; ++c) Z[c][ob] = f4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int kb = 0; kb < NRB; ++kb) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob) {
        const float w = wt[(16 * kb + 4 * g + ks) * (H + 4) + 16 * ob + pl];
#pragma unroll
        for (int c = 0; c < J; ++c) Z[c][ob] = mfma16(w, A[c][kb][ks], Z[c][ob]);
      }
    }
  }
}

// tanh jets of one element: zc[J] -> ac[J].
template <class S>
__device__ inline void tanh_jet(const float* zc, float* ac) {
  const float h = tanh_f(zc[0]);
  const float d1 = 1.f - h * h;
  const float d2 = -2.f * h * d1;
  ac[0] = h;
#pragma unroll
  for (int i = 0; i < S::XD; ++i) {
    if (S::has1(i)) {
      const float zi = zc[S::col1(i)];
      ac[S::col1(i)] = d1 * zi;
      if (S::has2(i)) ac[S::col2(i)] = d1 * zc[S::col2(i)] + d2 * zi * zi;
    }
  }
}

__device__ inline float sum_groups(float v) {
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

template <int CTRL>
__device__ inline float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Sum over the 16 lanes of a DPP row (lanes 16k..16k+15); every lane of the
// row receives the sum.  quad_perm[1,0,3,2], quad_perm[2,3,0,1],
// row_half_mirror, row_mirror.
__device__ inline float row_sum16(float v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  v += dpp<0x140>(v);
  return v;
}

// vals[rb][r] holds this lane's contribution for unit 16rb + 4g + r.  Sums
// every entry over the 16 pair-lanes of the row and adds the total for unit
// 16*(pl/4) + 4g + pl%4 into dst (lane pl owns that unit; no two lanes of a
// wave share a destination).
template <int NRB>
__device__ inline void reduce_units_to_lds(const float (*vals)[4], float* dst, int pl, int g) {
  float mine = 0.f;
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float s = row_sum16(vals[rb][r]);
      if (pl == rb * 4 + r) mine = s;
    }
  if (pl < NRB * 4) dst[16 * (pl >> 2) + 4 * g + (pl & 3)] += mine;
}

// ------------------------------------------------------------------ streaming helpers
//
// Layer-1 jets are never stored: with h, d1 = 1 - h^2, d2 = -2 h d1 of layer
// 1 and zt_i = W1[:, i] / sd_i (per subdomain),
//   a1_0 = h,  a1_i = d1 zt_i,  a1_ii = d2 zt_i^2,
// so the B operand of every layer-2 MFMA is rebuilt with one multiply.  Layer
// 2 is then evaluated one axis at a time (Z_0, then Z_i and Z_ii per axis),
// which keeps 3 x 16 accumulators live instead of J x 16.

// tanh derivatives from h = tanh(z): d1 = 1 - h^2, d2 = -2 h d1, d3 = -2 (d1^2 + h d2)
__device__ inline float tanh_d1(float h) { return 1.f - h * h; }
__device__ inline float tanh_d2(float h) { return -2.f * h * (1.f - h * h); }
__device__ inline float tanh_d3(float h) {
  const float d1 = 1.f - h * h;
  return -2.f * (d1 * d1 - 2.f * h * h * d1);
}

// zt[i * H + u] = W1[u][i] / sd_i for the chunk's subdomain, in LDS.
template <int XD, int H, class L>
__device__ inline void stage_zt(const float* sp, const float* isd, float* zt) {
  for (int t = threadIdx.x; t < XD * H; t += blockDim.x) {
    const int i = t / H, u = t % H;
    float s = isd[0];
#pragma unroll
    for (int j = 1; j < XD; ++j)
      if (i == j) s = isd[j];
    zt[t] = sp[L::W1 + u * XD + i] * s;
  }
}

// Layer 1 for this lane's units (rows 16rb + 4g + r): h = tanh(W1 xh + b1).
template <int XD, int H, class L>
__device__ inline void layer1_stream(const float* sp, const float* xh, float (*h1)[4], int g) {
#pragma unroll
  for (int rb = 0; rb < H / 16; ++rb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * rb + 4 * g + r;
      float z = 0.f;
#pragma unroll
      for (int i = 0; i < XD; ++i) z = fmaf(sp[L::W1 + u * XD + i], xh[i], z);
      z += sp[L::B1 + u];
      h1[rb][r] = tanh_f(z);
    }
  }
}

// One layer-2 MFMA pass: Z[ob] (+)= W2 . B with B[kb][ks] = base * scale.
// kind 0: B = h (Z initialised with b2), 1: B = d1 zt_i, 2: B = d2 zt_i^2,
// zti = the LDS row zt[i * H + .].
template <int H, class L>
__device__ inline void layer2_pass(const float* sp, const float* wt, const float (*h1)[4], const float* zti,
                                   int kind, f4* Z, int pl, int g) {
  constexpr int NRB = H / 16;
#pragma unroll
  for (int ob = 0; ob < NRB; ++ob) {
    if (kind == 0) {
      const int o = 16 * ob + 4 * g;
      Z[ob] = f4{sp[L::B2 + o], sp[L::B2 + o + 1], sp[L::B2 + o + 2], sp[L::B2 + o + 3]};
    } else {
      Z[ob] = f4{0.f, 0.f, 0.f, 0.f};
    }
  }
#pragma unroll
  for (int kb = 0; kb < NRB; ++kb) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const float h = h1[kb][ks];
      float b = h;
      if (kind == 1) {
        b = tanh_d1(h) * zti[16 * kb + 4 * g + ks];
      } else if (kind == 2) {
        const float z = zti[16 * kb + 4 * g + ks];
        b = tanh_d2(h) * z * z;
      }
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob) {
        const float w = wt[(16 * kb + 4 * g + ks) * (H + 4) + 16 * ob + pl];
        Z[ob] = mfma16(w, b, Z[ob]);
      }
    }
  }
}

// ------------------------------------------------------------------ forward

template <int XD, int H, int NH, int D1M, int D2M>
__global__ void __launch_bounds__(256)
fwd_kernel(const float* __restrict__ theta, const float* __restrict__ dom, const float* __restrict__ x,
           fbpx_worklist wl, float* __restrict__ pairbuf) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, NH>;
  constexpr int J = S::J, NRB = H / 16;
  __shared__ float sp[L::P];
  __shared__ float szt[XD * H];
  __shared__ float swt[NH == 2 ? H * (H + 4) : 1];

  const int chunk = blockIdx.x;
  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  const float* th = theta + (size_t)model * L::P;
  for (int i = threadIdx.x; i < L::P; i += blockDim.x) sp[i] = th[i];
  if constexpr (NH == 2) stage_w2<H>(th + L::W2, swt, nullptr);
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    isd[i] = 1.f / ((d.xmax[i] - d.xmin[i]) / 2.f);
  }
  __syncthreads();
  stage_zt<XD, H, L>(sp, isd, szt);
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int pl = lane & 15, g = lane >> 4;
  for (int base = beg + wave * 16; base < end; base += nw * 16) {
    const int k = base + pl;
    const bool valid = k < end;
    const int p = wl.pair_point[valid ? k : beg];
    float xv[XD], xh[XD];
#pragma unroll
    for (int i = 0; i < XD; ++i) {
      xv[i] = x[(size_t)p * XD + i];
      xh[i] = (xv[i] - mu[i]) * isd[i];
    }
    float h1[NRB][4];
    layer1_stream<XD, H, L>(sp, xh, h1, g);
    float ur[J];
#pragma unroll
    for (int c = 0; c < J; ++c) ur[c] = 0.f;
    if constexpr (NH == 1) {
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * rb + 4 * g + r;
          const float wlu = sp[L::WL + u];
          const float h = h1[rb][r];
          ur[0] = fmaf(wlu, h, ur[0]);
#pragma unroll
          for (int i = 0; i < XD; ++i) {
            const float z = szt[i * H + u];
            if (S::has1(i)) ur[S::col1(i)] = fmaf(wlu, tanh_d1(h) * z, ur[S::col1(i)]);
            if (S::has2(i)) ur[S::col2(i)] = fmaf(wlu, tanh_d2(h) * z * z, ur[S::col2(i)]);
          }
        }
    } else {
      f4 Z0[NRB];
      layer2_pass<H, L>(sp, swt, h1, szt, 0, Z0, pl, g);
      float h2[NRB][4];
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          h2[ob][r] = tanh_f(Z0[ob][r]);
          ur[0] = fmaf(sp[L::WL + 16 * ob + 4 * g + r], h2[ob][r], ur[0]);
        }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        if (!S::has1(i)) continue;
        f4 Zi[NRB], Zii[NRB];
        layer2_pass<H, L>(sp, swt, h1, szt + i * H, 1, Zi, pl, g);
        if (S::has2(i)) layer2_pass<H, L>(sp, swt, h1, szt + i * H, 2, Zii, pl, g);
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float wlu = sp[L::WL + 16 * ob + 4 * g + r];
            const float h = h2[ob][r];
            const float d1 = tanh_d1(h);
            const float zi = Zi[ob][r];
            ur[S::col1(i)] = fmaf(wlu, d1 * zi, ur[S::col1(i)]);
            if (S::has2(i)) ur[S::col2(i)] = fmaf(wlu, d1 * Zii[ob][r] + tanh_d2(h) * zi * zi, ur[S::col2(i)]);
          }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    float uj[J];
#pragma unroll
    for (int c = 0; c < J; ++c) uj[c] = sum_groups(ur[c]);
    uj[0] = (uj[0] + sp[L::BL]) * d.usd + d.umu;
#pragma unroll
    for (int c = 1; c < J; ++c) uj[c] *= d.usd;
    float wj[J], uw[J];
    window_jets<S>(xv, d, wj);
    jet_mul<S>(uj, wj, uw);
    if (valid) {
#pragma unroll
      for (int c = 0; c < J; ++c)
        if ((c & 3) == g) pairbuf[(size_t)k * J + c] = uw[c];
    }
  }
}

// ------------------------------------------------------------------ backward
//
// Reverse pass per pair, given gU = dL/d(segment sum of u*w jets):
//   gu = (d uw / d u)^T gU (jet_mul_T; u*w is linear in u, so u itself is not
//   needed), gur = sd_u * gu, ga_c[u] = WL[u] gur_c, dWL[u] += sum_c gur_c a2_c[u].
//   tanh-jet transpose of layer 2 (d3 = -2 (d1^2 + h d2)):
//     G_0  = d1 ga_0 + sum_i d2 z_i ga_i + sum_i (d2 z_ii + d3 z_i^2) ga_ii
//     G_i  = d1 ga_i + 2 d2 z_i ga_ii,   G_ii = d1 ga_ii
//   dW2 += sum_c G_c a1_c^T (MFMA over the 16 pairs after an LDS transpose),
//   db2 += G_0, GA_c = W2^T G_c (MFMA; registers chain),
//   layer 1 (z1_i = zt_i, z1_ii = 0):
//     gz1_0 = d1 GA_0 + sum_i d2 zt_i GA_i + sum_i d3 zt_i^2 GA_ii
//     gz1_i = d1 GA_i + 2 d2 zt_i GA_ii
//     dW1[u, k] += gz1_0 xh_k + gz1_k / sd_k,  db1 += gz1_0.
// Layer 2 is streamed one axis at a time (Z_0 first, then Z_i, Z_ii), and
// every per-axis term above is consumed before the next axis.  Small
// per-unit gradients are summed over the 16 pairs of a sub-tile with DPP row
// reductions and accumulated per wave in LDS; waves are reduced in a fixed
// order at the end of the chunk (deterministic, no atomics).

template <int XD, int H, int NH>
struct BwdSmem {
  using L = Layout<XD, H, NH>;
  static constexpr int TP = H + 4;                        // padded transpose row
  static constexpr int NSMALL = (NH == 2 ? 3 : 2) + XD;   // per-unit small grads
  static constexpr int SP = L::P;
  static constexpr int STRANS = (NH == 2) ? 4 * 4 * 16 * TP : 1;   // [wave][G, h, d1, d2][pair][unit]
  static constexpr int SSMALL = 4 * NSMALL * H;           // [wave][kind][u]
  static constexpr int SBL = 4 * 16;                      // dbL per wave per pl
};

// dW2 += G A^T over the 16 pairs of the sub-tile.  tg holds G^T and tb the
// base column (h, d1 or d2) as [pair][unit] rows of TP floats; the B operand
// is base * scale with scale[ib] the lane's column scaling (1, zt_i or zt_i^2
// for unit 16 ib + pl).
template <int H>
__device__ inline void dw2_accumulate(const float* tg, const float* tb, const float* scale, f4 (*dW2)[H / 16],
                                      int pl, int g) {
  constexpr int NRB = H / 16, TP = H + 4;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    float gv[NRB], av[NRB];
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob) gv[ob] = tg[(4 * ks + g) * TP + 16 * ob + pl];
#pragma unroll
    for (int ib = 0; ib < NRB; ++ib) av[ib] = tb[(4 * ks + g) * TP + 16 * ib + pl] * scale[ib];
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
      for (int ib = 0; ib < NRB; ++ib) dW2[ob][ib] = mfma16(gv[ob], av[ib], dW2[ob][ib]);
  }
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Write a D-layout register tile (rows = units, cols = pairs) as [pair][unit].
template <int H>
__device__ inline void store_T(float* t, const f4* v, int pl, int g) {
#pragma unroll
  for (int ob = 0; ob < H / 16; ++ob) *reinterpret_cast<f4*>(t + pl * (H + 4) + 16 * ob + 4 * g) = v[ob];
}

// GA[ib] = W2^T G  (A operand W2^T[i = 16ib+pl][o = 16ob+4g+ks] from the padded
// row-major copy wr; B operand = G in D layout).
template <int H>
__device__ inline void w2t_pass(const float* wr, const f4* G, f4* GA, int pl, int g) {
  constexpr int NRB = H / 16;
#pragma unroll
  for (int ib = 0; ib < NRB; ++ib) GA[ib] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ib = 0; ib < NRB; ++ib) {
        const float w = wr[(16 * ob + 4 * g + ks) * (H + 4) + 16 * ib + pl];
        GA[ib] = mfma16(w, G[ob][ks], GA[ib]);
      }
}

template <int XD, int H, int NH, int D1M, int D2M>
__global__ void __launch_bounds__(256, 1)
bwd_kernel(const float* __restrict__ theta, const float* __restrict__ dom, const float* __restrict__ x,
           fbpx_worklist wl, const uint8_t* __restrict__ chunk_train, const float* __restrict__ gseg,
           float* __restrict__ partial) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, NH>;
  using M = BwdSmem<XD, H, NH>;
  constexpr int J = S::J, NRB = H / 16, TP = M::TP, NS = M::NSMALL;
  // small-grad kinds: 0 = dWL, 1 = db1, 2 .. 2+XD-1 = dW1[:, i], 2+XD = db2 (NH == 2)
  constexpr int K_WL = 0, K_B1 = 1, K_W1 = 2, K_B2 = 2 + XD;
  __shared__ float sp[M::SP];
  __shared__ float sacc[M::SP];
  __shared__ float ssmall[M::SSMALL];
  __shared__ float sbl[M::SBL];
  __shared__ float szt[XD * H];
  __shared__ __attribute__((aligned(16))) float strans[M::STRANS];
  __shared__ float swt[NH == 2 ? H * (H + 4) : 1];
  __shared__ float swr[NH == 2 ? H * (H + 4) : 1];

  const int chunk = blockIdx.x;
  if (chunk_train != nullptr && chunk_train[chunk] == 0) return;
  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  const float* th = theta + (size_t)model * L::P;
  for (int i = threadIdx.x; i < L::P; i += blockDim.x) {
    sp[i] = th[i];
    sacc[i] = 0.f;
  }
  for (int i = threadIdx.x; i < M::SSMALL; i += blockDim.x) ssmall[i] = 0.f;
  if (threadIdx.x < M::SBL) sbl[threadIdx.x] = 0.f;
  if constexpr (NH == 2) stage_w2<H>(th + L::W2, swt, swr);
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    isd[i] = 1.f / ((d.xmax[i] - d.xmin[i]) / 2.f);
  }
  __syncthreads();
  stage_zt<XD, H, L>(sp, isd, szt);
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int pl = lane & 15, g = lane >> 4;
  float* mysmall = ssmall + wave * NS * H;
  f4 dW2[NH == 2 ? NRB : 1][NRB];
  if constexpr (NH == 2) {
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
      for (int ib = 0; ib < NRB; ++ib) dW2[ob][ib] = f4{0.f, 0.f, 0.f, 0.f};
  }
  float dbl = 0.f;
  float* tg = strans + (NH == 2 ? (wave * 4 + 0) * 16 * TP : 0);
  float* th1 = strans + (NH == 2 ? (wave * 4 + 1) * 16 * TP : 0);
  float* td1 = strans + (NH == 2 ? (wave * 4 + 2) * 16 * TP : 0);
  float* td2 = strans + (NH == 2 ? (wave * 4 + 3) * 16 * TP : 0);

  for (int base = beg + wave * 16; base < end; base += nw * 16) {
    const int k = base + pl;
    const bool valid = k < end;
    const int kk = valid ? k : beg;
    const int p = wl.pair_point[kk];
    const int seg = wl.pair_seg[kk];
    float xv[XD], xh[XD];
#pragma unroll
    for (int i = 0; i < XD; ++i) {
      xv[i] = x[(size_t)p * XD + i];
      xh[i] = (xv[i] - mu[i]) * isd[i];
    }
    // upstream gradient of this pair (independent of u)
    float gur[J];
    {
      float gU[J], wj[J];
#pragma unroll
      for (int c = 0; c < J; ++c) gU[c] = valid ? gseg[(size_t)seg * J + c] : 0.f;
      window_jets<S>(xv, d, wj);
      jet_mul_T<S>(gU, wj, gur);
#pragma unroll
      for (int c = 0; c < J; ++c) gur[c] *= d.usd;
    }
    if (g == 0) dbl += gur[0];

    float h1[NRB][4];
    layer1_stream<XD, H, L>(sp, xh, h1, g);
    float vWL[NRB][4], gz10[NRB][4], v[NRB][4];

    if constexpr (NH == 1) {
      // single hidden layer: ga_c = WL gur_c acts on layer 1 directly
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * rb + 4 * g + r;
          const float wlu = sp[L::WL + u];
          const float h = h1[rb][r];
          const float d1 = tanh_d1(h), d2 = tanh_d2(h), d3 = tanh_d3(h);
          float dwl = gur[0] * h;
          float gz0 = d1 * wlu * gur[0];
#pragma unroll
          for (int i = 0; i < XD; ++i) {
            if (!S::has1(i)) continue;
            const float z = szt[i * H + u];
            dwl = fmaf(gur[S::col1(i)], d1 * z, dwl);
            gz0 = fmaf(d2 * z, wlu * gur[S::col1(i)], gz0);
            if (S::has2(i)) {
              dwl = fmaf(gur[S::col2(i)], d2 * z * z, dwl);
              gz0 = fmaf(d3 * z * z, wlu * gur[S::col2(i)], gz0);
            }
          }
          vWL[rb][r] = dwl;
          gz10[rb][r] = gz0;
        }
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        // dW1[:, i] += gz1_0 xh_i + gz1_i / sd_i
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float t = gz10[rb][r] * xh[i];
            if (S::has1(i)) {
              const int u = 16 * rb + 4 * g + r;
              const float h = h1[rb][r];
              const float z = szt[i * H + u];
              const float wlu = sp[L::WL + u];
              float gzi = tanh_d1(h) * wlu * gur[S::col1(i)];
              if (S::has2(i)) gzi = fmaf(2.f * tanh_d2(h) * z, wlu * gur[S::col2(i)], gzi);
              t = fmaf(gzi, isd[i], t);
            }
            v[rb][r] = t;
          }
        reduce_units_to_lds<NRB>(v, mysmall + (K_W1 + i) * H, pl, g);
      }
    } else {
      // ---- column 0 of layer 2 (value): Z_0, h2, start of G_0
      f4 G0[NRB];
      float h2[NRB][4];
      {
        f4 Z0[NRB];
        layer2_pass<H, L>(sp, swt, h1, szt, 0, Z0, pl, g);
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float h = tanh_f(Z0[ob][r]);
            h2[ob][r] = h;
            const float wlu = sp[L::WL + 16 * ob + 4 * g + r];
            vWL[ob][r] = gur[0] * h;
            G0[ob][r] = tanh_d1(h) * wlu * gur[0];
          }
      }
      // transposed layer-1 columns for dW2 (h, d1, d2 as [pair][unit])
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        const f4 hv = {h1[rb][0], h1[rb][1], h1[rb][2], h1[rb][3]};
        *reinterpret_cast<f4*>(th1 + pl * TP + 16 * rb + 4 * g) = hv;
        if (S::N1 > 0) {
          const f4 dv = {tanh_d1(hv[0]), tanh_d1(hv[1]), tanh_d1(hv[2]), tanh_d1(hv[3])};
          *reinterpret_cast<f4*>(td1 + pl * TP + 16 * rb + 4 * g) = dv;
        }
        if (S::N2 > 0) {
          const f4 ev = {tanh_d2(hv[0]), tanh_d2(hv[1]), tanh_d2(hv[2]), tanh_d2(hv[3])};
          *reinterpret_cast<f4*>(td2 + pl * TP + 16 * rb + 4 * g) = ev;
        }
      }
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) gz10[rb][r] = 0.f;
      __builtin_amdgcn_sched_barrier(0);

      // ---- per axis: Z_i, Z_ii -> G_i, G_ii -> dW2, W2^T G -> layer-1 terms
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        if (!S::has1(i)) continue;
        const int ca = S::col1(i);
        const bool two = S::has2(i);
        const int cb = two ? S::col2(i) : 0;
        const float* zti = szt + i * H;
        f4 Gi[NRB], Gii[NRB];
        {
          f4 Zi[NRB], Zii[NRB];
          layer2_pass<H, L>(sp, swt, h1, zti, 1, Zi, pl, g);
          if (two) layer2_pass<H, L>(sp, swt, h1, zti, 2, Zii, pl, g);
#pragma unroll
          for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float h = h2[ob][r];
              const float d1 = tanh_d1(h);
              const float d2 = -2.f * h * d1;
              const float wlu = sp[L::WL + 16 * ob + 4 * g + r];
              const float zi = Zi[ob][r];
              const float gai = wlu * gur[ca];
              float dwl = fmaf(gur[ca], d1 * zi, vWL[ob][r]);
              float g0 = fmaf(d2 * zi, gai, G0[ob][r]);
              float gi = d1 * gai;
              if (two) {
                const float d3 = -2.f * (d1 * d1 + h * d2);
                const float zii = Zii[ob][r];
                const float gaii = wlu * gur[cb];
                dwl = fmaf(gur[cb], d1 * zii + d2 * zi * zi, dwl);
                g0 = fmaf(d2 * zii + d3 * zi * zi, gaii, g0);
                gi = fmaf(2.f * d2 * zi, gaii, gi);
                Gii[ob][r] = d1 * gaii;
              }
              vWL[ob][r] = dwl;
              G0[ob][r] = g0;
              Gi[ob][r] = gi;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // dW2 += G_i (d1 zt_i)^T [+ G_ii (d2 zt_i^2)^T]; the lane's dW2
        // columns are units 16 ib + pl
        float sc1[NRB], sc2[NRB];
#pragma unroll
        for (int ib = 0; ib < NRB; ++ib) {
          sc1[ib] = zti[16 * ib + pl];
          sc2[ib] = sc1[ib] * sc1[ib];
        }
        store_T<H>(tg, Gi, pl, g);
        wave_sync_lds();
        dw2_accumulate<H>(tg, td1, sc1, dW2, pl, g);
        wave_sync_lds();
        if (two) {
          store_T<H>(tg, Gii, pl, g);
          wave_sync_lds();
          dw2_accumulate<H>(tg, td2, sc2, dW2, pl, g);
          wave_sync_lds();
        }
        __builtin_amdgcn_sched_barrier(0);
        // GA = W2^T G, then the layer-1 transpose terms of this axis
        {
          f4 GAi[NRB], GAii[NRB];
          w2t_pass<H>(swr, Gi, GAi, pl, g);
          if (two) w2t_pass<H>(swr, Gii, GAii, pl, g);
#pragma unroll
          for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float h = h1[rb][r];
              const float z = zti[16 * rb + 4 * g + r];
              const float d1 = tanh_d1(h), d2 = tanh_d2(h);
              float gzi = d1 * GAi[rb][r];
              float g0 = fmaf(d2 * z, GAi[rb][r], gz10[rb][r]);
              if (two) {
                gzi = fmaf(2.f * d2 * z, GAii[rb][r], gzi);
                g0 = fmaf(tanh_d3(h) * z * z, GAii[rb][r], g0);
              }
              gz10[rb][r] = g0;
              v[rb][r] = gzi * isd[i];   // dW1[:, i] through z1_i = zt_i
            }
        }
        reduce_units_to_lds<NRB>(v, mysmall + (K_W1 + i) * H, pl, g);
        __builtin_amdgcn_sched_barrier(0);
      }
      // ---- column 0: dW2 += G_0 h^T, db2 += G_0, GA_0 = W2^T G_0
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[ob][r] = G0[ob][r];
      reduce_units_to_lds<NRB>(v, mysmall + K_B2 * H, pl, g);
      {
        float one[NRB];
#pragma unroll
        for (int ib = 0; ib < NRB; ++ib) one[ib] = 1.f;
        store_T<H>(tg, G0, pl, g);
        wave_sync_lds();
        dw2_accumulate<H>(tg, th1, one, dW2, pl, g);
        wave_sync_lds();
      }
      {
        f4 GA0[NRB];
        w2t_pass<H>(swr, G0, GA0, pl, g);
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) gz10[rb][r] = fmaf(tanh_d1(h1[rb][r]), GA0[rb][r], gz10[rb][r]);
      }
      // dW1[:, i] += gz1_0 xh_i (the gz1_i parts were added per axis)
#pragma unroll
      for (int i = 0; i < XD; ++i) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) v[rb][r] = gz10[rb][r] * xh[i];
        reduce_units_to_lds<NRB>(v, mysmall + (K_W1 + i) * H, pl, g);
      }
    }
    reduce_units_to_lds<NRB>(vWL, mysmall + K_WL * H, pl, g);
    reduce_units_to_lds<NRB>(gz10, mysmall + K_B1 * H, pl, g);
  }

  // ---- end of chunk: deterministic reduction into sacc, then one row write
  if (g == 0) sbl[wave * 16 + pl] = dbl;
  __syncthreads();
  for (int w = 0; w < nw; ++w) {
    if (wave == w) {
      if constexpr (NH == 2) {
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
          for (int ib = 0; ib < NRB; ++ib)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              sacc[L::W2 + (16 * ob + 4 * g + r) * H + 16 * ib + pl] += dW2[ob][ib][r];
      }
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NS * H; t += blockDim.x) {
    const int kind = t / H, u = t % H;
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += ssmall[(w * NS + kind) * H + u];
    int dst;
    if (kind == K_WL) dst = L::WL + u;
    else if (kind == K_B1) dst = L::B1 + u;
    else if (kind < K_W1 + XD) dst = L::W1 + u * XD + (kind - K_W1);
    else dst = L::B2 + u;
    sacc[dst] += s;
  }
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int q = 0; q < nw * 16; ++q) s += sbl[q];
    sacc[L::BL] = s;
  }
  __syncthreads();
  float* out = partial + (size_t)chunk * L::P;
  for (int i = threadIdx.x; i < L::P; i += blockDim.x) out[i] = sacc[i];
}

}  // namespace fbpx
