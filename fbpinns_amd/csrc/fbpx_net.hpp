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

// Hidden layer 2 pre-activation jets via MFMA: Z[c][ob] (D layout).
template <class S, int H, class L>
__device__ inline void layer2_mfma(const float* sp, const float (*A)[H / 16][4], f4 (*Z)[H / 16], int pl, int g) {
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
        const float w = sp[L::W2 + (16 * ob + pl) * H + 16 * kb + 4 * g + ks];
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

// ------------------------------------------------------------------ forward

template <int XD, int H, int NH, int D1M, int D2M>
__global__ void __launch_bounds__(256, 1)
fwd_kernel(const float* __restrict__ theta, const float* __restrict__ dom, const float* __restrict__ x,
           fbpx_worklist wl, float* __restrict__ pairbuf) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, NH>;
  constexpr int J = S::J, NRB = H / 16;
  __shared__ float sp[L::P];

  const int chunk = blockIdx.x;
  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  const float* th = theta + (size_t)model * L::P;
  for (int i = threadIdx.x; i < L::P; i += blockDim.x) sp[i] = th[i];
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], sd[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    sd[i] = (d.xmax[i] - d.xmin[i]) / 2.f;
    isd[i] = 1.f / sd[i];
  }
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
      xh[i] = (xv[i] - mu[i]) / sd[i];
    }
    float A[J][NRB][4];
    layer1<S, H, L>(sp, xh, isd, A, g);
    if constexpr (NH == 2) {
      f4 Z[J][NRB];
      layer2_mfma<S, H, L>(sp, A, Z, pl, g);
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float zc[J], ac[J];
#pragma unroll
          for (int c = 0; c < J; ++c) zc[c] = Z[c][ob][r];
          tanh_jet<S>(zc, ac);
#pragma unroll
          for (int c = 0; c < J; ++c) A[c][ob][r] = ac[c];
        }
      }
    }
    float uj[J];
#pragma unroll
    for (int c = 0; c < J; ++c) {
      float s = 0.f;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) s = fmaf(sp[L::WL + 16 * rb + 4 * g + r], A[c][rb][r], s);
      uj[c] = sum_groups(s);
    }
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
//   gu  = (d uw / d u)^T gU (jet_mul_T), gur = sd_u * gu
//   output layer: dbL += gur0, dWL[u] += sum_c gur_c a_c[u], ga_c[u] = WL[u] gur_c
//   tanh jet transpose (d3 = -2 (d1^2 + h d2)):
//     gz0  = d1 ga0 + sum_i d2 z_i ga_i + sum_i (d2 z_ii + d3 z_i^2) ga_ii
//     gz_i = d1 ga_i + 2 d2 z_i ga_ii,   gz_ii = d1 ga_ii
//   hidden layer 2: dW2 += sum_c gz_c a1_c^T (MFMA over pairs after an LDS
//     transpose), db2 += gz0, ga1_c = W2^T gz_c (MFMA, registers chain)
//   layer 1 (z_i = W1[:,i]/sd_i, z_ii = 0): dW1[u,i] += gz0 xh_i + gz_i/sd_i,
//     db1 += gz0.
// Small per-unit gradients are summed over the 16 pairs of a sub-tile with
// DPP row reductions and accumulated per wave in LDS (no atomics); waves are
// reduced in a fixed order at the end of the chunk, so results are
// deterministic.

template <int XD, int H, int NH>
struct BwdSmem {
  using L = Layout<XD, H, NH>;
  static constexpr int TP = H + 4;                        // padded transpose row
  static constexpr int NSMALL = (NH == 2 ? 3 : 2) + XD;   // per-unit small grads
  static constexpr int SP = L::P;
  static constexpr int STRANS = (NH == 2) ? 4 * 2 * 16 * TP : 1;
  static constexpr int SSMALL = 4 * NSMALL * H;           // [wave][kind][u]
  static constexpr int SBL = 4 * 16;                      // dbL per wave per pl
};

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
  __shared__ __attribute__((aligned(16))) float strans[M::STRANS];

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
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], sd[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    sd[i] = (d.xmax[i] - d.xmin[i]) / 2.f;
    isd[i] = 1.f / sd[i];
  }
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
      xh[i] = (xv[i] - mu[i]) / sd[i];
    }
    float A1[J][NRB][4];
    layer1<S, H, L>(sp, xh, isd, A1, g);

    // pre-activations of the last hidden layer (Z) and its input (A1)
    f4 Z[J][NRB];
    if constexpr (NH == 2) {
      layer2_mfma<S, H, L>(sp, A1, Z, pl, g);
    }

    // forward through the output layer
    float ur[J];
#pragma unroll
    for (int c = 0; c < J; ++c) ur[c] = 0.f;
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float ac[J];
        if constexpr (NH == 2) {
          float zc[J];
#pragma unroll
          for (int c = 0; c < J; ++c) zc[c] = Z[c][ob][r];
          tanh_jet<S>(zc, ac);
        } else {
#pragma unroll
          for (int c = 0; c < J; ++c) ac[c] = A1[c][ob][r];
        }
        const float wl_u = sp[L::WL + 16 * ob + 4 * g + r];
#pragma unroll
        for (int c = 0; c < J; ++c) ur[c] = fmaf(wl_u, ac[c], ur[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < J; ++c) ur[c] = sum_groups(ur[c]);
    float uj[J];
    uj[0] = (ur[0] + sp[L::BL]) * d.usd + d.umu;
#pragma unroll
    for (int c = 1; c < J; ++c) uj[c] = ur[c] * d.usd;
    (void)uj;

    // upstream gradient of this pair
    float gU[J], wj[J], gu[J], gur[J];
#pragma unroll
    for (int c = 0; c < J; ++c) gU[c] = valid ? gseg[(size_t)seg * J + c] : 0.f;
    window_jets<S>(xv, d, wj);
    jet_mul_T<S>(gU, wj, gu);
#pragma unroll
    for (int c = 0; c < J; ++c) gur[c] = gu[c] * d.usd;
    if (g == 0) dbl += gur[0];

    // output layer + tanh-jet transpose of the last hidden layer.
    // For NH == 2 the result gz overwrites Z; for NH == 1 it is the layer-1
    // gradient and is consumed directly below.
    float G1[J][NRB][4];  // d loss / d z1 jets (first-order columns and value)
    float vWL[NRB][4], vB2[NRB][4];
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int u = 16 * ob + 4 * g + r;
        float zc[J];
        if constexpr (NH == 2) {
#pragma unroll
          for (int c = 0; c < J; ++c) zc[c] = Z[c][ob][r];
        } else {
          // layer-1 pre-activation jets: z0 from h, z_i = W1[u,i]/sd_i, z_ii = 0
          zc[0] = 0.f;
#pragma unroll
          for (int i = 0; i < XD; ++i) {
            if (S::has1(i)) {
              zc[S::col1(i)] = sp[L::W1 + u * XD + i] * isd[i];
              if (S::has2(i)) zc[S::col2(i)] = 0.f;
            }
          }
        }
        const float h = (NH == 2) ? tanh_f(zc[0]) : A1[0][ob][r];
        const float d1 = 1.f - h * h;
        const float d2 = -2.f * h * d1;
        const float d3 = -2.f * (d1 * d1 + h * d2);
        const float wl_u = sp[L::WL + u];
        // dWL[u] += sum_c gur_c a_c
        float dwl = gur[0] * h;
        float gz0 = d1 * wl_u * gur[0];
        float gzc[J];
#pragma unroll
        for (int i = 0; i < XD; ++i) {
          if (S::has1(i)) {
            const int a = S::col1(i);
            const float zi = zc[a];
            const float gai = wl_u * gur[a];
            dwl = fmaf(gur[a], d1 * zi, dwl);
            gz0 = fmaf(d2 * zi, gai, gz0);
            float gzi = d1 * gai;
            if (S::has2(i)) {
              const int b = S::col2(i);
              const float zii = zc[b];
              const float gaii = wl_u * gur[b];
              dwl = fmaf(gur[b], d1 * zii + d2 * zi * zi, dwl);
              gz0 = fmaf(d2 * zii + d3 * zi * zi, gaii, gz0);
              gzi = fmaf(2.f * d2 * zi, gaii, gzi);
              gzc[b] = d1 * gaii;
            }
            gzc[a] = gzi;
          }
        }
        gzc[0] = gz0;
        vWL[ob][r] = dwl;
        vB2[ob][r] = gz0;
        if constexpr (NH == 2) {
#pragma unroll
          for (int c = 0; c < J; ++c) Z[c][ob][r] = gzc[c];
        } else {
#pragma unroll
          for (int c = 0; c < J; ++c) G1[c][ob][r] = gzc[c];
        }
      }
    }

    reduce_units_to_lds<NRB>(vWL, mysmall + K_WL * H, pl, g);
    if constexpr (NH == 2) {
      reduce_units_to_lds<NRB>(vB2, mysmall + K_B2 * H, pl, g);
      // dW2 += sum_c gz_c a1_c^T, contraction over the 16 pairs of the sub-tile
      float* tg = strans + (wave * 2 + 0) * 16 * TP;
      float* ta = strans + (wave * 2 + 1) * 16 * TP;
#pragma unroll
      for (int c = 0; c < J; ++c) {
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob) {
          *reinterpret_cast<f4*>(tg + pl * TP + 16 * ob + 4 * g) = Z[c][ob];
          f4 av = {A1[c][ob][0], A1[c][ob][1], A1[c][ob][2], A1[c][ob][3]};
          *reinterpret_cast<f4*>(ta + pl * TP + 16 * ob + 4 * g) = av;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          float gv[NRB], av[NRB];
#pragma unroll
          for (int ob = 0; ob < NRB; ++ob) gv[ob] = tg[(4 * ks + g) * TP + 16 * ob + pl];
#pragma unroll
          for (int ib = 0; ib < NRB; ++ib) av[ib] = ta[(4 * ks + g) * TP + 16 * ib + pl];
#pragma unroll
          for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
            for (int ib = 0; ib < NRB; ++ib) dW2[ob][ib] = mfma16(gv[ob], av[ib], dW2[ob][ib]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      // ga1_c = W2^T gz_c (A operand: W2^T[i = 16ib+pl][o = 16ob+4g+ks])
      f4 GA[J][NRB];
#pragma unroll
      for (int c = 0; c < J; ++c)
#pragma unroll
        for (int ib = 0; ib < NRB; ++ib) GA[c][ib] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob) {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
          for (int ib = 0; ib < NRB; ++ib) {
            const float w = sp[L::W2 + (16 * ob + 4 * g + ks) * H + 16 * ib + pl];
#pragma unroll
            for (int c = 0; c < J; ++c) GA[c][ib] = mfma16(w, Z[c][ob][ks], GA[c][ib]);
          }
        }
      }
      // layer-1 tanh-jet transpose
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * rb + 4 * g + r;
          const float h = A1[0][rb][r];
          const float d1 = 1.f - h * h;
          const float d2 = -2.f * h * d1;
          const float d3 = -2.f * (d1 * d1 + h * d2);
          float gz0 = d1 * GA[0][rb][r];
#pragma unroll
          for (int i = 0; i < XD; ++i) {
            if (S::has1(i)) {
              const int a = S::col1(i);
              const float zi = sp[L::W1 + u * XD + i] * isd[i];
              gz0 = fmaf(d2 * zi, GA[a][rb][r], gz0);
              float gzi = d1 * GA[a][rb][r];
              if (S::has2(i)) {
                const int b = S::col2(i);
                gz0 = fmaf(d3 * zi * zi, GA[b][rb][r], gz0);
                gzi = fmaf(2.f * d2 * zi, GA[b][rb][r], gzi);
              }
              G1[a][rb][r] = gzi;
            }
          }
          G1[0][rb][r] = gz0;
        }
      }
    }

    // layer-1 parameter gradients from G1 (d loss / d z1 jets)
    {
      float v[NRB][4];
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) v[rb][r] = G1[0][rb][r];
      reduce_units_to_lds<NRB>(v, mysmall + K_B1 * H, pl, g);
#pragma unroll
      for (int i = 0; i < XD; ++i) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float t = G1[0][rb][r] * xh[i];
            if (S::has1(i)) t = fmaf(G1[S::col1(i)][rb][r], isd[i], t);
            v[rb][r] = t;
          }
        reduce_units_to_lds<NRB>(v, mysmall + (K_W1 + i) * H, pl, g);
      }
    }
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
  // small grads: sum over waves and pair-lanes in fixed order
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
