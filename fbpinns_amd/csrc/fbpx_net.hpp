// Forward and backward jet kernels over (point, subdomain) pairs.
//
// Work decomposition: one workgroup (4 waves) per bin, a bin being a run of
// chunks (a chunk = a run of pairs of a single subdomain) processed in order;
// each wave processes 16 pairs at a time ("sub-tile").
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
//
// Layer-1 jets are never stored: with h, d1, d2 of layer 1 and
// zt_i = W1[:, i] / sd_i (per subdomain), a1_0 = h, a1_i = d1 zt_i and
// a1_ii = d2 zt_i^2, so the B operand of every layer-2 MFMA is rebuilt from h
// with one or two multiplies.  Layer 2 is evaluated one axis at a time (Z_0,
// then Z_i and Z_ii per axis), which keeps 3 x H/4 accumulators live instead
// of J x H/4.
#pragma once

#include "fbpx_common.hpp"

namespace fbpx {

// LDS copy of a subdomain's parameters except W2 (which is staged separately
// in the padded layouts below): W1 [u][i], B1, B2 (NH == 2), WL, BL.
template <int XD, int H, int NH>
struct Small {
  static constexpr int W1 = 0;
  static constexpr int B1 = H * XD;
  static constexpr int B2 = B1 + H;
  static constexpr int WL = B2 + (NH == 2 ? H : 0);
  static constexpr int BL = WL + H;
  static constexpr int N = BL + 1;
};

// The threads that run one chunk: the workgroup, or with WAVE one wave (a
// workgroup's four waves then run four chunks of their own).  A wave's LDS
// operations are performed in order, so a wavefront-scope fence orders its
// writes before its reads where the workgroup needs a barrier.
template <bool WAVE>
struct Team {
  __device__ static int tid() { return WAVE ? (int)(threadIdx.x & 63) : (int)threadIdx.x; }
  __device__ static int size() { return WAVE ? 64 : (int)blockDim.x; }
  __device__ static int wave() { return WAVE ? 0 : (int)(threadIdx.x >> 6); }   // within the team
  __device__ static int nw() { return WAVE ? 1 : (int)(blockDim.x >> 6); }
  __device__ static int index() { return WAVE ? (int)(threadIdx.x >> 6) : 0; }  // team of the workgroup
  __device__ static void sync() {
    if constexpr (WAVE)
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    else
      __syncthreads();
  }
};

// stage_small in two halves, so that a chunk's other first loads can be
// issued between them: load() issues every global load of the team's share
// (registers), store() writes them to LDS.  Teams of 64 or more threads (IT
// covers a 64-thread team; wider teams leave later slots unused).  The
// one-hidden-layer chunks use it; the two-hidden-layer ones keep the plain
// loop, whose register schedule their main loops are tuned against.
template <int XD, int H, int NH, bool WAVE = false>
struct SmallStage {
  using L = Layout<XD, H, NH>;
  using SM = Small<XD, H, NH>;
  static constexpr int IT = (SM::N + 63) / 64;
  float r[IT];
  __device__ void load(const float* th) {
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = Team<WAVE>::tid() + j * Team<WAVE>::size();
      if (i < SM::N) r[j] = th[i < SM::B2 ? i : (i < SM::WL ? L::B2 + (i - SM::B2) : L::WL + (i - SM::WL))];
    }
  }
  __device__ void store(float* ss) const {
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int i = Team<WAVE>::tid() + j * Team<WAVE>::size();
      if (i < SM::N) ss[i] = r[j];
    }
  }
};

template <int XD, int H, int NH, bool WAVE = false>
__device__ inline void stage_small(const float* th, float* ss) {
  using L = Layout<XD, H, NH>;
  using SM = Small<XD, H, NH>;
  for (int i = Team<WAVE>::tid(); i < SM::N; i += Team<WAVE>::size())
    ss[i] = th[i < SM::B2 ? i : (i < SM::WL ? L::B2 + (i - SM::B2) : L::WL + (i - SM::WL))];
}

// Lane-major LDS fragments of W2: the MFMA A operands of one K-step for all
// 64 lanes, NRB = H/16 contiguous floats per lane (one ds_read_b128 at
// H = 64).  With lane l = 16g + pl and K-step q:
//   wt[(q*64 + l)*NRB + ob] = W2[16ob + pl][16kb + 4g + ks],        q = 4kb + ks
//     (A of W2 . b: rows o = 16ob + pl, contraction units 16kb + 4g + ks)
//   wr[(q*64 + l)*NRB + ib] = c_o W2[o][16ib + pl], o = 16ob + 4g + ks, q = 4ob + ks
//     (A of (diag(c) W2)^T . g), c = WL (the backward carries G / WL, see
//     bwd_two) or 1 when wl is null.
// Consecutive lanes read consecutive 16-byte slots: conflict-free.
template <int H>
__device__ inline void stage_w2(const float* w2, const float* wl, float* wt, float* wr) {
  constexpr int NRB = H / 16;
  for (int t = threadIdx.x; t < 4 * NRB * 64; t += blockDim.x) {
    const int q = t >> 6, l = t & 63, g = l >> 4, pl = l & 15;
    const int kb = q >> 2, ks = q & 3;
    const int k = 16 * kb + 4 * g + ks;   // contraction unit (wt) / row o (wr)
    float a[NRB], c[NRB];
    const float cw = (wr && wl) ? wl[k] : 1.f;
#pragma unroll
    for (int j = 0; j < NRB; ++j) {
      if (wt) a[j] = w2[(16 * j + pl) * H + k];
      if (wr) c[j] = w2[k * H + 16 * j + pl] * cw;
    }
#pragma unroll
    for (int j = 0; j < NRB; ++j) {
      if (wt) wt[t * NRB + j] = a[j];
      if (wr) wr[t * NRB + j] = c[j];
    }
  }
}

// One K-step of lane-major fragments (see stage_w2).
template <int NRB>
struct Frag {
  float v[NRB];
};

template <int NRB>
__device__ inline Frag<NRB> ld_frag(const float* p) {
  Frag<NRB> f;
  if constexpr (NRB == 4) {
    const f4 t = *reinterpret_cast<const f4*>(p);
    f.v[0] = t[0];
    f.v[1] = t[1];
    f.v[2] = t[2];
    f.v[3] = t[3];
  } else if constexpr (NRB == 2) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 t = *reinterpret_cast<const f2*>(p);
    f.v[0] = t[0];
    f.v[1] = t[1];
  } else {
    f.v[0] = p[0];
  }
  return f;
}

// Sum over the four 16-lane rows (lane groups g) for every pl, in registers:
// v_permlane32_swap exchanges the upper half of its first operand with the
// lower half of its second, so with both holding v the two results add to
// v[l] + v[l ^ 32]; v_permlane16_swap does the same between odd and even
// rows.  Written as asm: the compiler builtins for these instructions read
// one result register twice on this toolchain.  The s_nop covers the
// VALU-write -> permlane-read hazard.
__device__ inline float sum_groups(float v) {
  float a = v, b = v;
  asm("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
  float s = a + b;
  a = s;
  b = s;
  asm("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
  return a + b;
}

template <int CTRL>
__device__ inline float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// One reduce-scatter stage over a DPP row: lanes whose stage bit is set keep
// the upper half of v[0..N), the others the lower half; each adds the half
// its partner (DPP control CTRL, a lane with the opposite bit) does not keep.
template <int N, int CTRL>
__device__ inline void split_stage(float* v, bool upper) {
#pragma unroll
  for (int k = 0; k < N / 2; ++k) {
    const float lo = v[k], hi = v[k + N / 2];
    const float send = upper ? lo : hi;
    const float keep = upper ? hi : lo;
    v[k] = keep + dpp<CTRL>(send);
  }
}

// The two mirror stages with whole DPP banks (4 lanes) on each side: lanes
// of the banks in the low mask keep lo + partner's lo, the others hi +
// partner's hi, written by two bank-masked DPP adds into one register (no
// selects).  row_mirror pairs pl with 15 - pl (banks 0,1 | 2,3),
// row_half_mirror pl with pl ^ 7 (banks 0,2 | 1,3).  One asm block per
// stage; its leading s_nop covers the VALU-write -> DPP-read hazard on the
// operands (the later pairs' operands are older still).
#define FBPX_DPP_PAIR(CTRL, LO, HI, D, A, B)                               \
  "v_add_f32_dpp %" #D ", %" #A ", %" #A " " CTRL " row_mask:0xf bank_mask:" LO "\n\t" \
  "v_add_f32_dpp %" #D ", %" #B ", %" #B " " CTRL " row_mask:0xf bank_mask:" HI "\n\t"

__device__ inline void mirror_stage(float* v) {
  float d[8];
  asm("s_nop 1\n\t"
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 0, 8, 16)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 1, 9, 17)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 2, 10, 18)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 3, 11, 19)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 4, 12, 20)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 5, 13, 21)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 6, 14, 22)
      FBPX_DPP_PAIR("row_mirror", "0x3", "0xc", 7, 15, 23)
      : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3]), "=&v"(d[4]), "=&v"(d[5]), "=&v"(d[6]), "=&v"(d[7])
      : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]), "v"(v[6]), "v"(v[7]),
        "v"(v[8]), "v"(v[9]), "v"(v[10]), "v"(v[11]), "v"(v[12]), "v"(v[13]), "v"(v[14]), "v"(v[15]));
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = d[k];
}

__device__ inline void half_mirror_stage(float* v) {
  float d[4];
  asm("s_nop 1\n\t"
      FBPX_DPP_PAIR("row_half_mirror", "0x5", "0xa", 0, 4, 8)
      FBPX_DPP_PAIR("row_half_mirror", "0x5", "0xa", 1, 5, 9)
      FBPX_DPP_PAIR("row_half_mirror", "0x5", "0xa", 2, 6, 10)
      FBPX_DPP_PAIR("row_half_mirror", "0x5", "0xa", 3, 7, 11)
      : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3])
      : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]), "v"(v[6]), "v"(v[7]));
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = d[k];
}
#undef FBPX_DPP_PAIR

// vals[rb][r] holds this lane's contribution for unit 16rb + 4g + r.  Sums
// every entry over the 16 pair-lanes of the row and adds the total for unit
// 16*(pl/4) + 4g + pl%4 into dst (lane pl owns that unit; no two lanes of a
// wave share a destination).  Butterfly reduce-scatter: partners
// row_mirror (pl ^ 15), row_half_mirror (pl ^ 7), quad_perm[3,2,1,0]
// (pl ^ 3), quad_perm[1,0,3,2] (pl ^ 1) flip bits 3, 2, 1, 0 of pl in turn
// while sharing the bits above; with fewer than 16 values the top stages
// are plain additions.  After the last stage lane pl holds the total of
// value index pl & (V - 1), which is multiplied by post before it is added.
template <int NRB>
__device__ inline void reduce_units_to_lds(const float (*vals)[4], float* dst, int pl, int g, float post = 1.f) {
  constexpr int V = NRB * 4;
  float v[V];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) v[4 * rb + r] = vals[rb][r];
  if constexpr (V <= 8) {
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] += dpp<0x140>(v[k]);
  }
  if constexpr (V <= 4) {
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] += dpp<0x141>(v[k]);
  }
  if constexpr (V == 16) mirror_stage(v);
  if constexpr (V >= 8) half_mirror_stage(v);
  split_stage<4, 0x1B>(v, (pl >> 1) & 1);
  split_stage<2, 0xB1>(v, pl & 1);
  if (pl < V) dst[16 * (pl >> 2) + 4 * g + (pl & 3)] += v[0] * post;
}

// Hide the values of a register tile from the optimiser (an empty asm that
// "may modify" each element), so that quantities derived from them (tanh
// derivatives) are recomputed where used instead of being kept alive across
// the axis phases, which would cost 16 registers per derived tile.
template <int NRB>
__device__ inline void opaque(float (*a)[4]) {
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) asm volatile("" : "+v"(a[rb][r]));
}

// v[c] for a wave-uniform runtime index c < N (selects, no register indexing).
template <int N>
__device__ inline float pick(const float* v, int c) {
  float r = v[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (c == k) ? v[k] : r;
  return r;
}

// tanh derivatives from h = tanh(z): d1 = 1 - h^2, d2 = -2 h d1, d3 = -2 (d1^2 + h d2)
__device__ inline float tanh_d1(float h) { return 1.f - h * h; }
__device__ inline float tanh_d2(float h) { return -2.f * h * (1.f - h * h); }
__device__ inline float tanh_d3(float h) {
  const float d1 = 1.f - h * h;
  return -2.f * (d1 * d1 - 2.f * h * h * d1);
}

// Per-subdomain layer-1 tangent tables in LDS (4 * XD * H floats): with
// zt_i[u] = W1[u][i] / sd_i,
//   zt[i*H + u] = zt_i[u],  zt[XD*H + i*H + u] = -2 zt_i[u],
//   zt[2*XD*H + i*H + u] = -2 zt_i[u]^2,  zt[3*XD*H + i*H + u] = zt_i[u]^2.
template <int XD, int H, class SM, bool WAVE = false>
__device__ inline void stage_zt(const float* ss, const float* isd, float* zt) {
  for (int t = Team<WAVE>::tid(); t < XD * H; t += Team<WAVE>::size()) {
    const int i = t / H, u = t % H;
    float s = isd[0];
#pragma unroll
    for (int j = 1; j < XD; ++j)
      if (i == j) s = isd[j];
    const float z = ss[SM::W1 + u * XD + i] * s;
    zt[t] = z;
    zt[XD * H + t] = -2.f * z;
    zt[2 * XD * H + t] = -2.f * z * z;
    zt[3 * XD * H + t] = z * z;
  }
}

// Per-lane LDS base pointers (lane group g = lane >> 4, pl = lane & 15).
// Every LDS read in the pair loops is one of these plus a compile-time
// offset, so an array costs one address register (ds_read immediate
// offsets) rather than one per access pattern.
//   w1  = ss + W1 + 4g*XD      W1[16rb + 4g + r][i]        at [(16rb + r)*XD + i]
//   vec = ss + 4g              B1/B2/WL[16rb + 4g + r]     at [off + 16rb + r]
//   zt  = szt + 4g             zt[i][16rb + 4g + r]        at [i*H + 16rb + r]
//   zm, zq, z2: the -2 zt, -2 zt^2 and zt^2 tables, same indexing
//   wt  = swt + lane*NRB       K-step q of the W2 fragments at [q*64*NRB]
//   wr  = swr + lane*NRB       K-step q of the transposed fragments, same
struct LaneLds {
  const float* w1;
  const float* vec;
  const float* zt;
  const float* zm;
  const float* zq;
  const float* z2;
  const float* wt;
  const float* wr;
};

template <int XD, int H, class SM>
__device__ inline LaneLds lane_lds(const float* ss, const float* szt, const float* swt, const float* swr, int pl,
                                   int g) {
  LaneLds v;
  v.w1 = ss + SM::W1 + 4 * g * XD;
  v.vec = ss + 4 * g;
  v.zt = szt + 4 * g;
  v.zm = szt + XD * H + 4 * g;
  v.zq = szt + 2 * XD * H + 4 * g;
  v.z2 = szt + 3 * XD * H + 4 * g;
  v.wt = swt + (16 * g + pl) * (H / 16);
  v.wr = swr + (16 * g + pl) * (H / 16);
  return v;
}

// Z[ob] += sum over K-steps q of the fragment A operands (LDS, lane-major,
// K-step q at frag + q*64*NRB) times B[q >> 2][q & 3] (registers).  A
// scheduling region of its own in which each fragment is read two K-steps
// ahead of the MFMAs that use it (group barriers: two reads, then per
// K-step its NRB MFMAs and the next read), so the LDS latency is hidden
// behind MFMAs instead of stalling each K-step.
template <int H, class BOp>
__device__ inline void frag_pass(const float* frag, BOp b, f4* Z) {
  constexpr int NRB = H / 16, NQ = 4 * NRB;
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const Frag<NRB> w = ld_frag<NRB>(frag + q * 64 * NRB);
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob) Z[ob] = mfma16(w.v[ob], b[q >> 2][q & 3], Z[ob]);
  }
  __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    __builtin_amdgcn_sched_group_barrier(0x008, NRB, 0);
    if (q + 2 < NQ) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Layer 1 for this lane's units (rows 16rb + 4g + r): h = tanh(W1 xh + b1).
template <int XD, int H, class SM>
__device__ inline void layer1_stream(const LaneLds& v, const float* xh, float (*h1)[4]) {
#pragma unroll
  for (int rb = 0; rb < H / 16; ++rb) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float z = v.vec[SM::B1 + 16 * rb + r];
#pragma unroll
      for (int i = 0; i < XD; ++i) z = fmaf(v.w1[(16 * rb + r) * XD + i], xh[i], z);
      h1[rb][r] = tanh_f(z);
    }
  }
}

// One layer-2 MFMA pass: Z[ob] (+)= W2 . B with B[kb][ks] = base * scale.
// kind 0: B = h (Z initialised with b2), 1: B = d1 zt_i, 2: B = d2 zt_i^2,
// 3: B = base zt_i with base = d1 supplied in place of h;
// zrow = i * H selects the zt row of axis i.
template <int H, class SM>
__device__ inline void layer2_pass(const LaneLds& v, const float (*h1)[4], int zrow, int kind, f4* Z) {
  constexpr int NRB = H / 16;
  float bv[NRB][4];
#pragma unroll
  for (int kb = 0; kb < NRB; ++kb)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const float h = h1[kb][ks];
      float b = h;
      if (kind == 1) {
        b = fmaf(-h, h, 1.f) * v.zt[zrow + 16 * kb + ks];            // d1 zt
      } else if (kind == 2) {
        b = (h * fmaf(-h, h, 1.f)) * v.zq[zrow + 16 * kb + ks];      // d2 zt^2 = h d1 (-2 zt^2)
      } else if (kind == 3) {
        b = h * v.zt[zrow + 16 * kb + ks];                           // (given d1) zt
      }
      bv[kb][ks] = b;
    }
#pragma unroll
  for (int ob = 0; ob < NRB; ++ob) {
    if (kind == 0) {
      const float* b2 = v.vec + SM::B2 + 16 * ob;
      Z[ob] = f4{b2[0], b2[1], b2[2], b2[3]};
    } else {
      Z[ob] = f4{0.f, 0.f, 0.f, 0.f};
    }
  }
  frag_pass<H>(v.wt, bv, Z);
}

// One layer-2 MFMA pass with the B operand given in registers (no bias):
// Z[ob] = W2 . b, b[kb][ks] the value for unit 16kb + 4g + ks of this lane's pair.
template <int H>
__device__ inline void layer2_pass_raw(const LaneLds& v, const float (*b)[4], f4* Z) {
#pragma unroll
  for (int ob = 0; ob < H / 16; ++ob) Z[ob] = f4{0.f, 0.f, 0.f, 0.f};
  frag_pass<H>(v.wt, b, Z);
}

// The pair's layer-1 combination of every tangent the upstream gradient
// asks for, s = sum_i (gur_i a1_i + gur_ii a1_ii) with a1_i = e1 zt_i and
// a1_ii = e2 zt_i^2 (e1 = 1 - h1^2, e2 = -2 h1 e1, e3 = e1 (4 - 6 e1)):
//   s = e1 (A - 2 h1 B),  A = sum_i gur_i zt_i,  B = sum_i gur_ii zt_i^2,
// and the factor that the W2^T vector vt multiplies in the layer-1 gradient,
//   rv = e2 A + e3 B = e1 (-2 h1 A + (4 - 6 e1) B).
template <class S, int H>
__device__ inline void pair_s(const LaneLds& v, const float (*h1)[4], const float* gur, float (*sv)[4],
                              float (*rv)[4]) {
#pragma unroll
  for (int rb = 0; rb < H / 16; ++rb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float A = 0.f, B = 0.f;
#pragma unroll
      for (int i = 0; i < S::XD; ++i) {
        if (S::has1(i)) A = fmaf(gur[S::col1(i)], v.zt[i * H + 16 * rb + r], A);
        if (S::has2(i)) B = fmaf(gur[S::col2(i)], v.z2[i * H + 16 * rb + r], B);
      }
      const float h = h1[rb][r];
      const float e1 = fmaf(-h, h, 1.f);
      sv[rb][r] = e1 * fmaf(-2.f * h, B, A);
      rv[rb][r] = e1 * fmaf(h * A, -2.f, fmaf(-6.f, e1, 4.f) * B);
    }
}

// ------------------------------------------------------------------ forward

// One chunk of the forward (a run of pairs of one subdomain).
template <int XD, int H, int NH, int D1M, int D2M, bool WAVE = false>
__device__ inline void fwd_chunk(const float* __restrict__ theta, const float* __restrict__ dom,
                                 const float* __restrict__ x, const fbpx_worklist& wl, float* __restrict__ pairbuf,
                                 int chunk) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, NH>;
  using SM = Small<XD, H, NH>;
  using T = Team<WAVE>;
  static_assert(!WAVE || NH == 1, "one wave per chunk: one-hidden-layer networks");
  constexpr int J = S::J, NRB = H / 16, NT = WAVE ? 4 : 1;
  __shared__ float ss_[NT][SM::N];
  __shared__ float szt_[NT][4 * XD * H];
  __shared__ __attribute__((aligned(16))) float swt[NH == 2 ? H * H : 1];
  __shared__ __attribute__((aligned(16))) float swr[NH == 2 && S::N1 > 0 ? H * H : 1];
  float* const ss = ss_[T::index()];
  float* const szt = szt_[T::index()];

  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  const float* th = theta + (size_t)model * L::P;
  const int lane = threadIdx.x & 63, wave = T::wave(), nw = T::nw();
  const int pl = lane & 15, g = lane >> 4;
  // the point of each sub-tile is fetched one sub-tile ahead (pair index at
  // the top of the previous sub-tile, coordinates in its middle), so the two
  // dependent global loads overlap the previous sub-tile's compute; with one
  // hidden layer the first sub-tile's pair index is loaded with the
  // parameter staging and its coordinates before the staging's LDS writes
  // and barriers
  float xn[XD];
  const int k0 = beg + wave * 16 + pl;
  if constexpr (NH == 1) {
    const int p0 = wl.pair_point[k0 < end ? k0 : beg];
    SmallStage<XD, H, NH, WAVE> st;
    st.load(th);
#pragma unroll
    for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)p0 * XD + i];
    st.store(ss);
  } else {
    stage_small<XD, H, NH, WAVE>(th, ss);
    stage_w2<H>(th + L::W2, nullptr, swt, S::N1 > 0 ? swr : nullptr);
  }
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    isd[i] = 1.f / ((d.xmax[i] - d.xmin[i]) / 2.f);
  }
  T::sync();
  stage_zt<XD, H, SM, WAVE>(ss, isd, szt);
  T::sync();

  const LaneLds v = lane_lds<XD, H, SM>(ss, szt, swt, swr, pl, g);
  if constexpr (NH == 2) {
    const int p0 = wl.pair_point[k0 < end ? k0 : beg];
#pragma unroll
    for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)p0 * XD + i];
  }
  for (int base = beg + wave * 16; base < end; base += nw * 16) {
    const int k = base + pl;
    const bool valid = k < end;
    const int slot = wl.pair_slot[valid ? k : beg];
    const int kn = k + nw * 16;
    const int pn = wl.pair_point[kn < end ? kn : beg];
    float xv[XD], xh[XD];
#pragma unroll
    for (int i = 0; i < XD; ++i) {
      xv[i] = xn[i];
      xh[i] = (xv[i] - mu[i]) * isd[i];
    }
    float h1[NRB][4];
    layer1_stream<XD, H, SM>(v, xh, h1);
    float ur[J];
#pragma unroll
    for (int c = 0; c < J; ++c) ur[c] = 0.f;
    if constexpr (NH == 1) {
#pragma unroll
      for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)pn * XD + i];
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float wlu = v.vec[SM::WL + 16 * rb + r];
          const float h = h1[rb][r];
          ur[0] = fmaf(wlu, h, ur[0]);
#pragma unroll
          for (int i = 0; i < XD; ++i) {
            const float z = v.zt[i * H + 16 * rb + r];
            if (S::has1(i)) ur[S::col1(i)] = fmaf(wlu, tanh_d1(h) * z, ur[S::col1(i)]);
            if (S::has2(i)) ur[S::col2(i)] = fmaf(wlu, tanh_d2(h) * z * z, ur[S::col2(i)]);
          }
        }
    } else {
      f4 Z0[NRB];
      layer2_pass<H, SM>(v, h1, 0, 0, Z0);
      // output-layer weights times the layer-2 tanh derivatives, per element:
      // wd1 = WL f1 and wd2 = WL f1 h2 = WL f2 / (-2), so that u_i = sum wd1 Z_i
      // and the Z_i^2 part of u_ii = sum WL (f1 Z_ii + f2 Z_i^2) costs two FMAs;
      // every second-derivative column is accumulated without its factor -2,
      // which is applied once to the pair's totals
      float wd1[NRB][4], wd2[NRB][4];
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float wlu = v.vec[SM::WL + 16 * ob + r];
          const float h = tanh_f(Z0[ob][r]);
          ur[0] = fmaf(wlu, h, ur[0]);
          wd1[ob][r] = wlu * fmaf(-h, h, 1.f);
          wd2[ob][r] = wd1[ob][r] * h;
        }
#pragma unroll
      for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)pn * XD + i];
      __builtin_amdgcn_sched_barrier(0);
      // sum_o WL f1 Z_c = sum_u vt[u] a1_c[u] with vt = W2^T (WL f1): one
      // W2^T pass per sub-tile replaces the Z_ii passes of every axis and the
      // Z_i passes of axes without a second derivative.  With e1 = 1 - h1^2,
      // a1_i = e1 zt_i and a1_ii = -2 h1 e1 zt_i^2, so the per-axis dot
      // products take ve = vt e1 (first order) and w = vt e1 h1 (second order).
      float e1[NRB][4], ve[NRB][4], w[NRB][4];
      if constexpr (S::N1 > 0) {
        f4 vt[NRB];
#pragma unroll
        for (int ib = 0; ib < NRB; ++ib) vt[ib] = f4{0.f, 0.f, 0.f, 0.f};
        frag_pass<H>(v.wr, wd1, vt);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float h = h1[rb][r];
            e1[rb][r] = fmaf(-h, h, 1.f);
            ve[rb][r] = vt[rb][r] * e1[rb][r];
            w[rb][r] = ve[rb][r] * h;
          }
      }
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        if (!S::has1(i)) continue;
        if (!S::has2(i)) {
          // u_i = sum_u vt[u] e1[u] zt_i[u]
#pragma unroll
          for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) ur[S::col1(i)] = fmaf(ve[rb][r], v.zt[i * H + 16 * rb + r], ur[S::col1(i)]);
          continue;
        }
        f4 Zi[NRB];
        layer2_pass<H, SM>(v, e1, i * H, 3, Zi);
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float zi = Zi[ob][r];
            ur[S::col1(i)] = fmaf(wd1[ob][r], zi, ur[S::col1(i)]);
            ur[S::col2(i)] = fmaf(wd2[ob][r] * zi, zi, ur[S::col2(i)]);
          }
        // + sum_u vt[u] h1 e1 zt_i^2 (the a1_ii column, factor -2 deferred)
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) ur[S::col2(i)] = fmaf(w[rb][r], v.z2[i * H + 16 * rb + r], ur[S::col2(i)]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    float uj[J];
#pragma unroll
    for (int c = 0; c < J; ++c) uj[c] = sum_groups(ur[c]);
    uj[0] = (uj[0] + ss[SM::BL]) * d.usd + d.umu;
#pragma unroll
    for (int c = 1; c < J; ++c) uj[c] *= (NH == 2 && c > S::N1) ? -2.f * d.usd : d.usd;
    float wj[J], uw[J];
    window_jets<S>(xv, d, wj);
    jet_mul<S>(uj, wj, uw);
    if (valid) {
#pragma unroll
      for (int c = 0; c < J; ++c)
        if ((c & 3) == g) pairbuf[(size_t)slot * J + c] = uw[c];
    }
  }
}

// Persistent forward launch: workgroup b runs the chunks of forward bin b
// in order.
// WAVE: wave w of workgroup b runs forward bin 4b + w (one-hidden-layer
// networks; a wave past the last bin has nothing to do).
template <int XD, int H, int NH, int D1M, int D2M, bool WAVE = false>
__global__ void __launch_bounds__(256, NH == 2 ? 3 : 2)
fwd_kernel(const float* __restrict__ theta, const float* __restrict__ dom, const float* __restrict__ x,
           fbpx_worklist wl, float* __restrict__ pairbuf) {
  const int bin = WAVE ? (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6) : (int)blockIdx.x;
  if (WAVE && bin >= wl.n_fbins) return;
  const int c0 = wl.fbin_chunk_start[bin], c1 = wl.fbin_chunk_start[bin + 1];
  for (int chunk = c0; chunk < c1; ++chunk) {
    if (chunk > c0) Team<WAVE>::sync();   // the team is done with the previous chunk's LDS
    fwd_chunk<XD, H, NH, D1M, D2M, WAVE>(theta, dom, x, wl, pairbuf, chunk);
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
//   dW2 += sum_c G_c a1_c^T (MFMA, contraction over pairs),
//   db2 += G_0, GA_c = W2^T G_c (MFMA; registers chain),
//   layer 1 (z1_i = zt_i, z1_ii = 0):
//     gz1_0 = d1 GA_0 + sum_i d2 zt_i GA_i + sum_i d3 zt_i^2 GA_ii
//     gz1_i = d1 GA_i + 2 d2 zt_i GA_ii
//     dW1[u, k] += gz1_0 xh_k + gz1_k / sd_k,  db1 += gz1_0.
// The two-layer kernel does not form G_i and G_ii column by column: every
// linear use of the layer-1 tangents collapses into s = sum_i (gur_i a1_i +
// gur_ii a1_ii) (pair_s), so with S = W2 s and vt = W2^T (WL f1)
//   G_0 gains f2 S, the G_i a1_i^T and G_ii a1_ii^T columns of dW2 become
//   (WL f1) s^T plus 2 gur_ii WL f2 Z_i a1_i^T per second-derivative axis,
//   and GA_i = gur_i vt (+ W2^T of that axis's column), GA_ii = gur_ii vt.
// The two-layer kernel carries G~ = G / WL (row-wise; no division is
// performed: G~ is formed without the WL factor) and restores WL where G is
// consumed: W2^T G = (diag(WL) W2)^T G~ through a row-scaled LDS copy of W2,
// dW2 = diag(WL) sum_c G~_c a1_c^T and db2 = WL sum G~_0 at the end.
// Small per-unit gradients are summed over the 16 pairs of a sub-tile with
// DPP row reductions and accumulated per wave in LDS; at the end of the chunk
// everything is reduced in a fixed order (deterministic, no atomics).

// Upstream gradient of one pair w.r.t. its network output jets:
// gur = sd_u * (d uw / d u)^T gU, gU = gseg[seg] (zero for padding lanes).
template <class S>
__device__ inline void pair_upstream(const float* gU, const float* xv, const Dom& d, float* gur) {
  constexpr int J = S::J;
  float wj[J];
  window_jets<S>(xv, d, wj);
  jet_mul_T<S>(gU, wj, gur);
#pragma unroll
  for (int c = 0; c < J; ++c) gur[c] *= d.usd;
}

// Chunk epilogue shared by both network depths: sums the per-wave small
// gradients (ssmall[wave][kind][u]; kinds WL, B1, W1[:, i], B2) and the
// per-lane output-bias partials (sbl) into sacc, then writes the row.
template <int XD, int H, int NH, bool WAVE = false, bool ADAM = false>
__device__ inline void finish_row(float* sacc, const float* ssmall, const float* sbl, int nw, float* out,
                                  float gs, const AdamFold& af, int model, const float* bc) {
  using L = Layout<XD, H, NH>;
  using T = Team<WAVE>;
  constexpr int NS = (NH == 2 ? 3 : 2) + XD;
  constexpr int K_WL = 0, K_B1 = 1, K_W1 = 2;
  for (int t = T::tid(); t < NS * H; t += T::size()) {
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
  if (T::tid() == 0) {
    float s = 0.f;
    for (int q = 0; q < nw * 16; ++q) s += sbl[q];
    sacc[L::BL] = s;
  }
  T::sync();
  // gs: d L / d loss for a gseg at unit upstream gradient (the row is linear
  // in gseg; 1 leaves it unchanged)
  if constexpr (!ADAM) {
    for (int i = T::tid(); i < L::P; i += T::size()) out[i] = sacc[i] * gs;
    return;
  }
  // the row is its subdomain's whole gradient: Adam on it in place, as
  // fbpx_adam_rows_partial updates a row from its one chunk row
  const size_t r0 = (size_t)model * L::P;
  for (int i = T::tid(); i < L::P; i += T::size()) {
    const float row = sacc[i] * gs;
    float gi = 0.f;
    gi += row;
    float p = af.theta[r0 + i], m = af.mu[r0 + i], v = af.nu[r0 + i];
    adam_elem(p, m, v, gi, bc[0], bc[1], af.k);
    af.theta[r0 + i] = p;
    af.mu[r0 + i] = m;
    af.nu[r0 + i] = v;
  }
}

// [xd, h, 1]: no hidden-to-hidden layer; waves work independently.
// WAVE: the chunk is one wave's (its own LDS copies, its slices of the
// per-wave sums).
// ADAM: bc holds the bias corrections, formed by the team's first lane in
// its first chunk (first) while the chunk's domain loads are in flight.
template <int XD, int H, int D1M, int D2M, bool WAVE = false, bool ADAM = false>
__device__ inline void bwd_one(const float* theta, const float* dom, const float* x, const fbpx_worklist& wl,
                               const float* gseg, float* partial, int chunk, float gs, const AdamFold& af,
                               float* bc, bool first = false) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, 1>;
  using SM = Small<XD, H, 1>;
  using T = Team<WAVE>;
  constexpr int J = S::J, NRB = H / 16, NS = 2 + XD, NT = WAVE ? 4 : 1;
  constexpr int K_WL = 0, K_B1 = 1, K_W1 = 2;
  __shared__ float ss_[NT][SM::N];
  __shared__ float sacc_[NT][L::P];
  __shared__ float ssmall_[4 * NS * H];   // per wave (four teams of one wave with WAVE)
  __shared__ float sbl_[64];
  __shared__ float szt_[NT][4 * XD * H];
  float* const ss = ss_[T::index()];
  float* const sacc = sacc_[T::index()];
  float* const szt = szt_[T::index()];
  float* const ssmall = ssmall_ + T::index() * NS * H;
  float* const sbl = sbl_ + T::index() * 16;

  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  int32_t cnt = 0;
  if constexpr (ADAM) cnt = af.count[0];   // issued with the chunk's first loads
  const float* th = theta + (size_t)model * L::P;
  const int lane = threadIdx.x & 63, wave = T::wave(), nw = T::nw();
  const int pl = lane & 15, g = lane >> 4;
  // pair point and segment of this lane's pair in the next sub-tile: loaded
  // one sub-tile ahead, the first with the parameter staging
  int pn, sn;
  {
    const int k0 = beg + wave * 16 + pl;
    const int kk = k0 < end ? k0 : beg;
    pn = wl.pair_point[kk];
    sn = wl.pair_seg[kk];
  }
  stage_small<XD, H, 1, WAVE>(th, ss);
  for (int i = T::tid(); i < L::P; i += T::size()) sacc[i] = 0.f;
  for (int i = T::tid(); i < T::nw() * NS * H; i += T::size()) ssmall[i] = 0.f;
  if (T::tid() < T::nw() * 16) sbl[T::tid()] = 0.f;
  Dom d;
  load_dom(dom, model, d);
  if constexpr (ADAM) {
    if (first && T::tid() == 0) adam_bias_n(cnt, af.k, bc);
  }
  float mu[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    isd[i] = 1.f / ((d.xmax[i] - d.xmin[i]) / 2.f);
  }
  T::sync();
  stage_zt<XD, H, SM, WAVE>(ss, isd, szt);
  T::sync();

  const LaneLds v = lane_lds<XD, H, SM>(ss, szt, szt, szt, pl, g);
  float* mysmall = ssmall + wave * NS * H;
  float dbl = 0.f;
  for (int base = beg + wave * 16; base < end; base += nw * 16) {
    // one wave per chunk: the weights and window tangents are read from LDS
    // per sub-tile rather than held in registers across the loop (95 VGPRs
    // instead of 120 at [1, 32, 1]: five waves per SIMD instead of four)
    if constexpr (WAVE) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int k = base + pl;
    const bool valid = k < end;   // else the pair at beg with a zero upstream gradient
    const int p = pn, seg = sn;
    {
      const int kn = k + nw * 16;
      const int kk = kn < end ? kn : beg;
      pn = wl.pair_point[kk];
      sn = wl.pair_seg[kk];
    }
    float xv[XD], xh[XD];
#pragma unroll
    for (int i = 0; i < XD; ++i) {
      xv[i] = x[(size_t)p * XD + i];
      xh[i] = (xv[i] - mu[i]) * isd[i];
    }
    float gur[J], gU[J];
#pragma unroll
    for (int c = 0; c < J; ++c) gU[c] = valid ? gseg[(size_t)seg * J + c] : 0.f;
    pair_upstream<S>(gU, xv, d, gur);
    if (g == 0) dbl += gur[0];
    float h1[NRB][4];
    layer1_stream<XD, H, SM>(v, xh, h1);
    float vWL[NRB][4], gz10[NRB][4], vv[NRB][4];
    // ga_c = WL gur_c acts on layer 1 directly
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float wlu = v.vec[SM::WL + 16 * rb + r];
        const float h = h1[rb][r];
        const float d1 = tanh_d1(h), d2 = tanh_d2(h), d3 = tanh_d3(h);
        float dwl = gur[0] * h;
        float gz0 = d1 * wlu * gur[0];
#pragma unroll
        for (int i = 0; i < XD; ++i) {
          if (!S::has1(i)) continue;
          const float z = v.zt[i * H + 16 * rb + r];
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
            const float h = h1[rb][r];
            const float z = v.zt[i * H + 16 * rb + r];
            const float wlu = v.vec[SM::WL + 16 * rb + r];
            float gzi = tanh_d1(h) * wlu * gur[S::col1(i)];
            if (S::has2(i)) gzi = fmaf(2.f * tanh_d2(h) * z, wlu * gur[S::col2(i)], gzi);
            t = fmaf(gzi, isd[i], t);
          }
          vv[rb][r] = t;
        }
      reduce_units_to_lds<NRB>(vv, mysmall + (K_W1 + i) * H, pl, g);
    }
    reduce_units_to_lds<NRB>(vWL, mysmall + K_WL * H, pl, g);
    reduce_units_to_lds<NRB>(gz10, mysmall + K_B1 * H, pl, g);
  }
  if (g == 0) sbl[wave * 16 + pl] = dbl;
  T::sync();
  finish_row<XD, H, 1, WAVE, ADAM>(sacc, ssmall, sbl, nw, partial + (size_t)chunk * L::P, gs, af, model, bc);
}

// W2^T G for one sub-tile: GA[ib] (A operand W2^T[i = 16ib+pl][o = 16ob+4g+ks]
// from the padded row-major copy wr; B operand = G in D layout).
template <int H>
__device__ inline void w2t_pass(const LaneLds& v, const f4* G, f4* GA) {
#pragma unroll
  for (int ib = 0; ib < H / 16; ++ib) GA[ib] = f4{0.f, 0.f, 0.f, 0.f};
  frag_pass<H>(v.wr, G, GA);
}

// Workgroup-shared dW2 accumulation.  The 4 waves of a workgroup process
// sub-tiles in lockstep (64 pairs per tile) and write two transposes to LDS:
//   stb[rb*BS + w*WS + (p >> 2)*PG + (p & 3)*16 + s] = layer-1 h of unit 16rb+s,
//                                                      pair p of wave w
//   stg[ob*BS + ...same...]                          = G_c of unit 16ob+s
// For h <= 32 each group of 4 pairs is padded from 64 to PG = 68 floats, so
// the 16 lanes of a 128-bit store phase (pairs 0..15, 16 bytes each) start
// in 16 distinct 4-bank groups instead of sharing 4 of them; a read (one pair
// group, 64 lanes) stays 64 consecutive floats.  At h = 64 the unpadded
// offsets (multiples of 64 floats) let the reads pair up as ds_read2st64,
// and the registers that saves are worth more than the store conflicts.
// dW2 blocks (16 x 16, rows o, columns u): H = 64: wave w owns column block
// w (all 4 row blocks) and contracts it over the pairs of all four waves, so
// each accumulator is held once per workgroup instead of once per wave (the
// whole 64 x 64 block set would not fit a wave's registers); the waves run
// their tiles in lockstep, with workgroup barriers around each column.
// H <= 32 (PRIV): every wave holds all NRB^2 blocks (16 registers at H = 32)
// over its own pairs, reading back only what it wrote itself, so the tile
// loop needs no workgroup barrier: LDS operations of one wavefront are
// performed in order, and a wavefront-scope fence keeps the compiler from
// moving the reads above the writes (or the next writes above the reads).
template <int H>
struct DW2Own {
  static constexpr int NRB = H / 16;
  static constexpr bool PRIV = H <= 32;
  static constexpr int NOB = PRIV ? NRB * NRB : 4;  // blocks held by a wave
  static constexpr int NSRC = PRIV ? 1 : 4;        // waves whose pairs it contracts over
  static constexpr int PG = H == 64 ? 64 : 68;      // floats per group of 4 pairs
  static constexpr int WS = 4 * PG;                 // per wave (16 pairs)
  static constexpr int BS = 4 * WS;                 // per 16-unit block (4 waves)
  static constexpr int TSZ = NRB * BS;              // floats per transpose buffer
  // float offset of pair 4 gp + pp, units 4 cu .. 4 cu + 3, in a wave's WS
  // floats.  SWZ: the 16-byte chunk c = 4 pp + cu is stored at
  // c ^ (gp ^ (pp & 2)) within its 64-float group, so that the 128-bit
  // stores of the 16 pairs do not collide in the banks (the unpadded h = 64
  // groups alias); the reads of a pair group stay 64 distinct floats
  template <bool SWZ>
  __device__ static int chunk_off(int gp, int pp, int cu) {
    const int c = 4 * pp + cu;
    return gp * PG + 4 * (SWZ ? c ^ (gp ^ (pp & 2)) : c);
  }
  __device__ static int col_block(int wave, int j) { return PRIV ? j % NRB : wave; }
  __device__ static int row_block(int wave, int j) { return PRIV ? j / NRB : j; }
};

// Orders a wave's LDS transposes around the dW2 reads: a workgroup barrier
// when the waves share the accumulators, else a wavefront-scope fence.
template <bool PRIV>
__device__ inline void tile_sync() {
  if constexpr (PRIV)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  else
    __syncthreads();
}

// dW2 += G_c a1_c^T over this wave's own 16 pairs (PRIV), all NRB^2 blocks.
// stgp, stbp: this wave's G and h transposes at its lane offset (+ 16 g +
// pl); pair group ks at + ks * PG, 16-unit block b at + b * BS.  a1_c of unit
// u = 16 ib + pl is rebuilt from h as in dw2_shared, with sc[16 ib] for the
// column block ib (KIND 1, 2).
template <int H, int KIND>
__device__ inline void dw2_private(const float* stgp, const float* stbp, const float* sc, f4* dW2) {
  using O = DW2Own<H>;
  constexpr int NRB = H / 16;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    float a[NRB], b[NRB];
#pragma unroll
    for (int ib = 0; ib < NRB; ++ib) {
      const float h = stbp[ib * O::BS + ks * O::PG];
      float v = h;
      if (KIND == 1) v = fmaf(-h, h, 1.f) * sc[16 * ib];
      if (KIND == 2) v = h * fmaf(-h, h, 1.f) * sc[16 * ib];
      b[ib] = v;
    }
#pragma unroll
    for (int ob = 0; ob < NRB; ++ob) a[ob] = stgp[ob * O::BS + ks * O::PG];
#pragma unroll
    for (int j = 0; j < O::NOB; ++j) dW2[j] = mfma16(a[j / NRB], b[j % NRB], dW2[j]);
  }
}

// dW2[j] += G_c a1_c^T over the tile.  a1_c of unit u = 16 ibw + pl is
// rebuilt from h: KIND 0: h, 1: (1 - h^2) sc, 2: h (1 - h^2) sc with sc
// supplied by the caller (zt_i and -2 zt_i^2 for unit u).
// stgl[ks] = stg + (NRB == 2 ? row block (wave >> 1) * BS : 0) +
// chunk_off<SWZ>(ks, g, pl >> 2) + (pl & 3) is this lane's G read base for
// pair 4 ks + g, unit pl (source wave w at + w * WS, row block ob at
// + ob * BS); its h read base is stbl + ks * PG with stbl = stb + ibw * BS +
// 16 g + pl (plain layout) or stgl[ks] + db (swizzled: both buffers share
// the layout, db = stb - stg + the row-block offsets).
template <int H, int KIND, bool SWZ>
__device__ inline void dw2_shared(const float* const* stgl, const float* stbl, int db, float sc, f4* dW2,
                                  int wave) {
  using O = DW2Own<H>;
  // swizzled buffers: the h buffer's read base for pair group ks is
  // stgl[ks] + db; db is kept opaque so that the four sums are formed at
  // their reads instead of held in four more registers through the tile
  if (SWZ) asm volatile("" : "+v"(db));
#pragma unroll
  for (int s = 0; s < O::NSRC; ++s) {
    const int w = O::NSRC == 1 ? wave : s;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const float h = SWZ ? stgl[ks][db + w * O::WS] : stbl[w * O::WS + ks * O::PG];
      float b = h;
      if (KIND == 1) b = fmaf(-h, h, 1.f) * sc;
      if (KIND == 2) b = h * fmaf(-h, h, 1.f) * sc;
#pragma unroll
      for (int j = 0; j < O::NOB; ++j) {
        const float a = stgl[ks][(O::NOB == 1 ? 0 : j * O::BS) + w * O::WS];
        dW2[j] = mfma16(a, b, dW2[j]);
      }
    }
  }
}

// Write this wave's D-layout tile (units x 16 pairs) into a transpose buffer;
// tl = t + wave * WS + (pl >> 2) * PG + (pl & 3) * 16 + 4 g.
template <int H>
__device__ inline void store_tile(float* tl, const f4* v) {
#pragma unroll
  for (int ob = 0; ob < H / 16; ++ob) *reinterpret_cast<f4*>(tl + ob * DW2Own<H>::BS) = v[ob];
}

// [xd, h, h, 1]: hidden-to-hidden layer on MFMA, streamed one axis at a time.
template <int XD, int H, int D1M, int D2M>
__device__ inline void bwd_two(const float* theta, const float* dom, const float* x, const fbpx_worklist& wl,
                               const float* gseg, float* partial, int chunk, float gs, const AdamFold& af,
                               const float* bc) {
  using S = Schema<XD, D1M, D2M>;
  using L = Layout<XD, H, 2>;
  using SM = Small<XD, H, 2>;
  using O = DW2Own<H>;
  constexpr int J = S::J, NRB = H / 16, NS = 3 + XD;
  constexpr int K_WL = 0, K_B1 = 1, K_W1 = 2, K_B2 = 2 + XD;
  // H <= 32 and XD <= 2 (wave-private dW2, registers to spare): the small
  // per-unit gradients (dW1, dWL, db1, db2) are summed per lane over the
  // chunk's tiles and reduced over the pair lanes once per chunk instead of
  // once per tile (with three axes the unrolled axis loop would spill)
  constexpr bool ACC = DW2Own<H>::PRIV && XD <= 2;
  constexpr int AXIS_UNROLL = H == 64 || ACC ? XD : 1;
  static_assert(H >= 16 && H <= 64, "hidden width 16, 32 or 64");
  static_assert(2 * O::TSZ >= L::P, "the gradient row reuses the transpose buffers");
  __shared__ float ss[SM::N];
  __shared__ float szt[4 * XD * H];
  __shared__ __attribute__((aligned(16))) float swt[H * H];
  __shared__ __attribute__((aligned(16))) float swr[H * H];
  __shared__ __attribute__((aligned(16))) float strans[2 * O::TSZ];
  __shared__ float ssmall[4 * NS * H];
  __shared__ float sbl[64];
  float* stb = strans;
  float* stg = strans + O::TSZ;

  const int model = wl.chunk_model[chunk];
  const int beg = wl.chunk_start[chunk];
  const int end = wl.chunk_start[chunk + 1];
  const float* th = theta + (size_t)model * L::P;
  stage_small<XD, H, 2>(th, ss);
  stage_w2<H>(th + L::W2, th + L::WL, swt, swr);
  for (int i = threadIdx.x; i < 4 * NS * H; i += blockDim.x) ssmall[i] = 0.f;
  if (threadIdx.x < 64) sbl[threadIdx.x] = 0.f;
  Dom d;
  load_dom(dom, model, d);
  float mu[XD], isd[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    mu[i] = (d.xmax[i] + d.xmin[i]) / 2.f;
    isd[i] = 1.f / ((d.xmax[i] - d.xmin[i]) / 2.f);
  }
  __syncthreads();
  stage_zt<XD, H, SM>(ss, isd, szt);
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int pl = lane & 15, g = lane >> 4;
  const int ucol = 16 * O::col_block(wave, 0) + pl;   // this lane's dW2 column unit (shared blocks)
  const float wl_red = ss[SM::WL + 16 * (pl >> 2) + 4 * g + (pl & 3)];   // WL of the unit a reduction leaves here
  const LaneLds v = lane_lds<XD, H, SM>(ss, szt, swt, swr, pl, g);
  // both transpose buffers are swizzled at h = 64 where the four extra read
  // bases fit the register budget (every first-derivative axis also carries
  // its second derivative; the h buffer's reads add an offset to them)
  constexpr bool SWZ = H == 64 && D1M == D2M;
  float* const stbw = stb + wave * O::WS + O::template chunk_off<SWZ>(pl >> 2, pl & 3, g);   // store bases
  float* const stgw = stg + wave * O::WS + O::template chunk_off<SWZ>(pl >> 2, pl & 3, g);
  // dW2 read bases: stgl[ks] for the G buffer; for the h buffer stbl (plain
  // layout) or stgl[ks] + stb_off (swizzled)
  const float* const stbl = stb + O::col_block(wave, 0) * O::BS + 16 * g + pl;
  const int stb_off = (int)(stb - stg) + O::col_block(wave, 0) * O::BS;
  const float* stgl[4];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) stgl[ks] = stg + O::template chunk_off<SWZ>(ks, g, pl >> 2) + (pl & 3);
  // PRIV: this wave's own transposes at its lane offset
  const float* const stgp = stg + wave * O::WS + 16 * g + pl;
  const float* const stbp = stb + wave * O::WS + 16 * g + pl;
  float* mysmall = ssmall + wave * NS * H;
  float aW1[ACC ? XD : 1][NRB][4], aWL[NRB][4], aB1[NRB][4], aB2[NRB][4];
  if constexpr (ACC) {
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int i = 0; i < XD; ++i) aW1[i][rb][r] = 0.f;
        aWL[rb][r] = aB1[rb][r] = aB2[rb][r] = 0.f;
      }
  }
  f4 dW2[O::NOB];
#pragma unroll
  for (int j = 0; j < O::NOB; ++j) dW2[j] = f4{0.f, 0.f, 0.f, 0.f};
  float dbl = 0.f;

  // with shared dW2 blocks every wave runs the same number of tiles (barriers
  // below); with wave-private blocks (H <= 32) a wave runs only the tiles in
  // which its 16-pair sub-tile holds a pair of the chunk.  Lanes past the
  // chunk end carry zero upstream gradient and so contribute nothing.
  const int ntile = O::PRIV ? (end - beg - 16 * wave + 63) >> 6 : (end - beg + 63) >> 6;
  // each tile's point and upstream jets are fetched during the previous tile
  // (indices at the start of its column-0 phase, data before its last dW2
  // column), so the two dependent global loads do not stall the tile start
  float xn[XD], gUn[J];
  {
    const int k = beg + wave * 16 + pl;
    const int kk = k < end ? k : beg;
    const int p = wl.pair_point[kk];
    const int seg = wl.pair_seg[kk];
#pragma unroll
    for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)p * XD + i];
#pragma unroll
    for (int c = 0; c < J; ++c) gUn[c] = k < end ? gseg[(size_t)seg * J + c] : 0.f;
  }
  for (int t = 0; t < ntile; ++t) {
    float xv[XD], xh[XD];
#pragma unroll
    for (int i = 0; i < XD; ++i) {
      xv[i] = xn[i];
      xh[i] = (xv[i] - mu[i]) * isd[i];
    }
    float gur[J];
    pair_upstream<S>(gUn, xv, d, gur);
    // materialise the pair's values here, so that the prefetched data dies
    // at the tile start instead of being carried into the axis phases
#pragma unroll
    for (int c = 0; c < J; ++c) asm volatile("" : "+v"(gur[c]));
#pragma unroll
    for (int i = 0; i < XD; ++i) asm volatile("" : "+v"(xh[i]));
    if (g == 0) dbl += gur[0];

    float h1[NRB][4];
    layer1_stream<XD, H, SM>(v, xh, h1);
    float vWL[NRB][4], gz10[NRB][4], vv[NRB][4];

    tile_sync<O::PRIV>();   // the previous tile's transposes have been consumed

    // ---- Z_0 = W2 h1 + b2 -> h2
    float h2[NRB][4];
    {
      f4 Z0[NRB];
      layer2_pass<H, SM>(v, h1, 0, 0, Z0);
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) h2[ob][r] = tanh_f(Z0[ob][r]);
    }
    __builtin_amdgcn_sched_barrier(0);

    // ---- S = W2 s,  G_S = f1,  vt = W2~^T G_S (d L / d s),  dW2 += G_S s^T
    //      (column S first, so that s is dead before the axes).  With
    //      Q = sum_i gur_ii Z_i^2 (accumulated per axis):
    //        G_0 = f1 gur_0 + f2 S + f3 Q,  dWL = gur_0 h2 + f1 S + f2 Q
    //      (f1 = 1 - h2^2, f2 = -2 h2 f1, f3 = f1 (4 - 6 f1)).  The layer-1
    //      gradient starts as gz1_0 = vt rv (see pair_s); the axes add
    //      e2 zt_i GA_i and column 0 adds e1 GA_0.
    f4 G0[NRB], vt[NRB];
    float Sx[NRB][4], Q[NRB][4];
    if constexpr (S::N1 == 0) {
      // value only: no tangents, s = 0
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = h2[ob][r];
          G0[ob][r] = fmaf(-h, h, 1.f) * gur[0];
          vWL[ob][r] = gur[0] * h;
          gz10[ob][r] = 0.f;
        }
    } else {
      float sv[NRB][4], rv[NRB][4];
      pair_s<S, H>(v, h1, gur, sv, rv);
      {
        f4 Sv[NRB];
        layer2_pass_raw<H>(v, sv, Sv);
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            Sx[ob][r] = Sv[ob][r];
            Q[ob][r] = 0.f;
          }
      }
      f4 gs[NRB];
      {
        f4 sw[NRB];
#pragma unroll
        for (int ob = 0; ob < NRB; ++ob) {
          sw[ob] = f4{sv[ob][0], sv[ob][1], sv[ob][2], sv[ob][3]};
#pragma unroll
          for (int r = 0; r < 4; ++r) gs[ob][r] = fmaf(-h2[ob][r], h2[ob][r], 1.f);
        }
        store_tile<H>(stbw, sw);
        store_tile<H>(stgw, gs);
      }
      w2t_pass<H>(v, gs, vt);
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) gz10[rb][r] = vt[rb][r] * rv[rb][r];
      tile_sync<O::PRIV>();
      if constexpr (O::PRIV)
        dw2_private<H, 0>(stgp, stbp, nullptr, dW2);
      else
        dw2_shared<H, 0, SWZ>(stgl, stbl, stb_off, 1.f, dW2, wave);
      tile_sync<O::PRIV>();
    }
    // layer-1 h for the B operands of the remaining dW2 columns (published
    // with the first column's G; every column has a barrier before its reads)
    {
      f4 hv[NRB];
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) hv[rb] = f4{h1[rb][0], h1[rb][1], h1[rb][2], h1[rb][3]};
      store_tile<H>(stbw, hv);
    }
    // from here on layer 1 enters only through e1 = 1 - h1^2 and h1 e1, and
    // the axes take layer 2 through p2 = h2 - h2^3 = -f2 / 2
    float e1[NRB][4], he[NRB][4], p2[NRB][4];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        e1[rb][r] = fmaf(-h1[rb][r], h1[rb][r], 1.f);
        he[rb][r] = h1[rb][r] * e1[rb][r];
        p2[rb][r] = fmaf(-h2[rb][r] * h2[rb][r], h2[rb][r], h2[rb][r]);
      }
    __builtin_amdgcn_sched_barrier(0);

    // ---- per axis.  d L / d a1_i = gur_i vt (+ W2~^T G_i when the axis has a
    //      second derivative), d L / d a1_ii = gur_ii vt;  G_i = 2 gur_ii f2 Z_i
    //        gz1_i = e1 (gur_i vt + GA_i) + 2 gur_ii e2 zt_i vt,
    //        gz1_0 += e2 zt_i GA_i.
    //      Unrolled at h = 64; a real loop below that (one register allocation
    //      serves every axis), which the compiler allocates with fewer spills.
#pragma unroll AXIS_UNROLL
    for (int i = 0; i < XD; ++i) {
      if (!S::has1(i)) continue;
      const bool two = S::has2(i);
      const float gur_a = pick<J>(gur, S::col1(i));
      const float gur_b = two ? pick<J>(gur, S::col2(i)) : 0.f;
      const float isd_i = pick<XD>(isd, i);
      const int zrow = i * H;
      float gzi[NRB][4];
      if (two) {
        f4 Gi[NRB];
        {
          f4 Zi[NRB];
          layer2_pass<H, SM>(v, e1, zrow, 3, Zi);
          const float gm4 = -4.f * gur_b;
#pragma unroll
          for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float z = Zi[ob][r];
              Q[ob][r] = fmaf(gur_b * z, z, Q[ob][r]);
              Gi[ob][r] = gm4 * (p2[ob][r] * z);   // 2 gur_b f2 z
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // column i: publish G_i, then W2~^T G_i while the other waves catch up
        store_tile<H>(stgw, Gi);
        {
          f4 GA[NRB];
          w2t_pass<H>(v, Gi, GA);
          const float gb2 = 2.f * gur_b;
#pragma unroll
          for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float p = he[rb][r] * v.zm[zrow + 16 * rb + r];   // e2 zt_i
              const float q = fmaf(gb2, p, gur_a * e1[rb][r]);
              gzi[rb][r] = fmaf(vt[rb][r], q, e1[rb][r] * GA[rb][r]);
              gz10[rb][r] = fmaf(p, GA[rb][r], gz10[rb][r]);
            }
        }
        tile_sync<O::PRIV>();
        if constexpr (O::PRIV)
          dw2_private<H, 1>(stgp, stbp, szt + zrow + pl, dW2);
        else
          dw2_shared<H, 1, SWZ>(stgl, stbl, stb_off, szt[zrow + ucol], dW2, wave);
        tile_sync<O::PRIV>();
      } else {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) gzi[rb][r] = e1[rb][r] * (gur_a * vt[rb][r]);
      }
      // dW1[:, i] through z1_i = zt_i (the 1 / sd_i factor after the sum)
      if constexpr (ACC) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) aW1[i][rb][r] = fmaf(gzi[rb][r], isd_i, aW1[i][rb][r]);
      } else {
        reduce_units_to_lds<NRB>(gzi, mysmall + (K_W1 + i) * H, pl, g, isd_i);
      }
    }
    if constexpr (S::N1 > 0) {
      opaque<NRB>(h2);
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float h = h2[ob][r];
          const float f1 = fmaf(-h, h, 1.f);
          const float f2 = -2.f * p2[ob][r];
          const float f3 = f1 * fmaf(-6.f, f1, 4.f);
          G0[ob][r] = fmaf(f3, Q[ob][r], fmaf(f2, Sx[ob][r], f1 * gur[0]));
          vWL[ob][r] = fmaf(f2, Q[ob][r], fmaf(f1, Sx[ob][r], gur[0] * h));
        }
    }
    // ---- column 0: db2 += G_0, dW2 += G_0 h^T, W2~^T G_0 -> layer 1
    const int kn = beg + (t + 1) * 64 + wave * 16 + pl;   // next tile's pair
    const bool valid_n = kn < end;
    const int pn = wl.pair_point[valid_n ? kn : beg];
    const int segn = wl.pair_seg[valid_n ? kn : beg];
    if constexpr (ACC) {
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) aB2[ob][r] += G0[ob][r];
    } else {
#pragma unroll
      for (int ob = 0; ob < NRB; ++ob)
#pragma unroll
        for (int r = 0; r < 4; ++r) vv[ob][r] = G0[ob][r];
      reduce_units_to_lds<NRB>(vv, mysmall + K_B2 * H, pl, g, wl_red);
    }
    store_tile<H>(stgw, G0);
    {
      f4 GA[NRB];
      w2t_pass<H>(v, G0, GA);
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) gz10[rb][r] = fmaf(e1[rb][r], GA[rb][r], gz10[rb][r]);
    }
    // dW1[:, i] += gz1_0 xh_i (the gz1_i parts were added per axis)
    if constexpr (ACC) {
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
          for (int i = 0; i < XD; ++i) aW1[i][rb][r] = fmaf(gz10[rb][r], xh[i], aW1[i][rb][r]);
          aWL[rb][r] += vWL[rb][r];
          aB1[rb][r] += gz10[rb][r];
        }
    } else {
#pragma unroll
      for (int i = 0; i < XD; ++i) {
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
          for (int r = 0; r < 4; ++r) vv[rb][r] = gz10[rb][r] * xh[i];
        reduce_units_to_lds<NRB>(vv, mysmall + (K_W1 + i) * H, pl, g);
      }
      reduce_units_to_lds<NRB>(vWL, mysmall + K_WL * H, pl, g);
      reduce_units_to_lds<NRB>(gz10, mysmall + K_B1 * H, pl, g);
    }
    // the next tile's point and upstream jets, in flight during the last dW2 column
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < XD; ++i) xn[i] = x[(size_t)pn * XD + i];
#pragma unroll
    for (int c = 0; c < J; ++c) gUn[c] = valid_n ? gseg[(size_t)segn * J + c] : 0.f;
    tile_sync<O::PRIV>();
    if constexpr (O::PRIV)
      dw2_private<H, 0>(stgp, stbp, nullptr, dW2);
    else
      dw2_shared<H, 0, SWZ>(stgl, stbl, stb_off, 1.f, dW2, wave);
    // (the next tile's leading barrier orders these reads before new writes)
  }

  // ---- end of chunk: the transpose buffers become the gradient row
  if constexpr (ACC) {
#pragma unroll
    for (int i = 0; i < XD; ++i) reduce_units_to_lds<NRB>(aW1[i], mysmall + (K_W1 + i) * H, pl, g);
    reduce_units_to_lds<NRB>(aWL, mysmall + K_WL * H, pl, g);
    reduce_units_to_lds<NRB>(aB1, mysmall + K_B1 * H, pl, g);
    reduce_units_to_lds<NRB>(aB2, mysmall + K_B2 * H, pl, g, wl_red);
  }
  if (g == 0) sbl[wave * 16 + pl] = dbl;
  __syncthreads();
  float* sacc = strans;
  for (int i = threadIdx.x; i < L::P; i += blockDim.x) sacc[i] = 0.f;
  __syncthreads();
  // H = 64: the waves hold disjoint dW2 blocks and write them at once;
  // H <= 32: every wave holds every block over its own pairs, summed wave by
  // wave in a fixed order
  for (int w = 0; w < (O::PRIV ? nw : 1); ++w) {
    if (!O::PRIV || wave == w) {
#pragma unroll
      for (int j = 0; j < O::NOB; ++j) {
        const int ob = O::row_block(wave, j);
        const int col = 16 * O::col_block(wave, j) + pl;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          sacc[L::W2 + (16 * ob + 4 * g + r) * H + col] += dW2[j][r] * v.vec[SM::WL + 16 * ob + r];
      }
    }
    __syncthreads();
  }
  finish_row<XD, H, 2>(sacc, ssmall, sbl, nw, partial + (size_t)chunk * L::P, gs, af, model, bc);
}

// g_scale: NULL, or d L / d loss when gseg holds the segment gradients at
// unit upstream gradient; it multiplies each chunk's gradient row as the row
// is written (finish_row).
// ADAM (one-hidden-layer networks): every trained subdomain row is one
// chunk, and each chunk's row is applied to theta by Adam (af) in place of
// being written to partial; the team (workgroup, or wave with WAVE) that
// finishes last advances the count.
template <int XD, int H, int NH, int D1M, int D2M, bool WAVE = false, bool ADAM = false>
__global__ void __launch_bounds__(256, 2)
bwd_kernel(const float* __restrict__ theta, const float* __restrict__ dom, const float* __restrict__ x,
           fbpx_worklist wl, const uint8_t* __restrict__ chunk_train, const float* __restrict__ gseg,
           const float* __restrict__ g_scale, float* __restrict__ partial, AdamFold af) {
  static_assert(!WAVE || NH == 1, "one wave per chunk: one-hidden-layer networks");
  static_assert(!ADAM || NH == 1, "the folded optimizer step: one-hidden-layer networks");
  const float gs = g_scale ? g_scale[0] : 1.f;
  // the bias corrections, formed in LDS by each team in its first chunk from
  // the count as it reads it before the last team advances it (the read
  // completes before the team's own count below), read where a row is
  // applied (after the chunk's barriers)
  __shared__ float sbc[2];
  bool first = true;
  // persistent: workgroup b runs the chunks of backward bin b in order (WAVE:
  // wave w runs bin 4b + w)
  const int bin = WAVE ? (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6) : (int)blockIdx.x;
  if (WAVE && bin >= wl.n_bbins) return;
  const int c0 = wl.bbin_chunk_start[bin], c1 = wl.bbin_chunk_start[bin + 1];
  for (int chunk = c0; chunk < c1; ++chunk) {
    if (chunk_train != nullptr && chunk_train[chunk] == 0) continue;
    Team<WAVE>::sync();   // the team is done with the previous chunk's LDS
    if constexpr (NH == 1)
      bwd_one<XD, H, D1M, D2M, WAVE, ADAM>(theta, dom, x, wl, gseg, partial, chunk, gs, af, sbc, first);
    else
      bwd_two<XD, H, D1M, D2M>(theta, dom, x, wl, gseg, partial, chunk, gs, af, sbc);
    first = false;
  }
  if constexpr (ADAM) {
    if (Team<WAVE>::tid() == 0 && atomicAdd(af.count_done, 1) == wl.n_bbins - 1) {
      atomicAdd(af.count, 1);
      atomicExch(af.count_done, 0);
    }
  }
}

}  // namespace fbpx
