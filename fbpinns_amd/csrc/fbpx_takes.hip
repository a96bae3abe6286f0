// Take and work-list builder on the device: the reference's get_inputs
// (trainers.py:332-419) -- inside test, active remap, all_ims = [active;
// fixed], m_take re-index, the (point, pou) unique of jnp.unique
// (trainers.py:380-388) -- followed by this library's kernel layout (tile
// order, segments and slots, chunks and launch bins; fbpx_worklist).
//
// Everything is integer work over the pair list.  The stages:
//   count    per point, the boxes holding it (inside_kernel) -> exclusive scan
//            -> offsets; the host reads the pair count s (one sync).
//   fill     point-major pairs (n_take, global model), ascending box id per
//            point: the row-major nonzero of the inside matrix
//            (decompositions_base.py:50-68).
//   used     models holding at least one point (LDS flags per workgroup).
//   remap    one workgroup: active 0 -> 1, times used (trainers.py:351-357),
//            all_ims = ascending active ids then ascending fixed ids, its
//            inverse, and the number of distinct pou values.
//   segments per point, the distinct pou values of its pairs (a 32-bit mask:
//            pou values lie in [0, 32)); segment = (point, pou) in ascending
//            order, so its index is jnp.unique's inverse p_take; the pairs
//            of a segment keep the point-major (reference) order, which
//            gives each pair its slot.
//   tiles    pairs grouped by model position, ascending point within: for
//            every (model position, block of 2048 points) a workgroup
//            repeats the inside test of the block's points against that one
//            box (skipped when the block's bounding box misses the box), so
//            a count pass, one exclusive scan in (model, block) order and a
//            fill pass place every pair stably without sorting; each tile
//            pair finds its point-major twin by binary search in the point's
//            ascending box list.
//   chunks   cut flags at every model start, every `chunk` pairs inside a
//            model and at both kernels' bin edges (round(linspace) over the
//            pairs, numpy's half-to-even rounding), compacted by a scan.
// All sums are exclusive scans with a fixed order; no atomics except the
// idempotent "used" flags, so the result is a function of the input alone.
#include <algorithm>

#include "fbpx_common.hpp"

namespace fbpx {
namespace {

constexpr int kThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;   // elements per scan workgroup
constexpr int kTileItems = 8;
constexpr int kTilePoints = kThreads * kTileItems;  // points per tile workgroup
constexpr int kUsedLds = 32768;                     // models flagged in LDS
constexpr int64_t kUsedPairs = 16384;               // pairs per "used" workgroup
constexpr int kMaxPou = 32;

// ------------------------------------------------------------------ scan

template <class T>
__device__ inline T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Exclusive scan over the 256 threads of the workgroup (thread order); total
// is the sum over all threads.  Every thread must call it.
template <class T>
__device__ inline T block_excl_scan(T v, T& total) {
  __shared__ T wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const T inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  T base = 0;
  total = 0;
#pragma unroll
  for (int i = 0; i < kThreads / 64; ++i) {
    if (i < w) base += wsum[i];
    total += wsum[i];
  }
  __syncthreads();
  return base + inc - v;
}

// tsum[tile] = sum of in[i] over the tile's elements i < n.
template <class In>
__global__ void __launch_bounds__(kThreads) scan_tile_sums(const In* __restrict__ in, int64_t n,
                                                           int64_t* __restrict__ tsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += (int64_t)in[base + k];
  int64_t tot;
  block_excl_scan<int64_t>(s, tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// Exclusive scan of tsum[0 .. nt) in place, one workgroup.
__global__ void __launch_bounds__(kThreads) scan_tile_prefix(int64_t* __restrict__ tsum, int64_t nt) {
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nt; b0 += kScanTile) {
    const int64_t base = b0 + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < nt ? tsum[base + k] : 0;
      s += v[k];
    }
    int64_t tot;
    int64_t run = carry + block_excl_scan<int64_t>(s, tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < nt) tsum[base + k] = run;
      run += v[k];
    }
    carry += tot;
  }
}

// out[i] = sum of in[j] for j < i, for i <= n (out[n] = total).  in and out
// may be the same array (each workgroup reads its elements before writing).
template <class In, class Out>
__global__ void __launch_bounds__(kThreads) scan_apply(const In* in, int64_t n, const int64_t* __restrict__ tsum,
                                                       Out* out) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? (int64_t)in[base + k] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t run = tsum[blockIdx.x] + block_excl_scan<int64_t>(s, tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k <= n) out[base + k] = (Out)run;
    run += v[k];
  }
}

int64_t scan_tiles(int64_t n) { return (n + 1 + kScanTile - 1) / kScanTile; }

// Exclusive scan of n elements into n + 1 outputs; tsum: scan_tiles(n) int64.
template <class In, class Out>
void scan(hipStream_t s, const In* in, int64_t n, Out* out, int64_t* tsum) {
  const int64_t nt = scan_tiles(n);
  hipLaunchKernelGGL(scan_tile_sums<In>, dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tsum);
  hipLaunchKernelGGL(scan_tile_prefix, dim3(1), dim3(kThreads), 0, s, tsum, nt);
  hipLaunchKernelGGL((scan_apply<In, Out>), dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, tsum, out);
}

// ------------------------------------------------------------------ helpers

template <class T>
__global__ void __launch_bounds__(kThreads) fill_kernel(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) p[i] = v;
}

template <class T>
void fill(hipStream_t s, T* p, int64_t n, T v) {
  if (n <= 0) return;
  const int64_t nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(kThreads), 0, s, p, n, v);
}

unsigned grid_for(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// used[j] = 1 for every model j holding a pair.  Each workgroup flags its
// pairs' models in LDS (when m fits) and writes each flag once.
__global__ void __launch_bounds__(kThreads) used_kernel(const int32_t* __restrict__ mg, int64_t s, int m,
                                                        uint8_t* __restrict__ used) {
  __shared__ uint8_t f[kUsedLds];
  const bool lds = m <= kUsedLds;
  if (lds) {
    for (int j = threadIdx.x; j < m; j += kThreads) f[j] = 0;
    __syncthreads();
  }
  const int64_t b0 = (int64_t)blockIdx.x * kUsedPairs;
  const int64_t b1 = b0 + kUsedPairs < s ? b0 + kUsedPairs : s;
  for (int64_t q = b0 + threadIdx.x; q < b1; q += kThreads) {
    const int j = mg[q];
    if (lds)
      f[j] = 1;
    else
      used[j] = 1;
  }
  if (lds) {
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += kThreads)
      if (f[j]) used[j] = 1;
  }
}

// info slots (device int32, read back by the host at the end)
enum { kInfoActive = 0, kInfoUsed, kInfoNpou, kInfoBadActive, kInfoBadPou, kInfoLen = 8 };

// One workgroup: the active remap of trainers.py:347-357 and all_ims.
__global__ void __launch_bounds__(kThreads) remap_kernel(const int32_t* __restrict__ active,
                                                         const uint8_t* __restrict__ used,
                                                         const int32_t* __restrict__ pou, int m,
                                                         int32_t* __restrict__ act_out, int32_t* __restrict__ all_ims,
                                                         int32_t* __restrict__ inv, int32_t* __restrict__ info) {
  __shared__ uint32_t pmask;
  __shared__ int bad_a, bad_p;
  if (threadIdx.x == 0) {
    pmask = 0;
    bad_a = 0;
    bad_p = 0;
  }
  __syncthreads();
  int carry = 0;
  for (int j0 = 0; j0 < m; j0 += kThreads) {
    const int j = j0 + threadIdx.x;
    int a = 0;
    if (j < m) {
      a = active[j];
      if (a < 0 || a > 2) atomicOr(&bad_a, 1);
      const int pv = pou[j];
      if (pv < 0 || pv >= kMaxPou)
        atomicOr(&bad_p, 1);
      else
        atomicOr(&pmask, 1u << pv);
      if (a == 0) a = 1;
      if (!used[j]) a = 0;
      act_out[j] = a;
      inv[j] = -1;
    }
    int tot;
    const int ex = block_excl_scan<int>(a == 1, tot);
    if (a == 1) {
      all_ims[carry + ex] = j;
      inv[j] = carry + ex;
    }
    carry += tot;
  }
  const int n_active = carry;
  for (int j0 = 0; j0 < m; j0 += kThreads) {
    const int j = j0 + threadIdx.x;
    const int a = j < m ? act_out[j] : 0;
    int tot;
    const int ex = block_excl_scan<int>(a == 2, tot);
    if (a == 2) {
      all_ims[carry + ex] = j;
      inv[j] = carry + ex;
    }
    carry += tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    info[kInfoActive] = n_active;
    info[kInfoUsed] = carry;
    info[kInfoNpou] = __popc(pmask);
    info[kInfoBadActive] = bad_a;
    info[kInfoBadPou] = bad_p;
  }
}

// Per point: the mask of pou values among its pairs (pmask) and the number
// of distinct values (its segment count, written to nseg for the scan).
__global__ void __launch_bounds__(kThreads) seg_count_kernel(int n, const int64_t* __restrict__ off,
                                                             const int32_t* __restrict__ mg,
                                                             const int32_t* __restrict__ pou,
                                                             uint32_t* __restrict__ pmask, int32_t* __restrict__ nseg) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  uint32_t mask = 0;
  for (int64_t q = off[p]; q < off[p + 1]; ++q) {
    const int v = pou[mg[q]];
    if ((unsigned)v < (unsigned)kMaxPou) mask |= 1u << v;   // others: reported by remap_kernel
  }
  pmask[p] = mask;
  nseg[p] = __popc(mask);
}

// Per point: the pair count of each of its segments (ascending pou).
__global__ void __launch_bounds__(kThreads) seg_size_kernel(int n, const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ mg,
                                                            const int32_t* __restrict__ pou,
                                                            const uint32_t* __restrict__ pmask,
                                                            const int32_t* __restrict__ point_seg_start,
                                                            int32_t* __restrict__ segcnt) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int seg = point_seg_start[p];
  const int64_t q0 = off[p], q1 = off[p + 1];
  for (uint32_t mm = pmask[p]; mm; mm &= mm - 1, ++seg) {
    const int v = __ffs(mm) - 1;
    int c = 0;
    for (int64_t q = q0; q < q1; ++q) c += pou[mg[q]] == v;
    segcnt[seg] = c;
  }
}

// Per point: p_take, the slot of each point-major pair (segment start + its
// rank among the segment's pairs in point-major order), np_take and the
// re-indexed m_take.
__global__ void __launch_bounds__(kThreads) seg_fill_kernel(int n, const int64_t* __restrict__ off,
                                                            const int32_t* __restrict__ mg,
                                                            const int32_t* __restrict__ pou,
                                                            const uint32_t* __restrict__ pmask,
                                                            const int32_t* __restrict__ point_seg_start,
                                                            const int32_t* __restrict__ seg_pair_start,
                                                            const int32_t* __restrict__ inv,
                                                            int32_t* __restrict__ p_take, int32_t* __restrict__ slot_pm,
                                                            int32_t* __restrict__ np_take,
                                                            int32_t* __restrict__ m_take) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= n) return;
  int seg = point_seg_start[p];
  const int64_t q0 = off[p], q1 = off[p + 1];
  for (int64_t q = q0; q < q1; ++q) m_take[q] = inv[mg[q]];
  for (uint32_t mm = pmask[p]; mm; mm &= mm - 1, ++seg) {
    const int v = __ffs(mm) - 1;
    int w = seg_pair_start[seg];
    np_take[seg] = p;
    for (int64_t q = q0; q < q1; ++q) {
      if (pou[mg[q]] == v) {
        p_take[q] = seg;
        slot_pm[q] = w++;
      }
    }
  }
}

// Bounding box of each block of kTilePoints points: lo[b][i], hi[b][i].
template <int XD>
__global__ void __launch_bounds__(kThreads) block_bbox_kernel(const float* __restrict__ x, int n,
                                                              float* __restrict__ lo, float* __restrict__ hi) {
  __shared__ float slo[kThreads][XD], shi[kThreads][XD];
  const int b = blockIdx.x;
  float l[XD], h[XD];
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    l[i] = INFINITY;
    h[i] = -INFINITY;
  }
  for (int k = 0; k < kTileItems; ++k) {
    const int p = b * kTilePoints + threadIdx.x * kTileItems + k;
    if (p < n) {
#pragma unroll
      for (int i = 0; i < XD; ++i) {
        const float v = x[(size_t)p * XD + i];
        l[i] = fminf(l[i], v);   // a NaN coordinate is never inside a box; fminf skips it
        h[i] = fmaxf(h[i], v);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    slo[threadIdx.x][i] = l[i];
    shi[threadIdx.x][i] = h[i];
  }
  __syncthreads();
  if (threadIdx.x < XD) {
    const int i = threadIdx.x;
    float a = INFINITY, c = -INFINITY;
    for (int t = 0; t < kThreads; ++t) {
      a = fminf(a, slo[t][i]);
      c = fmaxf(c, shi[t][i]);
    }
    lo[b * XD + i] = a;
    hi[b * XD + i] = c;
  }
}

// The inside test of point p against one box (decompositions.py:225-226).
template <int XD>
__device__ inline bool inside_box(const float* __restrict__ x, int p, const float* blo, const float* bhi) {
  bool in = true;
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    const float v = x[(size_t)p * XD + i];
    in = in && (v >= blo[i]) && (v <= bhi[i]);
  }
  return in;
}

// Workgroup (model position k, point block b), flattened k * nblk + b.
// FILL = false: blkcnt[k * nblk + b] = the block's points inside box
// all_ims[k] (0 for k >= n_used).  FILL = true: writes those points' tile
// pairs at base[k * nblk + b] onwards, in ascending point order.
template <int XD, bool FILL>
__global__ void __launch_bounds__(kThreads) tile_kernel(const float* __restrict__ x, int n, int nblk,
                                                        const float* __restrict__ xmin, const float* __restrict__ xmax,
                                                        const float* __restrict__ blo, const float* __restrict__ bhi,
                                                        const int32_t* __restrict__ all_ims,
                                                        const int32_t* __restrict__ info, int32_t* __restrict__ blkcnt,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ mg,
                                                        const int32_t* __restrict__ p_take,
                                                        const int32_t* __restrict__ slot_pm,
                                                        int32_t* __restrict__ pair_point, int32_t* __restrict__ pair_seg,
                                                        int32_t* __restrict__ pair_slot) {
  const int64_t wg = blockIdx.x;
  const int k = (int)(wg / nblk), b = (int)(wg % nblk);
  const int n_used = info[kInfoUsed];
  if (k >= n_used) {
    if (!FILL && threadIdx.x == 0) blkcnt[wg] = 0;
    return;
  }
  const int j = all_ims[k];
  float lo[XD], hi[XD];
  bool miss = false;
#pragma unroll
  for (int i = 0; i < XD; ++i) {
    lo[i] = xmin[(size_t)j * XD + i];
    hi[i] = xmax[(size_t)j * XD + i];
    miss = miss || hi[i] < blo[b * XD + i] || lo[i] > bhi[b * XD + i];
  }
  if (miss) {   // uniform across the workgroup: no point of the block can be inside
    if (!FILL && threadIdx.x == 0) blkcnt[wg] = 0;
    return;
  }
  const int p0 = b * kTilePoints + threadIdx.x * kTileItems;
  uint32_t bits = 0;
#pragma unroll
  for (int t = 0; t < kTileItems; ++t) {
    const int p = p0 + t;
    if (p < n && inside_box<XD>(x, p, lo, hi)) bits |= 1u << t;
  }
  int tot;
  const int ex = block_excl_scan<int>(__popc(bits), tot);
  if (!FILL) {
    if (threadIdx.x == 0) blkcnt[wg] = tot;
    return;
  }
  int t_out = blkcnt[wg] + ex;   // blkcnt holds the scanned bases here
  for (; bits; bits &= bits - 1) {
    const int p = p0 + __ffs(bits) - 1;
    // the pair's point-major twin: box j in the point's ascending box list
    const int64_t q_end = off[p + 1];
    int64_t a = off[p], c = q_end;
    while (a < c) {
      const int64_t mid = (a + c) >> 1;
      if (mg[mid] < j)
        a = mid + 1;
      else
        c = mid;
    }
    const bool found = a < q_end && mg[a] == j;   // always, for offsets of these points
    pair_point[t_out] = p;
    pair_seg[t_out] = found ? p_take[a] : -1;
    pair_slot[t_out] = found ? slot_pm[a] : -1;
    ++t_out;
  }
}

// model_start[k] = base[k * nblk] for k <= n_used (base[n_used * nblk] = s).
__global__ void __launch_bounds__(kThreads) model_start_kernel(const int32_t* __restrict__ base, int nblk, int m,
                                                               const int32_t* __restrict__ info,
                                                               int32_t* __restrict__ model_start) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k <= info[kInfoUsed] && k <= m) model_start[k] = base[(int64_t)k * nblk];
}

// Cut flags at model starts and every `chunk` pairs inside a model.
__global__ void __launch_bounds__(kThreads) cut_models_kernel(const int32_t* __restrict__ model_start, int m,
                                                              const int32_t* __restrict__ info, int chunk,
                                                              uint8_t* __restrict__ flags) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int n_used = info[kInfoUsed];
  if (k > n_used || k > m) return;
  const int t0 = model_start[k];
  flags[t0] = 1;
  if (chunk > 0 && k < n_used) {
    const int cnt = model_start[k + 1] - t0;
    for (int t = chunk; t < cnt; t += chunk) flags[t0 + t] = 1;
  }
}

// Cut flags at round(linspace(0, S, nb + 1)) (numpy: i * (S / nb) in float64,
// the last point exactly S, rounded half to even), S = s or the trained
// prefix model_start[n_active].
__global__ void __launch_bounds__(kThreads) cut_bins_kernel(int nb, int64_t s, const int32_t* __restrict__ model_start,
                                                            const int32_t* __restrict__ info, bool trained,
                                                            uint8_t* __restrict__ flags, int32_t* __restrict__ edges) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i > nb) return;
  const int64_t S = trained ? (int64_t)model_start[info[kInfoActive]] : s;
  const double step = (double)S / (double)nb;
  const int64_t t = i == nb ? S : (int64_t)rint((double)i * step);
  flags[t] = 1;
  edges[i] = (int32_t)t;
}

// chunk_start[cut[t]] = t for every flagged position t <= s.
__global__ void __launch_bounds__(kThreads) chunk_start_kernel(const uint8_t* __restrict__ flags,
                                                               const int32_t* __restrict__ cut, int64_t s,
                                                               int32_t* __restrict__ chunk_start) {
  const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (t <= s && flags[t]) chunk_start[cut[t]] = (int32_t)t;
}

// Per chunk: its model position (the last model start <= its start), model
// id and trained flag.
__global__ void __launch_bounds__(kThreads) chunk_info_kernel(const int32_t* __restrict__ chunk_start,
                                                              const int32_t* __restrict__ cut, int64_t s,
                                                              const int32_t* __restrict__ model_start,
                                                              const int32_t* __restrict__ all_ims,
                                                              const int32_t* __restrict__ info,
                                                              int32_t* __restrict__ chunk_model,
                                                              uint8_t* __restrict__ chunk_train) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  const int n_chunks = cut[s + 1] - 1;
  if (c >= n_chunks) return;
  const int t = chunk_start[c];
  int a = 0, e = info[kInfoUsed] + 1;   // upper bound of t in model_start[0 .. n_used]
  while (a < e) {
    const int mid = (a + e) >> 1;
    if (model_start[mid] <= t)
      a = mid + 1;
    else
      e = mid;
  }
  const int pos = a - 1;
  chunk_model[c] = all_ims[pos];
  chunk_train[c] = pos < info[kInfoActive] ? 1 : 0;
}

// row_chunk_start[k] = the chunk starting at model k's first pair (k <= n_active).
__global__ void __launch_bounds__(kThreads) rows_kernel(const int32_t* __restrict__ model_start,
                                                        const int32_t* __restrict__ cut,
                                                        const int32_t* __restrict__ info, int m,
                                                        int32_t* __restrict__ row_chunk_start) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k <= info[kInfoActive] && k <= m) row_chunk_start[k] = cut[model_start[k]];
}

// bin_chunk_start[i] = the chunk starting at bin edge i (nb > 0), or i for
// every chunk (nb == 0: one bin per chunk).
__global__ void __launch_bounds__(kThreads) bins_kernel(int nb, const int32_t* __restrict__ edges,
                                                        const int32_t* __restrict__ cut, int64_t s, int max_chunks,
                                                        int32_t* __restrict__ bin_chunk_start) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (nb > 0) {
    if (i <= nb) bin_chunk_start[i] = cut[edges[i]];
  } else if (i <= cut[s + 1] - 1 && i <= max_chunks) {
    bin_chunk_start[i] = i;
  }
}

// ------------------------------------------------------------------ work space

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

struct Carve {
  char* base;
  size_t off, cap;
  template <class T>
  T* take(int64_t count) {
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += align_up((size_t)(count > 0 ? count : 1) * sizeof(T));
    return p;
  }
};

int64_t nblocks(int n) { return (n + kTilePoints - 1) / kTilePoints; }

struct Work {
  int32_t* counts;
  int64_t* tsum;
  int32_t* mg;
  uint8_t* used;
  int32_t* inv;
  int32_t* info;
  uint32_t* pmask;
  int32_t* slot_pm;
  float* blo;
  float* bhi;
  int32_t* blkcnt;
  int32_t* model_start;
  uint8_t* flags;
  int32_t* cut;
  int32_t* fedges;
  int32_t* bedges;
};

Work carve(void* base, int n, int xd, int m, int64_t s, int fbins, int bbins, size_t* bytes) {
  Carve c{static_cast<char*>(base), 0, 0};
  Work w;
  const int64_t nb = nblocks(n);
  const int64_t nmax = std::max<int64_t>(std::max<int64_t>(n, s + 1), (int64_t)m * nb + 1);
  w.counts = c.take<int32_t>(n);
  w.tsum = c.take<int64_t>(scan_tiles(nmax) + 1);
  w.mg = c.take<int32_t>(s);
  w.used = c.take<uint8_t>(m);
  w.inv = c.take<int32_t>(m);
  w.info = c.take<int32_t>(kInfoLen);
  w.pmask = c.take<uint32_t>(n);
  w.slot_pm = c.take<int32_t>(s);
  w.blo = c.take<float>(nb * xd);
  w.bhi = c.take<float>(nb * xd);
  w.blkcnt = c.take<int32_t>((int64_t)m * nb + 1);
  w.model_start = c.take<int32_t>((int64_t)m + 1);
  w.flags = c.take<uint8_t>(s + 1);
  w.cut = c.take<int32_t>(s + 2);
  w.fedges = c.take<int32_t>((int64_t)fbins + 1);
  w.bedges = c.take<int32_t>((int64_t)bbins + 1);
  *bytes = c.off;
  return w;
}

template <int XD>
void launch_tiles(hipStream_t st, bool fill_pass, const float* x, int n, int nblk, int m, const float* xmin,
                  const float* xmax, const Work& w, const int32_t* all_ims, const int64_t* off,
                  const int32_t* p_take, int32_t* pair_point, int32_t* pair_seg, int32_t* pair_slot) {
  const dim3 grid((unsigned)((int64_t)m * nblk));
  if (fill_pass)
    hipLaunchKernelGGL((tile_kernel<XD, true>), grid, dim3(kThreads), 0, st, x, n, nblk, xmin, xmax, w.blo, w.bhi,
                       all_ims, w.info, w.blkcnt, off, w.mg, p_take, w.slot_pm, pair_point, pair_seg, pair_slot);
  else
    hipLaunchKernelGGL((tile_kernel<XD, false>), grid, dim3(kThreads), 0, st, x, n, nblk, xmin, xmax, w.blo, w.bhi,
                       all_ims, w.info, w.blkcnt, off, w.mg, p_take, w.slot_pm, pair_point, pair_seg, pair_slot);
}

int check_sizes(const char* what, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                int32_t m) {
  if (n < 0 || m < 1 || xd < 1 || xd > FBPX_MAX_XD || (n > 0 && !x) || !xmin || !xmax) {
    set_error("%s: bad arguments n=%d m=%d xd=%d", what, n, m, xd);
    return FBPX_EINVAL;
  }
  return FBPX_OK;
}

}  // namespace
}  // namespace fbpx

using namespace fbpx;

extern "C" {

int64_t fbpx_takes_max_chunks(int32_t m, int64_t n_pairs, int32_t chunk, int32_t fbins, int32_t bbins) {
  return (chunk > 0 ? n_pairs / chunk : 0) + 2 * (int64_t)m + (int64_t)fbins + (int64_t)bbins + 3;
}

int64_t fbpx_takes_work_bytes(int32_t n, int32_t xd, int32_t m, int64_t n_pairs, int32_t fbins, int32_t bbins) {
  if (n < 0 || m < 1 || xd < 1 || xd > FBPX_MAX_XD || n_pairs < 0 || fbins < 0 || bbins < 0) return -1;
  size_t bytes = 0;
  carve(nullptr, n, xd, m, n_pairs, fbins, bbins, &bytes);
  return (int64_t)bytes;
}

int fbpx_takes_count(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                     int32_t m, int64_t* offsets, void* work, int64_t work_bytes, int64_t* n_pairs) {
  int rc = check_sizes("takes_count", x, n, xd, xmin, xmax, m);
  if (rc) return rc;
  size_t need = 0;
  carve(nullptr, n, xd, m, 0, 0, 0, &need);
  if (!offsets || !n_pairs || !work || work_bytes < (int64_t)need) {
    set_error("takes_count: null output or work space of %lld bytes (need %zu)", (long long)work_bytes, need);
    return FBPX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  Work w = carve(work, n, xd, m, 0, 0, 0, &need);
  if (n > 0 && (rc = fbpx_inside_count(stream, x, n, xd, xmin, xmax, m, nullptr, w.counts))) return rc;
  scan<int32_t, int64_t>(s, w.counts, n, offsets, w.tsum);
  if ((rc = hip_check(hipGetLastError(), "takes_count"))) return rc;
  int64_t total = 0;
  if ((rc = hip_check(hipMemcpyAsync(&total, offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, s),
                      "takes_count: copy")))
    return rc;
  if ((rc = hip_check(hipStreamSynchronize(s), "takes_count: sync"))) return rc;
  if (total > INT32_MAX - 2) {
    set_error("takes_count: %lld (point, subdomain) pairs exceed the int32 work-list indices", (long long)total);
    return FBPX_EUNSUPPORTED;
  }
  *n_pairs = total;
  return FBPX_OK;
}

int fbpx_build_worklist(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                        int32_t m, const int32_t* active, const int32_t* pou, const int64_t* offsets, int64_t n_pairs,
                        int32_t chunk, int32_t fbins, int32_t bbins, int64_t max_chunks, void* work,
                        int64_t work_bytes, const fbpx_takes_out* out, fbpx_takes_info* info) {
  int rc = check_sizes("build_worklist", x, n, xd, xmin, xmax, m);
  if (rc) return rc;
  if (!active || !pou || !offsets || !out || !info || n_pairs < 0 || n_pairs > INT32_MAX - 2 || chunk < 0 ||
      fbins < 0 || bbins < 0) {
    set_error("build_worklist: bad arguments (n_pairs=%lld chunk=%d fbins=%d bbins=%d)", (long long)n_pairs, chunk,
              fbins, bbins);
    return FBPX_EINVAL;
  }
  const int64_t s = n_pairs;
  const int64_t need_chunks = fbpx_takes_max_chunks(m, s, chunk, fbins, bbins);
  if (max_chunks < need_chunks || max_chunks > INT32_MAX - 2) {
    set_error("build_worklist: chunk arrays for %lld chunks, need %lld (fbpx_takes_max_chunks)",
              (long long)max_chunks, (long long)need_chunks);
    return FBPX_EINVAL;
  }
  size_t need = 0;
  carve(nullptr, n, xd, m, s, fbins, bbins, &need);
  if (!work || work_bytes < (int64_t)need) {
    set_error("build_worklist: work space of %lld bytes, need %zu", (long long)work_bytes, need);
    return FBPX_EINVAL;
  }
  const fbpx_takes_out& o = *out;
  if ((s > 0 && (!o.m_take || !o.n_take || !o.p_take || !o.np_take || !o.pair_point || !o.pair_seg ||
                 !o.pair_slot)) ||
      !o.active || !o.all_ims || !o.point_seg_start || !o.seg_pair_start || !o.chunk_model || !o.chunk_start ||
      !o.chunk_train || !o.row_chunk_start || !o.fbin_chunk_start || !o.bbin_chunk_start) {
    set_error("build_worklist: null output array");
    return FBPX_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  Work w = carve(work, n, xd, m, s, fbins, bbins, &need);
  const int64_t nblk = nblocks(n);

  // point-major pairs and the models they use
  fill<uint8_t>(st, w.used, m, 0);
  if (s > 0) {
    if ((rc = fbpx_inside_fill(stream, x, n, xd, xmin, xmax, m, nullptr, offsets, o.n_take, w.mg))) return rc;
    hipLaunchKernelGGL(used_kernel, dim3((unsigned)((s + kUsedPairs - 1) / kUsedPairs)), dim3(kThreads), 0, st, w.mg,
                       s, m, w.used);
  }
  hipLaunchKernelGGL(remap_kernel, dim3(1), dim3(kThreads), 0, st, active, w.used, pou, m, o.active, o.all_ims, w.inv,
                     w.info);

  // segments: (point, pou) rows in ascending order and each pair's slot
  if (n > 0) {
    hipLaunchKernelGGL(seg_count_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, n, offsets, w.mg, pou, w.pmask,
                       o.point_seg_start);
  }
  scan<int32_t, int32_t>(st, o.point_seg_start, n, o.point_seg_start, w.tsum);
  fill<int32_t>(st, o.seg_pair_start, s, 0);
  if (n > 0)
    hipLaunchKernelGGL(seg_size_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, n, offsets, w.mg, pou, w.pmask,
                       o.point_seg_start, o.seg_pair_start);
  scan<int32_t, int32_t>(st, o.seg_pair_start, s, o.seg_pair_start, w.tsum);
  if (n > 0)
    hipLaunchKernelGGL(seg_fill_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, n, offsets, w.mg, pou, w.pmask,
                       o.point_seg_start, o.seg_pair_start, w.inv, o.p_take, w.slot_pm, o.np_take, o.m_take);

  // tile order: count per (model position, point block), scan, fill
  const int64_t nwg = (int64_t)m * nblk;
  if (nwg > 0) {
    switch (xd) {
      case 1: hipLaunchKernelGGL(block_bbox_kernel<1>, dim3((unsigned)nblk), dim3(kThreads), 0, st, x, n, w.blo, w.bhi); break;
      case 2: hipLaunchKernelGGL(block_bbox_kernel<2>, dim3((unsigned)nblk), dim3(kThreads), 0, st, x, n, w.blo, w.bhi); break;
      default: hipLaunchKernelGGL(block_bbox_kernel<3>, dim3((unsigned)nblk), dim3(kThreads), 0, st, x, n, w.blo, w.bhi); break;
    }
    for (int pass = 0; pass < 2; ++pass) {
      switch (xd) {
        case 1: launch_tiles<1>(st, pass, x, n, (int)nblk, m, xmin, xmax, w, o.all_ims, offsets, o.p_take, o.pair_point, o.pair_seg, o.pair_slot); break;
        case 2: launch_tiles<2>(st, pass, x, n, (int)nblk, m, xmin, xmax, w, o.all_ims, offsets, o.p_take, o.pair_point, o.pair_seg, o.pair_slot); break;
        default: launch_tiles<3>(st, pass, x, n, (int)nblk, m, xmin, xmax, w, o.all_ims, offsets, o.p_take, o.pair_point, o.pair_seg, o.pair_slot); break;
      }
      if (pass == 0) scan<int32_t, int32_t>(st, w.blkcnt, nwg, w.blkcnt, w.tsum);
    }
  } else {
    fill<int32_t>(st, w.blkcnt, 1, 0);
  }
  hipLaunchKernelGGL(model_start_kernel, dim3(grid_for((int64_t)m + 1)), dim3(kThreads), 0, st, w.blkcnt,
                     (int)(nblk > 0 ? nblk : 1), m, w.info, w.model_start);

  // chunks and bins
  fill<uint8_t>(st, w.flags, s + 1, 0);
  hipLaunchKernelGGL(cut_models_kernel, dim3(grid_for((int64_t)m + 1)), dim3(kThreads), 0, st, w.model_start, m, w.info,
                     chunk, w.flags);
  if (fbins > 0)
    hipLaunchKernelGGL(cut_bins_kernel, dim3(grid_for((int64_t)fbins + 1)), dim3(kThreads), 0, st, fbins, s,
                       w.model_start, w.info, false, w.flags, w.fedges);
  if (bbins > 0)
    hipLaunchKernelGGL(cut_bins_kernel, dim3(grid_for((int64_t)bbins + 1)), dim3(kThreads), 0, st, bbins, s,
                       w.model_start, w.info, true, w.flags, w.bedges);
  scan<uint8_t, int32_t>(st, w.flags, s + 1, w.cut, w.tsum);
  hipLaunchKernelGGL(chunk_start_kernel, dim3(grid_for(s + 1)), dim3(kThreads), 0, st, w.flags, w.cut, s,
                     o.chunk_start);
  hipLaunchKernelGGL(chunk_info_kernel, dim3(grid_for(max_chunks)), dim3(kThreads), 0, st, o.chunk_start, w.cut, s,
                     w.model_start, o.all_ims, w.info, o.chunk_model, o.chunk_train);
  hipLaunchKernelGGL(rows_kernel, dim3(grid_for((int64_t)m + 1)), dim3(kThreads), 0, st, w.model_start, w.cut, w.info,
                     m, o.row_chunk_start);
  hipLaunchKernelGGL(bins_kernel, dim3(grid_for((fbins > 0 ? fbins : max_chunks) + 1)), dim3(kThreads), 0, st, fbins,
                     w.fedges, w.cut, s, (int)max_chunks, o.fbin_chunk_start);
  hipLaunchKernelGGL(bins_kernel, dim3(grid_for((bbins > 0 ? bbins : max_chunks) + 1)), dim3(kThreads), 0, st, bbins,
                     w.bedges, w.cut, s, (int)max_chunks, o.bbin_chunk_start);
  if ((rc = hip_check(hipGetLastError(), "build_worklist"))) return rc;

  // sizes back to the host
  int32_t h_info[kInfoLen] = {};
  int32_t h_tail[3] = {};   // n_segs, cuts, s from the tile scan
  if ((rc = hip_check(hipMemcpyAsync(h_info, w.info, sizeof(h_info), hipMemcpyDeviceToHost, st), "copy info")) ||
      (rc = hip_check(hipMemcpyAsync(&h_tail[0], o.point_seg_start + n, 4, hipMemcpyDeviceToHost, st), "copy")) ||
      (rc = hip_check(hipMemcpyAsync(&h_tail[1], w.cut + s + 1, 4, hipMemcpyDeviceToHost, st), "copy")) ||
      (rc = hip_check(hipMemcpyAsync(&h_tail[2], w.blkcnt + nwg, 4, hipMemcpyDeviceToHost, st), "copy")) ||
      (rc = hip_check(hipStreamSynchronize(st), "build_worklist: sync")))
    return rc;
  if (h_info[kInfoBadActive]) {
    set_error("build_worklist: active values must be 0, 1 or 2 (trainers.py:348)");
    return FBPX_EINVAL;
  }
  if (h_info[kInfoBadPou]) {
    set_error("build_worklist: pou values must lie in [0, %d)", kMaxPou);
    return FBPX_EUNSUPPORTED;
  }
  if ((int64_t)h_tail[2] != s) {
    set_error("build_worklist: tile pass placed %d pairs of %lld (offsets not from fbpx_takes_count on these inputs?)",
              h_tail[2], (long long)s);
    return FBPX_EINVAL;
  }
  info->n_pairs = s;
  info->n_segs = h_tail[0];
  info->n_used = h_info[kInfoUsed];
  info->n_active = h_info[kInfoActive];
  info->n_chunks = h_tail[1] - 1;
  info->n_fbins = fbins > 0 ? fbins : h_tail[1] - 1;
  info->n_bbins = bbins > 0 ? bbins : h_tail[1] - 1;
  info->npou = h_info[kInfoNpou];
  return FBPX_OK;
}

}  // extern "C"
