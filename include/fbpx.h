/*
 * fbpx.h -- C ABI of libfbpx.so, the MI355X (gfx950) FBPINN train-step engine.
 *
 * Every pointer is a DEVICE pointer owned by the caller unless stated.  No
 * entry point allocates device memory or calls back into the host, and none
 * synchronises the stream except the take builder (fbpx_takes_count,
 * fbpx_build_worklist, whose output sizes are data dependent), so the
 * per-step sequence of calls can be captured into a hipGraph.
 * Every call returns 0 on success and a negative code on error; the text of
 * the last error on the calling thread is returned by fbpx_last_error().
 * Nothing throws across the ABI.  `stream` is a hipStream_t passed as void*.
 *
 * The derivative jet of a quantity q at a point is stored as J contiguous
 * floats: [q, dq/dx_i for each i in d1mask (ascending i),
 *          d2q/dx_i^2 for each i in d2mask (ascending i)],
 * J = 1 + popcount(d1mask) + popcount(d2mask), d2mask a subset of d1mask.
 *
 * Subdomain parameter rows (theta, [m][P] floats) pack the FCN layers of one
 * subdomain in the reference's layout (networks.py:50-58: W is (out, in),
 * row-major): W1 (h*xd), b1 (h), [W2 (h*h), b2 (h)] if nh == 2, WL (h), bL (1).
 *
 * Subdomain static rows (dom, [m][FBPX_DOM_STRIDE] floats):
 *   xmin[0..3], xmax[0..3], unnorm_mu, unnorm_sd, flag, pad
 * (decompositions.py:127-128,181: xmins, xmaxs, flags, unnorms).
 */
#ifndef FBPX_H
#define FBPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBPX_DOM_STRIDE 12
#define FBPX_MAX_XD 3

enum {
  FBPX_OK = 0,
  FBPX_EINVAL = -1,      /* bad argument (size, null pointer, mask) */
  FBPX_EUNSUPPORTED = -2,/* network / jet shape not compiled in */
  FBPX_EHIP = -3         /* HIP runtime error at launch */
};

/* Network and jet schema.  Layer sizes are [xd, h x nh, 1]. */
typedef struct {
  int32_t xd;      /* input dimension, 1..3 */
  int32_t h;       /* hidden width: 16, 32 or 64 */
  int32_t nh;      /* number of hidden layers: 1 or 2 */
  int32_t d1mask;  /* bit i set: first derivative along x_i is carried */
  int32_t d2mask;  /* bit i set: d2/dx_i^2 is carried (subset of d1mask) */
} fbpx_net;

/*
 * Work list for one set of takes.  Pairs are (point, subdomain) memberships.
 * "Tile order": pairs grouped by subdomain (ascending chunk index), within a
 * subdomain in ascending point order.  Each chunk is a contiguous run of
 * pairs of ONE subdomain (any length >= 1) and produces one partial gradient
 * row.  "Bins": the forward and the backward each launch one workgroup per
 * bin; a bin is a contiguous range of chunks (fbin_chunk_start /
 * bbin_chunk_start), processed in order by that workgroup.  The engine cuts
 * the pairs into as many equal bins as the GPU holds workgroups (one
 * persistent launch per kernel, no tail), with chunk boundaries at the
 * subdomain boundaries and both kernels' bin edges; one bin per chunk of at
 * most FBPX_CHUNK pairs is the plain one-workgroup-per-chunk launch.
 * "Segment": a unique (point, pou) row of the reference's jnp.unique
 * (trainers.py:380-385), ordered by (point, pou).
 */
#define FBPX_CHUNK 512
typedef struct {
  int32_t n_pairs;
  int32_t n_chunks;
  int32_t n_points;
  int32_t n_segs;
  int32_t npou;
  int32_t n_fbins;                /* forward workgroups */
  int32_t n_bbins;                /* backward workgroups */
  const int32_t* pair_point;      /* [n_pairs] tile order -> point index */
  const int32_t* pair_seg;        /* [n_pairs] tile order -> segment index */
  const int32_t* chunk_model;     /* [n_chunks] subdomain row of the chunk */
  const int32_t* chunk_start;     /* [n_chunks+1] tile-order start, last = n_pairs */
  const int32_t* point_seg_start; /* [n_points+1] segments of each point */
  const int32_t* seg_pair_start;  /* [n_segs+1] slot range of each segment */
  const int32_t* pair_slot;       /* [n_pairs] tile order -> slot in segment order: the
                                     pairs of segment s occupy slots
                                     seg_pair_start[s] .. seg_pair_start[s+1]-1, in the
                                     reference's pair order */
  const int32_t* fbin_chunk_start; /* [n_fbins+1] chunks of forward workgroup b:
                                      fbin_chunk_start[b] .. fbin_chunk_start[b+1]-1 */
  const int32_t* bbin_chunk_start; /* [n_bbins+1] chunks of backward workgroup b */
  int32_t fwd_wave_bins;          /* 1: each wave of a forward workgroup runs a bin of its own
                                     (four bins per workgroup, ceil(n_fbins / 4) workgroups),
                                     for one-hidden-layer networks; 0: one workgroup per bin */
  int32_t bwd_wave_bins;          /* the same for the backward */
} fbpx_worklist;

/* ---------------------------------------------------------------- takes */

/* Number of boxes containing each point (inclusive test xmin <= x <= xmax on
 * every axis, FP32), over models with model_mask[j] != 0 (model_mask may be
 * NULL = all).  Replaces _inside_sum_batch (decompositions_base.py:23-48) as
 * used by inside_points / inside_models (decompositions.py:201-215,217-227).
 * x: [n][xd]; xmin, xmax: [m][xd]; counts: [n] int32. */
int fbpx_inside_count(void* stream, const float* x, int32_t n, int32_t xd,
                      const float* xmin, const float* xmax, int32_t m,
                      const uint8_t* model_mask, int32_t* counts);

/* Row-major nonzero of the inside matrix: for point p, the ids of the boxes
 * containing it, ascending, at offsets[p]..offsets[p+1].  Bit-exact with
 * _inside_take_batch (decompositions_base.py:50-68, order pinned by its
 * self-test :100-128).  offsets: [n+1] int64 exclusive scan of counts. */
int fbpx_inside_fill(void* stream, const float* x, int32_t n, int32_t xd,
                     const float* xmin, const float* xmax, int32_t m,
                     const uint8_t* model_mask, const int64_t* offsets,
                     int32_t* n_take, int32_t* m_take);

/*
 * Takes and work list built on the device: the reference's get_inputs
 * (trainers.py:332-419: inside test, active remap :347-357, all_ims =
 * [active; fixed], m_take re-index, the (point, pou) jnp.unique :380-388)
 * and this library's fbpx_worklist layout for them, without host work.  Two
 * calls, each ending with one stream synchronisation (the sizes are data
 * dependent; the reference syncs here too, decompositions_base.py:74):
 *
 *   fbpx_takes_count    inside counts of the n points over all m boxes and
 *                       their exclusive scan into offsets[n + 1] (int64);
 *                       returns the pair count s = offsets[n] to the host.
 *   fbpx_build_worklist everything else, into caller-allocated arrays sized
 *                       from s (below); returns the exact sizes in *info.
 *
 * x: [n][xd] the training points (trainers.py:482-507 selects them);
 * xmin, xmax: [m][xd] f32 boxes; active: [m] int32 scheduler state in
 * {0, 1, 2}; pou: [m] int32 level of each subdomain, in [0, 32).  chunk: cut
 * chunks every `chunk` pairs inside a subdomain (0: only at subdomain
 * starts); fbins / bbins: forward / backward launch bins, each an equal
 * share of the pairs (the backward's of the trained prefix), cut at
 * round(linspace) like numpy (0: one bin per chunk).  work: device work space
 * of fbpx_takes_work_bytes(n, xd, m, s, fbins, bbins) bytes (with s = 0 for
 * fbpx_takes_count).  Results are a function of the inputs alone (exclusive
 * scans in a fixed order, no ordering by atomics); takes are bit-exact with
 * the reference's, tile order / segments / chunks with the fbpx_worklist
 * description.  No counterpart in the reference beyond get_inputs: its XLA
 * program re-gathers per pair instead of keeping a work list.
 */
typedef struct {
  int32_t* m_take;           /* [s]   position in all_ims of each pair's subdomain */
  int32_t* n_take;           /* [s]   point of each pair (point-major, ascending box id) */
  int32_t* p_take;           /* [s]   segment of each pair (jnp.unique inverse) */
  int32_t* np_take;          /* [s]   point of each segment (first n_segs written) */
  int32_t* active;           /* [m]   the remapped scheduler state (0, 1, 2) */
  int32_t* all_ims;          /* [m]   [active; fixed] subdomain ids (first n_used written) */
  int32_t* pair_point;       /* [s]   fbpx_worklist arrays, tile order */
  int32_t* pair_seg;         /* [s] */
  int32_t* pair_slot;        /* [s] */
  int32_t* point_seg_start;  /* [n + 1] */
  int32_t* seg_pair_start;   /* [s + 1] (first n_segs + 1 meaningful) */
  int32_t* chunk_model;      /* [max_chunks] */
  int32_t* chunk_start;      /* [max_chunks + 1] */
  uint8_t* chunk_train;      /* [max_chunks] 1 for chunks of trained subdomains */
  int32_t* row_chunk_start;  /* [m + 1] chunks of trained row k (first n_active + 1 written) */
  int32_t* fbin_chunk_start; /* [fbins + 1], or [max_chunks + 1] when fbins == 0 */
  int32_t* bbin_chunk_start; /* [bbins + 1], or [max_chunks + 1] when bbins == 0 */
} fbpx_takes_out;

typedef struct {
  int64_t n_pairs;
  int32_t n_segs, n_used, n_active, n_chunks, n_fbins, n_bbins, npou;
} fbpx_takes_info;

/* Upper bound of the chunk count for the chunk arrays (max_chunks). */
int64_t fbpx_takes_max_chunks(int32_t m, int64_t n_pairs, int32_t chunk, int32_t fbins, int32_t bbins);
/* Work-space bytes of the two calls (-1 for bad sizes). */
int64_t fbpx_takes_work_bytes(int32_t n, int32_t xd, int32_t m, int64_t n_pairs, int32_t fbins, int32_t bbins);
int fbpx_takes_count(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                     int32_t m, int64_t* offsets, void* work, int64_t work_bytes, int64_t* n_pairs);
/* offsets and n_pairs: from fbpx_takes_count on the same x and boxes. */
int fbpx_build_worklist(void* stream, const float* x, int32_t n, int32_t xd, const float* xmin, const float* xmax,
                        int32_t m, const int32_t* active, const int32_t* pou, const int64_t* offsets, int64_t n_pairs,
                        int32_t chunk, int32_t fbins, int32_t bbins, int64_t max_chunks, void* work,
                        int64_t work_bytes, const fbpx_takes_out* out, fbpx_takes_info* info);

/* ---------------------------------------------------------------- forward */

/* Window jets summed per segment: wsum[seg][J] = sum over the segment's pairs
 * of the jets of window_fn (decompositions.py:196-199, windows.py:25-35).
 * Parameter-independent; computed once per work list.  pair_scratch:
 * [n_pairs][J] caller-provided work space (e.g. the forward pair buffer),
 * written by slot. */
int fbpx_pou_prepare(void* stream, const fbpx_net* net, const float* dom,
                     const float* x, const fbpx_worklist* wl, float* pair_scratch,
                     float* wsum);

/* Per-pair jets of u*w, where u = unnorm(FCN(norm(x))) and w = window(x),
 * i.e. the vmapped FBPINN_model_inner (trainers.py:113-118,158) with the
 * nested-jvp tangents of _get_ujs (trainers.py:213-247) carried in registers.
 * theta: [m][P]; pairbuf: [n_pairs][J], row = the pair's slot (segment
 * order), so that fbpx_pou_reduce reads each segment's rows contiguously.
 * One workgroup per forward bin; the bins must cover every chunk. */
int fbpx_forward(void* stream, const fbpx_net* net, const float* theta,
                 const float* dom, const float* x, const fbpx_worklist* wl,
                 float* pairbuf);

/* POU reduction: per point, sum over pou of (sum_pairs u*w) / (sum_pairs w)
 * with the quotient rule on jets, divided by npou (trainers.py:162-170).
 * Sorted, atomic-free.  ujet: [J][n_points] (column-major jets).  A, B: NULL,
 * or the [J][n_points] jets of an affine constraining_fn A + B u (see
 * fbpx_affine_jets), which is then applied to every column of every point
 * (constraining_fn inside the model, trainers.py:174) before ujet is written. */
int fbpx_pou_reduce(void* stream, const fbpx_net* net, const fbpx_worklist* wl,
                    const float* pairbuf, const float* wsum, const float* A, const float* B,
                    float* ujet);

/* ---------------------------------------------------------------- backward */

/* Transpose of fbpx_pou_reduce: from d loss / d ujet ([J][n_points]) to
 * d loss / d (segment sum of u*w jets), gseg: [n_segs][J].  B: NULL, or the
 * affine map's B jets as given to fbpx_pou_reduce (its transpose is applied
 * first).  g_scale: NULL, or a DEVICE float by which g_ujet is multiplied as
 * it is read (d L / d loss, when g_ujet was computed at unit upstream
 * gradient by fbpx_residual_loss_grad). */
int fbpx_pou_backward(void* stream, const fbpx_net* net, const fbpx_worklist* wl,
                      const float* wsum, const float* B, const float* g_ujet, const float* g_scale,
                      float* gseg);

/* Reverse pass through the jets to the subdomain weights (the part of
 * value_and_grad, trainers.py:292, below the segment sums).  Recomputes
 * activations.  One workgroup per backward bin; chunks outside the bins or
 * with chunk_train[c] == 0 (fixed subdomains, trainers.py:252-256 -- no
 * gradient) are skipped and their partial row is left untouched.
 * partial: [n_chunks][P] per-chunk gradient rows.  g_scale: NULL, or a
 * DEVICE float by which gseg is multiplied (d L / d loss when gseg was
 * written at unit upstream gradient by fbpx_pou_loss). */
int fbpx_backward(void* stream, const fbpx_net* net, const float* theta,
                  const float* dom, const float* x, const fbpx_worklist* wl,
                  const uint8_t* chunk_train, const float* gseg, const float* g_scale, float* partial);

/* fbpx_backward with the optimizer step of fbpx_adam_rows_partial folded
 * in, for one-hidden-layer networks whose every trained subdomain row is
 * exactly one chunk (a caller precondition: rows of the row list with no
 * chunk or several would be updated differently): the workgroup that
 * finishes a chunk applies optax Adam (as fbpx_adam_rows, lr, b1, b2, eps)
 * to its subdomain's row of theta, mu and nu in place, with the chunk's
 * gradient row as the row's gradient, instead of writing a partial row, and
 * the last workgroup (or wave, with bwd_wave_bins) to finish adds 1 to count
 * (count_done: int32 0, left 0).  theta, mu, nu and count equal
 * fbpx_backward followed by fbpx_adam_rows_partial bit for bit.  Untrained
 * chunks (chunk_train 0) are left alone.  Replaces what those two calls
 * replace: the reverse mode of value_and_grad(FBPINN_loss) below the
 * segment sums and optax.adam's update (fbpinns/trainers.py:292-295), in
 * one launch instead of two. */
int fbpx_backward_adam(void* stream, const fbpx_net* net, float* theta, const float* dom, const float* x,
                       const fbpx_worklist* wl, const uint8_t* chunk_train, const float* gseg, const float* g_scale,
                       float* mu, float* nu, int32_t* count, int32_t* count_done, double lr, double b1, double b2,
                       double eps);

/* Resident 256-thread workgroups per compute unit of fbpx_forward's and
 * fbpx_backward's kernel for this network and jet schema (register and LDS
 * limits), for sizing the persistent launches (one workgroup per bin).  No
 * counterpart in the reference: its XLA compilation picks launch shapes
 * itself. */
int fbpx_occupancy(const fbpx_net* net, int32_t* fwd_per_cu, int32_t* bwd_per_cu);

/* Per-row gradient from the per-chunk partial rows: for k < n_rows,
 *   grad[rows[k]][e] = sum over c in row_chunk_start[k] .. row_chunk_start[k+1]-1
 *                      (ascending) of partial[c][e].
 * Rows not listed are not written.  grad: [m][P]. */
int fbpx_reduce_rows(void* stream, int32_t P, const float* partial,
                     const int32_t* rows, const int32_t* row_chunk_start,
                     int32_t n_rows, float* grad);

/* ---------------------------------------------------------------- Adam */

/* optax.adam on subdomain rows (trainers.py:294-295), for each listed row r
 * (rows[k], k < n_rows) with gradient grad[r]:
 *   mu = (1-b1) g + b1 mu;  nu = (1-b2) g^2 + b2 nu;
 *   p += -lr * (mu/(1-b1^c)) / (sqrt(nu/(1-b2^c)) + eps),  c = *count + 1.
 * Rows not listed (fixed or discarded subdomains) are untouched.
 * n_rows * P < 2^31 - 255 (FBPX_EINVAL otherwise).
 * theta, mu, nu, grad: [m][P].  The hyper-parameters are doubles so that
 * f32(1 - b1) and f32(-lr) are formed as the reference forms them (Python
 * floats meeting FP32 arrays).  count_done: NULL, or a DEVICE int32 that is 0
 * before the call (the call leaves it 0): then the call also does
 * *count += 1 after every workgroup has read the count, inside the launch
 * when it has at most 256 workgroups (its last workgroup), else by a second
 * launch (the step's last Adam call carries it, in place of
 * fbpx_count_increment).  Calls that may run at the same time (other
 * streams) need distinct count_done counters. */
int fbpx_adam_rows(void* stream, float* theta, float* mu, float* nu,
                   int32_t P, const float* grad, const int32_t* rows, int32_t n_rows,
                   int32_t* count, int32_t* count_done, double lr, double b1, double b2, double eps);

/* fbpx_reduce_rows followed by fbpx_adam_rows in one launch: the gradient of
 * each listed row is summed from the per-chunk partial rows (partial,
 * row_chunk_start as for fbpx_reduce_rows, same order) and applied at once,
 * without a gradient array.  Same results bit for bit.  No counterpart in the
 * reference beyond the two steps it fuses (trainers.py:292-295). */
int fbpx_adam_rows_partial(void* stream, float* theta, float* mu, float* nu, int32_t P,
                           const float* partial, const int32_t* rows, const int32_t* row_chunk_start,
                           int32_t n_rows, int32_t* count, int32_t* count_done, double lr, double b1,
                           double b2, double eps);

/* The same update on a flat vector p[n] with gradient g[n] (the non-subdomain
 * trainable leaves, e.g. a learnable PDE coefficient, problems.py:232-234). */
int fbpx_adam_flat(void* stream, float* p, float* mu, float* nu, const float* g,
                   int32_t n, int32_t* count, int32_t* count_done, double lr, double b1, double b2,
                   double eps);

/* *count += 1 (optax's shared int32 step count, trainers.py:531,640). */
int fbpx_count_increment(void* stream, int32_t* count);

/* Deterministic FP32 mean of x[n] into out[0] (jnp.mean in the problems'
 * loss_fn, problems.py:123,204,265,326,405).  Two fixed-shape stages with
 * an LDS tree, no atomics and no memset (safe inside a captured graph).
 * scratch: [256] floats of caller work space. */
int fbpx_mean(void* stream, const float* x, int64_t n, float* scratch, float* out);

/* ---------------------------------------------------------------- constraining function */

/*
 * Affine constraining function on jets.  When the problem's
 * constraining_fn(x, u) = A(x) + B(x) u with A, B independent of trainable
 * parameters (the hard-constraint ansatzes of problems.py:160-175, 289-300,
 * 359-370), the derivatives that the reference's nested jvp takes of the
 * constrained solution (trainers.py:174 inside the model differentiated by
 * _get_ujs/jacfwd, trainers.py:213-247) follow from the jets of A, B and of
 * the POU-combined output u~ by the Leibniz rule:
 *   value         A0 + B0 q0
 *   d/dx_i        Aa + Ba q0 + B0 qa
 *   d2/dx_i^2     Ab + Bb q0 + 2 Ba qa + B0 qb      (a = col1(i), b = col2(i)).
 * A, B, q: jets in the fbpx_pou_reduce layout, column c of point k at
 * c * ld + k (pass the pointers offset to the first point of the constraint).
 * req: HOST array of n_req <= FBPX_MAX_REQ codes, 0 = value, 1 + i = d/dx_i,
 * 1 + xd + i = d2/dx_i^2 (the schema must carry it).  out: HOST array of
 * n_req device pointers, out[r][k] for k < n.
 */
#define FBPX_MAX_REQ 16
int fbpx_affine_jets(void* stream, const fbpx_net* net, int32_t n, int32_t ld, const float* A,
                     const float* B, const float* q, int32_t n_req, const int32_t* req,
                     float* const* out);

/* Transpose of fbpx_affine_jets w.r.t. q: gq[c * ld + k] = sum_r
 * d out_r / d q_c * gout[r][k] for every column c < J of the n points
 * (columns that no request touches are written as 0). */
int fbpx_affine_jets_backward(void* stream, const fbpx_net* net, int32_t n, int32_t ld,
                              const float* B, int32_t n_req, const int32_t* req,
                              const float* const* gout, float* gq);

/*
 * Fused pointwise-residual loss, the form of the problems' loss_fn
 * (problems.py:123-124, 204, 265, 326, 405: a weighted mean of squared
 * pointwise residuals, replacing their chain of jnp ops and jnp.mean):
 *   r[k]  = sum over t < n_terms of coef[t] * F(fa[t], k) * F(fb[t], k),
 *   F(f, k) = col[f][k] for f >= 0 (col[f][col_idx[f][k]] when col_idx and
 *             col_idx[f] are not NULL), 1 for f < 0 (constant terms use
 *             fa = fb = -1),
 *   loss  = weight / n * sum over k < n of r[k]^2  (+ loss[0] on entry when
 *           accumulate != 0: the sum of several residuals without a kernel).
 * col: HOST array of n_cols <= FBPX_MAX_REQ device pointers to n floats (the
 * constraint's jets and per-point data such as sources or targets) or, for a
 * gathered column, to the vector it gathers from (a per-instance trainable
 * coefficient indexed by each point's instance); col_idx: NULL or a HOST
 * array of n_cols device int32 index pointers (NULL entries: plain columns,
 * the index is the point's; the device does not check the index values,
 * each of which must address an element of its gathered vector); fa, fb:
 * HOST arrays of n_terms <= FBPX_MAX_TERMS column indices; coef: DEVICE array of n_terms floats (trainable coefficients stay
 * on the device).  Writes r[n] and loss[0]; scratch: FBPX_RESIDUAL_BLOCKS floats.
 * done: NULL (a second one-workgroup launch sums the per-workgroup partials)
 * or a DEVICE int32 that is 0 before the call and is left 0: then a launch of
 * at most 256 workgroups sums them in its last workgroup (one launch instead
 * of two; larger grids keep the second launch).  Either way the sum has a
 * fixed order (deterministic, no memset).  Calls that may run at the same
 * time (other streams) need distinct done counters and scratch arrays.
 */
#define FBPX_MAX_TERMS 8
#define FBPX_RESIDUAL_BLOCKS 1024
int fbpx_residual_loss(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                       const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                       const float* coef, float weight, float* r, float* scratch, float* loss, int32_t accumulate,
                       int32_t* done);

/* fbpx_residual_loss and, in the same launch, the loss's gradient with
 * respect to its columns at unit upstream gradient (d L / d loss = 1, the
 * train step's seed): gcol[f][k] = 2 * weight / n * r[k] * d r[k] / d F(f, k)
 * for every f with gcol[f] != NULL, the arithmetic of
 * fbpx_residual_loss_backward with g_loss[0] = 1, bit for bit.  The caller
 * applies the actual d L / d loss where gcol is consumed (fbpx_pou_backward's
 * g_scale, fbpx_segment_sum's scale), so one launch replaces the loss and its
 * transpose when no coefficient is trained.  Same arguments as
 * fbpx_residual_loss, plus gcol: HOST array of n_cols device pointers to n
 * floats or NULL entries (the table itself must not be NULL). */
int fbpx_residual_loss_grad(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                            const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                            const float* coef, float weight, float* r, float* scratch, float* loss,
                            int32_t accumulate, int32_t* done, float* const* gcol);

/*
 * The POU reduction, the affine constraining function, the fused residual
 * losses and the POU transpose at unit upstream gradient in one launch, one
 * thread per point: fbpx_pou_reduce, then fbpx_residual_loss_grad for every
 * residual, then fbpx_pou_backward of the jet gradients, with the same
 * per-point arithmetic (ujet, r, the gathered columns' gradients and gseg
 * equal theirs bit for bit; the loss sums the same squares in another fixed
 * order).  Every (residual, jet column) pair reads a column of its own
 * constraint, and no two residuals read the same jet column of a constraint
 * (its gradient is written, not added).  The caller applies d L / d loss to
 * gseg in fbpx_backward (g_scale) and to the gathered columns' gradients in
 * fbpx_segment_sum (scale).
 *
 * Residual i covers the points [offset, offset + n) of the work list:
 *   r[k] = sum over t < n_terms of coef[t] F(fa[t], k) F(fb[t], k)  (k < n),
 *   F(f, k) = column jet[f] of point offset + k's constrained jets when
 *             jet[f] >= 0, else col[f][k] (col[f][col_idx[f][k]] when
 *             col_idx[f] is not NULL); 1 for f < 0,
 *   loss += weight / n * sum over k of r[k]^2  (residuals in order; the
 *           first adds to loss[0] on entry when accumulate != 0),
 *   gcol[f][k] = 2 weight / n * r[k] * d r[k] / d F(f, k) for memory columns
 *             (jet[f] < 0) with gcol[f] != NULL.
 * ujet: [J][n_points] out (the constrained jets); gseg: [n_segs][J] out;
 * scratch: ceil(n_points / 256) * n_res floats; done: as in
 * fbpx_residual_loss (the last workgroup sums the partials when the grid has
 * at most 256 workgroups, one workgroup writes them directly, else a second
 * one-workgroup launch does).
 */
#define FBPX_POU_LOSS_RES 4
#define FBPX_POU_LOSS_COLS 8
typedef struct {
  int32_t offset, n;                       /* points [offset, offset + n) */
  int32_t n_cols, n_terms;                 /* <= FBPX_POU_LOSS_COLS, <= FBPX_MAX_TERMS */
  float weight;
  const float* coef;                       /* DEVICE, n_terms */
  int32_t fa[FBPX_MAX_TERMS], fb[FBPX_MAX_TERMS];
  int32_t jet[FBPX_POU_LOSS_COLS];         /* jet column, or -1: col[f] */
  const float* col[FBPX_POU_LOSS_COLS];    /* DEVICE per-point data (n) or gathered vector */
  const int32_t* col_idx[FBPX_POU_LOSS_COLS]; /* DEVICE gather index (n, each value inside col[f]) or NULL */
  float* gcol[FBPX_POU_LOSS_COLS];         /* DEVICE per-point gradient (n) or NULL */
  float* r;                                /* DEVICE, n */
} fbpx_residual;
int fbpx_pou_loss(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* pairbuf,
                  const float* wsum, const float* A, const float* B, int32_t n_res, const fbpx_residual* res,
                  float* ujet, float* gseg, float* scratch, int64_t scratch_len, float* loss, int32_t accumulate,
                  int32_t* done);

/*
 * Several residual losses in one launch: a problem's loss_fn summing
 * mean-squared terms over its constraints, e.g. the inverse oscillator's
 * physics and data terms (fbpinns/problems.py:258-271).  Residual i runs on
 * workgroups of its own, as fbpx_residual_loss_grad (some gcol[f] given) or
 * fbpx_residual_loss (every gcol[f] NULL, in every residual) runs it, and r
 * and gcol equal those launches' bit for bit; so does the loss, which is
 * formed as those launches would form it one after another (the first adds
 * to loss[0] on entry when accumulate != 0, each later one adds its term).
 * res[i]: fbpx_residual with offset 0 and every jet[f] = -1 (all columns
 * are device memory, col[f] or col[f][col_idx[f][k]]).  scratch: one float
 * per workgroup, at most n_res * FBPX_RESIDUAL_BLOCKS; done: as in
 * fbpx_residual_loss (the last workgroup sums the partials when the grid has
 * at most 256 workgroups, else a second one-workgroup launch does).
 */
int fbpx_residual_losses(void* stream, int32_t n_res, const fbpx_residual* res, float* scratch,
                         int64_t scratch_len, float* loss, int32_t accumulate, int32_t* done);

/* Its transpose.  With s = g_loss[0] * 2 * weight / n (g_loss: DEVICE
 * pointer to d L / d loss):
 *   gcol[f][k] = s * r[k] * d r[k] / d F(f, k)     for every f with gcol[f] != NULL
 *                (per point, also for a gathered column: fbpx_segment_sum
 *                then sums it per vector entry),
 *   gcoef[t]   = s * sum over k of r[k] * (product of term t at k)   (gcoef may be NULL).
 * gcol: HOST array of n_cols device pointers (NULL entries: per-point data,
 * no gradient); scratch: FBPX_RESIDUAL_BLOCKS * n_terms floats; done as in
 * fbpx_residual_loss (for the gcoef sums). */
int fbpx_residual_loss_backward(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                                const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa,
                                const int32_t* fb, const float* coef, const float* r, float weight,
                                const float* g_loss, float* const* gcol, float* scratch, float* gcoef,
                                int32_t* done);

/* fbpx_residual_loss without the caller's work space: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops).  Same arguments and results as
 * fbpx_residual_loss with done = NULL (the partial sums are added by a second
 * one-workgroup launch), bit for bit.  The work space is a file-scope device
 * array of FBPX_RESIDUAL_BLOCKS floats (4 KiB per GPU, held until the process
 * ends), one partial sum per workgroup; the launch caps its grid at that
 * many workgroups, so every n >= 1 fits.  No allocation, no memset, nothing
 * that must start at zero.  Calls that overlap on one GPU (other streams or
 * threads) share the array and race; calls in one stream's order do not. */
int fbpx_residual_loss_noalloc(void* stream, int64_t n, int32_t n_cols, const float* const* col,
                               const int32_t* const* col_idx, int32_t n_terms, const int32_t* fa, const int32_t* fb,
                               const float* coef, float weight, float* r, float* loss, int32_t accumulate);

/*
 * Deterministic segmented sum, the transpose of a per-point gather
 * y[k] = v[idx[k]] of a trainable vector (a per-instance coefficient such as
 * each oscillator's learnt damping; the reverse mode of jax's gather,
 * a scatter-add, which the reference would run for such a loss):
 *   out[s] = sum over j in [seg_start[s], seg_start[s+1]) of x[perm[j]]
 * (perm NULL: x[j]) for s < n_seg, in increasing j, no atomics, times
 * scale[0] when scale (DEVICE float) is not NULL.  With perm a stable
 * argsort of idx and seg_start its segment offsets, out = gather^T x.
 * seg_start: DEVICE, n_seg + 1 int32; perm: DEVICE int32 or NULL.
 */
int fbpx_segment_sum(void* stream, int32_t n_seg, const int32_t* seg_start, const int32_t* perm, const float* x,
                     const float* scale, float* out);

/* fbpx_pou_backward and fbpx_segment_sum in one launch, each on workgroups
 * of its own with its own arithmetic (outputs equal the two calls' bit for
 * bit), both scaled by g_scale: the backward of a fused loss with a
 * gathered trainable vector, whose two transposes are independent.  No
 * counterpart in the reference beyond those of the two calls. */
int fbpx_pou_backward_segment_sum(void* stream, const fbpx_net* net, const fbpx_worklist* wl, const float* wsum,
                                  const float* B, const float* g_ujet, const float* g_scale, float* gseg,
                                  int32_t n_seg, const int32_t* seg_start, const int32_t* perm, const float* x,
                                  float* out);

/* ---------------------------------------------------------------- misc */

/* Parameters per subdomain row for a network shape (0 if unsupported). */
int32_t fbpx_param_count(const fbpx_net* net);
/* 1 if kernels for this network/jet shape are compiled in. */
int32_t fbpx_supported(const fbpx_net* net);
/* Text of the last error on this thread ("" if none). */
const char* fbpx_last_error(void);
/* Library version string. */
const char* fbpx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FBPX_H */
