/*
 * mivis.h — C ABI of libmivis.so, the MI355X (gfx950) hot path of mlx-vis:
 * exact brute-force kNN, the NNDescent local join, and the UMAP fuzzy graph
 * and edge-wise layout SGD.
 *
 * Conventions (every entry point):
 *   - returns 0 on success, a negative mv_status on failure; the message of
 *     the last failure on the calling thread is returned by mv_last_error();
 *   - every array pointer is a caller-owned DEVICE pointer, row-major and
 *     contiguous; the library never frees caller memory;
 *   - scratch memory ("ws") is caller-provided, sized by the matching
 *     *_ws_bytes query; it needs no initialisation and is not read back;
 *   - all work is enqueued on `stream` (a hipStream_t passed as void*; NULL is
 *     the default stream). Only outputs documented as HOST pointers imply a
 *     stream synchronisation.
 *   - calls that share a workspace must not overlap in time.
 * Citations name the reference function each entry point replaces
 * (paths relative to the mlx-vis source tree).
 */
#ifndef MIVIS_H
#define MIVIS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MV_ABI_VERSION 1

enum mv_status {
  MV_OK = 0,
  MV_EINVAL = -1,        /* invalid argument (maps to Python ValueError) */
  MV_EHIP = -2,          /* a HIP runtime call failed */
  MV_EWORKSPACE = -3,    /* workspace smaller than *_ws_bytes reported */
  MV_EUNSUPPORTED = -4,  /* valid request outside what the kernels cover */
  MV_EINJECTED = -5      /* fault injected through MIVIS_FAULT (tests only) */
};

enum mv_dtype {
  MV_F32 = 0,     /* FP32 MFMA distance screen */
  MV_BF16 = 1,    /* BF16 MFMA screen, FP32 accumulate */
  MV_FP8 = 2,     /* FP8 e4m3 MFMA screen, per-row absmax scale, FP32 accumulate */
  MV_BF16X3 = 3   /* split BF16 screen: x = hi + lo (two BF16 values), hi*hi + hi*lo + lo*hi
                     on BF16 MFMA with FP32 accumulate (error ~2^-16 relative) */
};

int mv_abi_version(void);
const char* mv_last_error(void);
/* Device properties of the current device: CU count, LDS bytes per block,
 * wavefront size. Host pointers. */
int mv_device_info(int* cu_count, int* lds_per_block, int* wave_size);

/* ------------------------------------------------------------------ kNN --
 * Replaces compute_knn(..., method="brute") — mlx_vis/_knn.py:47-88
 * (squared-distance tile _knn.py:60-63, self mask _knn.py:64-69, argsort
 * top-k _knn.py:71-72, sqrt _knn.py:86-87).
 *
 * For each query row q_r (r < nq) returns the k nearest rows of X in the
 * canonical order: ascending exact squared Euclidean distance, ties broken by
 * the lower index of X. The exact distance of a pair is the IEEE-754 double
 * sum, in order t = 0..d-1 and without fused multiply-add, of
 * ((double)q_t - (double)x_t)^2. The screen (dtype) only nominates
 * rerank*k candidates per row; an FP32 direct-difference re-rank orders them,
 * FP64 decides pairs that FP32 cannot separate, and a per-row certificate
 * proves that no point outside the candidates belongs to the top k. Rows whose
 * certificate fails are recomputed by an exhaustive FP64 scan, so the result
 * does not depend on dtype.
 *
 * Self exclusion (the +1e30 diagonal of _knn.py:64-69): if q_self != NULL,
 * row r never returns index q_self[r] (entries < 0 exclude nothing); else if
 * self_base >= 0 row r never returns index self_base + r.
 * out_dist: squared distance (euclid == 0) or its square root (euclid != 0).
 * out_stats (device, nullable, 4 x int32): [0] rows recomputed by the FP64
 * scan, [1] rows whose order needed the FP64 tie check, [2] rows that failed
 * the first certificate, [3] reserved.
 * Limits: k <= number of selectable rows; 1 <= k <= 64 for any queries,
 * 1 <= k <= 1024 for a self kNN (Q == X, nq == n) on the symmetric screen
 * (otherwise k > 64 returns MV_EUNSUPPORTED). The candidate count is
 * min(rerank*k, 128) (at least k) with the symmetric self-kNN screen (see
 * mv_knn_set_symmetric), min(rerank*k, 64) with the fused screen; a row whose
 * k-th neighbour is not separated from the last candidate by the screen's
 * error bound goes to the FP64 scan (BF16 / FP8: first to a split-BF16
 * re-screen with 64 candidates). From k = 57 on, a self kNN on the symmetric
 * screen instead selects from its n x n matrix: the k-th screen value by radix
 * select, then the canonical distances of every column within twice the
 * screen's error bound of it (no certificate, no scan; BF16 and FP8 requests
 * run the split-BF16 screen there; out_stats stays zero). d <= 3 with k <= 64
 * uses an exact grid search instead of a screen (dtype and rerank are then
 * ignored).
 */
int mv_knn_ws_bytes(int64_t nq, int64_t n, int64_t d, int k, int dtype, int rerank,
                    int same_qx, size_t* out_bytes);
int mv_knn_brute(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k,
                 int dtype, int rerank, const int32_t* q_self, int64_t self_base, int euclid,
                 int32_t* out_idx, float* out_dist, int32_t* out_stats,
                 void* ws, size_t ws_bytes, void* stream);

/* mv_knn_brute without a workspace argument and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, results and errors as mv_knn_brute,
 * always with the fused screen (the symmetric screen's n x n matrix does not
 * fit). The workspace is a static 128 MiB device array of the library, one per
 * GPU, held until the process ends: a call needing more (mv_knn_ws_bytes with
 * mv_knn_set_symmetric(0) > 134217728, e.g. nq = n = 8192 at d = 784 fits,
 * n = 100,000 at d = 784 does not) fails with MV_EUNSUPPORTED and a message.
 * Calls on one GPU share the array: calls that overlap in time on the same
 * GPU (other streams, other host threads) race. Phase timing is not
 * recorded. */
int mv_knn_brute_noalloc(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k,
                         int dtype, int rerank, const int32_t* q_self, int64_t self_base, int euclid,
                         int32_t* out_idx, float* out_dist, int32_t* out_stats, void* stream);

/* Phase timing of mv_knn_brute on the calling thread: with timing on, each
 * call records HIP events on its stream; mv_knn_last_phase_ms waits for the
 * last timed call and writes milliseconds of [prepass, screen, rerank,
 * exact scan] to the host array out4. */
int mv_knn_set_timing(int on);
int mv_knn_last_phase_ms(float* out4);
/* The screen phase of the last timed call split by kernel: out2[0] = the
 * distance kernel (k_dist_sym for the symmetric screen, else k_screen),
 * out2[1] = the row selection (k_select; 0 for the fused screen), in ms;
 * *out_sym = 1 if the symmetric screen ran. */
int mv_knn_last_screen_ms(float* out2, int32_t* out_sym);

/* Symmetric screen for self kNN (Q == X, every dtype, 4,096 <= n and
 * n_pad^2 * 4 bytes <= 48 GiB; on by default): only the tiles
 * on and above the diagonal of the distance matrix are multiplied, the full
 * matrix is kept in the workspace (mv_knn_ws_bytes includes it) and each row's
 * candidate list is selected from it; the returned neighbours and distances
 * are those of the fused screen.
 * Off: every self kNN uses the fused screen. Process-wide; set it before
 * mv_knn_ws_bytes and keep it until the matching mv_knn_brute has run.
 * Returns the previous setting (1 on, 0 off), not a status. */
int mv_knn_set_symmetric(int on);

/* Row squared norms: out[i] = sum_t X[i,t]^2 in FP32 (mx.sum(X*X, 1) at
 * mlx_vis/_knn.py:49 and mlx_vis/_nndescent/nndescent.py:51,77). */
int mv_row_sqnorms(const float* X, int64_t n, int64_t d, float* out, void* stream);


/* ----------------------------------------------------------- NNDescent --
 * Replaces NNDescent.build (mlx_vis/_nndescent/nndescent.py:34-403), the
 * whole iteration schedule included: init distances and sort (53-66), the
 * 100-D projection for iteration 0 and the 300-D one for iteration 1 (68-82,
 * 227-235), FP16 dots in iterations 1-4 (236, 615-641), reverse candidates
 * while update_frac >= 0.10 (241, 406-436), the full path while
 * update_frac >= 0.5 and the NEW-first converging path below it (103-220),
 * dedup/top-k (439-486), the FP32 recompute after iteration 4 (376-378),
 * positional change flags and the early stop at update_frac < delta (380-398).
 * The caller supplies the random draws of the reference's NumPy stream:
 * init_idx (n, k) (nndescent.py:54-56), P1 (d, p1) and P2 (d, p2) (NULL with
 * p = 0 when the projection is not smaller than d). out_changed (HOST, n_iters
 * entries, nullable): changed slots per iteration; out_iters (HOST, nullable):
 * iterations run. out_dist: Euclidean. Limits: k <= 128 (any max_candidates;
 * rows with more than 1,024 candidate slots are merged 1,024 at a time; from
 * k = 65 each lane holds two of the row's neighbours), n * k < 2^31,
 * d <= 4096 (d <= 3968 for k > 64).
 * One device-to-host read of a counter per iteration synchronises `stream`.
 */
int mv_nnd_ws_bytes(int64_t n, int64_t d, int k, int64_t p1, int64_t p2, size_t* out_bytes);
int mv_nnd_build(const float* X, int64_t n, int64_t d, int k, int n_iters, int max_candidates, float delta,
                 const int32_t* init_idx, const float* P1, int64_t p1, const float* P2, int64_t p2, int verbose,
                 int32_t* out_idx, float* out_dist, int64_t* out_changed, int* out_iters, void* ws,
                 size_t ws_bytes, void* stream);
/* Reverse candidates alone (_compute_reverse_candidates, nndescent.py:406-436):
 * rev[j, t] = source row of the t-th edge pointing at j in flat (row-major)
 * order, t < k; slots beyond the in-degree hold 0. k <= 128, n * k < 2^31.
 * ws sized by mv_nnd_ws_bytes(n, 1, k, 0, 0). */
int mv_nnd_reverse(const int32_t* idx, int64_t n, int k, int32_t* rev, void* ws, size_t ws_bytes, void* stream);
/* Squared distances of (i, idx[i, c]) as _gather_dists (half == 0,
 * nndescent.py:573-612) or _gather_dists_half (half != 0, 615-641: FP16
 * inputs, FP32 accumulation, FP16-rounded dot, FP32 norms). k <= 128. ws
 * sized by mv_nnd_ws_bytes(n, d, k, 0, 0). */
int mv_nnd_gather_dists(const float* X, int64_t n, int64_t d, const int32_t* idx, int k, int half,
                        float* out_dist, void* ws, size_t ws_bytes, void* stream);

/* ----------------------------------------------------------------- UMAP --
 * Fuzzy simplicial set, prune and schedule — replaces
 * UMAP._fuzzy_simplicial_set (mlx_vis/_umap/umap.py:199-270) and the
 * schedule precompute of UMAP._optimize (umap.py:399-404).
 * Inputs: the kNN graph (n, k) int32 indices and FP32 Euclidean distances.
 * Outputs over the dense (n, k) slot array (slot i*k + c is the forward edge
 * i -> knn_idx[i, c]): w_sym (symmetrised weight w + w' - w*w', umap.py:262),
 * eps (epochs_per_sample, -1 for slots pruned below max(w_sym)/n_epochs,
 * umap.py:265-270, 402-403) and epn (epochs_per_next, initialised to eps;
 * +inf for pruned slots). counters (device, 2 x uint32): [0] kept edges,
 * [1] bit pattern of max(w_sym). sigma_out / rho_out (device, n, nullable):
 * per-row bandwidth and rho (umap.py:208-236). k <= 256 (above 64 the
 * bisection runs one wavefront per row and sums in a different order).
 */
int mv_umap_graph_ws_bytes(int64_t n, int k, size_t* out_bytes);
int mv_umap_fuzzy_graph(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k, int n_epochs,
                        float* w_sym, float* eps, float* epn, uint32_t* counters, float* sigma_out,
                        float* rho_out, void* ws, size_t ws_bytes, void* stream);

/* In-edge lists of the kept slots (eps > 0), grouped by tail node: entries
 * t < (number of kept slots) hold the slot in_slots[t] and its tail
 * in_node[t]; nodes appear in ascending order, the slots of a node with at
 * most 16,384 in-edges in ascending order (of a node with more in no fixed
 * order). Used by mv_umap_epochs to sum tail updates without per-edge atomics,
 * in an order that makes a run repeat bit for bit.
 * ws: 3 * align256(4n) + align256(4 * ceil(n / 1024)) bytes. */
int mv_umap_in_edges(const int32_t* tail, const float* eps, int64_t n, int k, int32_t* in_slots,
                     int32_t* in_node, void* ws, size_t ws_bytes, void* stream);

/* Layout epochs [e0, e1) — replaces the epoch loop of UMAP._optimize
 * (umap.py:423-447) with UMAP._sgd_step (umap.py:359-387). Y (n, dim) is
 * updated in place. Each epoch e is synchronous (all gradients read Y as of
 * the start of e): slots with eps > 0 and epn <= e are active (then
 * epn += eps); an active edge moves head and tail with the clipped (+-4)
 * attractive gradient and the head with neg_rate clipped repulsive gradients
 * against nodes drawn uniformly from [0, n) by Philox-4x32-10 keyed by seed;
 * the step size is lr * (1 - e / n_epochs). tail = the kNN indices (n, k);
 * (in_slots, in_node, n_in) from mv_umap_in_edges. ws sized by
 * mv_umap_epochs_ws_bytes. dim <= 8, k <= 256. The same inputs give the same
 * Y to the bit on every run (unless a node has more than 16,384 in-edges). */
int mv_umap_epochs_ws_bytes(int64_t n, int k, int dim, size_t* out_bytes);
int mv_umap_epochs(float* Y, int64_t n, int dim, const int32_t* tail, int k, const float* eps, float* epn,
                   const int32_t* in_slots, const int32_t* in_node, int64_t n_in, int e0, int e1, int n_epochs,
                   int neg_rate, float lr, float a, float b, uint64_t seed, void* ws, size_t ws_bytes,
                   void* stream);

/* mv_umap_epochs with a per-node mask (no counterpart in the reference; used
 * to refine chosen points of an embedding against fixed ones): movable
 * (device, n bytes, nullable) — nodes with movable[i] == 0 keep their
 * position, every gradient is still computed from the full embedding. */
int mv_umap_epochs_masked(float* Y, int64_t n, int dim, const int32_t* tail, int k, const float* eps, float* epn,
                          const int32_t* in_slots, const int32_t* in_node, int64_t n_in, int e0, int e1,
                          int n_epochs, int neg_rate, float lr, float a, float b, uint64_t seed,
                          const uint8_t* movable, void* ws, size_t ws_bytes, void* stream);

/* One synchronous step with explicit edge and negative lists, exactly
 * UMAP._sgd_step (umap.py:359-387): Y_out = Y_in + updates, where every update
 * is computed from Y_in (Y_out must not alias Y_in). For parity checks. */
int mv_umap_sgd_step(float* Y_out, const float* Y_in, int64_t n, int dim, const int32_t* ef, const int32_t* et,
                     int64_t A, const int32_t* neg_from, const int32_t* neg_to, int64_t n_neg, float alpha,
                     float a, float b, void* stream);

/* Spectral initialisation — replaces UMAP._spectral_init's power iteration
 * (umap.py:298-340): `iters` products of D^-1/2 W D^-1/2 (degrees from the
 * heads of the kept slots, umap.py:307-314) with re-orthonormalisation of the
 * nc columns after each product. V (n, nc) holds the start vectors on entry
 * and the orthonormal result on return. 2 <= nc <= 9. */
int mv_umap_spectral_ws_bytes(int64_t n, int k, int nc, size_t* out_bytes);
int mv_umap_spectral(const float* w_sym, const float* eps, const int32_t* tail, int64_t n, int k, int nc,
                     int iters, float* V, void* ws, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Learned form: parametric MLP embedding din -> h1 -> h2 -> dout (GELU).
 * No counterpart in the reference, whose only MLP is the MMAE autoencoder
 * (mlx_vis/_mmae/mmae.py:22-60, AdamW loop 258-281): that is the precedent
 * for fitting a network to an embedding and encoding new points with it.
 *
 * Parameters live in one device blob (mv_mlp_param_bytes): FP32 master
 * W1 (h1 x din, stored with rows padded to dp = align64(din)), b1, W2
 * (h2 x h1), b2, W3 (dout x h2), b3, then Adam m and v, then BF16 copies of
 * W1, W2 and W2^T. Hidden widths are multiples of 64 up to 1024, dout <= 4.
 * Forward: H1 = gelu(X W1^T + b1), H2 = gelu(H1 W2^T + b2), Z = H2 W3^T + b3,
 * with X, W1, W2 and the activations rounded to BF16, FP32 accumulation, exact
 * erf GELU. ------------------------------------------------------------ */
int mv_mlp_param_bytes(int din, int h1, int h2, int dout, size_t* out_bytes, int64_t* out_n_params);
/* Workspace of a batch of `points` rows (training uses 2B points). */
int mv_mlp_batch_bytes(int din, int h1, int h2, int dout, int points, size_t* out_bytes);
/* Loads FP32 weights (device; W1 h1 x din row-major) into the blob, zeroes
 * the Adam moments and writes the BF16 copies. */
int mv_mlp_init(void* blob, int din, int h1, int h2, int dout, const float* W1, const float* b1, const float* W2,
                const float* b2, const float* W3, const float* b3, void* stream);
/* Gradient of one batch of B edges (B a multiple of 32): rows idx[0:B] are
 * heads and idx[B:2B] tails (rows of X, n x din FP32, and of the target
 * embedding Yp, n x dout). Loss (1/B) sum_e [log(1 + a d_e^(2b))
 * - sum_{s<neg} log(1 - 1/(1 + a d_es^(2b)))] + (lam/B) sum_rows ||z - y||^2,
 * negatives = tails of edges (e + 1 + 97 s) mod B; z-gradient = the layout's
 * law (umap.py:367-380) without the clip. grad_out (device, n_params FP32 in
 * blob order, nullable), loss_out (device, 2 floats: edge term, regression
 * term, nullable). ws: mv_mlp_batch_bytes(2B) + align256(2 n dp) bytes. */
int mv_mlp_grad(const void* blob, int din, int h1, int h2, int dout, const float* X, int64_t n, const float* Yp,
                const int32_t* idx, int B, int neg, float a, float b, float lam, float* grad_out, float* loss_out,
                void* ws, size_t ws_bytes, void* stream);
/* `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8, bias-corrected, step
 * numbers step0 + 1 ...) on the batches idx[t * 2B : (t + 1) * 2B].
 * loss_trace (device, 2 * steps floats, nullable). Same ws as mv_mlp_grad. */
int mv_mlp_train(void* blob, int din, int h1, int h2, int dout, const float* X, int64_t n, const float* Yp,
                 const int32_t* idx, int steps, int B, int neg, float a, float b, float lam, float lr, int step0,
                 float* loss_trace, void* ws, size_t ws_bytes, void* stream);
/* Inference Z (m x dout) of X (m x din FP32) in chunks of `chunk` rows;
 * ws: mv_mlp_batch_bytes(chunk). */
int mv_mlp_forward(const void* blob, int din, int h1, int h2, int dout, const float* X, int64_t m, float* Z,
                   int chunk, void* ws, size_t ws_bytes, void* stream);
/* Edge batches for mv_mlp_train from a fuzzy graph's slot arrays (eps from
 * mv_umap_fuzzy_graph, tail = kNN indices): `steps` batches of B edges drawn
 * in proportion to their weights (rejection on 1 / eps, Philox-4x32-10 keyed
 * by seed, counters from step0), written as heads then tails per batch. */
int mv_mlp_sample_edges(const float* eps, const int32_t* tail, int64_t n, int k, uint32_t seed, int step0, int steps,
                        int B, int32_t* out_idx, void* stream);
/* Per-row overlap of two neighbour lists, the 15-NN preservation count:
 * out_count[i] = |a[i, 0:ka] & b[i, 0:kb]|, out_flag[i] = out_count[i] <
 * min_count (each output nullable, not both). */
int mv_knn_overlap(const int32_t* a, int64_t lda, int ka, const int32_t* b, int64_t ldb, int kb, int64_t n,
                   int min_count, int32_t* out_count, uint8_t* out_flag, void* stream);

/* Trustworthiness (Venna & Kaski 2001) ranks; no counterpart in the reference,
 * which defines no metric. X (n, d) FP32 input space; rows (m) the query
 * points (NULL: rows 0 .. m-1); lo (m, ldl) the embedding's k neighbours of
 * each query. out_excess[q] = sum_t max(r(i, lo[q, t]) - k, 0), r the rank of
 * the neighbour among the n - 1 other points ordered by (FP32 squared
 * distance, index), 1 = nearest; a neighbour equal to the query point has
 * rank n; entries outside [0, n) count 0. T = 1 - 2 / (m k (2n - 3k - 1)) *
 * sum_q out_excess[q]. 1 <= k <= 64. */
int mv_trust_ws_bytes(int64_t m, int k, size_t* out);
int mv_trust_excess(const float* X, int64_t n, int64_t d, const int32_t* rows, int64_t m, const int32_t* lo,
                    int64_t ldl, int k, int64_t* out_excess, void* ws, size_t ws_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Edge-gradient family beyond UMAP: the contrastive losses of CNE and the
 * pair force laws of PaCMAP / LocalMAP / t-SNE, on the same embedding layout
 * (Y: n x dim FP32, dim <= 4). Gradients accumulate into `grad` (zeroed by
 * the caller) with atomics, as the reference's scatter-adds do.
 * ------------------------------------------------------------------------ */
/* CNE gradient over edges (ei[e], ej[e]) with m negatives each (m <= 32) —
 * replaces CNE._infonce_grad / _nce_grad / _neg_grad (mlx_vis/_cne/cne.py:
 * 219-313): loss 0 InfoNCE, 1 NCE, 2 NEG; Cauchy similarity 1 / (1 + d^2).
 * neg (device, E x m, nullable): explicit negatives (then batch must be 0);
 * else negatives are drawn uniformly from [0, n) by Philox keyed by seed and
 * counted by (edge, iter, slot). batch > 0: process `batch` edges drawn
 * uniformly with replacement (cne.py:337-339). Every term is multiplied by
 * scale (the reference's E / batch). */
int mv_cne_grad(const float* Y, int64_t n, int dim, const int32_t* ei, const int32_t* ej, int64_t E,
                const int32_t* neg, int m, int loss, float scale, uint64_t seed, int iter, int64_t batch,
                float* grad, void* stream);
/* CNE iterations [it0, it1) (1-based) of CNE._optimize (cne.py:315-372):
 * gradient (Philox negatives and batches, scale E / batch when batch < E)
 * and the bias-corrected Adam step (beta 0.9 / 0.999, eps 1e-8) on Y in
 * place. adam_m, adam_v, grad: n * dim FP32 each; grad must be zero on entry
 * and is zero on return. */
int mv_cne_optimize(float* Y, int64_t n, int dim, const int32_t* ei, const int32_t* ej, int64_t E, int m, int loss,
                    int it0, int it1, float lr, int64_t batch, uint64_t seed, float* adam_m, float* adam_v,
                    float* grad, void* stream);
/* One force law over P pairs (src, dst) — the per-pair terms of PaCMAP's step
 * functions (mlx_vis/_pacmap/pacmap.py:275-354) and of t-SNE's attractive
 * term (mlx_vis/_tsne/tsne.py:511-517). law: 0 PaCMAP neighbour
 * 20w / (10 + d)^2, 1 mid-near 20000w / (10000 + d)^2, 2 further pair
 * -2w / (1 + d)^2, 3 LocalMAP neighbour 20 w scale / ((10 + d)^2 sqrt(d)),
 * with d = |y_s - y_t|^2 + 1, each times (y_s - y_t) added to src and
 * subtracted from dst; 4 t-SNE 4w (y_s - y_t) / (1 + |y_s - y_t|^2) added to
 * src only. w_pair (device, P, nullable) overrides w per pair. */
int mv_pair_grad(const float* Y, int64_t n, int dim, const int32_t* src, const int32_t* dst, int64_t P,
                 const float* w_pair, float w, float scale, int law, float* grad, void* stream);
/* Adam on Y: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
 * Y -= step m / (sqrt(v) + eps), then g = 0 (the caller folds the bias
 * corrections into step and eps: PaCMAP's lr_t form, pacmap.py:300-303). */
int mv_adam_update(float* Y, float* grad, float* adam_m, float* adam_v, int64_t count, float b1, float b2, float step,
                   float eps, void* stream);

/* out[0:count] = standard normal draws (Philox-4x32-10 keyed by seed, Box-Muller):
 * the spectral start vectors, which the reference draws from MLX's generator
 * (umap.py:325), a stream that cannot be reproduced outside MLX. */
int mv_random_normal(float* out, int64_t count, uint64_t seed, void* stream);
/* Spectral embedding to [0, 10] (the tail of _spectral_init, umap.py:344-348):
 * e = V[:, 1:nc] * (10 / max |V[:, 1:nc]|) + 1e-4 * noise, then per column
 * Y = 10 (e - min) / (max - min + 1e-10). V: n x nc device (as mv_umap_spectral
 * leaves it), Y: n x (nc - 1) device. The noise is standard normal from
 * Philox-4x32-10 keyed by seed (the reference draws it from NumPy's legacy
 * generator). out_nonfinite (device int32, nullable) receives 1 if V holds a
 * non-finite value (the reference then falls back to a random init).
 * 2 <= nc <= 9. ws: >= 128 bytes of device memory. */
int mv_umap_spectral_embed(const float* V, int64_t n, int nc, uint64_t seed, float* Y, int32_t* out_nonfinite,
                           void* ws, size_t ws_bytes, void* stream);

/* ----------------------------------------------------------------- PCA --
 * The array-sized part of UMAP's PCA preprocessing (mlx_vis/_umap/umap.py:
 * 92-113): mean_out (d, FP64 device) = column means of X (n x d FP32),
 * cov_out (d x d, FP64 device) = (X - mean)^T (X - mean) / (n - 1) (FP32
 * products, FP64 sums over 1,024-row chunks) — replaces umap.py:95-98; the
 * host then takes the eigenvectors (umap.py:100-103, mx.linalg.eigh on the
 * CPU stream in the reference). d <= 8192. */
int mv_pca_ws_bytes(int64_t n, int64_t d, size_t* out_bytes);
int mv_pca_cov(const float* X, int64_t n, int64_t d, double* mean_out, double* cov_out, void* ws, size_t ws_bytes,
               void* stream);
/* out (n x p FP32) = (X - mean) V with mean (d, FP32) and V (d x p FP32
 * row-major), the projection of umap.py:104-105. 1 <= p <= d. */
int mv_pca_project(const float* X, int64_t n, int64_t d, const float* mean, const float* V, int64_t p, float* out,
                   void* stream);

/* p[0:count] = 0 on the stream (a fill kernel, safe under graph capture). */
int mv_zero_f32(float* p, int64_t count, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MIVIS_H */
