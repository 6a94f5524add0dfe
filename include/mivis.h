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
  MV_FP8 = 2      /* FP8 e4m3 MFMA screen, per-row absmax scale, FP32 accumulate */
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
 * Limits: 1 <= k <= 32, k <= rerank*k <= 64, k <= number of selectable rows.
 */
int mv_knn_ws_bytes(int64_t nq, int64_t n, int64_t d, int k, int dtype, int rerank,
                    int same_qx, size_t* out_bytes);
int mv_knn_brute(const float* Q, int64_t nq, const float* X, int64_t n, int64_t d, int k,
                 int dtype, int rerank, const int32_t* q_self, int64_t self_base, int euclid,
                 int32_t* out_idx, float* out_dist, int32_t* out_stats,
                 void* ws, size_t ws_bytes, void* stream);

/* Row squared norms: out[i] = sum_t X[i,t]^2 in FP32 (mx.sum(X*X, 1) at
 * mlx_vis/_knn.py:49 and mlx_vis/_nndescent/nndescent.py:51,77). */
int mv_row_sqnorms(const float* X, int64_t n, int64_t d, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MIVIS_H */
