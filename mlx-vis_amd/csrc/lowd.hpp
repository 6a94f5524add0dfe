// Exact kNN in one to three dimensions on a uniform grid (lowd.hip), used by
// mv_knn_brute for d <= LOWD_MAX_D.
#pragma once

#include "common.hpp"

namespace mv {

constexpr int LOWD_MAX_D = 3;

// Workspace of lowd_knn over n points.
size_t lowd_knn_ws_bytes(int64_t n);

// The k nearest rows of X (n x d, d <= 3) for each of the nq rows of Q, in
// the canonical order (FP64 sequential squared distance, lower index first);
// Q == X with nq == n searches the points in cell order. Self exclusion as
// mv_knn_brute. 1 <= k <= 64.
int lowd_knn(const float* Q, int64_t nq, const float* X, int64_t n, int d, int k, const int32_t* q_self,
             int64_t self_base, int euclid, int32_t* out_idx, float* out_dist, void* ws, size_t ws_bytes,
             hipStream_t s);

}  // namespace mv
