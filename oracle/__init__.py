"""CPU oracle for the mlx-vis hot path — TEST INFRASTRUCTURE ONLY.

A NumPy restatement of the reference's algorithms (mlx_vis/_knn.py,
mlx_vis/_nndescent/nndescent.py, mlx_vis/_umap/umap.py,
mlx_vis/_normalize.py), each function citing the reference lines it follows.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
it, and only as the checker (or the timed CPU baseline). The product package
(mlx_vis_amd) never imports it and has no CPU fallback.

How each part is pinned (the reference ships no golden vectors, fixtures or
tests, and its MLX dependency is not installable here):
  knn.exact_knn        sklearn brute-force kNN, scipy cdist, and brute-force
                       Python loops on tiny inputs (tests/test_oracle.py).
  umap.find_ab_params  scipy.optimize.curve_fit on the same 300 points.
  umap.fuzzy_sigmas    its defining equation sum_j exp(-(d_j-rho)/sigma) = log2 k.
  umap.symmetrize      the closed form w + w^T - w*w^T against a dense matrix.
  umap.schedule        the closed form of the number of samples per edge.
  umap.sgd_step        central finite differences of the UMAP edge loss and a
                       dense-matrix restatement.
  nndescent            invariants (no self, unique ids, sorted, exact stored
                       distances) and recall against exact_knn.
  metrics              sklearn.manifold.trustworthiness and set arithmetic.
"""
