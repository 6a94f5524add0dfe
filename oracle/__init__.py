"""CPU oracle for the mlx-vis hot path — TEST INFRASTRUCTURE ONLY.

A NumPy restatement of the reference's algorithms (mlx_vis/_knn.py,
mlx_vis/_nndescent/nndescent.py, mlx_vis/_umap/umap.py,
mlx_vis/_normalize.py), each function citing the reference lines it follows.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
it, and only as the checker (or the timed CPU baseline). The product package
(mlx_vis_amd) never imports it and has no CPU fallback.

How each part is pinned (the reference ships no golden vectors, fixtures or
tests, and its MLX dependency is not installable here; tests/golden holds
fixtures written by scikit-learn and SciPy, tests/golden/make_golden.py):
  knn.exact_knn        sklearn brute-force kNN (and the committed C0 fixture),
                       scipy cdist, brute-force Python loops on tiny inputs
                       (tests/test_oracle_knn.py, tests/test_golden.py).
  umap.find_ab_params  scipy.optimize.curve_fit on the same 300 points (and the
                       committed a/b fixture).
  umap.fuzzy_rows      its defining equation sum_j exp(-(d_j-rho)/sigma) = log2 k.
  umap.symmetrize      the closed form w + w^T - w*w^T against a dense matrix.
  umap.schedule        the closed form of the number of samples per edge.
  umap.sgd_step        central finite differences of the UMAP edge loss and a
                       dense-matrix restatement.
  umap.spectral_power  numpy.linalg.eigh on a symmetric graph.
  umap.spectral_scale  a hand-worked case, rescaling and column-0 invariance.
  nndescent            invariants (no self, unique ids, sorted, exact stored
                       distances), the reference's dedup rule on a hand-worked
                       case, and recall against exact_knn.
  mlp                  central finite differences and torch-CPU float64
                       autograd; torch.optim.Adam.
  cne, pacmap          finite differences of the losses their forces derive
                       from; samplers against their defining properties.
  metrics              sklearn.manifold.trustworthiness and set arithmetic.
"""
