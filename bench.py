#!/usr/bin/env python3
"""UMAP fit_transform benchmark on one MI355X.

Workload: X = 70,000 x 784 float32, a mixture of 10 isotropic Gaussians
(centres N(0, 4^2) per dimension, sigma 1, equal weights, seed 1), min-max
scaled to [0, 1], resident in HBM. One step = UMAP(n_neighbors=15,
min_dist=0.1, n_epochs=500, negative_sample_rate=5, knn_method="brute")
.fit_transform(X): exact brute-force kNN (k=15, FP32 MFMA screen of the
upper-triangle distance tiles + per-row selection + exact re-rank), fuzzy
graph, spectral init, 500 epochs of layout SGD, and the copy of the embedding
to the host.

Prints one JSON line. `value` is seconds per fit_transform (lower is better).
`roofline` covers the dominant kernel (the MFMA distance kernel of the kNN,
k_dist_sym: 2 d FLOP per pair on and above the diagonal), timed with HIP
events on its launch stream; `cpu_baseline` is the NumPy
restatement of the reference (the oracle) on a bounded sample of the same
workload, scaled to the whole job.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "mlx-vis_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

MFMA_FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
SCREEN_PROFILES = {"k_screen_f32": ROOT / "profiles" / "k_screen_f32_70000x784.md",
                   "k_dist_sym_f32": ROOT / "profiles" / "k_dist_sym_f32_70000x784.md"}


def screen_traffic(n, d, kernel):
    """HBM bytes per launch of the kNN distance kernel at 70,000 x 784, as
    measured by two rocprofv3 PMC passes (FETCH_SIZE x 2 + WRITE_SIZE x 1,
    factors from a 1 GiB copy) and recorded with their commands in the
    kernel's profile summary; None for other shapes or when the summary is
    absent."""
    prof = SCREEN_PROFILES.get(kernel)
    if (n, d) != (70_000, 784) or prof is None or not prof.exists():
        return None
    for line in prof.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch"):
            return int(float(line.split("=")[1]))
    return None


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=70_000, help="rows of the workload (default: the full 70,000)")
    ap.add_argument("--epochs", type=int, default=500)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--knn-reps", type=int, default=3)
    ap.add_argument("--learned-n", type=int, default=1_000_000,
                    help="fresh points mapped by the learned form (default 1,000,000)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-form line")
    ap.add_argument("--no-scale", action="store_true", help="skip the 1M x 128 NNDescent + UMAP line")
    ap.add_argument("--only", choices=["knn", "copy", "learned", "scale", "layout", "fit", "nndescent", "cpu-full"],
                    default=None,
                    help="profiling helpers: 'knn' runs only the brute kNN (--dtype), 'copy' a 1 GiB device copy, "
                         "'fit' --steps fit_transform calls of the workload, 'nndescent' --steps NNDescent builds "
                         "(k 15, 12 iterations, max_candidates 30) of the same data, 'cpu-full' one unsampled "
                         "run of the oracle's fit_transform on the host cores (no GPU; about five minutes)")
    ap.add_argument("--dtype", default="fp32", choices=["fp32", "bf16", "fp8", "bf16x3"])
    ap.add_argument("--nnd-k", type=int, default=15, help="k of the --only nndescent builds (default 15)")
    ap.add_argument("--rerank", type=int, default=None,
                    help="candidates per row = rerank * k for --only knn (default: the library's per-dtype default)")
    ap.add_argument("--no-extras", action="store_true", help="skip the precision-form and NNDescent lines")
    ap.add_argument("--neg-rate", type=int, default=5,
                    help="negative samples per edge for --only layout (default 5, the workload's)")
    ap.add_argument("--no-symmetric", action="store_true",
                    help="self kNN with the fused screen (k_screen) instead of the symmetric distance kernel")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def knn_phase_timing(Xd, reps, dtype="fp32", rerank=None):
    """Times mv_knn_brute's phases with HIP events on its stream."""
    import ctypes

    import torch

    from mlx_vis_amd import _capi, knn_brute_device

    lib = _capi.lib()
    lib.mv_knn_set_timing(1)
    phases, split = [], []
    sym = ctypes.c_int32(0)
    stats = torch.zeros(4, dtype=torch.int32, device=Xd.device)
    try:
        for _ in range(reps):
            idx, _ = knn_brute_device(Xd, 15, stats=stats, dtype=dtype, rerank=rerank)
            out = (ctypes.c_float * 4)()
            _capi.check(lib.mv_knn_last_phase_ms(out), "mv_knn_last_phase_ms")
            phases.append(list(out))
            sp = (ctypes.c_float * 2)()
            _capi.check(lib.mv_knn_last_screen_ms(sp, ctypes.byref(sym)), "mv_knn_last_screen_ms")
            split.append(list(sp))
    finally:
        lib.mv_knn_set_timing(0)
    n = len(phases)
    avg = [sum(p[i] for p in phases) / n for i in range(4)]
    # distance kernel and row selection inside the screen phase
    screen = {"symmetric": bool(sym.value), "kernel_ms": sum(p[0] for p in split) / n,
              "select_ms": sum(p[1] for p in split) / n}
    return avg, stats.cpu().tolist(), idx, screen


def knn_preservation_gpu(knn_hi, Y):
    """Mean overlap of the 15-NN sets in the input and the embedding."""
    import torch

    from mlx_vis_amd import knn_brute_device, knn_preservation_device

    Yd = torch.as_tensor(Y).cuda().contiguous()
    lo, _ = knn_brute_device(Yd, knn_hi.shape[1])
    counts, _ = knn_preservation_device(knn_hi, lo)
    return float(counts.double().mean().item()) / knn_hi.shape[1]


MLP_FLOP_PER_POINT = 2 * (784 * 512 + 512 * 512 + 512 * 2)   # forward of 784 -> 512 -> 512 -> 2
MFMA_BF16_PEAK_TFLOPS = 2500.0


def learned_form(Xd, um, fresh_n, reps=5):
    """Trains the MLP on the fitted UMAP (graph + embedding of the 70K set), maps
    `fresh_n` fresh points of the same mixture, times a plain layout of them,
    and scores both against the exact 15-NN of the fresh points; points below
    0.8 preservation get 50 plain epochs."""
    import numpy as np
    import torch

    from mlx_vis_amd import UMAP, ParametricUMAP, knn_brute_device, knn_preservation_device
    from mlx_vis_amd.datasets import mixture_fresh_784

    out = {}
    pu = ParametricUMAP(hidden=(512, 512), batch_size=4096, learning_rate=1e-3, random_state=42)
    log("learned form: training the MLP on the 70K graph and embedding")
    torch.cuda.synchronize()
    t = time.perf_counter()
    pu.fit(Xd, umap=um)
    torch.cuda.synchronize()
    out["train_s"] = round(time.perf_counter() - t, 4)
    out["train_steps"] = pu.steps_
    tr = pu.loss_trace_.cpu().numpy()
    out["loss_first_last"] = [round(float(tr[:20].sum(axis=1).mean()), 4), round(float(tr[-20:].sum(axis=1).mean()), 4)]
    z_train = pu.transform_device(Xd)
    out["train_knn15_preservation_mlp"] = round(knn_preservation_gpu(um.knn_[0], z_train), 4)

    log(f"learned form: generating {fresh_n} fresh points (seed 3)")
    Xf = mixture_fresh_784(fresh_n, seed=3)
    Xfd = torch.from_numpy(Xf).cuda()
    del Xf
    chunk = 1 << 20   # one launch pair per 1,048,576 points (H1 workspace 1 GiB)
    ws = pu.forward_workspace(min(chunk, fresh_n))
    Z = pu.transform_device(Xfd, chunk=chunk, ws=ws)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        Z = pu.transform_device(Xfd, chunk=chunk, ws=ws)
    ev[1].record()
    torch.cuda.synchronize()
    t_inf = ev[0].elapsed_time(ev[1]) / reps / 1e3
    flops = MLP_FLOP_PER_POINT * fresh_n
    out["infer_s"] = round(t_inf, 6)
    out["infer_points_per_s"] = round(fresh_n / t_inf, 1)
    out["infer_roofline"] = {"bound": "mfma", "achieved": round(flops / t_inf / 1e12, 2),
                             "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(flops / t_inf / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4)}

    log("learned form: plain layout of the fresh points (UMAP defaults: NNDescent kNN, 200 epochs)")
    plain = UMAP(n_neighbors=15, min_dist=0.1, negative_sample_rate=5, random_state=42)
    torch.cuda.synchronize()
    t = time.perf_counter()
    z_plain = plain.fit_transform(Xfd)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t
    out["plain_layout_s"] = round(t_plain, 4)
    out["speedup_infer_vs_plain"] = round(t_plain / t_inf, 1)

    log("learned form: exact 15-NN of the fresh points (split-BF16 screen, exact re-rank)")
    torch.cuda.synchronize()
    t = time.perf_counter()
    hi_idx, hi_dist = knn_brute_device(Xfd, 15, dtype="bf16x3")
    torch.cuda.synchronize()
    out["eval_exact_knn_s"] = round(time.perf_counter() - t, 3)
    out["eval_exact_knn_dtype"] = "bf16x3 (indices identical to the FP32 path by the certificate)"

    def pres(Zd):
        lo, _ = knn_brute_device(Zd.contiguous(), 15)
        c, _ = knn_preservation_device(hi_idx, lo)
        return c

    t = time.perf_counter()
    c_mlp = pres(Z)
    torch.cuda.synchronize()
    out["eval_lowd_knn_s"] = round(time.perf_counter() - t, 3)
    c_plain = pres(torch.from_numpy(z_plain).cuda())
    p_mlp = float(c_mlp.double().mean().item()) / 15
    p_plain = float(c_plain.double().mean().item()) / 15
    log("learned form: refining points below 0.8 preservation (50 plain epochs)")
    torch.cuda.synchronize()
    t = time.perf_counter()
    z_ref, _ = pu.refine_device(Xfd, Z, threshold=0.8, n_epochs=50, knn=(hi_idx, hi_dist))
    torch.cuda.synchronize()
    out["refine_s"] = round(time.perf_counter() - t, 3)
    c_ref = pres(z_ref)
    p_ref = float(c_ref.double().mean().item()) / 15
    flagged = (c_mlp < 12)
    out["knn15_preservation"] = {"mlp": round(p_mlp, 4), "plain_layout": round(p_plain, 4),
                                 "mlp_refined": round(p_ref, 4), "gap_mlp_vs_plain": round(p_plain - p_mlp, 4)}
    out["refine"] = {"flagged_frac": round(float(flagged.float().mean().item()), 4),
                     "flagged_before": round(float(c_mlp[flagged].double().mean().item()) / 15, 4)
                     if bool(flagged.any()) else None,
                     "flagged_after": round(float(c_ref[flagged].double().mean().item()) / 15, 4)
                     if bool(flagged.any()) else None}
    log("learned form: trustworthiness (k = 15) on 2,000 sampled fresh points")
    from mlx_vis_amd import trustworthiness

    rows = np.sort(np.random.default_rng(5).choice(fresh_n, min(2000, fresh_n), replace=False))
    out["trustworthiness_15"] = {
        name: round(trustworthiness(Xfd, Zd.contiguous(), 15, rows=rows), 4)
        for name, Zd in (("mlp", Z), ("plain_layout", torch.from_numpy(z_plain).cuda()), ("mlp_refined", z_ref))}
    out["trustworthiness_rows"] = int(rows.size)
    out["config"] = {"mlp": "784-512-512-2 GELU, BF16 weights + FP32 master, Adam lr 1e-3, batch 4096 edges",
                     "fresh_points": fresh_n, "threshold": 0.8, "refine_epochs": 50}
    return out


def scale_form(n=1_000_000, sample=10_000):
    """1,000,000 x 128 mixture of 50 Gaussians (Dirichlet weights, seed 2):
    NNDescent k=15, 12 iterations, recall on `sample` rows (drawn by
    default_rng(0) without replacement) against exact FP32 brute force over all
    rows; then UMAP 2-D, 200 epochs, 5 negatives, NNDescent kNN."""
    import numpy as np
    import torch

    from mlx_vis_amd import UMAP, NNDescent, knn_brute_device, knn_preservation_device
    from mlx_vis_amd.datasets import gaussian_mixture_128

    log(f"scale: generating the {n} x 128 mixture (seed 2)")
    X = gaussian_mixture_128(n, seed=2)
    Xd = torch.from_numpy(X).cuda()
    del X
    out = {}
    nnd = NNDescent(k=15, n_iters=12, random_state=42)
    nnd.build_device(Xd[:20000].contiguous())   # loads the kernels
    torch.cuda.synchronize()
    log("scale: NNDescent k=15, 12 iterations")
    t = time.perf_counter()
    gi, _ = nnd.build_device(Xd)
    torch.cuda.synchronize()
    out["nndescent_s"] = round(time.perf_counter() - t, 4)
    out["nndescent_iterations"] = nnd.iterations_
    rows = np.sort(np.random.default_rng(0).choice(n, sample, replace=False)).astype(np.int64)
    rows_d = torch.from_numpy(rows).cuda()
    Q = Xd.index_select(0, rows_d).contiguous()
    t = time.perf_counter()
    ex, _ = knn_brute_device(Xd, 15, Q=Q, q_self=rows_d.to(torch.int32))
    torch.cuda.synchronize()
    out["exact_sample_knn_s"] = round(time.perf_counter() - t, 4)
    c, _ = knn_preservation_device(ex, gi.index_select(0, rows_d).contiguous())
    out["recall_vs_exact"] = round(float(c.double().mean().item()) / 15, 4)
    out["recall_sample_rows"] = sample
    log("scale: UMAP 2-D, 200 epochs, NNDescent kNN")
    um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=200, negative_sample_rate=5, knn_method="nndescent",
              random_state=42)
    torch.cuda.synchronize()
    t = time.perf_counter()
    um.fit_transform(Xd)
    torch.cuda.synchronize()
    out["fit_transform_s"] = round(time.perf_counter() - t, 4)
    g = um.graph_
    from mlx_vis_amd._umap import find_ab_params, run_epochs

    a, b = find_ab_params(1.0, 0.1)
    Y = torch.from_numpy(um.embedding_).cuda()
    g.epn.copy_(g.eps.masked_fill(g.eps <= 0, float("inf")))
    torch.cuda.synchronize()
    t = time.perf_counter()
    run_epochs(Y, None, g, 1, 200, 200, 5, 1.0, a, b, 7)
    torch.cuda.synchronize()
    t_ep = time.perf_counter() - t
    # every kept edge is active in epoch e for floor(e / eps) - floor((e - 1) / eps) samples
    eps = g.eps[g.eps > 0].double()
    samples = float(torch.floor((199.0 / eps) + 1e-9).sum().item())
    out["layout"] = {"epochs_s": round(t_ep, 4), "kept_edges": g.n_edges,
                     "edge_samples": int(samples), "edge_samples_per_s": round(samples / t_ep, 1)}
    out["config"] = {"n": n, "d": 128, "k": 15, "nndescent_iters": 12, "n_epochs": 200, "negatives": 5}
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(X, knn_idx, knn_dist, n_epochs, budget_rows=1000, budget_epochs=16, spec_iters=5, c0_runs=3):
    """The oracle restatement (NumPy with BLAS threads, torch-CPU index_add_
    scatters), not MLX: a full run of the small plain-method workload
    (10,000 x 128, 200 epochs; one warm-up on 2,000 rows, then the median of
    c0_runs full runs), and the 70,000 x 784 fit_transform extrapolated from
    a bounded sample (kNN rows, spectral iterations and layout epochs scaled
    up)."""
    import numpy as np
    import torch

    from oracle import knn as oknn
    from oracle import umap as ou

    cores = min(16, len(os.sched_getaffinity(0)))
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # noqa: BLE001
        threadpool_limits = None
    torch_threads = torch.get_num_threads()
    # the layout's small per-epoch tensors run fastest on one thread (measured:
    # 17 ms per 56 K-edge step on 1 thread, 290 ms on 8)
    torch.set_num_threads(1)
    ctx = threadpool_limits(limits=cores) if threadpool_limits else None
    try:
        if ctx is not None:
            ctx.__enter__()
        # full runs of the small workload
        Xc0 = np.random.default_rng(0).standard_normal((10000, 128), dtype=np.float32)
        ou.fit_transform_random_init(Xc0[:2000], n_epochs=20)
        c0 = []
        for _ in range(c0_runs):
            t0 = time.perf_counter()
            ou.fit_transform_random_init(Xc0, n_epochs=200)
            c0.append(time.perf_counter() - t0)

        # the 70,000 x 784 workload on a bounded sample, scaled
        n = X.shape[0]
        rng = np.random.default_rng(0)
        rows = np.sort(rng.choice(n, budget_rows, replace=False))
        t0 = time.perf_counter()
        oknn.reference_knn_rows(X, rows, 15)
        t_knn = (time.perf_counter() - t0) * n / budget_rows

        t0 = time.perf_counter()
        r, c, v = ou.fuzzy_simplicial_set(knn_idx, knn_dist, n_epochs)
        ou.find_ab_params(1.0, 0.1)
        t_graph = time.perf_counter() - t0

        t0 = time.perf_counter()
        ou.spectral_power(r, c, v.astype(np.float64), n, rng.standard_normal((n, 3)), iters=spec_iters)
        t_spec = (time.perf_counter() - t0) * 100 / spec_iters

        # layout: budget_epochs epochs spread evenly over the schedule's
        # non-empty epochs, scaled by their count (late epochs hold few edges
        # but keep their fixed per-epoch cost, so scaling the first epochs by
        # edge samples would undercount)
        sets = ou.schedule(v, n_epochs)
        nonempty = [e for e in range(n_epochs) if sets[e].size]
        pick = np.unique(np.linspace(0, len(nonempty) - 1, budget_epochs).round().astype(int))
        ep = [nonempty[j] for j in pick]
        Y = torch.from_numpy(rng.uniform(-10, 10, size=(n, 2)).astype(np.float32))
        rt, ct = torch.from_numpy(r.astype(np.int64)), torch.from_numpy(c.astype(np.int64))
        done = 0
        t0 = time.perf_counter()
        for e in ep:
            act = torch.from_numpy(sets[e].astype(np.int64))
            A = act.numel()
            ef, et = rt[act], ct[act]
            nf = ef[torch.arange(5 * A) % A]
            nt = torch.from_numpy(rng.integers(0, n, size=5 * A))
            Y = ou.sgd_step_index_add(Y, ef, et, nf, nt, float(1.0 - e / n_epochs), 1.577, 0.895)
            done += A
        t_layout = (time.perf_counter() - t0) * len(nonempty) / len(ep)
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
        torch.set_num_threads(torch_threads)
    total_s = t_knn + t_graph + t_spec + t_layout
    return {
        "value": round(total_s, 2),
        "unit": "s",
        "cores": int(cores),
        "kind": "port",
        "sample": (f"NumPy restatement of the reference (oracle), not MLX, extrapolated: kNN (FP32 expansion + "
                   f"full-row argsort) on {budget_rows} of {n} query rows (x{n / budget_rows:.0f}); fuzzy graph + "
                   f"a/b in full; spectral init {spec_iters} of 100 iterations (x{100 / spec_iters:.0f}); layout "
                   f"{len(ep)} epochs spread over {ep[0]}-{ep[-1]} (torch-CPU index_add_, {done} edge samples, "
                   f"scaled by epochs {len(nonempty)}/{len(ep)})"),
        "parts_s": {"knn": round(t_knn, 2), "graph": round(t_graph, 2), "spectral": round(t_spec, 2),
                    "layout": round(t_layout, 2)},
        "cpu_model": cpu_model(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
        "threads": {"blas": int(cores), "torch_layout": 1},
        "small_workload_full_run": {
            "workload": "10,000 x 128 N(0,1) seed 0, FP32 brute kNN k=15 (full-row argsort), 200 epochs, "
                        "5 negatives, random init",
            "runs_s": [round(t, 2) for t in c0], "median_s": round(float(np.median(c0)), 2)},
    }


def cpu_full_run(X, n_epochs, block=1000):
    """One full, unsampled run of the 70,000 x 784 fit_transform by the oracle
    restatement on the host cores (same thread settings and parts as
    cpu_baseline): kNN of every row in blocks of `block` rows, fuzzy graph,
    a/b, 100 spectral power iterations and every layout epoch. It takes about
    five minutes, so it is a separate mode, not part of the default line."""
    import numpy as np
    import torch

    from oracle import knn as oknn
    from oracle import umap as ou

    cores = min(16, len(os.sched_getaffinity(0)))
    from threadpoolctl import threadpool_limits

    torch.set_num_threads(1)
    n = X.shape[0]
    rng = np.random.default_rng(0)
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        idx = np.empty((n, 15), np.int32)
        dist = np.empty((n, 15), np.float32)
        for b0 in range(0, n, block):
            rows = np.arange(b0, min(n, b0 + block))
            idx[rows], dist[rows] = oknn.reference_knn_rows(X, rows, 15)
            if (b0 // block) % 7 == 6:
                log(f"cpu-full: kNN {rows[-1] + 1} of {n} rows, {time.perf_counter() - t0:.1f} s")
        t_knn = time.perf_counter() - t0
        t0 = time.perf_counter()
        r, c, v = ou.fuzzy_simplicial_set(idx, dist, n_epochs)
        ou.find_ab_params(1.0, 0.1)
        t_graph = time.perf_counter() - t0
        t0 = time.perf_counter()
        ou.spectral_power(r, c, v.astype(np.float64), n, rng.standard_normal((n, 3)), iters=100)
        t_spec = time.perf_counter() - t0
        log(f"cpu-full: graph {t_graph:.1f} s, spectral {t_spec:.1f} s")
        sets = ou.schedule(v, n_epochs)
        Y = torch.from_numpy(rng.uniform(-10, 10, size=(n, 2)).astype(np.float32))
        rt, ct = torch.from_numpy(r.astype(np.int64)), torch.from_numpy(c.astype(np.int64))
        t0 = time.perf_counter()
        for e in range(n_epochs):
            if not sets[e].size:
                continue
            act = torch.from_numpy(sets[e].astype(np.int64))
            A = act.numel()
            ef, et = rt[act], ct[act]
            nf = ef[torch.arange(5 * A) % A]
            nt = torch.from_numpy(rng.integers(0, n, size=5 * A))
            Y = ou.sgd_step_index_add(Y, ef, et, nf, nt, float(1.0 - e / n_epochs), 1.577, 0.895)
            if e % 100 == 99:
                log(f"cpu-full: layout epoch {e + 1}, {time.perf_counter() - t0:.1f} s")
        t_layout = time.perf_counter() - t0
    return {"total_s": round(t_knn + t_graph + t_spec + t_layout, 2),
            "parts_s": {"knn": round(t_knn, 2), "graph": round(t_graph, 2), "spectral": round(t_spec, 2),
                        "layout": round(t_layout, 2)},
            "cores": int(cores), "threads": {"blas": int(cores), "torch_layout": 1},
            "cpu_model": cpu_model(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "workload": f"{n} x {X.shape[1]} mixture, oracle restatement (NumPy/torch-CPU, not MLX), one run, "
                        f"every row and epoch"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is not built",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mlx_vis_amd import UMAP, _capi
    from mlx_vis_amd.datasets import gaussian_mixture_784

    if args.no_symmetric:
        _capi.lib().mv_knn_set_symmetric(0)

    if args.only == "copy":
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4}), flush=True)
        return

    log("generating the 10-Gaussian 784-D mixture")
    X = gaussian_mixture_784(args.n, seed=1)
    if args.only == "cpu-full":
        print(json.dumps({"cpu_full_run": cpu_full_run(X, args.epochs)}), flush=True)
        return
    Xd = torch.from_numpy(X).cuda()
    n, d = X.shape
    torch.cuda.synchronize()

    if args.only == "knn":
        phases, stats, _, scr = knn_phase_timing(Xd, args.knn_reps, args.dtype, args.rerank)
        print(json.dumps({"knn_dtype": args.dtype, "ms_phases": phases, "stats": stats, "screen": scr,
                          "tflops_equiv_full_matrix": 2.0 * n * n * d / (phases[1] * 1e-3) / 1e12}), flush=True)
        return

    if args.only == "layout":
        # the layout epochs alone: the 70,000 x 784 graph built once, then
        # --knn-reps runs of the epochs from the same random init
        from mlx_vis_amd import find_ab_params, knn_brute_device
        from mlx_vis_amd._umap import fuzzy_graph_device, run_epochs

        idx, dist = knn_brute_device(Xd, 15)
        g = fuzzy_graph_device(idx, dist, args.epochs)
        a, b = find_ab_params(1.0, 0.1)
        Y0 = torch.from_numpy(np.random.RandomState(0).uniform(-10, 10, (n, 2)).astype(np.float32)).cuda()
        times = []
        for _ in range(args.knn_reps):
            Y = Y0.clone()
            g.epn.copy_(g.eps)
            g.epn[g.eps <= 0] = float("inf")
            torch.cuda.synchronize()
            t = time.perf_counter()
            run_epochs(Y, None, g, 1, args.epochs, args.epochs, args.neg_rate, 1.0, a, b, 42)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        print(json.dumps({"layout_epochs_s": times, "epochs": args.epochs - 1, "neg_rate": args.neg_rate}), flush=True)
        return

    if args.only == "scale":
        print(json.dumps({"scale": scale_form()}), flush=True)
        return

    if args.only == "learned":
        um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=args.epochs, negative_sample_rate=5,
                  knn_method="brute", random_state=42)
        um.fit_transform(Xd)
        print(json.dumps({"learned": learned_form(Xd, um, args.learned_n)}), flush=True)
        return

    def step():
        um = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=args.epochs, negative_sample_rate=5,
                  knn_method="brute", random_state=42)
        Y = um.fit_transform(Xd)
        return um, Y

    if args.only in ("fit", "nndescent"):
        # --steps calls of the headline fit_transform or of NNDescent on the same
        # data, each timed on the host clock after a synchronisation
        from mlx_vis_amd import NNDescent

        times = []
        for _ in range(max(args.steps, 1)):
            torch.cuda.synchronize()
            t = time.perf_counter()
            if args.only == "fit":
                step()
            else:
                nn = NNDescent(k=args.nnd_k, n_iters=12, max_candidates=30, random_state=0)
                gi, _ = nn.build_device(Xd)
            torch.cuda.synchronize()
            times.append(round((time.perf_counter() - t) * 1e3, 2))
        out = {args.only + "_ms": times}
        if args.only == "nndescent":
            # recall of the last build against the exact self kNN of the same k
            from mlx_vis_amd import knn_brute_device

            ei, _ = knn_brute_device(Xd, args.nnd_k)
            hits = 0
            for r0 in range(0, n, 8192):
                g_, e_ = gi[r0:r0 + 8192], ei[r0:r0 + 8192]
                hits += int((g_[:, :, None] == e_[:, None, :]).any(-1).sum().item())
            out.update({"k": args.nnd_k, "iterations": nn.iterations_, "recall_vs_exact": round(hits / (n * args.nnd_k), 4)})
        print(json.dumps(out), flush=True)
        return

    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        um, Y = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    log(f"fit_transform {dt * 1000:.1f} ms")
    # the same call with the NumPy array as input: the host -> device copy of
    # X (220 MB) and the device -> host copy of the embedding inside the call
    t = time.perf_counter()
    for _ in range(3):
        UMAP(n_neighbors=15, min_dist=0.1, n_epochs=args.epochs, negative_sample_rate=5, knn_method="brute",
             random_state=42).fit_transform(X)
    torch.cuda.synchronize()
    dt_numpy = (time.perf_counter() - t) / 3
    # peak device memory of one call (workspaces are torch allocations)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    step()
    torch.cuda.synchronize()
    peak_bytes = torch.cuda.max_memory_allocated()

    # phase breakdown of one more run (synchronised between phases; not timed above)
    from mlx_vis_amd import knn_brute_device
    from mlx_vis_amd._umap import find_ab_params, fuzzy_graph_device, run_epochs
    from mlx_vis_amd._spectral import spectral_init

    ph = {}
    torch.cuda.synchronize()
    t = time.perf_counter()
    ki, kd = knn_brute_device(Xd, 15)
    torch.cuda.synchronize()
    ph["knn"] = time.perf_counter() - t
    t = time.perf_counter()
    g = fuzzy_graph_device(ki, kd, args.epochs)
    torch.cuda.synchronize()
    ph["graph"] = time.perf_counter() - t
    t = time.perf_counter()
    Y0 = spectral_init(g, n, 2, 42)
    torch.cuda.synchronize()
    ph["init"] = time.perf_counter() - t
    a, b = find_ab_params(1.0, 0.1)
    t = time.perf_counter()
    run_epochs(Y0, None, g, 1, args.epochs, args.epochs, 5, 1.0, a, b, 1)
    torch.cuda.synchronize()
    ph["epochs"] = time.perf_counter() - t
    # Layout roofline: every kept edge is active floor((n_epochs - 1) / eps)
    # times; algorithmic bytes per epoch 12 E (slot scan) + 80 per active edge
    # (ids, head/tail and 5 negative gathers, head/tail updates), against HBM.
    eps_k = g.eps[g.eps > 0].double()
    samples = float(torch.floor((float(args.epochs - 1) / eps_k) + 1e-9).sum().item())
    lay_bytes = 12.0 * g.n_edges * (args.epochs - 1) + 80.0 * samples
    layout = {"epochs": args.epochs - 1, "kept_edges": g.n_edges, "edge_samples": int(samples),
              "edge_samples_per_s": round(samples / ph["epochs"], 1),
              "algorithmic_gbs": round(lay_bytes / ph["epochs"] / 1e9, 1),
              "frac_hbm": round(lay_bytes / ph["epochs"] / 1e9 / HBM_PEAK_GBS, 4)}

    log("kNN phase timing")
    phases, stats, ref_idx, scr = knn_phase_timing(Xd, args.knn_reps)
    screen_ms = phases[1]
    flops = 2.0 * n * n * d
    # the dominant kernel: the symmetric distance kernel multiplies the n(n+1)/2
    # pairs on and above the diagonal (the fused screen: all n^2)
    kern = "k_dist_sym_f32" if scr["symmetric"] else "k_screen_f32"
    kern_flops = 2.0 * d * n * (n + 1) / 2 if scr["symmetric"] else flops
    screen_tflops = kern_flops / (scr["kernel_ms"] * 1e-3) / 1e12
    knn_total_ms = sum(phases)
    quality = knn_preservation_gpu(um.knn_[0], Y)
    # trustworthiness of the embedding over all 70K points: exact ranks in the
    # input space for each embedding neighbour (m x n x d distance work, VALU)
    from mlx_vis_amd import trustworthiness
    torch.cuda.synchronize()
    t = time.perf_counter()
    trust = trustworthiness(Xd, torch.as_tensor(Y).cuda(), 15)
    torch.cuda.synchronize()
    t_trust = time.perf_counter() - t

    line = {
        "metric": "fit_transform s, 70K x 784 kNN+UMAP; kNN TFLOP/s; bf16/fp8 vs fp32 at exact kNN",
        "value": round(dt, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1000, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 10-Gaussian 784-D mixture, min-max scaled (seed 1), resident in HBM",
        "config": {"workload": f"UMAP fit_transform {n}x{d}, exact brute kNN k=15 (FP32), spectral init, "
                               f"{args.epochs} epochs, 5 negatives", "n": n, "d": d, "k": 15,
                   "n_epochs": args.epochs, "knn_method": "brute", "knn_dtype": "fp32"},
        "knn": {"screen_kernel": kern, "tflops_screen_kernel": round(screen_tflops, 2),
                "tflops_equiv_full_matrix": round(flops / (screen_ms * 1e-3) / 1e12, 2),
                "ms_total": round(knn_total_ms, 3),
                "ms_phases": {"prepass": round(phases[0], 3), "screen": round(phases[1], 3),
                              "rerank": round(phases[2], 3), "exact_scan": round(phases[3], 3)},
                "ms_screen_split": {"distance_kernel": round(scr["kernel_ms"], 3),
                                    "row_select": round(scr["select_ms"], 3)},
                "rows_fp64_scan": stats[0], "rows_fp64_tiecheck": stats[1],
                # 2 n^2 d over the whole kNN time: an equivalent rate (the symmetric
                # screen multiplies only half of the pairs), not a hardware rate
                "tflops_equiv_whole_knn": round(flops / (knn_total_ms * 1e-3) / 1e12, 2)},
        "phases_s": {k: round(v, 4) for k, v in ph.items()},
        "numpy_input_fit_transform_s": round(dt_numpy, 4),
        "peak_device_memory_gb": round(peak_bytes / 1e9, 2),
        "layout": layout,
        "quality": {"knn15_preservation": round(quality, 4), "trustworthiness_15": round(trust, 5),
                    "trustworthiness_s": round(t_trust, 4), "n_edges": g.n_edges},
        "roofline": {"kernel": kern, "bound": "mfma", "achieved": round(screen_tflops, 2),
                     "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(screen_tflops / MFMA_FP32_PEAK_TFLOPS, 4), "traffic": screen_traffic(n, d, kern)},
    }
    if not args.no_extras:
        from mlx_vis_amd import knn_brute_device
        # relative error of the returned Euclidean distances against FP64
        # distances of the same pairs, on a fixed sample of 2,000 rows
        err_rows = np.random.default_rng(5).choice(n, size=min(2000, n), replace=False)
        X64 = Xd.index_select(0, torch.as_tensor(err_rows, device=Xd.device)).double().cpu().numpy()
        Xall64 = Xd.double().cpu().numpy()

        def dist_err(dtype, rerank):
            idx_e, dist_e = knn_brute_device(Xd, 15, dtype=dtype, rerank=rerank)
            ii = idx_e.cpu().numpy()[err_rows]
            dd = dist_e.cpu().numpy()[err_rows].astype(np.float64)
            d64 = np.sqrt(((X64[:, None, :] - Xall64[ii]) ** 2).sum(-1))
            rel = np.abs(dd - d64) / np.maximum(d64, 1e-30)
            return {"max": float(f"{rel.max():.3e}"), "mean": float(f"{rel.mean():.3e}"), "rows": int(err_rows.size)}

        # The screens' own distance error: squared distances of the same pairs
        # (each sampled row and its 15 exact neighbours) from the operands as
        # each screen quantises them (centred rows; BF16 round to nearest
        # even; FP8 e4m3 with the per-row absmax scale 448 / max|x|; split BF16
        # hi + lo), products exact, against the FP64 distances. The FP32
        # accumulation of the MFMA adds ~1e-7 relative on top.
        ex_idx = ref_idx.cpu().numpy()[err_rows]
        mu64 = Xall64.mean(axis=0)
        Xc32 = torch.from_numpy((Xall64 - mu64).astype(np.float32))

        def quantise(dtype):
            if dtype == "bf16":
                return Xc32.to(torch.bfloat16).double()
            if dtype == "bf16x3":
                hi = Xc32.to(torch.bfloat16).float()
                return hi.double() + (Xc32 - hi).to(torch.bfloat16).double()
            amax = Xc32.abs().max(dim=1, keepdim=True).values.clamp_min(1e-30)
            sc = 448.0 / amax
            return ((Xc32 * sc).to(torch.float8_e4m3fn).float() / sc).double()

        ref2 = ((X64[:, None, :] - Xall64[ex_idx]) ** 2).sum(-1)

        def screen_err(dtype):
            Xq = quantise(dtype).numpy()
            got = ((Xq[err_rows][:, None, :] - Xq[ex_idx]) ** 2).sum(-1)
            rel = np.abs(got - ref2) / np.maximum(ref2, 1e-30)
            return {"max": float(f"{rel.max():.3e}"), "mean": float(f"{rel.mean():.3e}"), "pairs": int(rel.size)}

        prec = {"fp32_dist_rel_err_vs_fp64": dist_err("fp32", 2),
                "screen_sq_dist_rel_err_vs_fp64": {dt: screen_err(dt) for dt in ("bf16x3", "bf16", "fp8")}}

        # What the quantised screens' certificate can prove, per candidate
        # count C: for the sampled rows, the exact squared distances to every
        # point (FP64 on the device, centred rows, self dropped); tau = the
        # C-th smallest (the screen's C-th value up to its error); the bound
        # for every point past the candidates, (sqrt(tau - arith) - E_i -
        # E_max)^2, with E the norms of the rows' quantisation errors and
        # arith the screen's FP32 accumulation bound; a row certifies when its
        # 15th distance is below it. Reported: the median gap tau - d15, the
        # median margin the certificate needs (tau - bound), the fraction of
        # rows that can certify.
        def cert_gap():
            rows_t = torch.as_tensor(err_rows, device=Xd.device)
            Xc64 = torch.from_numpy(Xall64 - mu64).to(Xd.device)
            d2 = torch.cdist(Xc64.index_select(0, rows_t), Xc64).pow_(2)
            d2.scatter_(1, rows_t[:, None], float("inf"))
            top = torch.topk(d2, 256, dim=1, largest=False).values
            del d2
            sq = (Xc64 * Xc64).sum(dim=1)
            sqq, sqmax = sq.index_select(0, rows_t), sq.max()
            kth = top[:, 14]
            out = {}
            for dt, row_elems in (("bf16", 832), ("fp8", 896)):
                E = (Xc32.double() - quantise(dt)).norm(dim=1).to(Xd.device)
                Ei, Emax = E.index_select(0, rows_t), E.max()
                arith = (2 * row_elems + 10) * 2.0 ** -24 * 1.01 * (sqq + sqmax)
                res = {"E_median": round(float(E.median()), 5), "E_max": round(float(Emax), 5)}
                for C in (30, 60, 128, 256):
                    tau = top[:, C - 1]
                    lb = (torch.sqrt((tau - arith).clamp_min(0.0)) - Ei - Emax).clamp_min(0.0) ** 2
                    res[f"c{C}"] = {"gap_median": round(float((tau - kth).median()), 4),
                                    "margin_needed_median": round(float((tau - lb).median()), 4),
                                    "certifiable": round(float((kth < lb).double().mean()), 4)}
                out[dt] = res
            return out

        prec["certificate_gap"] = cert_gap()
        for dt, rr in (("bf16x3", 2), ("bf16", 2), ("bf16", 4), ("bf16", 9), ("fp8", 2), ("fp8", 9)):
            C = min(rr * 15, 128)
            log(f"kNN {dt} screen, {C} candidates")
            ph_dt, st_dt, idx_dt, _ = knn_phase_timing(Xd, args.knn_reps, dt, rr)
            tot = sum(ph_dt)
            prec[f"{dt}_c{C}"] = {"ms_total": round(tot, 3), "ms_phases": [round(v, 3) for v in ph_dt],
                        "speedup_vs_fp32": round(knn_total_ms / tot, 3),
                        "identical_indices": bool(torch.equal(idx_dt, ref_idx)),
                        "rows_uncertified_first_pass": st_dt[2], "rows_uncertified_after_rescreen": st_dt[3],
                        "rows_fp64_scan": st_dt[0], "dist_rel_err_vs_fp64": dist_err(dt, rr)}
        # the whole fit_transform with the split-BF16 screen (same certified kNN)
        def step_x3():
            um3 = UMAP(n_neighbors=15, min_dist=0.1, n_epochs=args.epochs, negative_sample_rate=5,
                       knn_method="brute", random_state=42, knn_dtype="bf16x3")
            um3.fit_transform(Xd)
            return um3
        step_x3()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            um3 = step_x3()
        torch.cuda.synchronize()
        prec["fit_transform_bf16x3_screen"] = {
            "s": round((time.perf_counter() - t) / max(args.steps, 1), 4),
            "identical_knn_indices": bool(torch.equal(um3.knn_[0], um.knn_[0]))}
        line["precision"] = prec
        log("small plain-method workload: 10,000 x 128 N(0,1), 200 epochs, random init")
        Xc0 = torch.from_numpy(np.random.default_rng(0).standard_normal((10000, 128), dtype=np.float32)).cuda()

        def step_c0():
            return UMAP(n_neighbors=15, min_dist=0.1, n_epochs=200, negative_sample_rate=5, init="random",
                        knn_method="brute", random_state=0).fit_transform(Xc0)
        step_c0()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            step_c0()
        torch.cuda.synchronize()
        line["small_workload"] = {"workload": "UMAP fit_transform 10000x128 N(0,1) seed 0, exact brute kNN k=15, "
                                              "random-uniform init in [-10, 10], 200 epochs, 5 negatives",
                                  "s": round((time.perf_counter() - t) / 3, 4)}
        log("NNDescent k=15, 12 iterations, max_candidates=30")
        from mlx_vis_amd import NNDescent

        nnd = NNDescent(k=15, n_iters=12, max_candidates=30, random_state=0)
        nnd.build_device(Xd)
        torch.cuda.synchronize()
        t = time.perf_counter()
        gi, _ = nnd.build_device(Xd)
        torch.cuda.synchronize()
        t_nnd = time.perf_counter() - t
        hi = ref_idx.long()
        inter = (gi.long()[:, :, None] == hi[:, None, :]).any(dim=2).sum(dim=1).float()
        line["nndescent"] = {"s": round(t_nnd, 4), "recall_vs_exact": round(float((inter / 15).mean().item()), 4),
                             "iterations": nnd.iterations_, "changed": nnd.changed_}
    if not args.no_extras and not args.no_learned:
        line["learned"] = learned_form(Xd, um, args.learned_n)
    if not args.no_extras and not args.no_scale:
        line["scale"] = scale_form()
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle restatement, bounded sample)")
        line["cpu_baseline"] = cpu_baseline(X, um.knn_[0].cpu().numpy(), um.knn_[1].cpu().numpy(), args.epochs)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
