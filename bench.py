"""Benchmark of the MI355X UMAP/densMAP fitting path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line.  Workloads:
  fit      (default) the full umap_gpu fit of the precision form: n=1,000,000,
           d=64 f32 (50 Gaussians on a random 8-D subspace + N(0, 0.05) noise),
           exhaustive kNN k=15 in BF16 on MFMA + exact FP32 re-rank of a
           64-candidate shortlist, fuzzy graph, spectral layout, 500 Adam epochs,
           2-D output.  One step = one fit with the data resident in HBM.
  epochs   the "adam_gpu" epoch loop alone (500 Adam epochs on a 1M-node
           k=15-style graph): one step = one full 500-epoch optimisation.
  serve    the served stream: Poisson arrivals of uneven fits into the fit
           server, fits/s and p50/p99 latency against sequential fits.
  fp32     the plain GPU form in FP32: n=100,000, d=50, exact FP32 kNN, 500
           epochs; trustworthiness@15 against the CPU optimisers.
  densmap  densMAP: n=500,000, d=50, k=30; local-radius Spearman against the
           CPU densMAP.
  plain    the reference's CPU-runnable case, n=10,000, d=50, k=15, 500
           epochs: the GPU fit beside the whole CPU run (kNN, graph, SGD),
           whose trustworthiness@15 is the reference.

The CPU baseline is the C restatement of the reference's parallel Adam
(optimise_embedding_adam_parallel) timed on this host's cores on a bounded
sample of the same workload.

`roofline` prices the dominant kernel (the epoch kernel; the FP32 shortlist
for `fp32`) against HBM bandwidth or MFMA peak.  The epoch kernel's random
8-byte gathers and atomics each touch a line of an embedding that does not
fit one XCD's L2, so `request_floor` (fit and epochs) also prices it against
the rate of uniform random gathers measured in the same run
(mrs_probe_random_gather): requests per launch / that rate, against the
measured launch time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "manifolds-rs_amd"))

METRIC = "UMAP fit sec @1M×64 k=15 (recall@15=1.0 vs FP64); edge-updates/s; served p99"
HBM_PEAK_GBS = 8000.0


def edge_kernel_bytes(E: int, A, r: int, dens_A=0.0):
    """Algorithmic bytes of the edge-parallel epoch kernel over launches with A
    active edges in total (dens_A of them in density epochs): per launch one
    bit per edge of the firing bitmask's plane (E / 8 B), and per active edge
    16 B of record (the 8-byte compact record {j, slot under j} and its run's
    {i, slot offset}), the 2 x 8 B endpoint gathers, 2r x 8 B negative-sample
    gathers and 2 x 8 B gradient scatters, plus in a density epoch the two
    endpoints' 16-byte packed density factors and the 4-byte edge weight.
    The bitmask planes themselves are written ahead, 128 epochs at a time, by
    k_fire_bits (timed apart: `fire_bits`)."""
    return E / 8.0 + A * (48.0 + 16.0 * r) + dens_A * 36.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def pmc_traffic(workload: str, kernel_prefix: str):
    """HBM bytes per launch of a kernel, from the rocprofv3 FETCH_SIZE/WRITE_SIZE
    summaries under profiles/ (copy-kernel calibrated); None if absent."""
    import re

    for f in sorted((ROOT / "profiles").glob("*.md")):
        for block in re.findall(r"```json\n(.*?)```", f.read_text(), re.S):
            try:
                entries = json.loads(block)
            except ValueError:
                continue
            for key, e in entries.items():
                if key.startswith(f"{workload}:{kernel_prefix}"):
                    return e["bytes_per_launch"], f"profiles/{f.name}"
    return None, None


def request_floor(ctx, n: int, firing_per_launch: float, neg_rate: int, dim: int,
                  avg_launch_us: float, workload: str, kernel_prefix: str):
    """The epoch kernel's bound in random requests, of two kinds measured apart
    on the same n-row table: gathers (y_j and 2*neg_rate negatives per firing
    edge; y_i is an L2 hit, the edges being sorted by i) at the rate of
    uniform random 8-byte gathers, and FP32 atomic adds (the larger endpoint's
    `dim` per firing edge plus one per run of the smaller endpoint; counted
    per launch by rocprofv3, TCP_TCC_ATOMIC_WITHOUT_RET_REQ, in the profile
    summary) at the rate of uniform random atomic adds.  The two kinds overlap
    in the kernel, so the floor is the larger of the two times."""
    rate_g = ctx.probe_random_gather(n, 1 << 24, 10)
    rate_a = ctx.probe_random_atomic(n, 1 << 24, 10)
    gathers = firing_per_launch * (1 + 2 * neg_rate)
    atom, src = pmc_field(workload, kernel_prefix, "atomic_req_per_launch")
    atom_src = src
    if atom is None:  # no counter pass: the larger endpoint's atomics only
        atom, atom_src = firing_per_launch * dim, "model: dim per firing edge (no counter pass)"
    t_g, t_a = gathers / rate_g * 1e6, atom / rate_a * 1e6
    floor_us = max(t_g, t_a)
    return dict(probe="uniform random float2 gathers and FP32 atomic adds on an n-row table, "
                      "2^24 per launch, 8 per thread (mrs_probe_random_gather / _atomic)",
                gather_rate_g_per_s=rate_g / 1e9, atomic_rate_g_per_s=rate_a / 1e9,
                gathers_per_firing_edge=1 + 2 * neg_rate, gathers_per_launch=gathers,
                atomics_per_launch=atom, atomics_source=atom_src,
                gather_floor_us=t_g, atomic_floor_us=t_a, floor_us=floor_us,
                bound="atomics" if t_a >= t_g else "gathers",
                achieved_us=avg_launch_us, frac=floor_us / avg_launch_us)


def pmc_field(workload: str, kernel_prefix: str, field: str):
    """A per-launch counter figure of a kernel from the profile summaries'
    JSON blocks (see pmc_traffic); None if absent."""
    import re

    for f in sorted((ROOT / "profiles").glob("*.md")):
        for block in re.findall(r"```json\n(.*?)```", f.read_text(), re.S):
            try:
                entries = json.loads(block)
            except ValueError:
                continue
            for key, e in entries.items():
                if key.startswith(f"{workload}:{kernel_prefix}") and field in e:
                    return e[field], f"profiles/{f.name}"
    return None, None


def synthetic_knn_graph(n: int, k: int, n_epochs: int, seed: int = 42):
    """A UMAP-shaped graph for the epoch workload: exact kNN (k) of n points drawn
    from 50 2-D Gaussian blobs, membership weights exp(-(d - rho)/sigma) with
    sigma the row's mean excess distance, fuzzy union W + W^T - W*W^T, weak
    edges (< max/n_epochs) dropped.  Returns a row-sorted adjacency CSR."""
    from scipy.sparse import csr_matrix
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    centres = rng.normal(0, 20, size=(50, 2))
    pts = centres[rng.integers(0, 50, size=n)] + rng.normal(size=(n, 2))
    d, idx = cKDTree(pts).query(pts, k=k + 1, workers=-1)
    d, idx = d[:, 1:] ** 2, idx[:, 1:]
    rho = d[:, :1]
    sig = np.maximum((d - rho).mean(axis=1, keepdims=True), 1e-6)
    w = np.exp(-(d - rho) / sig).astype(np.float32)
    W = csr_matrix((w.ravel(), idx.ravel(), np.arange(0, n * k + 1, k)), shape=(n, n))
    S = (W + W.T - W.multiply(W.T)).tocsr()
    S.data[S.data < S.data.max() / n_epochs] = 0
    S.eliminate_zeros()
    S.sort_indices()
    return S.indptr.astype(np.int64), S.indices.astype(np.int32), S.data.astype(np.float32)


def bench_epochs(args):
    import torch

    import mrs_amd as M

    n, k, n_epochs = args.n, 15, 500
    t0 = time.time()
    adj = synthetic_knn_graph(n, k, n_epochs)
    log(f"[bench] graph n={n} nnz={adj[0][-1]} built in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(42)
    embd0 = rng.uniform(-10, 10, size=(n, 2)).astype(np.float32)
    p = M.UmapOptimParams.default_2d()
    p.n_epochs = n_epochs

    ctx = M.Context(0)
    st = M.UmapGpuState(ctx, embd0, *adj, p)
    E = st.n_edges
    per = st.read(M.BUF_PER)
    counts = M.schedule_active_counts(per, n_epochs)
    A_total = int(counts.sum())
    d_embd0 = torch.from_numpy(embd0).cuda()
    stream = torch.cuda.ExternalStream(ctx.stream)

    def step():
        st.reset(d_embd0)
        st.run_epochs(0, n_epochs, 42, M.EPOCH_EDGE_ATOMIC)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    torch.cuda.synchronize()
    times = []
    st.set_profiling(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        ev0.record(stream)
        step()
        ev1.record(stream)
        ev1.synchronize()
        times.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    st.set_profiling(False)
    kt = st.kernel_times()
    ms_per_step = float(np.mean(times))

    # algorithmic bytes of the edge kernel (edge_kernel_bytes)
    r = p.neg_sample_rate
    edge_ms, edge_launches = kt["epoch_edges"]
    bytes_edge_total = edge_kernel_bytes(E, A_total, r) + (n_epochs - 1) * E / 8.0
    edge_avg_ms = edge_ms / edge_launches
    bytes_per_launch = bytes_edge_total / n_epochs
    achieved = bytes_per_launch / (edge_avg_ms * 1e-3) / 1e9
    adam_ms, adam_launches = kt["adam_clear"]
    traffic, traffic_src = pmc_traffic("epochs", "k_epoch_edges<2")

    res = dict(
        metric=METRIC,
        value=ms_per_step / 1e3,
        unit="s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=ms_per_step,
        higher_is_better=False,
        scaling="weak",
        vs_baseline=None,
        dtype="f32",
        data="synthetic",
        config=dict(workload=f"adam_gpu epoch loop only: {n_epochs} epochs, n={n}, "
                    f"2-D output, k={k} kNN graph of 50 2-D blobs, E={E}",
                    n=n, k=k, n_epochs=n_epochs, n_edges=E, edge_updates=A_total),
        edge_updates_per_s=A_total / (ms_per_step / 1e3),
        kernels={kk: dict(total_ms=v[0], launches=v[1], avg_us=1e3 * v[0] / v[1])
                 for kk, v in kt.items()},
        roofline=dict(bound="hbm", kernel="k_epoch_edges", achieved=achieved,
                      peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                      traffic=traffic, traffic_source=traffic_src,
                      avg_launch_us=edge_avg_ms * 1e3,
                      algorithmic_bytes_per_launch=bytes_per_launch),
        request_floor=request_floor(ctx, n, A_total / n_epochs, r, 2, edge_avg_ms * 1e3,
                                    "epochs", "k_epoch_edges<2"),
        wall_s=wall,
    )
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_epochs(embd0, adj, p, args)
    return res


def cpu_baseline_epochs(embd0, adj, p, args):
    """Reference CPU path (parallel Adam restated in C, LUT pow, OpenMP) on a
    bounded sample: the first S epochs of the same 1M-node fit, scaled to 500."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as orc

    import mrs_amd as M

    cores = os.cpu_count() or 1
    S = args.cpu_epochs
    q = M.UmapOptimParams.default_2d()
    q.n_epochs = p.n_epochs
    # run S epochs of the 500-epoch schedule: n_epochs stays 500 for the lr decay,
    # the oracle is stopped after S epochs by giving it a truncated copy
    q2 = M.UmapOptimParams.default_2d()
    q2.n_epochs = S
    t0 = time.perf_counter()
    orc.optimise_adam_parallel(embd0, *adj, q2, 42, use_lut=True, n_threads=cores)
    dt = time.perf_counter() - t0
    return dict(value=dt * p.n_epochs / S, unit="s", cores=cores, kind="port",
                sample=f"{S} of {p.n_epochs} epochs of the same graph, scaled x{p.n_epochs / S:g}",
                measured_s=dt)


def bench_fit(args):
    import torch

    import mrs_amd as M
    import synthetic

    n, d, k, n_epochs = args.n, 64, 15, 500
    t0 = time.time()
    x, _ = synthetic.subspace(n=n, d=d)
    log(f"[bench] data {x.shape} generated in {time.time() - t0:.1f}s")
    ctx = M.Context(0)
    xd = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    p = M.UmapParamsGpu.default()
    p.k = k
    p.knn_precision = M.KNN_BF16_RERANK
    p.knn_shortlist = 4 * k
    p.optim_params.n_epochs = n_epochs
    stream = torch.cuda.ExternalStream(ctx.stream)

    def step():
        return M.umap_gpu(ctx, xd, p, seed=42)

    for _ in range(args.warmup):
        step()
        log(f"[bench] warmup fit: {M.last_fit_timings()}")
    torch.cuda.synchronize()
    # the timed fits copy the kNN they compute into these (one device-to-device
    # copy of n*k entries per fit, ~0.1 ms at 1M): the recall check runs on the
    # timed fit's own neighbour lists
    fit_idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
    fit_dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    M.fit_knn_copy_to(fit_idx, fit_dist)
    times, stages = [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t_s = time.perf_counter()
        out = step()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t_s) * 1e3)
        stages.append(M.last_fit_timings())
        log(f"[bench] fit {times[-1]:.1f} ms: {stages[-1]}")
    M.fit_knn_copy_to(None, None)
    ms_per_step = float(np.mean(times))
    st = {kk: float(np.mean([s_[kk] for s_ in stages])) for kk in stages[0]}
    # the same fit from host memory to host memory, as the reference's umap_gpu
    # is called (data copied in, [2][n] result copied out)
    h_times = []
    for _ in range(max(1, min(args.steps, 3))):
        t_s = time.perf_counter()
        M.umap_gpu(ctx, x, p, seed=42)
        h_times.append((time.perf_counter() - t_s) * 1e3)
    host_ms = float(np.mean(h_times))
    log(f"[bench] host-to-host fit {host_ms:.1f} ms (device-resident {ms_per_step:.1f} ms)")

    # kNN: the shortlist sweeps only the (query block, stage) pairs its bounds
    # cannot rule out; the MFMA work done is that fraction of the 2 n^2 d FLOPs
    # of the brute force (BF16 dense MFMA peak 2.5 PFLOP/s)
    _, _, kt = M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k, timed=True)
    flops = 2.0 * n * n * d
    achieved = kt["swept"] * flops / (kt["shortlist_ms"] * 1e-3) / 1e12
    ctx.set_knn_pruning(False)
    # the first full sweep of this shape times every database split count (the
    # autotuner); the second is the measurement
    M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k)
    _, _, kfull = M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k, timed=True)
    ctx.set_knn_pruning(True)
    knn_prune = dict(pruned_ms=kt["pack_ms"] + kt["shortlist_ms"] + kt["rerank_ms"],
                     full_sweep_ms=kfull["pack_ms"] + kfull["shortlist_ms"] + kfull["rerank_ms"],
                     pruned_shortlist_ms=kt["shortlist_ms"], full_shortlist_ms=kfull["shortlist_ms"],
                     layout_ms=kt["pack_ms"], swept=kt["swept"], cells=kt["cells"],
                     full_sweep_tflops=flops / (kfull["shortlist_ms"] * 1e-3) / 1e12)

    # plain FP32 GPU kNN beside it (same shortlist/re-rank machinery, FP32 MFMA)
    _, _, kt32 = M.knn_exhaustive(ctx, xd, k, M.KNN_FP32, 0, timed=True)
    knn_cmp = dict(bf16_rerank_ms=kt["pack_ms"] + kt["shortlist_ms"] + kt["rerank_ms"],
                   fp32_ms=kt32["pack_ms"] + kt32["shortlist_ms"] + kt32["rerank_ms"])
    knn_cmp["speedup_bf16_over_fp32"] = knn_cmp["fp32_ms"] / knn_cmp["bf16_rerank_ms"]
    knn_cmp["fp32_shortlist_tflops"] = kt32["swept"] * flops / (kt32["shortlist_ms"] * 1e-3) / 1e12

    # exact edge-updates: schedule replay of this fit's graph
    idx, dist = M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
    e0 = torch.zeros((n, 2), device="cuda")
    stt = M.UmapGpuState(ctx, e0, params=p.optim_params, graph=g)
    per = stt.read(M.BUF_PER)
    A_total = int(M.schedule_active_counts(per, n_epochs).sum())
    E = stt.n_edges
    # epoch kernel: one profiled 500-epoch run
    stt.reset(e0)
    stt.set_profiling(True)
    stt.run_epochs(0, n_epochs, 42, M.EPOCH_EDGE_ATOMIC)
    stt.set_profiling(False)
    ktimes = stt.kernel_times()
    edge_ms, edge_launches = ktimes["epoch_edges"]
    r = p.optim_params.neg_sample_rate
    bytes_edge = (edge_kernel_bytes(E, A_total, r) + (n_epochs - 1) * E / 8.0) / n_epochs
    epoch_gbs = bytes_edge / (edge_ms / edge_launches * 1e-3) / 1e9
    fb_ms, fb_n = ktimes.get("fire_bits", (0.0, 0))
    sl_traffic, sl_src = pmc_traffic("fit", "k_shortlist<true")
    ep_traffic, ep_src = pmc_traffic("fit", "k_epoch_edges<2")

    res = dict(
        metric=METRIC, value=ms_per_step / 1e3, unit="s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=ms_per_step, higher_is_better=False, scaling="weak",
        vs_baseline=None, dtype="bf16-mfma/f32", data="synthetic",
        config=dict(workload="umap_gpu fit, n=1,000,000 x d=64 f32 (50 Gaussians on a random "
                    "8-D subspace + N(0,0.05)), exhaustive kNN k=15 BF16 MFMA + FP32 re-rank of "
                    "64 candidates, spectral init, 500 Adam epochs, 2-D output",
                    n=n, d=d, k=k, n_epochs=n_epochs, n_edges=E, edge_updates=A_total),
        stages_ms=st,
        host_to_host=dict(ms_per_fit=host_ms, fits=len(h_times),
                          note="data [n][64] f32 from host memory, [2][n] result to host "
                               "memory, as umap_gpu's callers pass them; value above has the "
                               "data resident in HBM"),
        knn_fp32_vs_bf16=knn_cmp,
        knn_pruned_vs_full_sweep=knn_prune,
        edge_updates_per_s=A_total / (st["epochs_ms"] * 1e-3),
        # the dominant kernel: 500 launches of the epoch kernel per fit
        roofline=dict(bound="hbm", kernel="k_epoch_edges", achieved=epoch_gbs,
                      peak=HBM_PEAK_GBS, unit="GB/s", frac=epoch_gbs / HBM_PEAK_GBS,
                      traffic=ep_traffic, traffic_source=ep_src,
                      avg_launch_us=1e3 * edge_ms / edge_launches,
                      algorithmic_bytes_per_launch=bytes_edge,
                      fire_bits_ms_per_fit=fb_ms, fire_bits_launches=fb_n),
        request_floor=request_floor(ctx, n, A_total / n_epochs, r, 2,
                                    1e3 * edge_ms / edge_launches, "fit", "k_epoch_edges<2"),
        knn_kernel=dict(bound="mfma", kernel="k_shortlist<bf16> (pruned sweep, both passes)",
                        achieved=achieved, peak=2500.0, unit="TFLOP/s", frac=achieved / 2500.0,
                        traffic=None, traffic_per_dispatch_mean=sl_traffic,
                        traffic_source=sl_src,
                        avg_launch_ms=kt["shortlist_ms"], swept=kt["swept"],
                        algorithmic_flops_per_launch=kt["swept"] * flops,
                        brute_force_flops=flops),
    )
    if not args.no_cpu:
        # the last timed fit's own kNN lists
        knn_host = (fit_idx.cpu().numpy(), fit_dist.cpu().numpy())
        res["cpu_baseline"] = cpu_baseline_fit(x, g, knn_host, p, args)
        res["recall"] = res["cpu_baseline"]["knn_recall_vs_fp64"]
    return res


def cpu_baseline_fit(x, g, knn_host, p, args):
    """Reference CPU path on bounded samples of the same fit: exhaustive kNN
    (FP32 brute force, OpenMP) for Q query rows against all n, scaled by n/Q;
    host graph (smooth_knn_dist, knn_to_coo, symmetrise, filter) on the full
    kNN; spectral init (components, Lanczos, layout) on the full graph;
    parallel Adam (LUT pow, OpenMP) for S of 500 epochs, scaled.  Also the
    checker: recall@k and distance error of the timed fit's kNN against an
    FP64 brute force on R seeded rows."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as orc

    import mrs_amd as M

    cores = os.cpu_count() or 1
    n = x.shape[0]
    k = p.k
    Q = args.cpu_queries
    t0 = time.perf_counter()
    orc.knn_exact(x, k, 0, Q)
    t_knn = (time.perf_counter() - t0) * n / Q
    # recall@k and distance error of the GPU (BF16 + re-rank) kNN against an
    # FP64 brute force on a seeded row sample (the CPU oracle is the checker)
    rows = np.sort(np.random.default_rng(7).choice(n, size=args.recall_rows, replace=False))
    idx64, d64 = orc.knn_exact_rows_f64(x, k, rows)
    idx_h, dist_h = knn_host
    hits = sum(len(np.intersect1d(idx_h[q], idx64[a])) for a, q in enumerate(rows))
    recall = dict(rows=int(rows.size), recall_at_k=hits / (rows.size * k),
                  exact_rows=int(sum(np.array_equal(idx_h[q], idx64[a]) for a, q in enumerate(rows))),
                  max_rel_dist_err=float(np.max(np.abs(dist_h[rows] - d64) / np.maximum(d64, 1e-30))))
    # host graph stages on the full kNN (the GPU's exact result, copied to the host)
    idx_h, dist_h = knn_host
    t0 = time.perf_counter()
    sg, rh = orc.smooth_knn_dist(dist_h, k)
    rr, cc, vv = orc.knn_to_coo(idx_h, dist_h, sg, rh)
    rr, cc, vv = orc.symmetrise(n, rr, cc, vv, 1.0)
    orc.filter_weak_edges(rr, cc, vv, p.optim_params.n_epochs)
    t_graph = time.perf_counter() - t0
    r, c, v = g.coo()
    t0 = time.perf_counter()
    orc.spectral_layout(n, r, c, v, n_comp=2, seed=42, data=x)
    t_spec = time.perf_counter() - t0
    sys.path.insert(0, str(ROOT / "tests"))
    from helpers import coo_to_adjacency

    adj = coo_to_adjacency(n, r, c, v)
    q2 = M.UmapOptimParams.default_2d()
    q2.n_epochs = args.cpu_epochs
    init = np.random.default_rng(0).uniform(-10, 10, size=(n, 2)).astype(np.float32)
    t0 = time.perf_counter()
    orc.optimise_adam_parallel(init, *adj, q2, 42, use_lut=True, n_threads=cores)
    t_adam = (time.perf_counter() - t0) * p.optim_params.n_epochs / args.cpu_epochs
    return dict(value=t_knn + t_graph + t_spec + t_adam, unit="s", cores=cores, kind="port",
                build="gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (oracle/Makefile)",
                extrapolated=True,
                knn_recall_vs_fp64=recall,
                sample=f"kNN: {Q} of {n} query rows x all {n} (scaled x{n / Q:g}); graph: "
                       f"full; spectral init: full; parallel Adam: {args.cpu_epochs} of "
                       f"{p.optim_params.n_epochs} epochs on the same graph (scaled x"
                       f"{p.optim_params.n_epochs / args.cpu_epochs:g})",
                knn_s=t_knn, graph_s=t_graph, spectral_s=t_spec, adam_s=t_adam)


def fit_steps(M, ctx, xd, p, args, dens=None):
    """W untimed and K timed umap_gpu/densmap_gpu fits with the data resident in
    HBM: (mean ms per fit, mean stage timings, the last output)."""
    import torch

    for _ in range(args.warmup):
        M.umap_gpu(ctx, xd, p, seed=42, dens=dens)
    torch.cuda.synchronize()
    times, stages, out = [], [], None
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t_s = time.perf_counter()
        out = M.umap_gpu(ctx, xd, p, seed=42, dens=dens)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t_s) * 1e3)
        stages.append(M.last_fit_timings())
        log(f"[bench] fit {times[-1]:.1f} ms: {stages[-1]}")
    st = {kk: float(np.mean([s_[kk] for s_ in stages])) for kk in stages[0]}
    return float(np.mean(times)), st, out


def oracle_graph(orc, knn, n, k, n_epochs):
    """The oracle's host graph stages on a kNN: (rows, cols, vals) filtered COO."""
    idx, dist = knn
    sg, rh = orc.smooth_knn_dist(dist, k)
    r, c, v = orc.knn_to_coo(idx, dist, sg, rh)
    r, c, v = orc.symmetrise(n, r, c, v, 1.0)
    return orc.filter_weak_edges(r, c, v, n_epochs)


def sample_trustworthiness(x, emb, m=5000, k=15, seed=11):
    """trustworthiness@k on a seeded m-point sample (the sample's own
    neighbourhoods in both spaces; the same sample for every embedding)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from metrics import trustworthiness

    rows = np.sort(np.random.default_rng(seed).choice(x.shape[0], size=min(m, x.shape[0]),
                                                      replace=False))
    return trustworthiness(x[rows], np.asarray(emb)[rows], k)


def bench_fp32(args):
    """The plain GPU form, FP32 throughout: n=100,000, d=50 (20 anisotropic
    Gaussians + 5 % uniform background), exact FP32 kNN k=15, 500 Adam epochs,
    2-D output.  Quality: trustworthiness@15 (5 000-point sample) of the GPU
    fit, of the CPU parallel Adam (the reference's CPU optimiser, restated) and
    of the CPU restatement of adam_gpu (identical hashed negatives), both
    started from the fit's own graph and spectral layout."""
    import torch

    import mrs_amd as M
    import synthetic

    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import pyoracle as orc
    from helpers import coo_to_adjacency

    n, d, k, n_epochs = 100_000, 50, 15, 500
    x, _ = synthetic.anisotropic(n=n, d=d)
    ctx = M.Context(0)
    xd = torch.from_numpy(x).cuda()
    p = M.UmapParamsGpu.default()
    p.k = k
    p.knn_precision = M.KNN_FP32
    p.optim_params.n_epochs = n_epochs
    ms, st, out = fit_steps(M, ctx, xd, p, args)
    e_gpu = out.T.cpu().numpy()
    _, _, kt = M.knn_exhaustive(ctx, xd, k, M.KNN_FP32, 0, timed=True)
    flops = 2.0 * n * n * d
    achieved = kt["swept"] * flops / (kt["shortlist_ms"] * 1e-3) / 1e12
    # the fit's graph and layout, then the CPU optimisers from them
    idx, dist = M.knn_exhaustive(ctx, x, k, M.KNN_FP32)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
    init = g.spectral_layout(2, seed=42, data=x)
    adj = coo_to_adjacency(n, *g.coo())
    q = M.UmapOptimParams.default_2d()
    q.n_epochs = n_epochs
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    e_par = orc.optimise_adam_parallel(init, *adj, q, 42, use_lut=True, n_threads=cores)
    t_par = time.perf_counter() - t0
    e_hash = orc.optimise_adam_gpu(init, *adj, q, 42)
    tw = dict(gpu=sample_trustworthiness(x, e_gpu), cpu_parallel_adam=sample_trustworthiness(x, e_par),
              cpu_adam_gpu_restated=sample_trustworthiness(x, e_hash))
    tw["gpu_minus_cpu_parallel"] = tw["gpu"] - tw["cpu_parallel_adam"]
    log(f"[bench] trustworthiness@15: {tw}")
    t0 = time.perf_counter()
    Q = args.cpu_queries
    orc.knn_exact(x, k, 0, Q)
    t_knn = (time.perf_counter() - t0) * n / Q
    t0 = time.perf_counter()
    oracle_graph(orc, (idx, dist), n, k, n_epochs)
    t_graph = time.perf_counter() - t0
    return dict(
        metric="plain FP32 GPU fit sec @100k×50 k=15 (trustworthiness@15 vs CPU parallel Adam)",
        value=ms / 1e3, unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=False, scaling="weak", vs_baseline=None,
        dtype="f32-mfma/f32", data="synthetic",
        config=dict(workload="umap_gpu fit, n=100,000 x d=50 f32 (20 anisotropic Gaussians, "
                    "5% uniform background), exhaustive FP32 kNN k=15, spectral init, 500 Adam "
                    "epochs, 2-D output", n=n, d=d, k=k, n_epochs=n_epochs),
        stages_ms=st,
        roofline=dict(bound="mfma", kernel="k_shortlist<f32>", achieved=achieved, peak=157.3,
                      unit="TFLOP/s", frac=achieved / 157.3, traffic=None,
                      avg_launch_ms=kt["shortlist_ms"], swept=kt["swept"],
                      algorithmic_flops_per_launch=kt["swept"] * flops),
        trustworthiness_at_15=tw,
        cpu_baseline=dict(value=t_knn + t_graph + t_par, unit="s", cores=cores, kind="port",
                          sample=f"kNN: {Q} of {n} query rows x all {n} (scaled x{n / Q:g}); "
                                 f"graph: full; parallel Adam: all {n_epochs} epochs; spectral "
                                 f"init not included",
                          knn_s=t_knn, graph_s=t_graph, adam_s=t_par),
    )


def bench_knn(args):
    """The exhaustive kNN of the precision configuration alone: 1M x 64 (50
    Gaussians on an 8-D subspace), k = 15, BF16 MFMA shortlist of 64 with the
    exact FP32 re-rank, data in HBM; --knn-full-sweep sweeps every database
    stage (the brute force) instead of the pruned sweep.  Reports the search
    time and the shortlist kernel's MFMA rate (the FLOPs of the stages swept)."""
    import torch

    import mrs_amd as M
    import synthetic

    n, d, k = args.n, 64, 15
    x, _ = synthetic.subspace(n=n, d=d, n_clusters=50, seed=42)
    xd = torch.from_numpy(x).cuda()
    ctx = M.Context(0)
    ctx.set_knn_pruning(not args.knn_full_sweep)
    for _ in range(max(args.warmup, 1)):  # the first call of a shape also autotunes
        M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k)
    torch.cuda.synchronize()
    times, kts = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, _, kt = M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 4 * k, timed=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        kts.append(kt)
    sl = float(np.mean([kt["shortlist_ms"] for kt in kts]))
    flops = 2.0 * n * n * d * kts[-1]["swept"]
    tflops = flops / (sl * 1e-3) / 1e12
    return dict(metric="exhaustive kNN seconds @1M x 64 k=15 (BF16 shortlist + FP32 re-rank)",
                value=float(np.mean(times)), unit="s", n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=1e3 * float(np.mean(times)),
                higher_is_better=False, scaling="weak", vs_baseline=None, dtype="bf16-mfma/f32",
                data="synthetic",
                config=dict(workload=f"kNN {n} x {d}, k=15, "
                                     f"{'full' if args.knn_full_sweep else 'pruned'} sweep"),
                roofline=dict(bound="mfma", kernel="k_shortlist<bf16>", achieved=tflops,
                              peak=2500.0, unit="TFLOP/s", frac=tflops / 2500.0, traffic=None,
                              avg_launch_ms=sl, swept=kts[-1]["swept"]),
                timings=kts[-1])


def bench_plain(args):
    """The reference's CPU-runnable case beside the GPU fit: n=10,000, d=50 (10
    isotropic Gaussians, centres ~N(0, 25 I)), exhaustive kNN k=15, spectral
    layout, 500 epochs, neg rate 5, a and b fitted to min_dist 0.1 / spread 1.0.
    The CPU run (the oracle's exhaustive kNN, host graph and plain SGD,
    optimise_embedding_sgd, src/training/umap_optimisers.rs:653-865, from the
    same spectral layout) is the trustworthiness reference for the GPU fit's
    Adam.  At this size the GPU fit is bound by launch latency: the epoch
    kernel's roofline is reported with its per-launch time."""
    import torch

    import mrs_amd as M
    import synthetic

    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import pyoracle as orc
    from helpers import coo_to_adjacency

    n, d, k, n_epochs = 10_000, 50, 15, 500
    x, _ = synthetic.plain_cpu(n=n, d=d)
    a, b = orc.fit_params(0.1, 1.0)
    ctx = M.Context(0)
    xd = torch.from_numpy(x).cuda()
    p = M.UmapParamsGpu.default()
    p.k = k
    p.knn_precision = M.KNN_FP32
    p.optim_params.n_epochs = n_epochs
    p.optim_params.a, p.optim_params.b = a, b
    ms, st, out = fit_steps(M, ctx, xd, p, args)
    e_gpu = out.T.cpu().numpy()
    # epoch kernel: algorithmic bytes per launch from the exact schedule replay
    idx, dist = M.knn_exhaustive(ctx, x, k, M.KNN_FP32)
    g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
    init = g.spectral_layout(2, seed=42, data=x)
    stt = M.UmapGpuState(ctx, torch.from_numpy(init).cuda(), params=p.optim_params, graph=g)
    A_total = int(M.schedule_active_counts(stt.read(M.BUF_PER), n_epochs).sum())
    E = stt.n_edges
    stt.set_profiling(True)
    stt.run_epochs(0, n_epochs, 42, M.EPOCH_EDGE_ATOMIC)
    stt.set_profiling(False)
    edge_ms, edge_launches = stt.kernel_times()["epoch_edges"]
    stt.close()
    r = p.optim_params.neg_sample_rate
    bytes_edge = (edge_kernel_bytes(E, A_total, r) + (n_epochs - 1) * E / 8.0) / n_epochs
    epoch_gbs = bytes_edge / (edge_ms / edge_launches * 1e-3) / 1e9
    # the reference's CPU run, whole: kNN, graph, SGD (same layout to start from)
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    cidx, cdist = orc.knn_exact(x, k)
    t_knn = time.perf_counter() - t0
    knn_rows_equal = float(np.mean(np.all(cidx == idx, axis=1)))
    t0 = time.perf_counter()
    rr, cc, vv = oracle_graph(orc, (cidx, cdist), n, k, n_epochs)
    t_graph = time.perf_counter() - t0
    q = M.UmapOptimParams.default_2d()
    q.n_epochs = n_epochs
    q.a, q.b = a, b
    t0 = time.perf_counter()
    e_sgd = orc.optimise_sgd(init, *coo_to_adjacency(n, rr, cc, vv), q, 42)
    t_sgd = time.perf_counter() - t0
    tw = dict(gpu=sample_trustworthiness(x, e_gpu), cpu_sgd=sample_trustworthiness(x, e_sgd))
    tw["gpu_minus_cpu_sgd"] = tw["gpu"] - tw["cpu_sgd"]
    log(f"[bench] trustworthiness@15: {tw}")
    return dict(
        metric="plain UMAP GPU fit sec @10k×50 k=15 (trustworthiness@15 vs the CPU SGD run)",
        value=ms / 1e3, unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=False, scaling="weak", vs_baseline=None,
        dtype="f32-mfma/f32", data="synthetic",
        config=dict(workload="umap_gpu fit, n=10,000 x d=50 f32 (10 isotropic Gaussians, "
                    "centres ~N(0, 25 I)), exhaustive FP32 kNN k=15, spectral init, 500 Adam "
                    "epochs, a/b from min_dist 0.1 spread 1.0, 2-D output",
                    n=n, d=d, k=k, n_epochs=n_epochs, n_edges=E, edge_updates=A_total),
        stages_ms=st,
        roofline=dict(bound="hbm", kernel="k_epoch_edges", achieved=epoch_gbs,
                      peak=HBM_PEAK_GBS, unit="GB/s", frac=epoch_gbs / HBM_PEAK_GBS,
                      traffic=None, avg_launch_us=1e3 * edge_ms / edge_launches,
                      algorithmic_bytes_per_launch=bytes_edge),
        trustworthiness_at_15=tw,
        knn_rows_equal_to_oracle=knn_rows_equal,
        cpu_baseline=dict(value=t_knn + t_graph + t_sgd, unit="s", cores=cores, kind="port",
                          sample=f"whole run: exhaustive kNN of all {n} rows, graph, SGD "
                                 f"{n_epochs} epochs (single thread); spectral init not "
                                 f"included",
                          knn_s=t_knn, graph_s=t_graph, sgd_s=t_sgd),
    )


def log_radii(emb, rows, cols, a, b):
    """The embedding's local log radii of densMAP (umap_dens_radii,
    src/training/umap_optimiser_gpu.rs:487-544, in float64): ln(1e-8 +
    sum_j phi_ij d_ij^2 / sum_j phi_ij), phi = 1 / (1 + a d^(2b)), over the
    graph's edges."""
    e = np.asarray(emb, np.float64)
    d2 = ((e[rows] - e[cols]) ** 2).sum(1)
    phi = 1.0 / (1.0 + a * d2 ** b)
    n = e.shape[0]
    num = np.bincount(rows, weights=phi * d2, minlength=n)
    den = np.maximum(np.bincount(rows, weights=phi, minlength=n), np.finfo(np.float32).eps)
    return np.log(1e-8 + num / den)


def bench_densmap(args):
    """densMAP on the GPU: n=500,000, d=50 (30 Gaussians with per-cluster sd
    log-uniform in [0.2, 5]), k=30, lambda 2.0, frac 0.3, var_shift 0.1, 500
    Adam epochs, BF16 kNN + exact re-rank.  Reports the fit time and the
    Spearman correlation of the original and embedded local radii (R_o from
    the density state, R_e from the final layout over the graph); the same
    correlation for the GPU and for the CPU densMAP (the oracle's parallel Adam
    with the density term) on a 20,000-point draw of the same distribution."""
    import torch

    import mrs_amd as M
    import synthetic

    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import pyoracle as orc
    from helpers import coo_to_adjacency
    from metrics import spearman

    k, n_epochs = 30, 500
    dp = M.DensParams.densmap_default()
    p = M.UmapParamsGpu.default()
    p.k = k
    p.knn_precision = M.KNN_BF16_RERANK
    p.optim_params.n_epochs = n_epochs
    a, b = p.optim_params.a, p.optim_params.b
    ctx = M.Context(0)

    def spearman_of(x, emb, idx, dist):
        g = M.FuzzyGraph(ctx, idx, dist, n_epochs=n_epochs)
        r_o, _, _ = g.dens_state(idx, dist)
        rows, cols, _ = g.coo()
        return spearman(r_o, log_radii(emb, rows, cols, a, b)), g

    n = 500_000
    x, _ = synthetic.uneven_density(n=n)
    xd = torch.from_numpy(x).cuda()
    ms, st, out = fit_steps(M, ctx, xd, p, args, dens=dp)
    idx, dist = M.knn_exhaustive(ctx, xd, k, M.KNN_BF16_RERANK, 0)
    rho_full, g_full = spearman_of(x, out.T.cpu().numpy(), idx, dist)
    # the dominant kernel: the edge epoch (350 plain and 150 density launches),
    # timed with HIP events on a state of the fit's graph and density state.
    # Algorithmic bytes per launch: edge_kernel_bytes (a density launch adds
    # the two endpoints' packed density factors and the edge weight)
    r_o, ms_o, tot_o = g_full.dens_state(idx, dist)
    ds = M.DensState(params=dp, r=r_o.ctypes.data, mu_sum=ms_o.ctypes.data, mu_tot=tot_o)
    e0 = torch.zeros((n, 2), device="cuda")
    sd = M.UmapGpuState(ctx, e0, params=p.optim_params, dens=ds, graph=g_full)
    E = sd.n_edges
    act = M.schedule_active_counts(sd.read(M.BUF_PER), n_epochs)
    rneg = p.optim_params.neg_sample_rate
    dens_ep = np.array([(e + 1) / n_epochs > 1.0 - dp.frac for e in range(n_epochs)])
    alg_bytes = float(np.sum(edge_kernel_bytes(E, act, rneg, dens_ep * act)))
    sd.reset(out.T.contiguous())
    sd.set_profiling(True)
    sd.run_epochs(0, n_epochs, 42, M.EPOCH_EDGE_ATOMIC)
    sd.set_profiling(False)
    kt = sd.kernel_times()
    edge_ms, edge_n = kt["epoch_edges"]
    dens_gbs = alg_bytes / (edge_ms * 1e-3) / 1e9
    # counter traffic: the plain and density instantiations' per-launch HBM
    # bytes from the profile summary, weighted by their launches in the fit
    span = 128 if E >= (1 << 20) else 64
    t_plain, src = pmc_traffic("densmap", f"k_epoch_edges<2, false, {span},")
    t_dens, _ = pmc_traffic("densmap", f"k_epoch_edges<2, true, {span},")
    n_dens = int(dens_ep.sum())
    traffic = (None if t_plain is None or t_dens is None
               else ((n_epochs - n_dens) * t_plain + n_dens * t_dens) / n_epochs)
    roof = dict(bound="hbm", kernel="k_epoch_edges (plain and density launches)",
                achieved=dens_gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=dens_gbs / HBM_PEAK_GBS,
                traffic=traffic, traffic_source=src, avg_launch_us=1e3 * edge_ms / edge_n,
                algorithmic_bytes_per_launch=alg_bytes / n_epochs,
                density_update_ms_per_call=kt["dens_update"][0] / max(kt["dens_update"][1], 1))
    log(f"[bench] densMAP epoch kernel {roof['avg_launch_us']:.1f} us/launch, "
        f"{dens_gbs:.0f} GB/s algorithmic")
    plain = M.umap_gpu(ctx, xd, p, seed=42).T.cpu().numpy()
    rho_plain, _ = spearman_of(x, plain, idx, dist)
    log(f"[bench] Spearman(R_o, R_e) at n={n}: densMAP {rho_full:.4f}, plain UMAP {rho_plain:.4f}")
    # GPU vs CPU densMAP on a smaller draw
    ns = 20_000
    xs, _ = synthetic.uneven_density(n=ns, seed=7)
    e_gpu = M.umap_gpu(ctx, xs, p, seed=42, dens=dp).T
    sidx, sdist = M.knn_exhaustive(ctx, xs, k, M.KNN_FP32)
    rho_gpu_s, gs = spearman_of(xs, e_gpu, sidx, sdist)
    rr, cc, vv = oracle_graph(orc, (sidx, sdist), ns, k, n_epochs)
    ok, R, _, mu_tot = orc.dens_state(ns, rr, cc, vv, sidx, sdist)
    init = gs.spectral_layout(2, seed=42, data=xs)
    q = M.UmapOptimParams.default_2d()
    q.n_epochs = n_epochs
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    e_cpu = orc.optimise_adam_parallel(init, *coo_to_adjacency(ns, rr, cc, vv), q, 42,
                                       n_threads=cores, dens_r=R,
                                       dens_params=(dp.lambda_, dp.frac, dp.var_shift),
                                       mu_tot=mu_tot)
    t_cpu = time.perf_counter() - t0
    rho_cpu_s, _ = spearman_of(xs, e_cpu, sidx, sdist)
    log(f"[bench] Spearman at n={ns}: GPU {rho_gpu_s:.4f}, CPU densMAP {rho_cpu_s:.4f}")
    return dict(
        metric="densMAP GPU fit sec @500k×50 k=30 (local-radius Spearman vs CPU densMAP)",
        value=ms / 1e3, unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=False, scaling="weak", vs_baseline=None,
        dtype="bf16-mfma/f32, f64 reductions", data="synthetic",
        config=dict(workload="densmap_gpu fit, n=500,000 x d=50 f32 (30 Gaussians, sd "
                    "log-uniform in [0.2, 5]), k=30, lambda 2.0, frac 0.3, var_shift 0.1, "
                    "BF16 kNN + exact re-rank, spectral init, 500 Adam epochs",
                    n=n, d=50, k=k, n_epochs=n_epochs),
        stages_ms=st,
        roofline=roof,
        local_radius_spearman=dict(densmap_gpu=rho_full, plain_umap_gpu=rho_plain,
                                   sample_n=ns, sample_gpu=rho_gpu_s,
                                   sample_cpu_densmap=rho_cpu_s),
        cpu_baseline=dict(value=t_cpu * n / ns, unit="s", cores=cores, kind="port",
                          sample=f"parallel Adam with the density term, {n_epochs} epochs on a "
                                 f"{ns}-point draw, scaled by n ({n / ns:g}x); kNN, graph and "
                                 f"init not included", adam_s_sample=t_cpu),
    )


def copy_probe(mib: int) -> None:
    """One device-to-device copy of `mib` MiB by a plain elementwise kernel: it
    reads and writes exactly mib MiB, so HBM byte counters read on it give the
    factor between counter units and bytes."""
    import torch

    src = torch.ones(mib * 2**20 // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: {mib} MiB read + {mib} MiB written")
    del src, dst


def bench_serve(args):
    """Served stream: requests with n log-uniform in [2000, 200000], d in {32, 50,
    100}, 5-40 Gaussian clusters, k = 15, 200 Adam epochs.  Throughput (value):
    every request queued at once, the server admitting them as HBM and slots
    free up.  Latency: Poisson arrivals at a fixed rate the server sustains
    (--load of the measured throughput), p50/p99 from submission to result.
    Both against the same requests fitted one after another with
    mrs_umap_gpu."""
    import threading

    if args.prep_workers > 8 or args.hw_queues > 16:
        # one run at 12 workers and 32 hardware queues ended in a device memory
        # fault with no cause found; settings past those that ran cleanly are refused
        raise SystemExit("serve: at most 8 prep workers and 16 hardware queues")
    if args.hw_queues:
        # hardware queues per process: each prep worker and the batcher get one,
        # so their small kernels overlap instead of queueing behind each other
        # (set before the HIP runtime starts)
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import mrs_amd as M

    try:  # a switched-off server refuses here, before any fit runs
        M.Server(0, 0, args.prep_workers, args.max_active).close()
    except M.ManifoldsError as e:
        raise SystemExit(f"serve workload unavailable: {e}")
    rng = np.random.default_rng(42)
    # a pool of distinct datasets covering the size range; requests draw from it
    # with their own fit seeds (every request recomputes kNN, graph and layout)
    pool = []
    for i in range(args.pool):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(200_000))))
        d = int(rng.choice([32, 50, 100]))
        c = int(rng.integers(5, 41))
        centres = rng.normal(0, 5, size=(c, d)).astype(np.float32)
        lab = rng.integers(0, c, size=n)
        pool.append((centres[lab] + rng.normal(size=(n, d)).astype(np.float32)).astype(np.float32))
    reqs = [(int(rng.integers(0, len(pool))), int(rng.integers(0, 2**31))) for _ in range(args.requests)]
    p = M.UmapParamsGpu.default()
    p.k = 15
    p.optim_params.n_epochs = 200
    p.knn_precision = M.KNN_BF16_RERANK
    log(f"[bench] serve: {len(pool)} datasets, {len(reqs)} requests, "
        f"{sum(pool[i].shape[0] for i, _ in reqs) / len(reqs):.0f} points per request on average")
    # sequential baseline: the same requests (a seeded sample), one mrs_umap_gpu at a time
    ctx = M.Context(0)
    M.umap_gpu(ctx, pool[0], p, seed=1)  # warm-up
    sample = reqs[: args.seq_sample]
    stage_sum = {}
    t0 = time.perf_counter()
    for i, sd in sample:
        M.umap_gpu(ctx, pool[i], p, seed=sd)
        for kk, v in M.last_fit_timings().items():
            stage_sum[kk] = stage_sum.get(kk, 0.0) + v
    seq_s = (time.perf_counter() - t0) / len(sample)
    seq_fps = 1.0 / seq_s
    seq_stages = {kk: v / len(sample) for kk, v in stage_sum.items()}
    log(f"[bench] sequential: {seq_s * 1e3:.1f} ms per fit ({seq_fps:.2f} fits/s); "
        f"mean stages {seq_stages}")
    with M.Server(0, 0, args.prep_workers, args.max_active) as sv:
        sv.wait(sv.submit(pool[0], p, seed=1))  # warm-up

        def run(reqs_run, arrivals):
            """Submit reqs_run at the given arrival offsets (None: all at once);
            returns (span s, latencies ms)."""
            tickets = [None] * len(reqs_run)
            lat = [None] * len(reqs_run)

            def collector():
                for q in range(len(reqs_run)):
                    while tickets[q] is None:
                        time.sleep(1e-4)
                    _, times = sv.wait(tickets[q])
                    lat[q] = times["latency_ms"]

            th = threading.Thread(target=collector)
            start = time.perf_counter()
            th.start()
            for q, (i, sd) in enumerate(reqs_run):
                if arrivals is not None:
                    delay = start + arrivals[q] - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                tickets[q] = sv.submit(pool[i], p, seed=sd)
            th.join()
            return time.perf_counter() - start, np.array(lat)

        # throughput: every request submitted at once, the server admits them
        # as memory and slots free up
        span, lat_burst = run(reqs, None)
        served_fps = len(reqs) / span
        st = sv.stats()
        # latency: Poisson arrivals at a fixed rate the server sustains (a share
        # of the measured throughput), so the latency is service, not backlog
        rate = args.rate if args.rate > 0 else args.load * served_fps
        n_lat = min(len(reqs), args.latency_requests)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_lat))
        span_l, lat = run(reqs[:n_lat], arrivals)
    log(f"[bench] served {served_fps:.1f} fits/s (sequential {seq_fps:.1f}); at {rate:.1f}/s "
        f"p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms")
    return dict(
        metric="served fits/s and p99 latency (UMAP fits, k=15, 200 epochs)",
        value=served_fps, unit="fits/s", n_gpus=1, steps=1, warmup=1,
        ms_per_step=span * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="bf16-mfma/f32", data="synthetic",
        config=dict(workload=f"served stream: {len(reqs)} requests, n log-uniform [2000, 200000], "
                    f"d in {{32, 50, 100}}, 5-40 clusters, k=15, 200 epochs, BF16 kNN; "
                    f"throughput with every request queued at once, latency over {n_lat} "
                    f"Poisson arrivals at {rate:.2f}/s", requests=len(reqs),
                    latency_requests=n_lat, rate_per_s=rate, load=rate / served_fps,
                    prep_workers=args.prep_workers, max_active=args.max_active,
                    hw_queues=os.environ.get("GPU_MAX_HW_QUEUES")),
        sequential_fits_per_s=seq_fps, sequential_sample=len(sample),
        sequential_stage_ms=seq_stages,
        speedup_vs_sequential=served_fps / seq_fps,
        latency_ms=dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                        mean=float(lat.mean()), max=float(lat.max()),
                        achieved_rate_per_s=n_lat / span_l),
        sequential_latency_ms=seq_s * 1e3,
        burst_latency_ms=dict(p50=float(np.percentile(lat_burst, 50)),
                              p99=float(np.percentile(lat_burst, 99))),
        server=st,
    )


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fit")
    ap.add_argument("--cpu-queries", type=int, default=512)
    ap.add_argument("--recall-rows", type=int, default=10000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-epochs", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--requests", type=int, default=2000, help="serve: number of requests")
    ap.add_argument("--pool", type=int, default=48, help="serve: distinct datasets")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serve: arrivals per second of the latency run (0 = --load x the "
                         "measured throughput)")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serve: latency run's arrival rate as a share of the throughput")
    ap.add_argument("--latency-requests", type=int, default=1000,
                    help="serve: requests in the latency run")
    ap.add_argument("--seq-sample", type=int, default=40, help="serve: sequential sample size")
    ap.add_argument("--prep-workers", type=int, default=8)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="serve: GPU_MAX_HW_QUEUES for the run (0 = leave as is; at most 16)")
    ap.add_argument("--max-active", type=int, default=256)
    ap.add_argument("--knn-full-sweep", action="store_true",
                    help="knn: sweep every database stage (no pruning)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy a device buffer of this many MiB (a plain copy kernel of "
                         "known size, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps(dict(error="this build measures one GPU only; multi-GPU runs are "
                                    "out of scope (the kNN would shard by query rows and the "
                                    "epochs by node ranges)", n_gpus=args.gpus)))
        sys.exit(2)
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    if args.workload == "fit":
        res = bench_fit(args)
    elif args.workload == "epochs":
        res = bench_epochs(args)
    elif args.workload == "serve":
        res = bench_serve(args)
    elif args.workload == "fp32":
        res = bench_fp32(args)
    elif args.workload == "densmap":
        res = bench_densmap(args)
    elif args.workload == "plain":
        res = bench_plain(args)
    elif args.workload == "knn":
        res = bench_knn(args)
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    if "autotune" not in res:
        import mrs_amd as M

        # launch shapes the run chose by measurement (family, shape class,
        # choice, milliseconds per candidate); decided in the warm-up
        res["autotune"] = M.autotune_records()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
