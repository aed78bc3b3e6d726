"""Benchmark: batched weighted leaf-ridge fits on one MI355X.

Workload: 4,096 independent leaves, d=8, full quadratic basis (p=45), leaf
sizes log-uniform in [64, 16384] rows (~12.2M rows), x ~ U[-1,1]^8, weights
products of 1-6 sigmoid gates, target a 6-bump Gaussian mixture, ridge 1e-6
(trace-scaled), seed 2.  One step = one fit of all 4,096 leaves with inputs
resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
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
sys.path.insert(0, str(ROOT / "adaptive-regression-trees-art-_amd"))

METRIC = ("leaf fits/s & tree builds/s at ≤1e-6 rel err vs FP64 plain; "
          "p50/p99 fit latency")
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 2500.0, "bf16": 2500.0}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="mixed", choices=("fp64", "mixed"))
    ap.add_argument("--leaves", type=int, default=4096)
    ap.add_argument("--chunk-rows", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", choices=("all", "ridge", "ridge_kernels", "sampler", "split",
                                       "tree", "serving", "serving_stream"),
                    default="all",
                    help="run one leg (profiling); the JSON line keeps the ridge metric; "
                         "ridge_kernels runs only --warmup + --steps fits of --precision "
                         "(the counter and trace driver); serving_stream serves the "
                         "--requests stream once at --precision (the serving trace driver)")
    ap.add_argument("--sampler-steps", type=int, default=3)
    ap.add_argument("--tree-steps", type=int, default=3)
    ap.add_argument("--polytopes", type=int, default=65536)
    ap.add_argument("--requests", type=int, default=1000,
                    help="requests of the serving stream (its first requests, in order)")
    ap.add_argument("--rate", type=float, default=200.0, help="serving arrivals per second")
    ap.add_argument("--split-nodes", type=int, default=2048,
                    help="nodes of the annealed split leg (64 restarts each)")
    ap.add_argument("--max-split-parts", type=int, default=0,
                    help="cap the workgroups per node of the device split search (0: the "
                         "library's policy, up to 8; 1: one workgroup per node)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after timing, run a plain 1 GiB device copy three times: the "
                         "calibration kernel for FETCH_SIZE / WRITE_SIZE counter runs")
    return ap.parse_args()


# --------------------------------------------------------------------------- CPU baselines
#
# Every CPU baseline runs before the first device call (so the worker
# processes are forked from a process that has not initialized the GPU):
# one thread, then all usable cores (os.sched_getaffinity, at most 16, the
# box's share for one GPU) as forked workers over balanced shards, each
# worker with one BLAS thread.  The full host build of the plain fit runs in
# a background process during the GPU legs and is collected at the end.

_CPU = {}  # inputs shared with forked workers (copy-on-write)


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _run_parallel(fn, shards):
    """Wall seconds and results of fn over shards on a fork pool."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    ctx = multiprocessing.get_context("fork")
    with ProcessPoolExecutor(max_workers=len(shards), mp_context=ctx) as pool:
        list(pool.map(_warm_worker, range(len(shards))))  # workers up before timing
        t0 = time.perf_counter()
        out = list(pool.map(fn, shards))
        dt = time.perf_counter() - t0
    return dt, out


def _warm_worker(_):
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)
    import oracle.ridge  # noqa: F401
    import oracle.sampling  # noqa: F401
    import oracle.split  # noqa: F401
    return os.getpid()


def _ridge_shard(leaves):
    from threadpoolctl import threadpool_limits

    from oracle.ridge import leaf_fit_fast

    wl = _CPU["ridge_sample"].subset(np.asarray(leaves))
    with threadpool_limits(1):
        leaf_fit_fast(wl.X, wl.y, wl.w, wl.offsets, _CPU["ridge"])
    return len(leaves)


def _sampler_shard(polys):
    from threadpoolctl import threadpool_limits

    from oracle import sampling as S

    wl, C, moves, thin, d = _CPU["polytopes"], 64, 1000, 10, 12
    rng = np.random.default_rng(int(polys[0]))
    with threadpool_limits(1):
        for q in polys:
            A, b = wl.polytope(int(q))
            S.batched_sample(A, b, np.repeat(wl.starts[q][None], C, axis=0), C * (moves // thin),
                             0, thin, rng.normal(size=(moves, C, d)), rng.random((moves, C)),
                             margin=S.DEVICE_CHORD_MARGIN)
    return len(polys)


def _split_shard(problems):
    from threadpoolctl import threadpool_limits

    from oracle import split as O

    X, y, starts = _CPU["split"]
    so = O.Settings(ridge=1e-4, degree=1, bias_last=True)
    with threadpool_limits(1):
        for q in problems:
            O.annealed(X[q], y[q], starts[q], so, iters=200, t0=0.2, t1=0.01, step=0.05)
    return len(problems)


def _full_build(conn):
    from threadpoolctl import threadpool_limits

    with threadpool_limits(1):
        t0 = time.perf_counter()
        res = cpu_plain_fit_builder(max_depth=6).build()
        dt = time.perf_counter() - t0
    conn.send({"seconds": dt, "nodes": res.tree.num_nodes(), "leaves": res.tree.num_leaves()})
    conn.close()


def _balanced(sizes, parts):
    """Index groups of roughly equal total size (largest first, greedy)."""
    groups, loads = [[] for _ in range(parts)], np.zeros(parts)
    for i in np.argsort(-np.asarray(sizes), kind="stable"):
        k = int(np.argmin(loads))
        groups[k].append(int(i))
        loads[k] += sizes[i]
    return [sorted(g) for g in groups if g]


def run_cpu_baselines(args):
    """All CPU baselines (before any device work); returns (results, the
    background full-build process and its pipe end)."""
    import multiprocessing

    from threadpoolctl import threadpool_limits

    from art.synthetic import leaf_ridge_workload, polytope_workload, split_workload
    from oracle import sampling as S
    from oracle import split as O

    W = _workers()
    out = {"workers": W}
    legs = args.only
    bg = None
    if legs in ("all", "ridge"):
        wl = leaf_ridge_workload(n_leaves=args.leaves, seed=2)
        _CPU["ridge_workload"] = wl
        one, sample, cref, cref_status, stride = cpu_baseline(wl)
        _CPU["ridge_sample"], _CPU["ridge"] = sample, wl.ridge
        sizes = np.diff(sample.offsets)
        dt, _ = _run_parallel(_ridge_shard, _balanced(sizes, W))
        out["ridge"] = (one, sample, cref, cref_status, stride,
                        {**one, "value": sample.n_leaves / dt, "cores": W,
                         "sample": one["sample"] + f"; all cores: {W} forked workers, "
                                                   f"{dt:.2f} s"})
        print(f"[bench] cpu ridge: 1 core {one['value']:.0f}, {W} cores "
              f"{sample.n_leaves / dt:.0f} leaf fits/s", file=sys.stderr, flush=True)
    if legs in ("all", "sampler"):
        P = args.polytopes
        wl = polytope_workload(P, d=12, seed=4)
        _CPU["polytopes"] = wl
        k = 16
        n = 64 * 100
        with threadpool_limits(1):
            t0 = time.perf_counter()
            _sampler_shard(list(range(k)))
            dt1 = time.perf_counter() - t0
        many = list(range(k * W))
        dtW, _ = _run_parallel(_sampler_shard, [many[i::W] for i in range(W)])
        out["sampler"] = (
            {"value": k * n / dt1, "unit": "samples/s", "cores": 1, "kind": "port",
             "sample": (f"first {k} polytopes, NumPy batched transitions (64 chains, 1000 "
                        f"moves each), {dt1:.2f} s")},
            {"value": k * W * n / dtW, "unit": "samples/s", "cores": W, "kind": "port",
             "sample": f"first {k * W} polytopes on {W} forked workers, {dtW:.2f} s"})
        print(f"[bench] cpu sampler: 1 core {k * n / dt1:.3g}, {W} cores "
              f"{k * W * n / dtW:.3g} samples/s", file=sys.stderr, flush=True)
    if legs in ("all", "split"):
        K = max(2, W)
        X, y = split_workload(K, seed=3)
        rng = np.random.default_rng(3)
        starts = np.empty((K, 17))
        for q in range(K):
            starts[q] = O.initial_theta(X[q], rng.normal(size=16))
        _CPU["split"] = (X, y, starts)
        with threadpool_limits(1):
            t0 = time.perf_counter()
            _split_shard([0, 1])
            dt1 = time.perf_counter() - t0
        dtW, _ = _run_parallel(_split_shard, [[q] for q in range(K)])
        out["split"] = (
            {"value": 2 / dt1 / 64, "unit": "node splits/s", "cores": 1, "kind": "port",
             "sample": (f"2 (node, restart) problems, NumPy FP64 annealed loop, {dt1:.2f} s; "
                        "node splits/s = problems/s / 64")},
            {"value": K / dtW / 64, "unit": "node splits/s", "cores": W, "kind": "port",
             "sample": f"{K} problems on {W} forked workers, {dtW:.2f} s"})
        print(f"[bench] cpu split: 1 core {2 / dt1 / 64:.3g}, {W} cores {K / dtW / 64:.3g} "
              "node splits/s", file=sys.stderr, flush=True)
    if legs in ("all", "tree"):
        with threadpool_limits(1):
            t0 = time.perf_counter()
            cut = cpu_plain_fit_builder(max_depth=2).build()
            dt = time.perf_counter() - t0
        out["tree_cut"] = (dt, cut.tree.num_nodes())
        # The full host build runs in the background during the GPU legs
        # (started after the measurements above, so it competes with none).
        ctx = multiprocessing.get_context("fork")
        here, there = ctx.Pipe(duplex=False)
        bg = (ctx.Process(target=_full_build, args=(there,), daemon=True), here)
        bg[0].start()
    return out, bg


def collect_full_build(bg, timeout=120.0):
    """The background full host build's figures (or None), and no process left."""
    if bg is None:
        return None
    proc, conn = bg
    got = None
    if conn.poll(timeout):
        got = conn.recv()
    proc.join(5.0)
    if proc.is_alive():
        proc.terminate()
        proc.join(5.0)
    conn.close()
    return got


def cpu_baseline(wl, max_seconds=30.0):
    """FP64 CPU restatement (SciPy per leaf) on the host, one BLAS thread."""
    from threadpoolctl import threadpool_limits

    from oracle.ridge import leaf_fit_fast

    sys.path.insert(0, str(ROOT))
    # Bounded sample: every leaf with stride s, keeping the size distribution.
    stride = 1
    while True:
        sample = wl.subset(np.arange(0, wl.n_leaves, stride))
        if sample.n_rows * stride <= wl.n_rows * 1.5 or stride >= 64:
            break
        stride *= 2
    with threadpool_limits(1):
        t0 = time.perf_counter()
        coef, status, _ = leaf_fit_fast(sample.X, sample.y, sample.w, sample.offsets, wl.ridge)
        dt = time.perf_counter() - t0
    return {
        "value": sample.n_leaves / dt,
        "unit": "leaf fits/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{sample.n_leaves} leaves (every {stride}th), {sample.n_rows} rows; "
                   f"SciPy cho_factor/cho_solve per leaf, 1 BLAS thread, {dt:.2f} s"),
    }, sample, coef, status, stride


def timed_fits(batch, lib, X, y, w, ridge, precision, steps, warmup):
    """Warm up, then time `steps` fits with per-step events on the fit stream and
    the library's own events around its first kernel and the rest."""
    import ctypes

    import torch

    for _ in range(warmup):
        batch.fit(X, y, w, ridge, precision=precision)
    torch.cuda.synchronize()
    lib.art_plan_timing_enable(batch.plan, steps)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    ev[0].record(stream)
    for i in range(steps):
        batch.fit(X, y, w, ridge, precision=precision)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)])
    total_ms = ev[0].elapsed_time(ev[-1])
    first = (ctypes.c_float * steps)()
    rest = (ctypes.c_float * steps)()
    n = lib.art_plan_timing_read(batch.plan, first, rest, steps)
    lib.art_plan_timing_enable(batch.plan, 0)
    return {"ms_per_step": total_ms / steps, "step_ms": step_ms,
            "first_kernel_ms": float(np.mean(first[:n])), "rest_ms": float(np.mean(rest[:n]))}


def bench_ridge_kernels(args):
    """The leaf-fit kernels alone (for rocprofv3 traces and counter runs):
    the 4096-leaf workload resident in HBM, then --warmup + --steps fits of
    --precision back to back, nothing else on the device."""
    import torch

    from art import LeafRidgeBatch
    from art.synthetic import leaf_ridge_workload

    wl = _CPU.get("ridge_workload") or leaf_ridge_workload(n_leaves=args.leaves, seed=2)
    dev = torch.device("cuda", 0)
    X, y, w = (torch.from_numpy(a).to(dev) for a in (wl.X, wl.y, wl.w))
    with LeafRidgeBatch(wl.offsets, 8, degree=2, chunk_rows=args.chunk_rows) as batch:
        for _ in range(args.warmup):
            batch.fit(X, y, w, wl.ridge, precision=args.precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.fit(X, y, w, wl.ridge, precision=args.precision)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / max(args.steps, 1) * 1e3
        n_chunks = batch.n_chunks
    if args.copy_probe:
        copy_probe(torch)
    print(f"[bench] {args.steps} {args.precision} fits, {ms:.3f} ms each (host clock)",
          file=sys.stderr, flush=True)
    return {"leg": "ridge_kernels", "precision": args.precision, "steps": args.steps,
            "warmup": args.warmup, "ms_per_fit_host_clock": ms, "chunks": n_chunks,
            "rows": int(wl.n_rows), "leaves": int(wl.n_leaves)}


def graph_replays(batch, X, y, w, ridge, precision, steps, reference):
    """The fit captured as one HIP graph: `steps` timed replays (events on the
    batch stream after one untimed replay), then a bitwise check of the
    coefficients the replays left against the eager fit's."""
    import torch

    graph = batch.capture(X, y, w, ridge, precision=precision)
    graph.replay()
    batch.synchronize()
    stream = graph.stream  # events on the stream the graph launches on
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    equal = True
    ev[0].record(stream)
    for i in range(steps):
        graph.replay()
        ev[i + 1].record(stream)
    batch.synchronize()
    equal &= bool(np.array_equal(batch.coefficients.cpu().numpy(), reference, equal_nan=True))
    step_ms = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)])
    return {"ms_per_step": float(ev[0].elapsed_time(ev[-1]) / steps),
            "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                           "p99": float(np.percentile(step_ms, 99))},
            "equal_to_eager_bitwise": equal,
            "note": "the whole mixed fit (Gram, solve, refinement rounds, FP64 fallback) as one "
                    "graph launch"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/hbm_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, each
    scaled by the factor a plain copy of known size reads), or None."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get("kernels", {}).get(kernel)
    return None if entry is None else float(entry["bytes_per_launch"])


def peak_probe(torch):
    """Achievable peaks on this box beside the published ones: a 4 GiB
    device copy (read + write bytes), and bf16 and FP64 dense GEMMs through
    torch.matmul (hipBLASLt / rocBLAS) on contiguous square operands."""
    def timed(fn, reps):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps * 1e-3

    out = {}
    src = torch.ones((4 << 30) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    sec = timed(lambda: dst.copy_(src), 5)
    out["hbm_copy_GBs"] = 2 * src.numel() * 8 / sec / 1e9
    del src, dst
    for name, dtype, n, reps in (("bf16_gemm_TFLOPs", torch.bfloat16, 16384, 5),
                                 ("fp64_gemm_TFLOPs", torch.float64, 8192, 3)):
        a = torch.randn((n, n), dtype=dtype, device="cuda")
        b = torch.randn((n, n), dtype=dtype, device="cuda").contiguous()
        sec = timed(lambda: torch.matmul(a, b), reps)
        out[name] = 2.0 * n ** 3 / sec / 1e12
        del a, b
    torch.cuda.empty_cache()
    out["published"] = {"hbm_GBs": HBM_PEAK_GBS, "bf16_dense_TFLOPs": PEAK_TFLOPS.get("bf16"),
                        "fp64_TFLOPs": PEAK_TFLOPS["fp64"]}
    print(f"[bench] measured peaks: copy {out['hbm_copy_GBs']:.0f} GB/s, bf16 GEMM "
          f"{out['bf16_gemm_TFLOPs']:.0f} TF/s, FP64 GEMM {out['fp64_gemm_TFLOPs']:.1f} TF/s",
          file=sys.stderr, flush=True)
    return out


def copy_probe(torch, nbytes=1 << 30):
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: 3 copies of {nbytes} bytes", file=sys.stderr, flush=True)


def all_samples_violation(torch, out, wl, d, block=512):
    """Largest a.x - b over every sample of every polytope, on the device:
    the general cuts as padded (P, kmax, d) rows (padding a = 0, b = 1e300)
    through batched FP64 products, and the box rows as max |x_j| - 1."""
    P = out.shape[0]
    cuts = np.diff(wl.offsets) - 2 * d
    kmax = int(cuts.max())
    Ac = np.zeros((P, kmax, d))
    bc = np.full((P, kmax), 1e300)
    for q in range(P):
        s0 = int(wl.offsets[q]) + 2 * d
        Ac[q, :cuts[q]] = wl.A[s0:s0 + cuts[q]]
        bc[q, :cuts[q]] = wl.b[s0:s0 + cuts[q]]
    worst, worst_box = -np.inf, -np.inf
    for q0 in range(0, P, block):
        q1 = min(q0 + block, P)
        x = out[q0:q1]
        at = torch.from_numpy(np.ascontiguousarray(Ac[q0:q1].transpose(0, 2, 1))).to(x.device)
        bt = torch.from_numpy(bc[q0:q1]).to(x.device)
        viol = torch.bmm(x, at.contiguous()) - bt[:, None, :]
        worst = max(worst, float(viol.max()))
        worst_box = max(worst_box, float(x.abs().max()) - 1.0)
        del viol, at, bt
    return worst, worst_box


def bench_sampler(args):
    """Hit-and-run leg: P polytopes in d=12 (box plus 4-40 random cuts), 64
    chains x 1000 moves, thinning 10 (6,400 samples per polytope), Philox
    directions; one step = one launch over every polytope, samples written
    to HBM."""
    import torch

    from art.sampling import PolytopeBatch
    from art.synthetic import polytope_workload

    P, C, moves, thin, d = args.polytopes, 64, 1000, 10, 12
    wl = _CPU.get("polytopes") or polytope_workload(P, d=d, seed=4)
    batch = PolytopeBatch.from_csr(wl.A, wl.b, wl.offsets)
    n = C * (moves // thin)
    x0 = torch.from_numpy(np.repeat(wl.starts[:, None, :], C, axis=1)).to("cuda")
    out = torch.empty((P, n, d), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    batch.sample(x0, n, first=thin, stride=thin, seed=1, out=out)
    torch.cuda.synchronize()
    steps = max(1, args.sampler_steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for i in range(steps):
        batch.sample(x0, n, first=thin, stride=thin, seed=100 + i, out=out)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(steps)])
    ms = float(step_ms.mean())
    m = np.diff(wl.offsets).astype(np.float64)
    flops = C * moves * float(np.sum(2 * m * d + 3 * m + 2 * d))
    worst, worst_box = all_samples_violation(torch, out, wl, d)
    print(f"[bench] sampler {ms:.2f} ms/step, largest A x - b over all {P * n} samples "
          f"{worst:.3e} (box rows {worst_box:.3e})", file=sys.stderr, flush=True)
    result = {
        "metric": "hit-and-run samples/s", "value": P * n / (ms * 1e-3), "unit": "samples/s",
        "ms_per_step": ms, "steps": steps,
        "polytopes_per_s": P / (ms * 1e-3),
        "dtype": "f64 chord and slacks, f32 Philox directions",
        "config": {"workload": (f"{P} polytopes, d=12, box plus 4-40 random cuts, 64 chains x "
                                "1000 moves, thinning 10, seed 4"),
                   "polytopes": P, "mean_rows": float(m.mean()), "samples_per_polytope": n},
        "max_constraint_violation": max(worst, worst_box),
        "max_cut_violation": worst, "max_box_violation": worst_box,
        "violation_checked_polytopes": P, "violation_checked_samples": P * n,
        "violation_tolerance": 0.0,
        "samples_satisfy_every_constraint": bool(max(worst, worst_box) <= 0.0),
        "roofline": {"bound": "fp64", "achieved": flops / (ms * 1e-3) / 1e12,
                     "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                     "frac": flops / (ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"],
                     "algorithmic_flops_per_launch": flops,
                     "flops_per_chain_move": "2 m d + 3 m + 2 d",
                     "samples_write_GBs": P * n * d * 8 / (ms * 1e-3) / 1e9,
                     "kernel": "har_kernel<12, true>"},
    }
    if "sampler" in _CPU_RESULTS:
        # The reference's NumPy batched transition (oracle restatement) on a
        # bounded sample of the same polytopes, measured before device work.
        result["cpu_baseline"], result["cpu_baseline_all_cores"] = _CPU_RESULTS["sampler"]
    del out
    torch.cuda.empty_cache()
    return result


def bench_split(args):
    """Annealed split leg: nodes of n=8192 points in d=16 (8-bump target,
    generated on the device), 64 random restarts each, affine children
    (p=17, ridge 1e-4), T from 0.2 to 0.01 over 200 iterations, step 0.05;
    one launch runs every (node, restart) problem and its hard refit.  The
    value is the split-bf16 child-Gram mode (the configuration's precision);
    the FP64 mode runs on the same starts beside it, and both are compared."""
    import torch

    from art import AffineRidgeModel, BatchedOptimizerWorkspace, SoftObliqueSplitter
    from art.synthetic import split_workload_device

    K, R, iters = args.split_nodes, 64, 200
    dev = torch.device("cuda", 0)
    X, y, _ = split_workload_device(K, dev, seed=3)
    n, d = int(X.shape[1]), int(X.shape[2])
    s = SoftObliqueSplitter(AffineRidgeModel(1e-4), 0.2, n_restarts=R, random_state=3)
    ws = BatchedOptimizerWorkspace.allocate(s, K, n, d)
    ws.load_device(X, y)
    init = []
    starts = ws._device_starts(np.tile(np.arange(K), R), init).reshape(R, K, d + 1)
    runs = {}
    for prec in ("mixed", "mixed", "fp64"):  # the first mixed launch is the warmup
        t = []
        runs[prec] = (ws.anneal(starts, iters=iters, timing=t, child_precision=prec), t[-1])
    (res, anneal_ms), (ref, fp64_ms) = runs["mixed"], runs["fp64"]
    init_ms = init[0]
    ms = init_ms + anneal_ms
    p = ws.p
    # Algorithmic flops per point, problem and iteration: two weighted Grams
    # and right-hand sides (2 x 2 (p+1)^2 / 2 symmetric) plus residuals,
    # gradient and gates (2 x 2p + 2d + 20), all counted at FP64 weight.
    per_point = 2 * (p + 1) * (p + 1) + 4 * p + 2 * d + 20
    flops = float(K) * R * iters * n * per_point
    ok = res["ok"] & ref["ok"]
    dtheta = np.abs(res["theta"] - ref["theta"]).max(axis=-1)[ok]
    dloss = (np.abs(res["soft_loss_history"] - ref["soft_loss_history"])
             / np.abs(ref["soft_loss_history"])).max(axis=-1)[ok]
    print(f"[bench] split {ms:.1f} ms ({init_ms:.1f} starts + {anneal_ms:.1f} annealer, "
          f"FP64 children {fp64_ms:.1f}) for {K} nodes x {R} restarts", file=sys.stderr,
          flush=True)
    result = {
        "metric": "annealed node splits/s", "value": K / (ms * 1e-3), "unit": "node splits/s",
        "ms_per_step": ms, "starts_ms": init_ms, "annealer_ms": anneal_ms, "problems": K * R,
        "warmup": 1, "steps": 1,
        "dtype": ("bf16 3-way split child Grams (6 MFMA products, f32 acc into f64); f64 gates, "
                  "predictions, gradients, Cholesky and hard refit"),
        "fp64_children": {"annealer_ms": fp64_ms, "node_splits_per_s": K / ((init_ms + fp64_ms)
                                                                           * 1e-3),
                          "speedup_of_mixed": fp64_ms / anneal_ms},
        "err_vs_fp64_children": {
            "theta_max_abs": float(dtheta.max()), "theta_p99_abs": float(np.quantile(dtheta, .99)),
            "soft_loss_max_rel": float(dloss.max()),
            "hard_loss_max_rel": float((np.abs(res["loss"] - ref["loss"]) / np.abs(ref["loss"]))[ok].max()),
            "same_partition": float(np.mean(res["n_right"][ok] == ref["n_right"][ok])),
            "same_best_restart": float(np.mean(res["best"] == ref["best"]))},
        "error_codes": {str(k): int(v) for k, v in
                        zip(*np.unique(res["info"][..., 0], return_counts=True))},
        "config": {"workload": (f"{K} nodes x 64 restarts, n=8192, d=16 (rows on the device, "
                                "8-bump targets), affine p=17, ridge 1e-4, T 0.2->0.01 over 200 "
                                "iterations, step 0.05, seed 3"),
                   "nodes": K, "restarts": R, "iterations": iters},
        "failed_problems": int((~res["ok"]).sum()),
        "roofline": {"bound": "mfma", "achieved": flops / (anneal_ms * 1e-3) / 1e12,
                     "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                     "frac": flops / (anneal_ms * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"],
                     "algorithmic_flops_per_launch": flops,
                     "flops_per_point_iteration": per_point,
                     "kernel": "anneal_kernel (16 starts per workgroup, split-bf16 child Grams)",
                     "note": "algorithmic flops at FP64 weight against the FP64 MFMA peak"},
    }
    if not args.no_cpu:
        from oracle import split as O

        so = O.Settings(ridge=1e-4, degree=1, bias_last=True)
        th, _ = O.annealed(X[1].cpu().numpy(), y[1].cpu().numpy(), starts[0, 1], so, iters=iters,
                           t0=0.2, t1=0.01, step=0.05)
        result["err_vs_oracle"] = {"theta_max_abs_first_start": float(np.max(np.abs(
            O.project_unit_w(th) - res["theta"][0, 1])))}
    if "split" in _CPU_RESULTS:
        result["cpu_baseline"], result["cpu_baseline_all_cores"] = _CPU_RESULTS["split"]
    ws.close()
    del X, y
    torch.cuda.empty_cache()
    return result


def cpu_plain_fit_builder(max_depth):
    """The plain-fit builder on the host: the oracle's FP64 normal-equation
    ridge for node fits, its soft oblique optimizer (one Gaussian-direction,
    median-offset start per node) for splits and its scalar hit-and-run chain
    with NumPy draws for top-ups -- the reference's CPU path restated."""
    from dataclasses import dataclass, field

    from art.sampling import HitAndRunSampler, as_rng
    from art.splitters import (SoftObliqueSplitter, SplitNotFoundError, SplitResult,
                               project_unit_w)
    from art.synthetic import plain_fit_builder
    from oracle import ridge as R
    from oracle import sampling as S
    from oracle import split as O

    @dataclass
    class HostQuadRidge:
        ridge: float = 1e-6
        coef_: np.ndarray = field(default=None, repr=False)

        def clone(self):
            return HostQuadRidge(self.ridge)

        def effective_dimension(self, d):
            return 1 + d + d * (d + 1) // 2

        def fit(self, X, y):
            return self.fit_weighted(X, y, np.ones(len(y)))

        def fit_weighted(self, X, y, weights, weight_floor=1e-12):
            phi = R.features(np.asarray(X, dtype=float), 2, True, False)
            self.coef_, _ = R.solve_weighted_ridge(phi, y, weights, self.ridge, weight_floor)
            return self

        def predict(self, X):
            return R.features(np.atleast_2d(X), 2, True, False) @ self.coef_

    class HostSplitter(SoftObliqueSplitter):
        def split(self, X, y, parent_model=None, parent_loss=None, prepared_design=None,
                  hard_refit=True, **kw):
            rng = self._rng()
            w = rng.normal(size=X.shape[1])
            w /= np.linalg.norm(w)
            theta0 = np.append(w, np.median(X @ w))
            so = O.Settings(temperature=self.temperature, ridge=self.model_template.ridge,
                            min_side_points=self.min_side_points, degree=2, bias_last=False)
            res = O.optimize(X, y, theta0, so)
            theta = project_unit_w(res["theta"])
            right = X @ theta[:-1] - theta[-1] >= 0.0
            n_right = int(right.sum())
            if min(n_right, len(y) - n_right) < so.required_side_points(len(y)):
                raise SplitNotFoundError("min_side_points", "A side has too few points.")
            return SplitResult(theta[:-1], float(theta[-1]), None, None, None,
                               float(parent_loss), None, None, len(y) - n_right, n_right,
                               res["converged"], res["n_iters"], res["stop_reason"], {})

    class HostHitAndRun(HitAndRunSampler):
        def sample(self, region, n, random_state=None, x0=None):
            rng = as_rng(random_state)
            steps = self.burn_in + n * self.thinning
            dirs = rng.normal(size=(steps, region.dimension))
            shape = self._shaping(region.dimension)
            if shape is not None:
                dirs = dirs @ shape.T
            return S.scalar_sample(region.A, region.b, x0, n, self.burn_in, self.thinning, dirs,
                                   rng.random(steps))

    model = HostQuadRidge(1e-6)
    return plain_fit_builder(
        seed=0, max_depth=max_depth, model_template=model,
        splitter=HostSplitter(model, 0.05, min_side_points=64),
        sampler=HostHitAndRun(burn_in=0, thinning=20), profile_build_timing=True)


def bench_tree(args):
    """Plain-fit leg: whole tree builds of the 2-D plain ART fit (4000 samples
    per node, soft oblique T=0.05 on the batched device optimizer, quadratic
    leaves, depth <= 6, >= 64 points per side), then 256 hit-and-run samples
    in every leaf polytope (one launch); a step is one build plus the leaf
    sampling."""
    import torch

    from art.sampling import PolytopeBatch
    from art.synthetic import PlainFitOracle, plain_fit_builder

    steps = max(1, args.tree_steps)

    # Leaf sampling starts: the leaf's own sample mean lies inside its convex
    # region (the region is the intersection of half-spaces containing them).
    def leaf_samples(res):
        leaves = list(res.tree.iter_leaves())
        x0 = np.stack([leaf.metadata["sample_centroid"] for leaf in leaves])
        batch = PolytopeBatch([leaf.region for leaf in leaves])
        return batch.sample(torch.from_numpy(x0[:, None, :]).to("cuda"), 256, first=1,
                            stride=20, seed=7), len(leaves)

    def build(seed):
        res = plain_fit_builder(seed=seed, store_leaf_representative=True,
                                profile_build_timing=True).build()
        samples, n_leaves = leaf_samples(res)
        torch.cuda.synchronize()
        return res, samples, n_leaves

    for w in range(max(0, args.warmup)):
        build(1000 + w)
    times, results = [], []
    for k in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, samples, n_leaves = build(k)
        times.append(time.perf_counter() - t0)
        results.append((res, samples, n_leaves))
    ms = 1e3 * float(np.mean(times))
    res, samples, n_leaves = results[0]
    tree = res.tree
    Xt = np.random.default_rng(11).uniform(-1, 1, (20000, 2))
    pred = tree.predict(Xt)
    host = tree._predict_models(Xt)
    rmse = float(np.sqrt(np.mean((pred - PlainFitOracle.clean(Xt)) ** 2)))
    leaves = list(tree.iter_leaves())
    inside = all(leaf.region.contains(samples[i].cpu().numpy()).all()
                 for i, leaf in enumerate(leaves))
    print(f"[bench] tree {ms:.1f} ms/build, {tree.num_leaves()} leaves, rmse {rmse:.2e}",
          file=sys.stderr, flush=True)
    result = {
        "metric": "tree builds/s", "value": 1e3 / ms, "unit": "tree builds/s",
        "ms_per_step": ms, "steps": steps,
        "latency_ms": {"p50": float(np.percentile(times, 50)) * 1e3,
                       "p99": float(np.percentile(times, 99)) * 1e3},
        "dtype": "f64",
        "config": {"workload": ("plain ART fit: d=2 on [-1,1]^2, 4000 uniform samples per node "
                                "(seed k for build k), f = sin(3x1)cos(2x2) + 0.5 exp(-8|x-(0.3,"
                                "-0.2)|^2) + N(0, 1e-6) noise, soft oblique T=0.05 (batched "
                                "device optimizer, BFS cohorts of 32), quadratic ridge 1e-6, "
                                "depth <= 6, >= 64 points per side, relative L2 stop 1e-2, "
                                "256 hit-and-run samples per leaf polytope"),
                   "nodes": tree.num_nodes(), "leaves": n_leaves,
                   "oracle_queries": res.oracle_queries},
        "build_timing": tree.metadata.get("build_timing"),
        "test_rmse_vs_clean_target": rmse,
        "gpu_vs_host_prediction_max_abs": float(np.max(np.abs(pred - host))),
        "leaf_samples_inside_regions": bool(inside),
    }
    if "tree_cut" in _CPU_RESULTS:
        # A bounded host sample measured before device work: the same build
        # cut at depth 2, one thread, scaled per node to the GPU tree's node
        # count (the full host build, run in the background, replaces it in
        # main when it finishes).
        dt, nodes = _CPU_RESULTS["tree_cut"]
        per_node = dt / nodes
        result["cpu_baseline"] = {
            "value": 1.0 / (per_node * tree.num_nodes()), "unit": "tree builds/s", "cores": 1,
            "kind": "port",
            "sample": (f"the same build cut at depth 2 ({nodes} nodes) on the host: FP64 normal-"
                       f"equation fits, the oracle's soft oblique optimizer, a scalar NumPy "
                       f"hit-and-run chain; {dt:.1f} s, {per_node * 1e3:.0f} ms per node, scaled "
                       f"to the GPU tree's {tree.num_nodes()} nodes")}
    return result


def bench_serving(args):
    """Serving leg: the first ``--requests`` requests of the seeded stream
    (Poisson arrivals at ``--rate``/s; d in {2,4,8,12,16,20}; N log-uniform in
    [1e3, 2e6]; 4-16 Gaussian bumps; 30% refits adding 1-10% new rows to an
    earlier tree; quadratic leaves for d <= 12, linear above; depth <= 10,
    >= 64 rows to split a node's half; 200 GB HBM budget).  Rows are generated
    on the device before timing.  Leg 1 serves the stream as it arrives
    (continuous batching) with mixed-precision statistics of open nodes, and
    again in FP64 for comparison (the served trees must agree); leg 2 serves
    the same requests one at a time, and its latencies are those of a FIFO
    single server fed the same arrivals."""
    import torch

    from art.serving import FitRequest, RefitRequest, ServingConfig, ServingEngine
    from art.synthetic import serving_rows, serving_stream

    events = serving_stream(args.requests, rate=args.rate)
    dev = torch.device("cuda", 0)
    data = [serving_rows(e, dev) for e in events]
    torch.cuda.synchronize()
    rows = sum(e.rows for e in events)

    def request(i, tag):
        e = events[i]
        X, y = data[i]
        if e.kind == "fit":
            return FitRequest(X, y, max_depth=10, min_leaf=64, seed=e.seed,
                              tree_id=f"{tag}{e.tree}")
        return RefitRequest(f"{tag}{e.tree}", X, y)

    def warm(eng):
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        for d in (2, 4, 8, 12, 16, 20):
            X = torch.rand((3000, d), generator=g, dtype=torch.float64, device=dev)
            eng.fit(FitRequest(X, X[:, 0].contiguous(), max_depth=2, tree_id=f"warm{d}"))
            eng.evict(f"warm{d}")

    def serve(precision, tag):
        """The stream as it arrives (continuous batching) at one precision."""
        cfg = ServingConfig(hbm_budget_bytes=200e9, precision=precision)
        with ServingEngine(cfg) as eng:
            warm(eng)
            t0 = time.perf_counter()
            futs = []
            for i, e in enumerate(events):
                delay = t0 + e.arrival - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                futs.append(eng.submit(request(i, tag)))
            res = []
            for i, f in enumerate(futs):
                res.append(f.result())
                if (i + 1) % 100 == 0:
                    print(f"[bench] serving ({precision}): {i + 1} results at "
                          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            t1 = time.perf_counter()
            makespan = t1 - t0
            # Engine time with no request to work on, inside [t0, t1]: the
            # stream's work took makespan - idle seconds of the engine.
            idle = sum(max(0.0, min(b, t1) - max(a, t0)) for a, b in list(eng.idle_periods))
            out = {"res": res, "makespan": makespan, "steps": eng.steps, "idle": idle,
                   "phases": {k: round(v, 4) for k, v in eng.phase_seconds.items()}}
        lat = np.array([r.timing.latency for r in res])
        out["lat"] = lat
        print(f"[bench] serving ({precision}) {len(events)} requests in {makespan:.2f} s, p50 "
              f"{np.percentile(lat, 50) * 1e3:.0f} ms p99 {np.percentile(lat, 99) * 1e3:.0f} ms",
              file=sys.stderr, flush=True)
        return out

    if args.only == "serving_stream":
        # Trace driver: the stream once, continuously batched, at --precision.
        out = serve(args.precision, "s")
        lat = out["lat"]
        return {"leg": "serving_stream", "precision": args.precision, "requests": len(events),
                "makespan_s": out["makespan"], "value": len(events) / out["makespan"],
                "unit": "requests/s", "engine_idle_s": out["idle"],
                "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3,
                               "p99": float(np.percentile(lat, 99)) * 1e3},
                "engine_levels": out["steps"], "engine_phase_s": out["phases"],
                "while_arriving": while_arriving(events, lat)}
    mixed = serve("mixed", "b")
    fp64 = serve("fp64", "f")
    res, makespan, lat = mixed["res"], mixed["makespan"], mixed["lat"]
    steps, phases = mixed["steps"], mixed["phases"]
    same = all(np.array_equal(a.coefficients, b.coefficients) and
               np.array_equal(a.status, b.status) for a, b in zip(res, fp64["res"]))
    differ = [i for i, (a, b) in enumerate(zip(res, fp64["res"]))
              if a.coefficients.shape != b.coefficients.shape]
    coef_diff = max((float(np.max(np.abs(a.coefficients - b.coefficients), initial=0.0))
                     for a, b in zip(res, fp64["res"])
                     if a.coefficients.shape == b.coefficients.shape), default=0.0)
    leaves = [len(r.coefficients) for r, e in zip(res, events) if e.kind == "fit"]
    bad = sum(int((r.status != 0).sum()) for r in res)
    one = ServingConfig(hbm_budget_bytes=200e9, max_active=1, precision="mixed")
    service = []
    with ServingEngine(one) as eng:
        warm(eng)
        for i in range(len(events)):
            t = time.perf_counter()
            eng.submit(request(i, "s")).result()
            service.append(time.perf_counter() - t)
            if (i + 1) % 100 == 0:
                print(f"[bench] one at a time: {i + 1} done, {sum(service):.1f} s",
                      file=sys.stderr, flush=True)
    finish, lat1 = 0.0, []
    for e, sv in zip(events, service):
        finish = max(finish, e.arrival) + sv
        lat1.append(finish - e.arrival)
    lat1 = np.array(lat1)
    print(f"[bench] one at a time: {sum(service):.2f} s of service, p50 "
          f"{np.percentile(lat1, 50) * 1e3:.0f} ms p99 {np.percentile(lat1, 99) * 1e3:.0f} ms",
          file=sys.stderr, flush=True)
    del data
    torch.cuda.empty_cache()
    return {
        "metric": "served requests/s and p50/p99 request latency",
        "value": len(events) / makespan, "unit": "requests/s",
        "arrival_rate": args.rate,
        "keeps_up_with_arrivals": bool(len(events) / (makespan - mixed["idle"]) >= args.rate),
        "service_capacity": {
            "value": len(events) / (makespan - mixed["idle"]), "unit": "requests/s",
            "engine_idle_s": mixed["idle"],
            "definition": "requests / (makespan - engine idle time in the window); the engine "
                          "keeps up when this reaches the arrival rate"},
        "drain_after_last_arrival_s": makespan - events[-1].arrival,
        "while_arriving": while_arriving(events, lat),
        "latency_ms_by_quarter": [float(np.mean(q)) * 1e3 for q in np.array_split(lat, 4)],
        "latency_ms": {"p50": float(np.percentile(lat, 50)) * 1e3,
                       "p99": float(np.percentile(lat, 99)) * 1e3,
                       "max": float(lat.max()) * 1e3},
        "makespan_s": makespan, "engine_levels": steps, "engine_phase_s": phases,
        "requests": len(events), "fits": len(leaves),
        "refits": sum(e.kind == "refit" for e in events), "rows": rows,
        "mean_leaves_per_tree": float(np.mean(leaves)) if leaves else 0.0,
        "failed_leaf_solves": bad,
        "dtype": ("bf16 3-way split (6 MFMA products)/f32-acc statistics of open nodes "
                  "(p <= 64, d <= 16); f64 leaf systems, leaf fits, split search, refits"),
        "fp64": {"value": len(events) / fp64["makespan"], "unit": "requests/s",
                 "latency_ms": {"p50": float(np.percentile(fp64["lat"], 50)) * 1e3,
                                "p99": float(np.percentile(fp64["lat"], 99)) * 1e3},
                 "engine_phase_s": fp64["phases"]},
        "mixed_equals_fp64": {"same_leaf_models_and_status": bool(same),
                              "requests_with_other_leaf_counts": len(differ),
                              "max_abs_coefficient_difference": coef_diff},
        "one_at_a_time": {"value": len(events) / sum(service), "unit": "requests/s",
                          "service_ms": {"p50": float(np.percentile(service, 50)) * 1e3,
                                         "p99": float(np.percentile(service, 99)) * 1e3},
                          "latency_ms": {"p50": float(np.percentile(lat1, 50)) * 1e3,
                                         "p99": float(np.percentile(lat1, 99)) * 1e3}},
        "speedup_vs_one_at_a_time": sum(service) / makespan,
        "config": {"workload": (f"serving stream, first {len(events)} requests of seed 5: "
                                f"Poisson {args.rate:g}/s, d in {{2,4,8,12,16,20}}, N log-uniform "
                                "[1e3, 2e6], 4-16 bumps, 30% refits of 1-10% new rows, quadratic "
                                "leaves d <= 12 else linear, depth <= 10, min_leaf 64, split "
                                "sample 4096 rows, 30 iterations, HBM budget 200 GB")},
    }


def while_arriving(events, lat, start=1.0):
    """Completions against arrivals over [start, last arrival] s of the stream:
    1000 / makespan includes the last requests' own latency after the last
    arrival, so it stays below the arrival rate even for an engine that keeps
    pace; the completion rate inside the arrival window and the backlog at
    the last arrival do not."""
    arrive = np.array([e.arrival for e in events])
    done = arrive + lat
    t1 = float(arrive[-1])
    span = t1 - start
    completions = int(((done >= start) & (done <= t1)).sum())
    arrivals = int(((arrive >= start) & (arrive <= t1)).sum())
    return {"window_s": [start, t1], "completions_per_s": completions / span,
            "arrivals_per_s": arrivals / span,
            "completions_per_arrival": completions / max(arrivals, 1),
            "backlog_at_last_arrival": int(len(events) - (done <= t1).sum()),
            "definition": "completions and arrivals inside the window (seconds from the first "
                          "arrival); backlog = requests not finished at the last arrival"}


TUNE_CANDIDATES = (2048, 4096, 8192, 16384)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    sys.path.insert(0, str(ROOT))
    background = None
    if not args.no_cpu and args.only not in ("ridge_kernels", "serving_stream"):
        results, background = run_cpu_baselines(args)
        _CPU_RESULTS.update(results)
    try:
        run_legs(args)
    finally:
        full = collect_full_build(background)
        if full is not None:
            print(f"[bench] cpu full plain-fit build: {full['seconds']:.1f} s, "
                  f"{full['nodes']} nodes", file=sys.stderr, flush=True)
    emit(args, full)


_CPU_RESULTS = {}
_LINE = {}


def emit(args, full):
    """Print the one JSON line (the full host build attached to the tree leg)."""
    if not _LINE:
        return
    line = _LINE["line"]
    tree = line.get("tree_builds") if isinstance(line, dict) else None
    if tree is not None and full is not None:
        if "cpu_baseline" in tree:
            tree["cpu_baseline_depth2_extrapolated"] = tree["cpu_baseline"]
        tree["cpu_baseline"] = {
            "value": 1.0 / full["seconds"], "unit": "tree builds/s", "cores": 1, "kind": "port",
            "sample": (f"the full plain-fit build on the host ({full['nodes']} nodes, "
                       f"{full['leaves']} leaves): FP64 normal-equation fits, the oracle's soft "
                       f"oblique optimizer, a scalar NumPy hit-and-run chain; "
                       f"{full['seconds']:.1f} s on one thread, run during the GPU legs")}
    if isinstance(line, dict):
        import multiprocessing

        # Every forked worker and the background build have been joined.
        line["child_processes_at_end"] = len(multiprocessing.active_children())
    print(json.dumps(line))


def run_legs(args):
    import torch

    from art import LeafRidgeBatch, leaf_ridge
    from art.synthetic import leaf_ridge_workload

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    sys.path.insert(0, str(ROOT))
    if args.max_split_parts > 0:
        from art import batched_optimizer
        batched_optimizer.MAX_PARTS = min(batched_optimizer.MAX_PARTS, args.max_split_parts)
    if args.only == "ridge_kernels":
        _LINE["line"] = bench_ridge_kernels(args)
        return
    if args.only == "sampler":
        _LINE["line"] = bench_sampler(args)
        return
    if args.only == "split":
        _LINE["line"] = bench_split(args)
        return
    if args.only == "tree":
        _LINE["line"] = bench_tree(args)
        return
    if args.only in ("serving", "serving_stream"):
        _LINE["line"] = bench_serving(args)
        return
    wl = _CPU.get("ridge_workload") or leaf_ridge_workload(n_leaves=args.leaves, seed=2)
    dev = torch.device("cuda", 0)
    X = torch.from_numpy(wl.X).to(dev)
    y = torch.from_numpy(wl.y).to(dev)
    w = torch.from_numpy(wl.w).to(dev)
    from art import autotune

    # Chunk sizes: each form at the size the autotuner measures fastest on this
    # workload (before any timing), unless --chunk-rows pins one.
    precisions = ["fp64"] + ([args.precision] if args.precision != "fp64" else [])
    chunk, tuned = {}, {}
    for prec in precisions:
        if args.chunk_rows:
            chunk[prec], tuned[prec] = args.chunk_rows, None
        else:
            chunk[prec], tuned[prec] = autotune.tune_chunk_rows(
                wl.offsets, 8, X, y, w, wl.ridge, precision=prec, force=True,
                candidates=TUNE_CANDIDATES, reps=5)
        print(f"[bench] {prec} chunk rows {chunk[prec]} ({tuned[prec]})", file=sys.stderr,
              flush=True)
    plain_batch = LeafRidgeBatch(wl.offsets, 8, degree=2, chunk_rows=chunk["fp64"])
    batch = plain_batch if args.precision == "fp64" else \
        LeafRidgeBatch(wl.offsets, 8, degree=2, chunk_rows=chunk[args.precision])
    p, N, K = batch.p, wl.n_rows, wl.n_leaves
    lib = batch.lib
    print(f"[bench] leaves={K} rows={N} p={p} chunks={batch.n_chunks}", file=sys.stderr, flush=True)

    # FP64 plain form: the accuracy reference and the comparison point.
    plain = timed_fits(plain_batch, lib, X, y, w, wl.ridge, "fp64", args.steps, args.warmup)
    ref = plain_batch.coefficients.cpu().numpy()
    ref_status = plain_batch.status.cpu().numpy()
    print(f"[bench] fp64 plain {plain['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    run = plain
    if args.precision == "mixed":
        run = timed_fits(batch, lib, X, y, w, wl.ridge, "mixed", args.steps, args.warmup)
        print(f"[bench] mixed {run['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    coef = batch.coefficients.cpu().numpy()
    st = batch.status.cpu().numpy()
    graphed = graph_replays(batch, X, y, w, wl.ridge, args.precision, args.steps, coef)
    print(f"[bench] graph replay {graphed['ms_per_step']:.3f} ms/step, equal to eager: "
          f"{graphed['equal_to_eager_bitwise']}", file=sys.stderr, flush=True)
    ok = (st == 0) & (ref_status == 0)
    err = rel_err(coef[ok], ref[ok])
    fallback = batch.fallback_count() if args.precision == "mixed" else 0

    ms_per_step = run["ms_per_step"]
    kernel_ms = run["first_kernel_ms"]
    if args.precision == "mixed":
        # Dominant kernel: the split-bf16 Gram pass, bound by HBM.  Algorithmic
        # bytes: X (8 doubles), y and w per row, read once.
        traffic_bytes = N * (8 + 2) * 8
        achieved = traffic_bytes / (kernel_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic("gram_leaf_mixed_kernel<3, 8>"),
                    "kernel": "gram_leaf_mixed_kernel<3, 8>", "algorithmic_bytes_per_launch": traffic_bytes,
                    "kernel_ms": kernel_ms}
    else:
        # FP64 Gram on MFMA: symmetric entries p(p+1)/2 plus p right-hand-side
        # entries, 2 flops each per row.
        flops = (p * (p + 1) + 2.0 * p) * N
        achieved = flops / (kernel_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["fp64"],
                    "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS["fp64"],
                    "traffic": measured_traffic("gram_leaf_f64_kernel<3>"),
                    "kernel": "gram_leaf_f64_kernel<3>", "algorithmic_flops_per_launch": flops,
                    "kernel_ms": kernel_ms}
    result = {
        "metric": METRIC,
        "value": K / (ms_per_step * 1e-3),
        "unit": "leaf fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "bf16 3-way split (6 MFMA products)/f32-acc/f64-refine",
        "data": "synthetic (seeded, generated on host, resident in HBM before timing)",
        "config": {
            "workload": ("batched weighted ridge: 4096 leaves, d=8, quadratic p=45, "
                         "log-uniform 64..16384 rows, ridge 1e-6, seed 2"),
            "leaves": K, "rows": N, "p": p, "precision": args.precision,
            "chunk_rows": chunk, "chunks": batch.n_chunks,
        },
        "latency_ms": {"p50": float(np.percentile(run["step_ms"], 50)),
                       "p99": float(np.percentile(run["step_ms"], 99)),
                       "unit": "one 4096-leaf batch fit"},
        "rel_err_vs_fp64_plain": {"max": float(err.max()) if err.size else 0.0,
                                  "p99": float(np.percentile(err, 99)) if err.size else 0.0},
        "fp64_plain": {"ms_per_step": plain["ms_per_step"],
                       "leaf_fits_per_s": K / (plain["ms_per_step"] * 1e-3),
                       "gram_ms": plain["first_kernel_ms"], "solve_ms": plain["rest_ms"]},
        "speedup_vs_fp64_plain": plain["ms_per_step"] / ms_per_step,
        "fallback_leaves": fallback,
        "kernels_ms": {"first": kernel_ms, "rest": run["rest_ms"]},
        "failed_leaves": int((st != 0).sum()),
        "graph_replay": graphed,
        "autotune": {"chunk_rows_ms": {prec: (None if ms is None else
                                              {str(k): v for k, v in ms.items()})
                                       for prec, ms in tuned.items()},
                     "chosen_chunk_rows": chunk,
                     "note": "each form runs at its own fastest chunk size (median of 5 fits "
                             "per candidate), chosen before timing",
                     "key": "leaf_fit: d, p, basis, precision, K and N power-of-two buckets"},
        "roofline": roofline,
    }
    if "ridge" in _CPU_RESULTS:
        cpu, sample, cref, cref_status, stride, all_cores = _CPU_RESULTS["ridge"]
        result["cpu_baseline"] = cpu
        result["cpu_baseline_all_cores"] = all_cores
        got = coef[::stride][: sample.n_leaves]
        both = (cref_status == 0) & (st[::stride][: sample.n_leaves] == 0)
        cerr = rel_err(got[both], cref[both])
        result["rel_err_vs_cpu_fp64"] = {"max": float(cerr.max()),
                                         "p99": float(np.percentile(cerr, 99))}
    if args.copy_probe:
        copy_probe(torch)
    batch.close()
    plain_batch.close()
    del X, y, w
    torch.cuda.empty_cache()
    result["measured_peaks"] = peak_probe(torch)
    if args.only == "all":
        result["split_optimizer"] = bench_split(args)
        result["hit_and_run"] = bench_sampler(args)
        result["tree_builds"] = bench_tree(args)
        result["serving"] = bench_serving(args)
    _LINE["line"] = result


if __name__ == "__main__":
    main()
