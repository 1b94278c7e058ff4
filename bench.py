#!/usr/bin/env python3
"""Benchmark of the full-batch t-SNE gradient pass on one MI355X.

A step is one iteration of the t-SNE schedule (attraction over the CSR P, FFT-interpolated repulsion,
gains / momentum update) over the whole synthetic data set. Default workload: the MNIST-shaped
precision run -- 70,000 x 784 float32 (10 prototypes with ~19% nonzero pixels, + N(0, 0.1^2) noise,
clipped, quantised to /255), perplexity-30 affinities (90-NN), 2-D, 750 iterations (250 early
exaggeration). Inputs (P, Y0) are resident in HBM before the timed region. value = iterations/s of the FP32
path (FP32 bulk, FP64 Z and reductions). With --steps K below the schedule length the timed iterations are K
iterations in up to five runs of consecutive iterations, one centred in each fifth of the schedule, each
started from the schedule's own state there (the schedule is run once, untimed, to make those states); with
K at least the schedule length the whole schedule is timed from iteration 0. The whole FP32 and FP64
schedules, FP32 against FP64 from the same states, and the FP32 gradient error against FP64 are reported
beside it, and the 1M configuration (gmm1m) the same way.

The per-stage roofline comes from a second, profiled pass of the same timed runs (HIP events on the
library's stream, algorithmic bytes from the engine), per padded FFT size for the FFT kernels. cpu_baseline
times the reference's own compiled module (oracle/_ref) on this host's cores on a bounded sample of
iterations at the same states.

Two more legs run with the default workload and are reported beside the headline:
- `serving`: a 200,000 x 50 MAGNET(loss="tsne") model kept resident, transform() requests of 1-4096 rows
  arriving by a Poisson process at 200 QPS through the continuous-batching server; p50 / p99 latency
  (`transform_p99_ms`) and the saturated rate;
- `batched`: 64 independent embeddings (2,000-50,000 points each, 500 iterations) through the batched engine,
  against the same instances one after another, with a bit-for-bit comparison of the results.
--workload gmm1m / gmm10k run the 1M and 10k configurations; --attraction-probe times the CSR attraction of a
workload alone (for PMC counters); --copy-probe times a plain copy (the PMC scale reference).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lib-magnet_amd"))

METRIC = "t-SNE iters/s at 70k & 1M pts (FP32 vs plain FP64 GPU); transform p99 ms"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "mnist70k": dict(n=70_000, d=784, n_iter=750, early=250, desc="MNIST-shaped 70,000 x 784, perplexity 30, 750 iterations"),
    "gmm1m": dict(n=1_000_000, d=50, n_iter=1000, early=250, desc="1,000,000 x 50 Gaussian mixture (100 clusters, Dirichlet sizes), perplexity 30, 1,000 iterations"),
    "gmm10k": dict(n=10_000, d=50, n_iter=500, early=166, desc="10,000 x 50 Gaussian mixture (10 clusters), perplexity 30, 500 iterations"),
}
BATCHED = dict(count=64, n_min=2000, n_max=50000, d=32, n_iter=500, early=166,
               desc="64 independent embeddings, n_i log-uniform in [2000, 50000] (seed 1000 + i), 32-d Gaussian mixtures "
                    "of 5-20 clusters, perplexity 30, 500 iterations (166 early), FP32")
SERVING = dict(n_train=200_000, d=50, clusters=50, pool=60_000, qps=200.0, duration=16.0, n_iter=100,
               desc="MAGNET(loss='tsne') on 200,000 x 50 (Gaussian mixture, 50 clusters), resident; transform() requests "
                    "of 1-4096 new rows (log-uniform sizes, same mixture) arriving by a Poisson process; 100 iterations each")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(name, seed=0):
    rng = np.random.default_rng(seed)
    if name == "mnist70k":
        n, d = 70_000, 784
        protos = (rng.random((10, d)) < 0.19) * rng.uniform(0.3, 1.0, (10, d))
        lab = rng.integers(0, 10, n)
        X = protos[lab] + rng.normal(0, 0.1, (n, d))
        X = np.round(np.clip(X, 0, 1) * 255) / 255
        return X.astype(np.float32), lab
    if name == "gmm1m":
        n, d, k = 1_000_000, 50, 100
        sizes = rng.multinomial(n, rng.dirichlet(np.ones(k)))
        lab = np.repeat(np.arange(k), sizes)
        X = rng.normal(0, 4, (k, d))[lab] + rng.normal(0, 1, (n, d))
        return X.astype(np.float32), lab
    n, d, k = 10_000, 50, 10
    lab = np.repeat(np.arange(k), n // k)
    X = rng.normal(0, 5, (k, d))[lab] + rng.normal(0, 1, (n, d))
    return X.astype(np.float32), lab


def schedule_segments(steps, n_iter, max_runs=5):
    """The timed iterations: the whole schedule when steps >= n_iter; else `steps` iterations in up to five runs of
    consecutive iterations (at least four each), one centred in each of equal strata of the schedule, so that the
    early phase, the switch, the clipped-grid stretch and the late phase are all timed."""
    if steps >= n_iter:
        return [(0, n_iter)]
    runs = max(1, min(max_runs, steps // 4))
    lens = [steps // runs + (1 if r < steps % runs else 0) for r in range(runs)]
    out = []
    for r, L in enumerate(lens):
        a = int(round((r + 0.5) * n_iter / runs - L / 2))
        a = max(0, min(n_iter - L, a))
        out.append((a, a + L))
    return out


class Loop:
    """One precision of the whole loop on resident inputs: the CSR on the device and a reusable mg_tsne_session."""

    def __init__(self, P, prec, n_iter, early):
        import torch
        from magnet import capi
        self.dt = np.float32 if prec == "fp32" else np.float64
        self.tdt = torch.float32 if prec == "fp32" else torch.float64
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self.csr = (dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32)), dev(P.data.astype(self.dt)))
        self.sess = capi.Session(*self.csr, n_iter=n_iter, early_iter=early)
        self.n_iter = n_iter

    def state(self, Y0=None, like=None):
        import torch
        if like is not None:
            return tuple(t.to(self.tdt).clone() for t in like)
        Y = torch.from_numpy(np.ascontiguousarray(Y0, dtype=self.dt)).cuda()
        return Y, torch.zeros_like(Y), torch.ones_like(Y)

    def run(self, st, a, b):
        return self.sess.run(*st, a, b)

    def timed(self, st, a, b):
        """seconds for iterations [a, b) on state st, bracketed by device synchronisations"""
        import torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.sess.run(*st, a, b)
        torch.cuda.synchronize()
        return time.perf_counter() - t0


def checkpoint_run(loop, Y0, starts):
    """The whole schedule from Y0 in runs split at `starts`, keeping a copy of the state at each; returns (snapshots,
    seconds of the iterations alone, per-iteration sumQ)."""
    st = loop.state(Y0)
    snaps, sec, sq, it = {}, 0.0, [], 0
    for c in sorted(set(starts)) + [loop.n_iter]:
        if c > it:
            sec += loop.timed(st, it, c)
            it = c
        if c < loop.n_iter:
            snaps[c] = tuple(t.clone() for t in st)
    return snaps, sec


def timed_segments(loop, snaps, segs, profile=False):
    """The timed region: each run of consecutive iterations from the schedule's own state at its start; the
    per-run times (device-synchronised on both sides) are summed."""
    from magnet import capi
    total = 0.0
    capi.profile(profile)
    for a, b in segs:
        st = tuple(t.clone() for t in snaps[a])
        total += loop.timed(st, a, b)
    stages = capi.profile_read() if profile else None
    if profile:
        stages["_grid"] = capi.profile_grid()
        stages["_sized"] = capi.profile_sized()
    capi.profile(False)
    return total, stages


def matched_precision(l32, l64, snaps, k):
    """FP32 and FP64 from the same states: k iterations of each from every snapshot (FP32 states, widened to
    float64 for the FP64 run); returns per-state (t, ms32, ms64)."""
    rows = []
    for t, st in sorted(snaps.items()):
        t1 = min(t + k, l32.n_iter)
        if t1 <= t:
            continue
        ms = []
        for loop in (l32, l64):
            s = loop.state(like=st)
            ms.append(1e3 * loop.timed(s, t, t1) / (t1 - t))
        rows.append((t, ms[0], ms[1]))
    return rows


def trace_same_states(l32, l64, snaps):
    """KL divergence and gradient norm of FP32 and FP64 (the loop's FP64 compensated diagnostics) at each snapshot,
    one traced iteration each from the same state: [(t, KL32, KL64, |g|32, |g|64)]."""
    rows = []
    for t, st in sorted(snaps.items()):
        if t >= l32.n_iter:
            continue
        vals = []
        for loop in (l32, l64):
            s = loop.state(like=st)
            _, kl, gn = loop.sess.run(*s, t, t + 1, trace=True)
            vals.append((float(kl[0]), float(gn[0])))
        rows.append((t, vals[0][0], vals[1][0], vals[0][1], vals[1][1]))
    return rows


def gradient_error(P, l32, l64, st, it, early):
    """Relative L2 of the FP32 path's gradient rho A - F / sumQ against the FP64 path's on the same layout (the FP32
    state widened), each through the per-call attraction and FFTConv2D objects of its precision."""
    import torch
    from magnet import capi
    from magnet.tsne import grid_params
    Y32 = st[0].to(torch.float32).contiguous()
    rho = 12.0 if it < early else 1.0
    grads = []
    for loop in (l32, l64):
        Y = Y32.to(loop.tdt).contiguous()
        A = capi.attraction(*loop.csr, Y)
        xmin, ymin, width, nb = grid_params(capi.bbox(Y32))
        conv = capi.FFTConv(capi.MG_F32 if loop.dt == np.float32 else capi.MG_F64, nb * 6, 3)
        conv.set_tsne(width)
        s, F = conv.compute(Y, xmin, ymin, width, nb)
        grads.append(rho * A.double() - F.double() / s)
    return float(torch.linalg.norm(grads[0] - grads[1]) / torch.linalg.norm(grads[1]))


from magnet.tune import SPACE as TUNE_SPACE  # noqa: E402


def precision_run(P, n_iter, early, steps, warmup, profile=True, cpu=False, threads=1):
    """The headline measurement of one workload: FP32 loop timed over the segments of the schedule, its profiled
    pass, the whole FP32 and FP64 schedules, FP32 against FP64 from the same states and the FP32 gradient error."""
    n = P.shape[0]
    Y0 = np.random.default_rng(1).normal(0, 1e-4, (n, 2))
    segs = schedule_segments(steps, n_iter)
    starts = [a for a, _ in segs]
    match_starts = sorted(set(starts + [int(round(f * (n_iter - 10))) for f in (0.1, 0.3, 0.5, 0.7, 0.9)]))
    l32 = Loop(P, "fp32", n_iter, early)
    if warmup > 0:
        l32.run(l32.state(Y0), 0, min(warmup, n_iter))
    snaps, full32 = checkpoint_run(l32, Y0, starts + match_starts)
    # launch shapes that change no result, picked by timing 10 iterations from the state at 30% of the schedule
    # (where the grid is largest while it clips) and 10 from the late phase, together (magnet.tune)
    from magnet.tune import autotune_loop
    t_tune = [int(round(0.3 * (n_iter - 10))), match_starts[-1]]
    chosen, tune_log = autotune_loop(l32.sess, [(snaps[t], t) for t in t_tune], None, iters=10, repeats=3)
    sec32, _ = timed_segments(l32, snaps, segs)
    out = {"segments": segs, "seconds": sec32, "full32_s": full32,
           "autotune": {"at_iterations": t_tune, "chosen": chosen, "sweep_s_per_2x10_iters": tune_log}}
    if profile:
        _, out["stages"] = timed_segments(l32, snaps, segs, profile=True)
    l64 = Loop(P, "fp64", n_iter, early)
    l64.run(l64.state(Y0), 0, min(10, n_iter))
    _, full64 = checkpoint_run(l64, Y0, [])
    out["full64_s"] = full64
    out["matched"] = matched_precision(l32, l64, {t: snaps[t] for t in match_starts}, 10)
    out["grad_err"] = [(t, gradient_error(P, l32, l64, snaps[t], t, early)) for t in match_starts]
    out["trace"] = trace_same_states(l32, l64, {t: snaps[t] for t in match_starts})
    out["snaps"] = snaps
    out["match_starts"] = match_starts
    return out


def serving_bench(qps=None, duration=None, seed=0, n_workers=3):
    """The serving configuration: fit, then Poisson transform() requests against the resident model through the
    continuous-batching server; also the saturated rate (all requests queued at once)."""
    import torch
    from magnet import MAGNET, TransformServer
    from magnet.serving import poisson_load, latency_report
    S = SERVING
    qps = qps or S["qps"]
    duration = duration or S["duration"]
    rng = np.random.default_rng(seed)
    k, d = S["clusters"], S["d"]
    ntot = S["n_train"] + S["pool"]
    lab = rng.integers(0, k, ntot)
    X = (rng.normal(0, 4, (k, d))[lab] + rng.normal(0, 1, (ntot, d))).astype(np.float32)
    Xtr, pool = X[:S["n_train"]], X[S["n_train"]:]
    t0 = time.perf_counter()
    m = MAGNET(loss="tsne", random_state=0, transform_epochs=S["n_iter"]).fit(Xtr)
    t_fit = time.perf_counter() - t0
    srv = TransformServer(m, n_workers=n_workers).start()

    def sampler(r):
        size = int(round(np.exp(r.uniform(0, np.log(4096)))))
        return pool[r.integers(0, len(pool), size)]
    try:
        # warm-up of every worker (its stream's memory pool, its k-NN scratch), with small and large requests
        for r in [srv.submit(pool[:sz]) for sz in (4096, 64, 1) * n_workers]:
            r.wait(120)
        t1 = time.perf_counter()
        reqs = poisson_load(srv, sampler, qps, duration, seed + 1)
        wall = time.perf_counter() - t1
        rep = latency_report(reqs, wall)
        r2 = np.random.default_rng(seed + 2)
        burst = [sampler(r2) for _ in range(200)]
        t2 = time.perf_counter()
        breqs = [srv.submit(b) for b in burst]
        for r in breqs:
            r.wait(600)
        sat = sum(len(b) for b in burst) / (time.perf_counter() - t2)
    finally:
        srv.stop()
    rep.update({"qps": qps, "duration_s": duration, "fit_s": round(t_fit, 2), "saturated_points_per_s": round(sat, 1),
                "batches": len(srv.batches), "n_workers": n_workers, "workload": S["desc"]})
    return rep


def batched_instances(seed0=1000):
    """The batched configuration: 64 instances, n_i = round(exp(U(ln 2000, ln 50000))) with instance seed 1000 + i,
    32-d Gaussian mixtures of 5-20 clusters (centres N(0, 25 I)), perplexity-30 affinities."""
    from magnet.affinity import perplexity_affinities
    B = BATCHED
    Ps, Ys = [], []
    for i in range(B["count"]):
        rng = np.random.default_rng(seed0 + i)
        n = int(round(np.exp(rng.uniform(np.log(B["n_min"]), np.log(B["n_max"])))))
        k = int(rng.integers(5, 21))
        X = (rng.normal(0, 5, (k, B["d"]))[rng.integers(0, k, n)] + rng.normal(0, 1, (n, B["d"]))).astype(np.float32)
        Ps.append(perplexity_affinities(X, 30.0))
        Ys.append(rng.normal(0, 1e-4, (n, 2)))
    return Ps, Ys


def batched_bench(tune=True):
    """All instances through mg_tsne_optimize_batch (stream count from the autotuner; one stream and one host thread
    with tune=False) against the same instances one after another through mg_tsne_optimize; both from resident
    inputs, results compared bit for bit."""
    import torch
    from magnet import capi
    from magnet.batch import device_instances, autotune_streams, LAST_SWEEP
    B = BATCHED
    t0 = time.perf_counter()
    Ps, Ys = batched_instances()
    t_build = time.perf_counter() - t0
    kw = dict(n_iter=B["n_iter"], early_iter=B["early"])
    seq = device_instances(Ps, Ys, "fp32")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for ip, ix, dv, Y, U, G in seq:
        capi.optimize(ip, ix, dv, Y, U, G, 0, B["n_iter"], **kw)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t1
    bat = device_instances(Ps, Ys, "fp32")
    t2 = time.perf_counter()
    k, nt = autotune_streams(bat) if tune else (1, 1)
    t_tune = time.perf_counter() - t2
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    capi.optimize_batch(bat, n_streams=k, n_threads=nt, **kw)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t3
    same = all(torch.equal(a[3], b[3]) for a, b in zip(seq, bat))
    npts = sum(P.shape[0] for P in Ps)
    inst_it = B["count"] * B["n_iter"]
    return {"workload": B["desc"], "instances": B["count"], "points": int(npts), "n_streams": k, "host_threads": nt,
            "instance_iters_per_s": round(inst_it / t_bat, 1), "sequential_instance_iters_per_s": round(inst_it / t_seq, 1),
            "speedup_vs_sequential": round(t_seq / t_bat, 3), "batch_s": round(t_bat, 3), "sequential_s": round(t_seq, 3),
            "autotune_s": round(t_tune, 2), "autotune_sweep_12_iters": [list(x) for x in LAST_SWEEP] if tune else None,
            "sort_items": capi.tune("sort_items"), "P_build_s": round(t_build, 1), "identical_to_single_runs": bool(same)}


def cpu_baseline(P, n_iter, early, states, threads, points):
    """The reference's compiled module on this host, in a process of its own (it registers the module name
    `_magnet`, like the product): per-iteration time at the given points of the schedule (GPU FP32 states),
    2 warm-up + 5 timed reference iterations at each; it/s from the mean."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cpu_baseline.npz")
        arrays = {"indptr": P.indptr.astype(np.int32), "indices": P.indices.astype(np.int32), "data": P.data.astype(np.float32),
                  "meta": np.array([P.shape[0], n_iter, early, threads]), "points": np.array(points)}
        for t in points:
            for name, a in zip("YUG", states[t]):
                arrays[f"{name}{t}"] = a
        np.savez(path, **arrays)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", path], capture_output=True, text=True)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        log(f"bench: cpu baseline failed: {r.stderr[-500:]}")
        return None
    return json.loads(lines[-1])


def cpu_baseline_worker(path):
    f = np.load(path)
    n, n_iter, early, threads = [int(v) for v in f["meta"]]
    points = [int(t) for t in f["points"]]
    print(json.dumps(cpu_baseline_run(f, n, n_iter, early, threads, points)), flush=True)


def cpu_baseline_run(f, n, n_iter, early, threads, points):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import scipy.fft as sfft
    R = oracle.ref_module()
    if R is None:
        return None
    ip, ix, d = f["indptr"], f["indices"], f["data"]
    gpu_state_fn = lambda t: (f[f"Y{t}"].copy(), f[f"U{t}"].copy(), f[f"G{t}"].copy())  # noqa: E731
    lr = max(200.0, n / 12.0)
    cache = {}

    def spectra(nb, width):   # magnet/tsne.py:48-71, cached per grid like the reference
        key = (nb, round(width, 6))
        if key not in cache:
            N = nb * 3
            M = sfft.next_fast_len(2 * N, real=True)
            off = np.arange(M)
            off = np.where(off <= M // 2, off, off - M).astype(np.float64) * (width / 3)
            r2 = off[:, None] ** 2 + off[None, :] ** 2
            K1 = (1.0 / (1.0 + r2)).astype(np.float32); K2 = (1.0 / (1.0 + r2) ** 2).astype(np.float32)
            conv = R.FFTConv2D(int(M), 3)
            conv.set_kernels(np.ascontiguousarray(sfft.rfft2(K1, workers=threads).real, dtype=np.float32),
                             np.ascontiguousarray(sfft.rfft2(K2, workers=threads).real, dtype=np.float32))
            cache[key] = conv
        return cache[key]

    from magnet.tsne import grid_params
    per_it = []
    for t in points:
        Y, U, G = gpu_state_fn(t)
        times = []
        for r in range(7):
            it = min(t + r, n_iter - 1)
            rho, mom = (12.0, 0.5) if it < early else (1.0, 0.8)
            t0 = time.perf_counter()
            attr = np.asarray(R.tsne_attraction(ip, ix, d, Y, 1.0, threads))
            xmin, ymin, width, nb = grid_params(R.bbox_2d(Y, threads))
            sumQ, F = spectra(nb, width).compute(Y, xmin, ymin, width, nb, threads)
            R.tsne_step(Y, U, G, attr, np.asarray(F), float(rho), float(1.0 / max(sumQ, 1e-12)), float(lr), float(mom), threads)
            if r >= 2:
                times.append(time.perf_counter() - t0)
        per_it.append(float(np.median(times)))
    mean = float(np.mean(per_it))
    return {"value": round(1.0 / mean, 3), "unit": "iters/s", "cores": threads, "kind": "reference",
            "sample": f"reference _magnet (oracle/_ref, n_threads={threads}) on the same P: median of 5 iterations "
                      f"(after 2 warm-up) at iterations {points} of the schedule, started from the GPU FP32 states; "
                      f"it/s = 1 / mean of the {len(per_it)} medians ({', '.join(f'{1e3 * v:.1f}' for v in per_it)} ms)"}


# single-kernel stages: the roofline's dominant kernel is the one of these with the largest total time
SINGLE_KERNEL = {"columns": "k_columns", "attraction": "k_attr2", "rows_forward": "k_rows_fwd",
                 "rows_inverse": "k_rows_inv", "gather": "k_gather"}


def traffic_of(workload, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters (FETCH_SIZE + WRITE_SIZE, each
    scaled by the reading of a plain copy of known size), from profiles/traffic_<workload>.json; None if absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(kernel)
    return None if rec is None else rec.get("bytes_per_launch")


FFT_COL_KERNELS = ("k_columns", "kw_columns", "k_spec1", "k_spec2", "k_spec2p", "kb_spec1", "kb_spec2", "kb_spec2p")
FFT_ROW_KERNELS = ("k_rows_fwd", "k_rows_inv", "kw_rows_fwd", "kw_rows_inv", "k_rows_fwd_fold", "k_rows_inv_fold",
                   "k_rows_fwd_fold_p", "k_rows_inv_fold_p")


def kernel_key(name):
    """A rocprofv3 kernel name as the traffic table keys it: the kernel's own name with its element type, and for the
    FFT kernels the padded FFT size M (the row kernels are compiled for M / 2)."""
    s = name.replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth, cut = 0, len(s)
    for i, ch in enumerate(s):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            cut = i
            break
    s = s[:cut]
    base, _, targs = s.partition("<")
    base = base.split("::")[-1]
    args, depth, cur = [], 0, ""
    for ch in targs[:-1] if targs.endswith(">") else targs:
        if ch == "," and depth == 0:
            args.append(cur.strip()); cur = ""
            continue
        depth += ch == "<"
        depth -= ch == ">"
        cur += ch
    if cur.strip():
        args.append(cur.strip())
    if base in FFT_ROW_KERNELS and len(args) > 1:
        return f"{base}<{args[0]}, {2 * int(args[1])}>"
    if base in FFT_COL_KERNELS and len(args) > 1:
        return f"{base}<{args[0]}, {args[1]}>"
    return f"{base}<{args[0]}>" if args else base


def pmc_per_kernel(directory, counter):
    """{kernel key: (dispatches, mean counter value)} from the rocprofv3 --pmc CSV files under `directory`."""
    import csv
    import glob
    acc = {}
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                k = kernel_key(row["Kernel_Name"])
                n, v = acc.get(k, (0, 0.0))
                acc[k] = (n + 1, v + float(row["Counter_Value"]))
    return {k: (n, v / n) for k, (n, v) in acc.items()}


def traffic_table(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib, out):
    """The HBM traffic table bench.py reads (`roofline.traffic`): FETCH_SIZE and WRITE_SIZE per dispatch (KiB) of
    every kernel from two --pmc runs, each scaled so that the same counter reads the known size of a plain device
    copy (`--copy-probe copy_mib`, profiled the same way)."""
    copy_bytes = copy_mib * (1 << 20)
    cf = max(pmc_per_kernel(copy_fetch_dir, "FETCH_SIZE").items(), key=lambda kv: kv[1][1])
    cw = max(pmc_per_kernel(copy_write_dir, "WRITE_SIZE").items(), key=lambda kv: kv[1][1])
    ff, wf = copy_bytes / (cf[1][1] * 1024), copy_bytes / (cw[1][1] * 1024)
    fetch, write = pmc_per_kernel(fetch_dir, "FETCH_SIZE"), pmc_per_kernel(write_dir, "WRITE_SIZE")
    table = {"_scale": {"copy_bytes": copy_bytes, "copy_kernel": cf[0], "FETCH_SIZE_KiB": cf[1][1],
                        "WRITE_SIZE_KiB": cw[1][1], "fetch_factor": ff, "write_factor": wf}}
    for k in sorted(set(fetch) & set(write), key=lambda k: -(fetch[k][0] * fetch[k][1] + write[k][0] * write[k][1])):
        fb, wb = fetch[k][1] * 1024 * ff, write[k][1] * 1024 * wf
        table[k] = {"dispatches": fetch[k][0], "fetch_bytes_per_launch": round(fb), "write_bytes_per_launch": round(wb),
                    "bytes_per_launch": round(fb + wb)}
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"traffic_table": out, "kernels": len(table) - 1, "scale": table["_scale"]}), flush=True)


def copy_probe(mib):
    """A plain device-to-device copy of `mib` MiB (the PMC scale reference): prints its size and time."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.copy_(a)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": n * 4, "seconds": dt, "GBs": round(2 * n * 4 / dt / 1e9, 1)}), flush=True)


def attraction_probe(workload, reps=5):
    """The CSR attraction alone (the PMC probe of mg_tsne_attraction): the workload's P, a random layout
    (N(0, 30^2) per coordinate, seed 1), `reps` launches; prints the mean launch time."""
    import torch
    from magnet import capi
    from magnet.affinity import perplexity_affinities
    X, _ = make_data(workload)
    P = perplexity_affinities(X, 30.0)
    del X
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    ip, ix = dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32))
    dv = dev((P.data / P.data.sum()).astype(np.float32))
    Y = dev(np.random.default_rng(1).normal(0, 30, (P.shape[0], 2)).astype(np.float32))
    capi.attraction(ip, ix, dv, Y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        capi.attraction(ip, ix, dv, Y)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "attraction", "workload": workload, "n": int(P.shape[0]), "nnz": int(P.nnz),
                      "mean_us": round(1e6 * (time.perf_counter() - t0) / reps, 1)}), flush=True)


def roofline_and_stages(stages, workload, steps):
    """Per-stage event timings of the profiled pass and the roofline of its dominant kernel: the (stage, FFT size)
    of a single-kernel stage with the largest total time (FFT kernels are compiled per size)."""
    gridstat = stages.pop("_grid")
    sized = stages.pop("_sized")
    rows = {k: dict(ms_total=round(v[0], 3), launches=v[1], avg_us=round(1e3 * v[0] / max(v[1], 1), 2),
                    GBs=round(v[2] / max(v[0], 1e-9) / 1e6, 1)) for k, v in stages.items() if v[1] > 0}
    (dname, dM), (dms, dcnt, dbytes) = max(((k, v) for k, v in sized.items() if k[0] in SINGLE_KERNEL and v[1] > 0),
                                           key=lambda kv: kv[1][0])
    kernel = SINGLE_KERNEL[dname] + (f"<float, {dM}>" if dM else "<float>")
    achieved = dbytes / (dms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kernel, "stage": dname, "fft_size": dM or None, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_of(workload, kernel), "avg_launch_us": round(1e3 * dms / dcnt, 2),
                "launches": dcnt, "algo_bytes_per_launch": round(dbytes / dcnt)}
    tot_ms = sum(v[0] for v in stages.values())
    tot_bytes = sum(v[2] for v in stages.values())
    extra = {"stages": rows, "stage_sizes": {f"{k}@{M}": dict(ms_total=round(v[0], 3), launches=v[1],
                                                               avg_us=round(1e3 * v[0] / max(v[1], 1), 2))
                                             for (k, M), v in sorted(sized.items()) if M and v[1] > 0},
             "pass_GBs": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1),
             "grid": {"mean_N": round(gridstat[0], 1), "max_N": gridstat[1], "max_M": gridstat[2]},
             "algo_bytes_per_iter": round(tot_bytes / steps)}
    return roofline, extra


def schedule_extra(R, n_iter):
    m32, m64 = sum(r[1] for r in R["matched"]), sum(r[2] for r in R["matched"])
    errs = [e for _, e in R["grad_err"]]
    return {
        "timed_runs": [list(x) for x in R["segments"]],
        "full_schedule": {"fp32_it_s": round(n_iter / R["full32_s"], 2), "fp64_it_s": round(n_iter / R["full64_s"], 2),
                          "note": "each precision's whole schedule from Y0 (its own, chaotic trajectory)"},
        "fp32_over_fp64_same_states": round(m64 / m32, 3),
        "matched_ms_per_iter": [[t, round(a, 4), round(b, 4)] for t, a, b in R["matched"]],
        "fp32_gradient_rel_l2": {"max": float(f"{max(errs):.3g}"), "at": [[t, float(f"{e:.3g}")] for t, e in R["grad_err"]]},
        "kl_and_grad_norm_same_states": {
            "columns": ["t", "KL fp32", "KL fp64", "|grad| fp32", "|grad| fp64"],
            "rows": [[t, round(a, 7), round(b, 7), float(f"{c:.7g}"), float(f"{d:.7g}")] for t, a, b, c, d in R["trace"]],
            "note": "FP64 compensated reductions in both precisions; one traced iteration from each same state"},
    }


def scale_leg(warmup):
    """The 1M configuration beside the headline: FP32 and FP64 whole schedules, FP32 against FP64 from the same
    states (10 iterations from each of five) and the FP32 gradient error there."""
    from magnet.affinity import perplexity_affinities
    W = WORKLOADS["gmm1m"]
    t0 = time.time()
    X, _ = make_data("gmm1m")
    P = perplexity_affinities(X, 30.0)
    del X
    t_p = time.time() - t0
    R = precision_run(P, W["n_iter"], W["early"], W["n_iter"], min(warmup, 10), profile=False)
    ex = schedule_extra(R, W["n_iter"])
    return {"workload": W["desc"], "n": int(P.shape[0]), "nnz": int(P.nnz), "P_build_s": round(t_p, 1),
            "fp32_it_s": ex["full_schedule"]["fp32_it_s"], "fp64_it_s": ex["full_schedule"]["fp64_it_s"],
            "fp32_over_fp64_same_states": ex["fp32_over_fp64_same_states"], "matched_ms_per_iter": ex["matched_ms_per_iter"],
            "fp32_gradient_rel_l2": ex["fp32_gradient_rel_l2"], "kl_and_grad_norm_same_states": ex["kl_and_grad_norm_same_states"]}


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--cpu-baseline-worker":
        cpu_baseline_worker(sys.argv[2])
        return
    if len(sys.argv) == 8 and sys.argv[1] == "--traffic-from":
        # --traffic-from FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR COPY_MIB OUT_JSON (host only)
        traffic_table(*sys.argv[2:6], int(sys.argv[6]), sys.argv[7])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="iterations timed (default: the whole schedule)")
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="mnist70k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB", help="only run a plain copy of MIB MiB")
    ap.add_argument("--attraction-probe", action="store_true", help="only time the CSR attraction of the workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the transform() serving leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-instance batched leg")
    ap.add_argument("--batched-probe", action="store_true",
                    help="only the batched leg, on one stream and one host thread (no autotuning sweep)")
    ap.add_argument("--no-scale", action="store_true", help="skip the 1M leg")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the t-SNE pass of one embedding runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from magnet import capi
    from magnet.affinity import perplexity_affinities
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.batched_probe:
        capi.lib()
        capi.use_torch_stream()
        print(json.dumps(batched_bench(tune=False)), flush=True)
        return
    if args.attraction_probe:
        capi.lib()
        capi.use_torch_stream()
        attraction_probe(args.workload)
        return
    capi.lib()
    capi.use_torch_stream()
    W = WORKLOADS[args.workload]
    steps = args.steps or W["n_iter"]
    steps = min(steps, W["n_iter"])
    t0 = time.time()
    X, lab = make_data(args.workload)
    P = perplexity_affinities(X, 30.0)
    del X
    log(f"bench: {W['desc']}: n={P.shape[0]} nnz={P.nnz} (P built in {time.time() - t0:.1f}s)")
    threads = min(16, os.cpu_count() or 1)
    R = precision_run(P, W["n_iter"], W["early"], steps, args.warmup)
    sec32, segs = R["seconds"], R["segments"]
    log(f"bench: fp32 {steps} iterations in runs {segs}: {sec32:.4f}s -> {steps / sec32:.1f} it/s; "
        f"whole schedule {W['n_iter'] / R['full32_s']:.1f} it/s")
    roofline, extra = roofline_and_stages(R["stages"], args.workload, steps)
    extra.update(schedule_extra(R, W["n_iter"]))
    extra["autotune"] = R["autotune"]
    for name, _ in TUNE_SPACE:   # later legs start from the defaults (the batched leg tunes its own)
        capi.tune(name, dict(TUNE_SPACE)[name][0])
    log(f"bench: stages {json.dumps(extra['stages'])}")
    log(f"bench: same states FP32 vs FP64 (t, ms32, ms64): {extra['matched_ms_per_iter']} -> "
        f"{extra['fp32_over_fp64_same_states']}x; FP32 gradient error {extra['fp32_gradient_rel_l2']}")
    cpu = None
    if not args.no_cpu:
        import torch
        torch.cuda.synchronize()
        pts = [a for a, _ in segs] if len(segs) > 1 else R["match_starts"]
        states = {t: tuple(np.ascontiguousarray(a.cpu().numpy()) for a in R["snaps"][t]) for t in pts}
        tc = time.time()
        cpu = cpu_baseline(P, W["n_iter"], W["early"], states, threads, pts)
        log(f"bench: cpu baseline {cpu and cpu['value']} it/s in {time.time() - tc:.1f}s")
    del R
    if args.workload == "mnist70k" and not args.no_scale:
        ts = time.time()
        extra["gmm1m"] = scale_leg(args.warmup)
        log(f"bench: gmm1m {json.dumps(extra['gmm1m'])} ({time.time() - ts:.1f}s)")
    serving = None
    if not args.no_serving and args.workload == "mnist70k":
        ts = time.time()
        serving = serving_bench()
        log(f"bench: serving {json.dumps(serving)} ({time.time() - ts:.1f}s)")
    batched = None
    if not args.no_batched and args.workload == "mnist70k":
        tb = time.time()
        batched = batched_bench()
        log(f"bench: batched {json.dumps(batched)} ({time.time() - tb:.1f}s)")
    line = {
        "metric": METRIC, "value": round(steps / sec32, 2), "unit": "iters/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * sec32 / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": W["desc"], "n": int(P.shape[0]), "nnz": int(P.nnz), "n_iter": W["n_iter"],
                   "early_iter": W["early"], "precision": "fp32 (FP64 Z and reductions)", "n_interp": 3,
                   "box_width": 0.7},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
        "transform_p99_ms": serving and serving["p99_ms"], "serving": serving, "batched": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
