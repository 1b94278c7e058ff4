"""Batched form of the t-SNE pass: many independent embeddings of uneven sizes on one schedule.

optimize_tsne_batch runs every instance's full-batch t-SNE loop (the loop of optimize_tsne, magnet/tsne.py:132-214)
through one C-ABI call, mg_tsne_optimize_batch. Every stage of an iteration is one launch for all instances; the
iteration is device-driven (a kernel decides each instance's grid from its last bounding box), so its launches are
captured once in a hipGraph and replayed, and the host only reads the instances' states every few iterations to
plan again when one outgrows its window of FFT sizes. Each instance's result is the one optimize_tsne gives it alone
(same kernels, same grids, same order of operations). The stream and thread counts are picked by autotune_streams:
short timed runs of each candidate on copies of the states, the fastest kept; then the launch choices that change
no result (magnet.tune: the radix sort's keys per thread; graph replay or launches one by one) are swept the same
way.

The reference fits one embedding per call; this form has no counterpart there.
"""
import time

import numpy as np
import scipy.sparse as sp

__all__ = ["optimize_tsne_batch", "autotune_streams", "device_instances"]

_TUNED = {}
LAST_SWEEP = []   # [(n_streams, n_threads, seconds)] of the last autotuner sweep


def device_instances(Ps, Ys, precision="fp32"):
    """(indptr, indices, data, Y, U, G) device tensors per instance; P normalised to sum 1 (tsne.py:149-151)."""
    import torch
    dt = np.float32 if precision == "fp32" else np.float64
    out = []
    for P, Y in zip(Ps, Ys):
        P = sp.csr_matrix(P, dtype=dt)
        P = sp.csr_matrix(P / P.sum())
        Y = np.asarray(Y)
        if Y.shape != (P.shape[0], 2):
            raise ValueError("every Y must be (n, 2) with n = P.shape[0]")
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        out.append((dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32)), dev(P.data.astype(dt)),
                    dev(Y.astype(dt)), dev(np.zeros(Y.shape, dt)), dev(np.ones(Y.shape, dt))))
    return out


CANDIDATES = ((1, 1), (4, 1), (8, 1), (4, 4), (8, 4), (8, 8), (16, 8), (16, 16))


def autotune_streams(instances, candidates=CANDIDATES, iters=12, early_exaggeration=12.0, **kw):
    """(n_streams, n_threads) with the shortest run of `iters` iterations of the schedule's late phase on copies
    of the instances' states (the results of a run do not depend on the choice). Cached per (instances, total
    points, dtype)."""
    import torch
    from . import capi
    key = (len(instances), sum(t[3].shape[0] for t in instances) // 4096, str(instances[0][3].dtype))
    if key in _TUNED:
        return _TUNED[key]
    best, best_t = candidates[0], float("inf")
    LAST_SWEEP.clear()
    for k, nt in candidates:
        if k > len(instances):
            continue
        trial = [(ip, ix, dv, Y.clone(), U.clone(), G.clone()) for ip, ix, dv, Y, U, G in instances]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.optimize_batch(trial, n_iter=iters, early_iter=0, early_exaggeration=early_exaggeration, n_streams=k,
                            n_threads=nt, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        LAST_SWEEP.append((k, nt, round(t, 4)))
        if t < best_t:
            best, best_t = (k, nt), t
    # then the sort's keys per thread at that choice (the batch sorts all its points at once)
    from .tune import autotune

    def run():
        trial = [(ip, ix, dv, Y.clone(), U.clone(), G.clone()) for ip, ix, dv, Y, U, G in instances]
        capi.optimize_batch(trial, n_iter=iters, early_iter=0, early_exaggeration=early_exaggeration, n_streams=best[0],
                            n_threads=best[1], **kw)
    chosen, log = autotune(run, (("sort_items", (0, 2, 4, 8)), ("batch_graph", (1, 0))), repeats=2)
    LAST_SWEEP.extend(log)
    _TUNED[key] = best
    return best


def optimize_tsne_batch(Ps, Ys, n_iter=500, early_iter=166, early_exaggeration=12.0, exaggeration=1.0,
                        momentum=(0.5, 0.8), n_interp=3, box_width=0.7, dof=1.0, precision="fp32", n_streams="auto", n_threads=1):
    """Full-batch t-SNE of every (P_b, Y_b) pair on one schedule; learning rate max(200, n_b / early_exaggeration)
    per instance. Returns the layouts (float32 for fp32, float64 for fp64) in the order given."""
    from . import capi
    if precision not in ("fp32", "fp64"):
        raise ValueError("precision must be 'fp32' or 'fp64'")
    if len(Ps) != len(Ys) or not Ps:
        raise ValueError("need as many layouts as affinity matrices, at least one")
    capi.lib()
    capi.use_torch_stream()
    inst = device_instances(Ps, Ys, precision)
    kw = dict(exaggeration=exaggeration, momentum=momentum, n_interp=n_interp, box_width=box_width, dof=dof)
    if n_streams == "auto":
        k, nt = autotune_streams(inst, early_exaggeration=early_exaggeration, **kw)
    else:
        k, nt = int(n_streams), int(n_threads)
    capi.optimize_batch(inst, n_iter=n_iter, early_iter=early_iter, early_exaggeration=early_exaggeration, n_streams=k,
                        n_threads=nt, **kw)
    return [t[3].cpu().numpy() for t in inst]
