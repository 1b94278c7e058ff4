"""MAGNET's sampled layout pass on the device (magnet/__init__.py:722-848 schedule, src/layout.cpp optimiser): the
low-dimensional kernel fit, the edge sampling schedule, and the epoch loop over mg_optimize_layout.

find_ab_params fits q(d) = 1 / (1 + a d^2b) to the UMAP target curve (1 below min_dist, exp(-(d - min_dist) /
spread) beyond) by least squares on 300 points of [0, 3 spread]; with b given, only a is fitted
(magnet/__init__.py:26-48 states the same curve). sampling_frequencies sets each edge's samples per epoch to
min(1, s w) with s such that they add up to the target (magnet/__init__.py:51-68), edge_schedule keeps the edges
sampled at least once over the run (magnet/__init__.py:727-737), and optimize_layout_sgd runs the epochs with the
learning rate falling linearly to zero and the mid-range weight annealed to zero over the first mid_phase of the run
(magnet/__init__.py:808-833).
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["find_ab_params", "sampling_frequencies", "edge_schedule", "reverse_pairs", "optimize_layout_sgd"]


def find_ab_params(spread=1.0, min_dist=0.1, b=None):
    """(a, b) of the power kernel closest, in least squares, to the UMAP membership curve."""
    from scipy.optimize import least_squares
    x = np.linspace(0.0, 3.0 * spread, 300)
    target = np.where(x < min_dist, 1.0, np.exp(-(x - min_dist) / spread))

    def curve(a_, b_):
        return 1.0 / (1.0 + a_ * np.power(x, 2.0 * b_))

    if b is None:
        fit = least_squares(lambda t: curve(t[0], t[1]) - target, x0=[1.0, 1.0], method="lm")
        return float(fit.x[0]), float(fit.x[1])
    fit = least_squares(lambda t: curve(t[0], b) - target, x0=[1.0], method="lm")
    return float(fit.x[0]), float(b)


def sampling_frequencies(w, target):
    """freq_i = min(1, s w_i) with sum(freq) = target (all ones when there are no more weights than the target).
    Sorted by weight, the first t saturate when s >= 1 / w_t, and the sum is t + s (sum of the rest): the t where
    that line crosses the target gives s exactly."""
    w = np.asarray(w, dtype=np.float64)
    n = len(w)
    if n == 0 or n <= target:
        return np.ones_like(w)
    desc = np.sort(w)[::-1]
    rest = np.concatenate([np.cumsum(desc[::-1])[::-1], [0.0]])   # rest[t] = sum of desc[t:]
    for t in range(n):
        # with desc[:t] saturated: total(s) = t + s rest[t], valid for s in [1 / desc[t - 1], 1 / desc[t])
        if rest[t] <= 0:
            break
        s = (target - t) / rest[t]
        if s * desc[t] <= 1.0:   # desc[t] stays unsaturated, so the first t are exactly the saturated ones
            return np.minimum(1.0, s * w)
    return np.minimum(1.0, w / desc[-1])


def edge_schedule(P, n_epochs, samples_per_node=5.0):
    """(head, tail, epochs_per_sample) of the symmetric weighted graph P: edge e is sampled every 1 / freq_e epochs,
    freq from sampling_frequencies with samples_per_node * n samples per epoch; edges sampled less than once over
    n_epochs are left out. Edges come in row order, so they are sorted by head."""
    C = sp.csr_matrix(P).tocoo()
    freq = sampling_frequencies(C.data.astype(np.float64), samples_per_node * P.shape[0])
    keep = freq >= 1.0 / n_epochs
    return C.row[keep].astype(np.int32), C.col[keep].astype(np.int32), (1.0 / freq[keep]).astype(np.float32)


def reverse_pairs(pairs, n):
    """CSR (indptr, indices) of who chose each node as a mid-range partner."""
    m = pairs.shape[1]
    R = sp.csr_matrix((np.ones(pairs.size, np.float32), (pairs.ravel(), np.repeat(np.arange(n), m))), shape=(n, n))
    R.sort_indices()
    return R.indptr.astype(np.int32), R.indices.astype(np.int32)


def optimize_layout_sgd(Y, head, tail, eps, n_epochs, a, b, rep_a=None, rep_b=None, learning_rate=1.0,
                        negative_sample_rate=5.0, repulsion_strength=2.0, exaggeration=1.0, mid_pairs=None,
                        mid_weight=0.0, mid_rate=0.25, mid_phase=0.5, clip="coord", n_fixed=0, seed=0, attraction_scale=1.5):
    """n_epochs of the sampled pass on Y (n, 2) float32 (returned, float32): learning rate learning_rate
    (1 - e / n_epochs), mid-range weight mid_weight max(0, 1 - e / (mid_phase n_epochs)), on the device.
    attraction_scale: 1.5 (the owner mode's factor, for graphs whose edges come in both directions) or 1 (edges
    without their reverse, as in the out-of-sample placement)."""
    import torch
    from . import capi
    capi.lib()
    n = Y.shape[0]
    dev = lambda x: torch.as_tensor(np.ascontiguousarray(x)).cuda()  # noqa: E731
    Yd = dev(np.asarray(Y, np.float32))
    dh, dt, de = dev(head.astype(np.int32)), dev(tail.astype(np.int32)), dev(eps.astype(np.float32))
    eons = de.clone()
    eonns = (de / float(negative_sample_rate)).contiguous()
    use_mid = mid_pairs is not None and mid_weight > 0 and mid_rate > 0
    pairs = dev(mid_pairs.astype(np.int32)) if use_mid else None
    rp, ri = (dev(x) for x in reverse_pairs(np.asarray(mid_pairs), n)) if use_mid else (None, None)
    total = int(n_epochs)
    mid_len = max(1.0, mid_phase * total)
    alpha = lambda e: learning_rate * (1.0 - e / total)  # noqa: E731
    mid = (lambda e: mid_weight * max(0.0, 1.0 - e / mid_len)) if use_mid else (lambda e: 0.0)  # noqa: E731
    # runs over which both rates are linear: split where the mid-range weight reaches zero
    cuts = sorted({0, total, min(total, int(np.ceil(mid_len)))})
    kw = dict(a=a, b=b, rep_a=rep_a if rep_a is not None else 0.0, rep_b=rep_b if rep_b is not None else 0.0,
              negative_sample_rate=negative_sample_rate, exaggeration=exaggeration, repulsion_strength=repulsion_strength,
              mid_rate=mid_rate if use_mid else 0.0, clip_mode={"coord": 0, "norm": 1}[clip], n_fixed=n_fixed, seed=seed,
              pairs=pairs, rev_indptr=rp, rev_indices=ri, attraction_scale=attraction_scale)
    for e0, e1 in zip(cuts[:-1], cuts[1:]):
        if e1 > e0:
            capi.optimize_layout(Yd, dh, dt, de, eons, eonns, e0, e1, alpha(e0), alpha(e1), mid(e0), mid(e1), **kw)
    torch.cuda.synchronize()
    return Yd.cpu().numpy()
