"""Affinity matrices P for the t-SNE pass (upstream of the hot path).

perplexity_affinities: exact k-NN (brute force, k = 3 * perplexity by default) and the textbook
perplexity calibration of each row's Gaussian bandwidth by bisection on the entropy, symmetrised as
(P + P^T) / (2n) -- the affinities the measurement configurations ask for (perplexity 30, about 90-NN).
The reference library itself builds a fuzzy 15-NN union instead (magnet/knn.py:215-230) and computes
perplexity affinities only in benchmark scripts, through openTSNE.

The k-NN search runs on torch (GPU when available) in tiles; it is input preparation, not the pass.
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["exact_knn", "perplexity_affinities", "calibrate_perplexity"]


def exact_knn(X, k, device=None, tile=4096):
    """(indices (n, k) int64, squared distances (n, k) float32) of the k nearest other rows of X."""
    import torch
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    Xt = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float32), device=device)
    n = Xt.shape[0]
    k = min(k, n - 1)
    sq = (Xt * Xt).sum(1)
    out_i = np.empty((n, k), np.int64)
    out_d = np.empty((n, k), np.float32)
    for s in range(0, n, tile):
        e = min(n, s + tile)
        D = sq[s:e, None] + sq[None, :] - 2.0 * (Xt[s:e] @ Xt.T)
        D.clamp_(min=0.0)
        D[torch.arange(e - s, device=device), torch.arange(s, e, device=device)] = float("inf")
        d, i = torch.topk(D, k, dim=1, largest=False, sorted=True)
        out_i[s:e] = i.cpu().numpy()
        out_d[s:e] = d.cpu().numpy()
    return out_i, out_d


def calibrate_perplexity(dist2, perplexity=30.0, tol=1e-5, n_iter=100):
    """Conditional affinities p_{j|i} (n, k) with entropy log(perplexity) per row, by bisection on beta."""
    d = np.asarray(dist2, dtype=np.float64)
    d = d - d[:, :1]   # shift by the nearest distance (cancels in the normalisation, avoids underflow)
    n = d.shape[0]
    target = np.log(perplexity)
    beta = np.ones(n)
    lo = np.full(n, -np.inf)
    hi = np.full(n, np.inf)
    for _ in range(n_iter):
        W = np.exp(-d * beta[:, None])
        Ws = W.sum(1)
        H = np.log(Ws) + beta * (d * W).sum(1) / Ws
        diff = H - target
        if np.abs(diff).max() < tol:
            break
        up = diff > 0   # entropy too high: narrow the kernel
        lo = np.where(up, beta, lo)
        hi = np.where(up, hi, beta)
        beta = np.where(up, np.where(np.isinf(hi), beta * 2, (beta + hi) / 2), np.where(np.isinf(lo), beta / 2, (beta + lo) / 2))
    W = np.exp(-d * beta[:, None])
    return W / W.sum(1, keepdims=True)


def perplexity_affinities(X, perplexity=30.0, k=None, device=None):
    """Symmetric CSR P (float32 values, int32 indices) summing to 1."""
    n = X.shape[0]
    k = int(min(n - 1, k if k is not None else 3 * perplexity))
    idx, d2 = exact_knn(X, k, device)
    Pc = calibrate_perplexity(d2, perplexity)
    rows = np.repeat(np.arange(n), k)
    P = sp.csr_matrix((Pc.ravel(), (rows, idx.ravel())), shape=(n, n))
    P = (P + P.T).tocsr()
    P.sort_indices()
    P = P / (2.0 * n)
    P = sp.csr_matrix(P, dtype=np.float32)
    P.indptr = P.indptr.astype(np.int32)
    P.indices = P.indices.astype(np.int32)
    return P
