"""Affinity matrices P for the t-SNE pass (upstream of the hot path), built on the device.

perplexity_affinities: the k = 3 * perplexity nearest neighbours of every row (exact search on the device), the
textbook perplexity calibration of each row's Gaussian precision by bisection on the entropy (mg_perplexity_rows),
and the symmetrisation (P + P^T) / 2n into CSR (mg_symmetrize_knn) -- the affinities the measurement configurations
ask for (perplexity 30, 90-NN). The reference library builds a fuzzy 15-NN union instead (magnet/knn.py:215-230)
and computes perplexity affinities only in its benchmark scripts, through openTSNE
(benchmarks/opentsne_audit.py:18-23).
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["perplexity_affinities"]


def perplexity_affinities(X, perplexity=30.0, k=None, tol=1e-5, n_iter=100, on_device=False, tile=4096):
    """Symmetric P (float32 values, int32 indices, rows sorted by column) summing to 1: a scipy CSR matrix, or with
    on_device=True the device tensors (indptr, indices, data). Needs a GPU."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("perplexity_affinities runs on the GPU, and no GPU is visible")
    from . import capi
    from .graph import knn_query
    n = X.shape[0]
    k = int(min(n - 1, k if k is not None else 3 * perplexity))
    idx, d2 = knn_query(X, X, k, exclude_self=True, squared=True, tile=tile)
    Pc = capi.perplexity_rows(d2.contiguous(), perplexity, tol, n_iter)
    del d2
    indptr, indices, data = capi.symmetrize_knn(idx.contiguous(), Pc, 0, 1.0 / (2.0 * n))
    if on_device:
        return indptr, indices, data
    return sp.csr_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
