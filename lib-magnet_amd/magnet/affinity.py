"""Affinity matrices P for the t-SNE pass (upstream of the hot path), built on the device.

perplexity_affinities: the k = 3 * perplexity nearest neighbours of every row (exact, mg_knn_exact_graph, or
NN-descent's approximation, mg_nn_descent, both on the device), the textbook perplexity calibration of each row's Gaussian precision by bisection
on the entropy (mg_perplexity_rows), and the symmetrisation (P + P^T) / 2n into CSR (mg_symmetrize_knn) -- the
affinities the measurement configurations ask for (perplexity 30, 90-NN). The reference library builds a fuzzy
15-NN union instead (magnet/knn.py:215-230) and computes perplexity affinities only in its benchmark scripts, through
openTSNE (benchmarks/opentsne_audit.py:18-23), whose default neighbour search is approximate as well.
"""
import numpy as np
import scipy.sparse as sp

__all__ = ["perplexity_affinities", "affinities_from_knn"]


def _csr(n, indptr, indices, data, on_device):
    if on_device:
        return indptr, indices, data
    return sp.csr_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))


def affinities_from_knn(idx, d2, perplexity=30.0, tol=1e-5, n_iter=100, on_device=False):
    """Symmetric P (float32, rows sorted, summing to 1) from neighbour lists: idx (n, k) int32 and squared distances
    d2 (n, k) float32, device tensors: calibration (mg_perplexity_rows) and (P + P^T) / 2n (mg_symmetrize_knn)."""
    from . import capi
    n = idx.shape[0]
    Pc = capi.perplexity_rows(d2.contiguous(), perplexity, tol, n_iter)
    indptr, indices, data = capi.symmetrize_knn(idx.contiguous(), Pc, 0, 1.0 / (2.0 * n))
    return _csr(n, indptr, indices, data, on_device)


def perplexity_affinities(X, perplexity=30.0, k=None, tol=1e-5, n_iter=100, on_device=False, knn="exact",
                          random_state=0):
    """Symmetric P (float32 values, int32 indices, rows sorted by column) summing to 1: a scipy CSR matrix, or with
    on_device=True the device tensors (indptr, indices, data). knn="exact" (default) takes the exact neighbour lists
    (mg_knn_exact_graph: NN-descent's lists bound an all-pairs scan); knn="nndescent" NN-descent's own lists (the
    reference's parameters: 50 candidates, 10 rounds, delta 0.001, its tree-count rule, leaf size 30). Needs a GPU."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("perplexity_affinities runs on the GPU, and no GPU is visible")
    from . import capi
    n = X.shape[0]
    k = int(min(n - 1, k if k is not None else 3 * perplexity))
    if knn == "nndescent":
        Xt = X if isinstance(X, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(X, np.float32), device="cuda")
        n_trees = int(max(2, min(8, round(np.sqrt(n) / 40))))
        idx, d2 = capi.nn_descent(Xt.contiguous(), k, 50, 10, 0.001, n_trees, 30, int(random_state))
    elif knn == "exact":
        Xt = X if isinstance(X, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(X, np.float32), device="cuda")
        idx, d2 = capi.knn_exact_graph(Xt.contiguous(), k)
    else:
        raise ValueError("knn must be 'nndescent' or 'exact'")
    return affinities_from_knn(idx, d2, perplexity, tol, n_iter, on_device)
