"""The sampled-layout oracle against the reference's compiled module (oracle/_ref), run by tests/test_oracle.py in a
process of its own: `python tests/ref_layout_checks.py exact|law` prints "ok" when the check holds.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def _blob_graph(n=3000, k=15, d=20, seed=0):
    """blobs (6 centres) and a symmetric exp(-d / mean d) k-NN graph, edges sorted by head"""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    y = rng.integers(0, 6, n)
    X = (rng.normal(0, 10, (6, d))[y] + rng.normal(0, 1, (n, d))).astype(np.float32)
    D = ((X[:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    nn = np.argsort(D, 1)[:, :k]
    dist = np.sqrt(np.take_along_axis(D, nn, 1))
    A = sp.csr_matrix((np.exp(-dist / dist.mean()).ravel(), (np.repeat(np.arange(n), k), nn.ravel())), shape=(n, n))
    A = ((A + A.T) * 0.5).tocsr()
    A.sort_indices()
    C = A.tocoo()
    w = C.data
    eps = (w.max() / w).astype(np.float32)
    return X, y, C.row.astype(np.int32), C.col.astype(np.int32), eps


def _ref_layout(R, Y, head, tail, eps, e0, e1, alpha0, alpha1, nsr, pairs=None, mid_w=0.0, mid_rate=0.0, rev=None, seed=3):
    eons = eps.copy(); eonns = (eps / nsr).astype(np.float32)
    z2 = np.zeros((0, 2), np.float32); zi = np.zeros(0, np.int32)
    R.optimize_layout(Y, head, tail, eps, eons, eonns, zi, zi, e0, e1, alpha0, alpha1, mid_w, mid_w, 1.577, 0.895, 1.577,
                      0.895, nsr, 0, 1.0, 1.0, 0, mid_rate, 0, 0, 0, z2, z2, zi, 0.9, 0.999, zi,
                      pairs if pairs is not None else np.zeros((0, 1), np.int32), zi, zi, 0.0, 0, np.zeros(0, np.float32),
                      seed, 1, 1, rev[0] if rev else zi, rev[1] if rev else zi)


def check_exact():
    """no random numbers drawn: the oracle equals the compiled reference bit for bit after 30 epochs"""
    R = O.ref_module()
    X, y, head, tail, eps = _blob_graph(1500)
    Y0 = np.random.default_rng(1).uniform(-10, 10, (len(X), 2)).astype(np.float32)
    Yr = Y0.copy()
    _ref_layout(R, Yr, head, tail, eps, 0, 30, 1.0, 0.4, 1e-6)
    Yo = Y0.copy()
    eons = eps.copy(); eonns = (eps / 1e-6).astype(np.float32)
    O.layout_epochs(Yo, head, tail, eps, eons, eonns, 0, 30, 1.0, 0.4, a=1.577, b=0.895, rep_a=1.577, rep_b=0.895,
                    negative_sample_rate=1e-6, repulsion_strength=1.0, seed=3)
    assert np.abs(Yo - Y0).max() > 1e-2 and np.array_equal(Yo, Yr)


def check_law():
    """negative samples and mid-range pairs: both separate the blobs after 200 epochs, at the same spread"""
    R = O.ref_module()
    from sklearn.neighbors import KNeighborsClassifier
    X, y, head, tail, eps = _blob_graph(3000)
    n = len(X)
    pairs = np.random.default_rng(2).integers(0, n, (n, 5)).astype(np.int32)
    import scipy.sparse as sp
    Rv = sp.csr_matrix((np.ones(pairs.size, np.float32), (pairs.ravel(), np.repeat(np.arange(n), 5))), shape=(n, n))
    Rv.sort_indices()
    rev = (Rv.indptr.astype(np.int32), Rv.indices.astype(np.int32))
    Y0 = np.random.default_rng(1).uniform(-10, 10, (n, 2)).astype(np.float32)
    Yr = Y0.copy()
    _ref_layout(R, Yr, head, tail, eps, 0, 200, 1.0, 0.005, 5.0, pairs, 0.05, 0.25, rev)
    Yo = Y0.copy()
    eons = eps.copy(); eonns = (eps / 5.0).astype(np.float32)
    O.layout_epochs(Yo, head, tail, eps, eons, eonns, 0, 200, 1.0, 0.005, 0.05, 0.05, a=1.577, b=0.895, rep_a=1.577,
                    rep_b=0.895, negative_sample_rate=5.0, repulsion_strength=1.0, mid_rate=0.25, seed=3, pairs=pairs,
                    rev_indptr=rev[0], rev_indices=rev[1])
    for Y in (Yr, Yo):
        assert np.isfinite(Y).all() and KNeighborsClassifier(10).fit(Y, y).score(Y, y) > 0.95
    assert abs(Yo.std() / Yr.std() - 1) < 0.15


if __name__ == "__main__":
    {"exact": check_exact, "law": check_law}[sys.argv[1]]()
    print("ok")
