"""Writes tests/golden/c1_layout.npz: the final layout of the small configuration (10,000 x 50 Gaussian mixture,
10 clusters, perplexity-30 affinities, seeded N(0, 1e-4) start, 500 iterations, 166 of early exaggeration) computed
by the reference's own native kernels, and its 15-NN preservation.

The reference module is the reference's `_magnet` compiled from its C++ sources into oracle/_ref (oracle/Makefile
target `ref`); nothing of the reference's Python package is imported. Its fused loop (magnet/tsne.py:148-184),
grid decision (magnet/tsne.py:30-46) and kernel spectra (magnet/tsne.py:48-71) are the few lines of glue that
make_golden.py restates; every native call runs with n_threads=1, where the reference is deterministic.
The 15-NN preservation is the reference's metric (benchmarks/metrics.py:53-59): over 5,000 query rows drawn with
seed 0, the mean fraction of each row's 15 nearest rows in the data that are among its 15 nearest in the layout.

Run from the repository root:  python tests/golden/make_c1_layout.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "lib-magnet_amd"))
sys.path.insert(0, HERE)
import oracle  # noqa: E402
from make_golden import repulsion  # noqa: E402

OUT = os.path.join(HERE, "c1_layout.npz")
N_ITER, EARLY = 500, 166


def c1_data():
    """The small configuration's data and affinities (as bench.py builds them for the 10k workload), on the host
    (the oracle's exact k-NN, perplexity calibration and symmetrisation)."""
    from affinity_oracle import perplexity_affinities
    rng = np.random.default_rng(0)
    n, d, k = 10_000, 50, 10
    lab = np.repeat(np.arange(k), n // k)
    X = (rng.normal(0, 5, (k, d))[lab] + rng.normal(0, 1, (n, d))).astype(np.float32)
    P = perplexity_affinities(X, 30.0)
    Y0 = np.random.default_rng(1).normal(0, 1e-4, (n, 2))
    return X, lab, P, Y0


def preservation15(X, Y, k=15, n_query=5000, seed=0):
    """benchmarks/metrics.py:53-59: mean |kNN_X(i) & kNN_Y(i)| / k over the query rows."""
    from sklearn.neighbors import NearestNeighbors
    n = len(Y)
    q = np.sort(np.random.default_rng(seed).choice(n, size=min(n_query, n), replace=False))
    hi = NearestNeighbors(n_neighbors=k + 1, algorithm="brute").fit(X).kneighbors(X[q], return_distance=False)
    lo = NearestNeighbors(n_neighbors=k + 1).fit(Y).kneighbors(Y[q], return_distance=False)
    hi = np.array([row[row != i][:k] for i, row in zip(q, hi)])
    lo = np.array([row[row != i][:k] for i, row in zip(q, lo)])
    return float(np.mean([len(np.intersect1d(a, b)) for a, b in zip(hi, lo)]) / k)


def main():
    R = oracle.ref_module()
    if R is None:
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    X, lab, P, Y0 = c1_data()
    n = P.shape[0]
    ip, ix = P.indptr.astype(np.int32), P.indices.astype(np.int32)
    d = (P.data / P.data.sum()).astype(np.float32)   # magnet/tsne.py:149-151
    Y = np.ascontiguousarray(Y0.astype(np.float32))
    U, G = np.zeros_like(Y), np.ones_like(Y)
    lr = max(200.0, n / 12.0)
    t0 = time.time()
    sq = []
    for it in range(N_ITER):
        rho, mom = (12.0, 0.5) if it < EARLY else (1.0, 0.8)
        attr = np.asarray(R.tsne_attraction(ip, ix, d, Y, 1.0, 1))
        _, sumQ, F = repulsion(R, Y, 3)
        sq.append(sumQ)
        R.tsne_step(Y, U, G, attr, F, float(rho), float(1.0 / max(sumQ, 1e-12)), float(lr), float(mom), 1)
    pres = preservation15(X, Y)
    print(f"reference loop: {N_ITER} iterations in {time.time() - t0:.1f}s, 15-NN preservation {pres:.4f}")
    np.savez_compressed(OUT, Y=Y, sumQ=np.array(sq), pres15=np.float64(pres), nnz=np.int64(P.nnz),
                        data_sum=np.float64(P.data.astype(np.float64).sum()))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
