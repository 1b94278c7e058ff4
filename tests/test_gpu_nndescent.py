"""NN-descent on the device (mg_nn_descent, magnet.graph.knn_nndescent) held to the reference's own tests of its
NN-descent (tests/test_magnet.py:143-173: recall against brute force, duplicates and tiny inputs, cosine zero rows),
to the compiled reference's graph in law (the same recall within 0.01 on the same data), to determinism, and at the
configurations' sizes (recall and time)."""
import os
import subprocess
import sys
import time

import numpy as np
import pytest

from magnet import capi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()
    capi.use_torch_stream()
    return torch


def blobs(n=3000, d=20, centers=5, seed=0):
    """tests/test_magnet.py:15-19 of the reference"""
    rng = np.random.default_rng(seed)
    y = rng.integers(0, centers, n)
    C = rng.normal(0, 10, (centers, d))
    return (C[y] + rng.normal(0, 1, (n, d))).astype(np.float32), y


def recall(ind, ei):
    return float(np.mean([len(np.intersect1d(a, b)) / len(b) for a, b in zip(ind, ei)]))


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_nndescent_matches_brute_force(metric):
    """tests/test_magnet.py:143-156: shape, no self, sorted distances, recall > 0.97 on 500 rows, the nearest
    distance within 1e-3 of brute force."""
    from sklearn.neighbors import NearestNeighbors
    from magnet.graph import knn_nndescent
    X, _ = blobs(4000, d=30, centers=8, seed=1)
    ind, dist = knn_nndescent(X, 15, metric=metric, random_state=0)
    assert ind.shape == (4000, 14) and dist.shape == (4000, 14)
    assert not (ind == np.arange(4000)[:, None]).any()
    assert (np.diff(dist, axis=1) >= -1e-5).all()
    nn = NearestNeighbors(n_neighbors=15, metric=metric, algorithm="brute").fit(X)
    ed, ei = nn.kneighbors(X[:500])
    ed, ei = ed[:, 1:], ei[:, 1:]
    r = recall(ind[:500], ei)
    print(f"{metric}: recall {r:.4f}")
    assert r > 0.97
    assert np.allclose(dist[:500, 0], ed[:, 0], atol=1e-3)


def test_nndescent_handles_duplicates_and_small_n():
    """tests/test_magnet.py:158-164: 500 rows of 10 distinct values (nearest distance 0), and 60 rows with k = 15."""
    from magnet.graph import knn_nndescent
    X = np.repeat(np.random.default_rng(0).normal(size=(10, 5)).astype(np.float32), 50, axis=0)
    ind, dist = knn_nndescent(X, 10)
    assert ind.shape == (500, 9) and np.isfinite(dist).all() and (dist[:, 0] == 0).all()
    X, _ = blobs(60, d=5)
    ind, dist = knn_nndescent(X, 15)
    assert ind.shape == (60, 14) and np.isfinite(dist).all()
    assert not (ind == np.arange(60)[:, None]).any()
    assert all(len(set(r)) == 14 for r in ind)


def test_nndescent_cosine_zero_rows_are_not_everyones_neighbours():
    """tests/test_magnet.py:167-173"""
    from magnet.graph import knn_nndescent
    X, _ = blobs(2000, d=20, centers=4, seed=2)
    X[:100] = 0.0
    ind, dist = knn_nndescent(X, 15, metric="cosine")
    assert (ind[100:] < 100).mean() < 0.01
    assert (dist[100:] < 0.99).all()


def test_nndescent_is_deterministic_and_the_pybind_entry_agrees():
    """Two runs with one seed give the same graph bit for bit; _magnet.nn_descent (the reference's signature, host
    arrays) gives the graph of the device entry; another seed gives another graph of the same recall."""
    import torch
    import _magnet
    X, _ = blobs(6000, d=16, centers=6, seed=4)
    Xt = torch.as_tensor(X, device="cuda")
    a = capi.nn_descent(Xt, 20, 50, 10, 0.001, 3, 30, 7)
    b = capi.nn_descent(Xt, 20, 50, 10, 0.001, 3, 30, 7)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    ip, dp = _magnet.nn_descent(X, 20, 50, 10, 0.001, 3, 30, 7, -1, False)
    assert np.array_equal(ip, a[0].cpu().numpy()) and np.array_equal(dp, a[1].cpu().numpy())
    c = capi.nn_descent(Xt, 20, 50, 10, 0.001, 3, 30, 8)
    assert not torch.equal(a[0], c[0])
    D = ((X[:300, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    D[np.arange(300), np.arange(300)] = np.inf
    ex = np.argsort(D, 1, kind="stable")[:, :20]
    ra, rc = recall(a[0].cpu().numpy()[:300], ex), recall(c[0].cpu().numpy()[:300], ex)
    print(f"recall seed 7 {ra:.4f}, seed 8 {rc:.4f}")
    assert ra > 0.97 and rc > 0.97
    with pytest.raises(RuntimeError):
        capi.nn_descent(Xt, 200, 50, 10, 0.001, 3, 30, 7)   # k above 128


def _reference_graph(X, k, n_trees, seed):
    """the compiled reference's nn_descent on X, in a process of its own (tests/ref_nndescent_graph.py)"""
    import glob
    import tempfile
    if not glob.glob(os.path.join(ROOT, "oracle", "_ref", "_magnet*.so")):
        pytest.skip("oracle/_ref not built")
    with tempfile.TemporaryDirectory() as td:
        src, dst = os.path.join(td, "x.npy"), os.path.join(td, "g.npz")
        np.save(src, X)
        r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "ref_nndescent_graph.py"), src, dst, str(k),
                            str(n_trees), str(seed)], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        g = np.load(dst)
        return g["ind"], g["dist"], float(g["seconds"])


@pytest.mark.parametrize("n,d,k", [(20000, 30, 20), (20000, 100, 30)])
def test_nndescent_recall_as_the_reference(n, d, k):
    """The same data through the compiled reference (its C++ nn_descent, 16 threads) and the device: the device's
    recall against exact search on 1000 rows, and the share of rows whose nearest neighbour it finds, are at least the
    reference's less 0.01 (both approximate the same sets; their random choices differ)."""
    import torch
    from magnet.graph import knn_query
    X, _ = blobs(n, d=d, centers=12, seed=5)
    n_trees = int(max(2, min(8, round(np.sqrt(n) / 40))))
    ri, rd, rs = _reference_graph(X, k, n_trees, 0)
    Xt = torch.as_tensor(X, device="cuda")
    torch.cuda.synchronize()
    t = time.time()
    gi, gd = capi.nn_descent(Xt, k, 50, 10, 0.001, n_trees, 30, 0)
    torch.cuda.synchronize()
    gs = time.time() - t
    q = np.random.default_rng(0).choice(n, 1000, replace=False)
    ei, ed = knn_query(Xt, Xt[torch.as_tensor(q, device="cuda")], k + 1, squared=True)
    ei, ed = ei.cpu().numpy(), ed.cpu().numpy()
    ex = np.array([[j for j in row if j != i][:k] for i, row in zip(q, ei)])
    r_ref, r_dev = recall(ri[q], ex), recall(gi.cpu().numpy()[q], ex)
    near_dev = np.isclose(gd.cpu().numpy()[q, 0], ed[:, 1], rtol=1e-4).mean()
    near_ref = np.isclose(rd[q].min(1), ed[:, 1], rtol=1e-4).mean()
    print(f"n={n} d={d} k={k}: recall device {r_dev:.4f} ({gs:.3f} s), reference {r_ref:.4f} ({rs:.2f} s); nearest "
          f"found on {near_dev:.3f} / {near_ref:.3f} of the rows")
    assert r_dev >= r_ref - 0.01 and near_dev >= near_ref - 0.01


@pytest.mark.parametrize("name", ["mnist70k", "gmm1m"])
def test_nndescent_at_the_configurations_sizes(name):
    """The MNIST-shaped input with k = 90 (the perplexity-30 lists) and the 1M input with k = 20 (the reference's
    k_internal): recall on 1000 rows against exact search at least the compiled reference's less 0.01, on the same
    data with the same trees and rounds, and the two times (the reference on 16 host threads)."""
    import torch
    from bench import make_data
    from magnet.graph import knn_query
    X, _ = make_data(name)
    n = len(X)
    k = 90 if name == "mnist70k" else 20
    n_trees = int(max(2, min(8, round(np.sqrt(n) / 40))))
    ri, _, rs = _reference_graph(X, k, n_trees, 0)
    Xt = torch.as_tensor(X, device="cuda")
    del X
    torch.cuda.synchronize()
    t = time.time()
    gi, gd = capi.nn_descent(Xt, k, 50, 10, 0.001, n_trees, 30, 0)
    torch.cuda.synchronize()
    gs = time.time() - t
    q = np.random.default_rng(0).choice(n, 1000, replace=False)
    ei, _ = knn_query(Xt, Xt[torch.as_tensor(q, device="cuda")], k + 1, tile=1000)
    ei = ei.cpu().numpy()
    ex = np.array([[j for j in row if j != i][:k] for i, row in zip(q, ei)])
    r_dev, r_ref = recall(gi.cpu().numpy()[q], ex), recall(ri[q], ex)
    print(f"{name}: k={k} recall device {r_dev:.4f} in {gs:.2f} s, reference {r_ref:.4f} in {rs:.1f} s")
    assert r_dev >= r_ref - 0.01


def test_mid_near_pairs_equal_the_oracle():
    """mg_mid_near_pairs against the oracle's restatement with the same hashed draws: the same partner for every pair
    but those whose two nearest draws tie within float32 rounding; the pybind entry (the reference's
    mid_near_pairs_dense signature) gives the same array."""
    import torch
    import _magnet
    import affinity_oracle as AO
    X, _ = blobs(5000, d=12, centers=5, seed=3)
    got = capi.mid_near_pairs(torch.as_tensor(X, device="cuda"), 5, 6, 9).cpu().numpy()
    want = AO.mid_near_pairs(X, 5, 6, 9)
    print(f"mid-near partners equal on {np.mean(got == want):.5f}")
    assert np.mean(got == want) > 0.999 and not (got == np.arange(5000)[:, None]).any()
    assert np.array_equal(_magnet.mid_near_pairs_dense(X, 5, 6, 9, -1), got)


@pytest.mark.parametrize("n,d,k,kind", [(3000, 20, 90, "normal"), (2000, 5, 30, "integer"), (4000, 784, 90, "normal"),
                                        (500, 5, 100, "duplicates"), (1000, 3, 5, "two_values")])
def test_exact_graph_equals_brute_force(n, d, k, kind):
    """mg_knn_exact_graph against a float64 brute force: distances to 1e-5 and the same neighbour set wherever the
    k-th and (k+1)-th distances are apart; with integer features (exact distances, many ties) the stable
    (distance, index) order exactly; duplicate rows (ties at distance 0 beyond the bound's buffer: the second scan)."""
    import torch
    rng = np.random.default_rng(n + d)
    if kind == "normal":
        X = rng.normal(0, 1, (n, d)).astype(np.float32)
    elif kind == "integer":
        X = rng.integers(0, 4, (n, d)).astype(np.float32)
    elif kind == "duplicates":
        X = np.repeat(rng.normal(size=(10, d)).astype(np.float32), n // 10, axis=0)
    else:
        X = np.repeat(np.array([[0, 0, 0], [1, 2, 3]], np.float32), n // 2, axis=0)
    idx, d2 = capi.knn_exact_graph(torch.as_tensor(X, device="cuda"), k)
    idx, d2 = idx.cpu().numpy(), d2.cpu().numpy()
    X64 = X.astype(np.float64)
    D = (X64 ** 2).sum(1)[:, None] + (X64 ** 2).sum(1)[None, :] - 2.0 * X64 @ X64.T
    if kind != "normal":
        D = ((X64[:, None, :] - X64[None, :, :]) ** 2).sum(-1) if n * n * d <= 30_000_000 else D
    np.fill_diagonal(D, np.inf)
    order = np.argsort(D, axis=1, kind="stable")
    ref = np.take_along_axis(D, order, 1)
    assert not (idx == np.arange(n)[:, None]).any() and (idx >= 0).all()
    if kind == "normal":
        assert np.allclose(d2, ref[:, :k], rtol=1e-5, atol=1e-5)
        clear = (ref[:, k] - ref[:, k - 1]) > 1e-4 * np.maximum(ref[:, k], 1e-12)
        bad = [q for q in np.nonzero(clear)[0] if set(idx[q]) != set(order[q, :k])]
        assert not bad, bad[:5]
    else:   # exact ties (integer features or repeated rows): the stable order itself
        assert np.array_equal(idx, order[:, :k].astype(np.int32))
        if kind == "duplicates":
            assert np.allclose(d2, ref[:, :k], rtol=1e-5, atol=1e-6)
        else:
            assert np.array_equal(d2.astype(np.float64), ref[:, :k])
