"""MAGNET(loss="tsne") end to end on the GPU: the reference's own quality tests for this loss
(tests/test_magnet.py:234-240 and :440-447), the transform's independence from the other points of a call, and
the continuous-batching service (each request's result the same alone and in the mix)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magnet import capi
    capi.lib()
    capi.use_torch_stream()
    return torch


def blobs(n=3000, d=20, centers=5, seed=0):
    """the reference's test data (tests/test_magnet.py:15-19)"""
    rng = np.random.default_rng(seed)
    y = rng.integers(0, centers, n)
    C = rng.normal(0, 10, (centers, d))
    return (C[y] + rng.normal(0, 1, (n, d))).astype(np.float32), y


def knn_purity(Y, y, k=10):
    from sklearn.neighbors import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=k + 1).fit(Y).kneighbors(Y, return_distance=False)[:, 1:]
    return (y[nn] == y[:, None]).mean()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_tsne_loss_separates_blobs(precision):
    from magnet import MAGNET
    X, y = blobs(2000, d=10, centers=5, seed=5)
    m = MAGNET(loss="tsne", n_epochs=300, random_state=0, precision=precision).fit(X)
    assert m.embedding_.shape == (2000, 2) and np.isfinite(m.embedding_).all()
    assert knn_purity(m.embedding_, y) > 0.95
    assert m.graph_.shape == (2000, 2000) and m.mid_pairs_.shape == (2000, 5)


@pytest.fixture(scope="module")
def fitted():
    from magnet import MAGNET
    X, y = blobs(2000, d=10, centers=5, seed=5)
    m = MAGNET(loss="tsne", random_state=0).fit(X[:1700])
    return m, X, y


def test_transform_after_fit(fitted):
    m, X, y = fitted
    Yq = m.transform(X[1700:])
    from sklearn.neighbors import KNeighborsClassifier
    assert Yq.shape == (300, 2) and np.isfinite(Yq).all()
    assert (KNeighborsClassifier(10).fit(m.embedding_, y[:1700]).predict(Yq) == y[1700:]).mean() > 0.95


def test_transform_of_wide_rows():
    """Rows of 100 features (past the fused k-NN's 64): the search goes through GEMM distance tiles and
    mg_topk_rows; the reference's transform test still holds."""
    from magnet import MAGNET
    from sklearn.neighbors import KNeighborsClassifier
    X, y = blobs(2000, d=100, centers=5, seed=6)
    m = MAGNET(loss="tsne", random_state=0).fit(X[:1700])
    Yq = m.transform(X[1700:])
    assert Yq.shape == (300, 2) and np.isfinite(Yq).all()
    assert (KNeighborsClassifier(10).fit(m.embedding_, y[:1700]).predict(Yq) == y[1700:]).mean() > 0.95


def test_transform_point_does_not_depend_on_its_batch(fitted):
    m, X, y = fitted
    Q = X[1700:]
    full = m.transform(Q)
    part = m.transform(Q[37:91])
    one = m.transform(Q[5:6])
    assert np.allclose(full[37:91], part, rtol=0, atol=1e-5 * np.abs(full).max())
    assert np.allclose(full[5:6], one, rtol=0, atol=1e-5 * np.abs(full).max())


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_transform_point_outside_the_grid_feels_the_far_field(precision):
    """A new point outside the training layout's grid feels the layout's far-field repulsion, the training points
    taken at their centroid, which decays like r^-3 (not an extrapolation of the field, which grows with distance).
    One iteration from rest moves a point by -lr * gain * g, gain 1.2 where g > 0 and 0.8 elsewhere (a learning rate
    of 1e6 makes the step large against the coordinates' rounding): the gradient recovered from that step matches the
    exact rho A - F / Z on the host (exact pairwise sums) to 1e-3 at 10 to 50 layout spans; a point inside the grid
    matches to the field's interpolation error."""
    import torch
    from magnet import MAGNET
    X, y = blobs(2000, d=10, centers=5, seed=5)
    m = MAGNET(loss="tsne", random_state=0, precision=precision).fit(X[:1700])
    conv, (Xtr, Ytr, dt, lr) = m._serving_state()
    Yt = Ytr.double().cpu().numpy()
    n = len(Yt)
    span = float(np.abs(Yt).max())
    k = 14
    nbr = np.tile(np.arange(k, dtype=np.int32), (5, 1))
    w = np.full((5, k), 1.0 / k)
    start = np.array([[10.0, 0.0], [0.0, -50.0], [30.0, 30.0], [-10.0, 7.0], [0.0, 0.0]]) * span
    start[4] = Yt[:k].mean(0) + 0.1
    start = start.astype(np.float32).astype(np.float64)   # representable in either precision
    Yn = torch.as_tensor(start, dtype=dt, device="cuda")
    lr_big = 1e6
    conv.transform(Ytr, torch.as_tensor(nbr, device="cuda"), torch.as_tensor(w, dtype=dt, device="cuda"), Yn, n_iter=1,
                   learning_rate=lr_big, momentum=(0.5, 0.8), momentum_switch=1, exaggeration=1.0, dof=1.0)
    step = Yn.double().cpu().numpy() - start
    g_dev = -step / (lr_big * np.where(step < 0, 1.2, 0.8))
    D2 = ((Yt[:, None] - Yt[None]) ** 2).sum(-1)
    Q = 1.0 / (1.0 + D2)
    np.fill_diagonal(Q, 0.0)
    Z = Q.sum()
    g_ref = []
    for i in range(5):
        d = start[i] - Yt
        q = 1.0 / (1.0 + (d ** 2).sum(1))
        A = ((w[i] / n) * q[:k])[:, None] * d[:k]
        F = ((q ** 2)[:, None] * d).sum(0)
        g_ref.append(A.sum(0) - F / Z)
    g_ref = np.array(g_ref)
    rel = np.linalg.norm(g_dev - g_ref, axis=1) / np.linalg.norm(g_ref, axis=1)
    print(f"{precision}: layout span {span:.1f}; gradient relative error far {rel[:4]}, inside {rel[4]:.2e}")
    assert (rel[:4] < 1e-3).all() and rel[4] < 2e-2


@pytest.mark.parametrize("n_workers", [1, 3])
def test_serving_same_alone_and_in_the_mix(fitted, n_workers):
    """Each request's result is its result alone, whichever requests share its batch and whichever batches run
    beside it on the other workers' streams."""
    import threading
    from magnet import TransformServer
    m, X, y = fitted
    Q = X[1700:]
    rng = np.random.default_rng(0)
    cuts = np.sort(rng.choice(np.arange(1, 300), 11, replace=False))
    parts = np.split(np.arange(300), cuts)
    alone = [m.transform(Q[p]) for p in parts]
    srv = TransformServer(m, max_batch_points=120, n_workers=n_workers).start()
    try:
        reqs = [None] * len(parts)

        def client(i):
            reqs[i] = srv.submit(Q[parts[i]])
        threads = [threading.Thread(target=client, args=(i,)) for i in range(len(parts))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        outs = [r.wait(120) for r in reqs]
    finally:
        srv.stop()
    scale = np.abs(np.concatenate(alone)).max()
    for a, o in zip(alone, outs):
        assert o.shape == a.shape and np.allclose(a, o, rtol=0, atol=1e-5 * scale)
    assert all(r.latency is not None and r.latency > 0 and r.batch is not None for r in reqs)
    assert sum(b[1] for b in srv.batches) == 300 and max(b[1] for b in srv.batches) <= max(120, max(map(len, parts)))


@pytest.mark.parametrize("k,cols", [(14, 200_003), (5, 37), (29, 10_000), (1, 7), (16, 16)])
def test_topk_rows_matches_a_full_sort(k, cols):
    """mg_topk_rows: the k smallest of each row, ordered by (value, column), as a stable full sort gives them;
    rows with ties included."""
    import torch
    from magnet import capi
    rng = np.random.default_rng(k + cols)
    D = rng.random((9, cols)).astype(np.float32)
    D[3] = np.round(D[3] * 4) / 4          # many ties
    D[5] = 0.5                             # all equal
    idx, val = capi.topk_rows(torch.as_tensor(D, device="cuda"), k)
    idx, val = idx.cpu().numpy(), val.cpu().numpy()
    order = np.argsort(D, axis=1, kind="stable")[:, :k]
    assert np.array_equal(idx, order.astype(np.int32))
    assert np.array_equal(val, np.take_along_axis(D, order, 1))


@pytest.mark.parametrize("d,k,exclude,m,nt", [(50, 14, False, 700, 3000), (7, 5, True, 700, 3000), (64, 29, False, 700, 3000),
                                               (1, 3, True, 700, 3000), (12, 14, True, 500, 500),
                                               (50, 14, False, 5, 200_000), (7, 29, True, 3, 150_000),
                                               (65, 14, False, 300, 3000), (100, 5, True, 300, 3000),
                                               (784, 29, False, 200, 5000), (300, 14, False, 4, 70_000)])
def test_fused_knn_is_exact(d, k, exclude, m, nt):
    """mg_knn against a float64 brute force: distances to 1e-5 relative, and the same neighbour set wherever the
    k-th and (k+1)-th distances are apart; exclude_self leaves each row out of its own list; a query's result does
    not depend on the other queries of the call. 500 training rows split into fewer chunks than k (the merge
    then writes k results with fewer threads); a few queries against 150k-200k rows split into 1024 chunks (more
    lists than merge threads: each thread merges four in turn before the tree)."""
    import torch
    from magnet import capi
    rng = np.random.default_rng(d * 100 + k)
    T = rng.normal(0, 1, (nt, d)).astype(np.float32)
    Q = T[:m].copy() if exclude else rng.normal(0, 1, (m, d)).astype(np.float32)
    Tt, Qt = torch.as_tensor(T, device="cuda"), torch.as_tensor(Q, device="cuda")
    idx, d2 = capi.knn(Tt, Qt, k, exclude)
    idx, d2 = idx.cpu().numpy(), d2.cpu().numpy()
    Q64, T64 = Q.astype(np.float64), T.astype(np.float64)
    if m * nt * d <= 50_000_000:
        D = ((Q64[:, None, :] - T64[None, :, :]) ** 2).sum(-1)
    else:
        D = (Q64 ** 2).sum(1)[:, None] + (T64 ** 2).sum(1)[None, :] - 2.0 * Q64 @ T64.T
    if exclude:
        D[np.arange(m), np.arange(m)] = np.inf
    order = np.argsort(D, axis=1, kind="stable")
    ref = np.take_along_axis(D, order, 1)
    assert np.allclose(d2, ref[:, :k], rtol=1e-5, atol=1e-6)
    clear = (ref[:, k] - ref[:, k - 1]) > 1e-4 * np.maximum(ref[:, k], 1e-12)
    for q in np.nonzero(clear)[0]:
        assert set(idx[q]) == set(order[q, :k]), q
    if exclude:
        assert not np.any(idx == np.arange(m)[:, None])
    q1 = min(123, m - 1)
    one_i, one_d = capi.knn(Tt, Qt[q1:q1 + 1].contiguous(), k, False)
    if not exclude:
        assert np.array_equal(one_i.cpu().numpy()[0], idx[q1]) and np.array_equal(one_d.cpu().numpy()[0], d2[q1])


@pytest.mark.parametrize("d,k,exclude,m,nt", [(12, 14, False, 300, 3000), (5, 29, True, 40, 2000), (12, 29, False, 3, 150_000),
                                               (3, 7, True, 4, 100_000), (100, 14, False, 300, 3000), (200, 29, True, 40, 2000)])
def test_knn_ties_break_by_index(d, k, exclude, m, nt):
    """Small integer features make every squared distance an exact integer, with many exact ties: mg_knn must
    give the stable sort's (distance, index) order exactly, across chunk boundaries and through the merge
    (3000 rows: 16 chunks; 100k-150k rows: 1024 chunks, four lists per merge thread)."""
    import torch
    from magnet import capi
    rng = np.random.default_rng(nt + d)
    T = rng.integers(0, 4, (nt, d)).astype(np.float32)
    Q = T[:m].copy() if exclude else rng.integers(0, 4, (m, d)).astype(np.float32)
    idx, d2 = capi.knn(torch.as_tensor(T, device="cuda"), torch.as_tensor(Q, device="cuda"), k, exclude)
    Qi, Ti = Q.astype(np.int64), T.astype(np.int64)
    D = ((Qi ** 2).sum(1)[:, None] + (Ti ** 2).sum(1)[None, :] - 2 * Qi @ Ti.T).astype(np.float64)
    if exclude:
        D[np.arange(m), np.arange(m)] = np.inf
    order = np.argsort(D, axis=1, kind="stable")[:, :k]
    assert np.array_equal(idx.cpu().numpy(), order.astype(np.int32))
    assert np.array_equal(d2.cpu().numpy().astype(np.float64), np.take_along_axis(D, order, 1))


def test_device_smooth_knn_matches_the_numpy_restatement():
    """mg_smooth_knn against the oracle's smooth_knn (the reference's numpy algorithm) on sorted distances,
    with leading zeros and an all-zero row (the fallback to the call's mean distance)."""
    import torch
    from magnet import capi
    from affinity_oracle import smooth_knn
    rng = np.random.default_rng(3)
    D = np.sort(rng.gamma(2.0, 1.0, (500, 14)).astype(np.float32), axis=1)
    D[7, :3] = 0.0
    D[11] = 0.0
    ref = smooth_knn(D, 15)
    mean_all = torch.full((500,), float(D.mean()), device="cuda")
    got = capi.smooth_knn(torch.as_tensor(D, device="cuda"), 15, mean_all).cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-6)
    assert np.all(got[11] == 1.0) and np.all(got[7, :3] == 1.0)
