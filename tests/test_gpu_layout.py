"""The sampled layout pass on the device (mg_optimize_layout, the reference's optimize_layout in its owner mode):
against the oracle epoch by epoch, deterministic, fixed nodes fixed; the reference's own tests of the owner mode, of
loss="hybrid" and of transform() after the full-batch modes; and the placement's quality axes."""
import numpy as np
import pytest
import scipy.sparse as sp

from magnet import capi
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()
    capi.use_torch_stream()
    return torch


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def blob_graph(n=3000, k=15, d=20, seed=0):
    """6 blobs and a symmetric exp(-d / mean d) k-NN graph (edges sorted by head); the graph of the reference's
    test_owner_parallel_mode (tests/test_magnet.py:388-392)"""
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
    return X, y, A


def test_device_epochs_equal_the_oracle():
    """30 epochs with negative samples, mid-range pairs and their reverse, learning rate and spring weight moving, 300
    fixed nodes: the device layout and sampling state equal the oracle's bit for bit (the same random streams, every
    operation rounded on its own; exponents b = 1 keep the power function, whose last bits differ between libraries,
    out of it)."""
    from magnet.layout import edge_schedule, reverse_pairs
    X, y, A = blob_graph(3000)
    n = A.shape[0]
    head, tail, eps = edge_schedule(A, 50, 5.0)
    pairs = np.random.default_rng(2).integers(0, n, (n, 5)).astype(np.int32)
    rp, ri = reverse_pairs(pairs, n)
    Y0 = np.random.default_rng(1).uniform(-10, 10, (n, 2)).astype(np.float32)
    kw = dict(a=1.78, b=1.0, rep_a=1.58, rep_b=1.0, negative_sample_rate=5.0, repulsion_strength=2.0, mid_rate=0.25,
              clip_mode=0, n_fixed=300, seed=12345)
    for clip in (0, 1):
        kw["clip_mode"] = clip
        Yo, eo, eno = Y0.copy(), eps.copy(), (eps / 5.0).astype(np.float32)
        O.layout_epochs(Yo, head, tail, eps, eo, eno, 0, 30, 1.0, 0.4, 0.2, 0.05, pairs=pairs, rev_indptr=rp, rev_indices=ri,
                        **kw)
        Yd, ed, end = dev(Y0), dev(eps), dev((eps / 5.0).astype(np.float32))
        capi.optimize_layout(Yd, dev(head), dev(tail), dev(eps), ed, end, 0, 30, 1.0, 0.4, 0.2, 0.05, pairs=dev(pairs),
                             rev_indptr=dev(rp), rev_indices=dev(ri), **kw)
        Yg = host(Yd)
        print(f"clip {clip}: device vs oracle after 30 epochs, max |dY| = {np.abs(Yg - Yo).max():.2e}")
        assert np.array_equal(Yg, Yo) and np.array_equal(host(ed), eo) and np.array_equal(host(end), eno)
        assert np.array_equal(Yg[:300], Y0[:300]) and np.abs(Yg[300:] - Y0[300:]).max() > 1.0


def test_owner_mode_layout_separates_blobs():
    """tests/test_magnet.py:388-398 of the reference: the owner mode on the blobs' k-NN graph at the default settings:
    10-NN accuracy of the layout > 0.95. The reference starts its 40 epochs from a spectral layout of the graph; from
    a random start the run takes 200 epochs."""
    from sklearn.neighbors import KNeighborsClassifier
    from magnet.layout import edge_schedule, find_ab_params, optimize_layout_sgd
    X, y, A = blob_graph(3000)
    a, b = find_ab_params(1.0, 0.05, 0.8)
    ra, rb = find_ab_params(1.0, 0.05, 1.0)
    head, tail, eps = edge_schedule(A, 200, 5.0)
    Y0 = np.random.default_rng(0).uniform(-10, 10, (3000, 2)).astype(np.float32)
    Y = optimize_layout_sgd(Y0, head, tail, eps, 200, a, b, ra, rb, repulsion_strength=2.0, seed=7)
    acc = KNeighborsClassifier(10).fit(Y, y).score(Y, y)
    print("owner-mode layout 10-NN accuracy", acc)
    assert np.isfinite(Y).all() and acc > 0.95


def test_pybind_optimize_layout_keeps_the_reference_signature():
    """_magnet.optimize_layout with the reference's arguments (src/bindings.cpp:11-23): owner mode and hogwild-with-
    fixed-nodes run on the device and update Y and the epoch arrays in place; the options this path lacks raise."""
    import _magnet
    from magnet.layout import edge_schedule
    X, y, A = blob_graph(1500)
    head, tail, eps = edge_schedule(A, 30, 5.0)
    Y = np.random.default_rng(0).uniform(-10, 10, (1500, 2)).astype(np.float32)
    Y0 = Y.copy()
    eons, eonns = eps.copy(), (eps / 5.0).astype(np.float32)
    z2, zi = np.zeros((0, 2), np.float32), np.zeros(0, np.int32)
    args = [head, tail, eps, eons, eonns, zi, zi, 0, 30, 1.0, 0.0, 0.0, 0.0, 1.78, 0.6, 1.58, 0.9, 5.0, 0, 1.0, 2.0, 0,
            0.25, 0, 0, 0, z2, z2, zi, 0.9, 0.999, zi, np.zeros((0, 1), np.int32), zi, zi, 0.0, 0, np.zeros(0, np.float32),
            7, 1, 1]
    _magnet.optimize_layout(Y, *args)
    assert np.abs(Y - Y0).max() > 1.0 and np.all(eons > eps)
    for pos, bad in ((18, 1), (36, 2), (23, 1)):   # loss 1 (InfoNCE), kernel 2 (sech), optimizer 1 (Adam)
        a = list(args)
        a[pos] = bad
        with pytest.raises(RuntimeError):
            _magnet.optimize_layout(Y.copy(), *a)


def test_hybrid_loss():
    """tests/test_magnet.py:370-384 of the reference: loss="hybrid" (hybrid_iter 60) on 3000 blobs: finite, mid-range
    weight 0.2 (the SGD pass of the hybrid) against 0.1 for loss="umap", a layout other than the SGD pass's, at the
    exact pass's scale 31 sqrt(n / 70000) within a factor 2; n_components=3 refused."""
    from sklearn.datasets import make_blobs
    from magnet import MAGNET
    X, y = make_blobs(n_samples=3000, centers=6, n_features=20, random_state=0)
    X = X.astype(np.float32)
    M = MAGNET(loss="hybrid", hybrid_iter=60, random_state=0).fit(X)
    Y = M.embedding_
    assert Y.shape == (3000, 2) and np.isfinite(Y).all()
    assert M.mid_weight_ == 0.2
    M_sgd = MAGNET(loss="umap", random_state=0).fit(X)
    assert M_sgd.mid_weight_ == 0.1 and MAGNET(mid_weight=0.3).mid_weight == 0.3
    assert not np.allclose(Y, M_sgd.embedding_)
    rule = 31.0 * np.sqrt(3000 / 70000.0)
    print("hybrid layout std", Y[:, 0].std(), "rule", rule)
    assert 0.5 * rule < Y[:, 0].std() < 2.0 * rule
    from sklearn.neighbors import KNeighborsClassifier
    for E in (Y, M_sgd.embedding_):
        assert KNeighborsClassifier(10).fit(E, y).score(E, y) > 0.95
    with pytest.raises(ValueError):
        MAGNET(loss="hybrid", n_components=3, random_state=0).fit(X)


@pytest.mark.parametrize("loss", ["tsne", "hybrid", "umap"])
def test_sgd_transform_after_fit(loss):
    """tests/test_magnet.py:440-447 of the reference with its placement (transform_mode "sgd"): 10-NN accuracy of the
    placed rows > 0.95; the quality axes of the reference's k48 study (benchmarks/theory/k48_transform.py:1-11) are
    printed: nn15 (overlap of the 15 input-space and layout-space nearest fitted points) and self (re-placed fitted
    points' median distance to their own position, in 1-NN spacings)."""
    from sklearn.neighbors import KNeighborsClassifier, NearestNeighbors
    from magnet import MAGNET
    rng = np.random.default_rng(5)
    yy = rng.integers(0, 5, 2000)
    X = (rng.normal(0, 10, (5, 10))[yy] + rng.normal(0, 1, (2000, 10))).astype(np.float32)
    m = MAGNET(loss=loss, random_state=0, transform_mode="sgd", hybrid_iter=100).fit(X[:1700])
    Yq = m.transform(X[1700:])
    assert Yq.shape == (300, 2) and np.isfinite(Yq).all()
    acc = (KNeighborsClassifier(10).fit(m.embedding_, yy[:1700]).predict(Yq) == yy[1700:]).mean()
    nn_in = NearestNeighbors(n_neighbors=15).fit(X[:1700]).kneighbors(X[1700:], return_distance=False)
    nn_out = NearestNeighbors(n_neighbors=15).fit(m.embedding_).kneighbors(Yq, return_distance=False)
    nn15 = float(np.mean([len(set(a) & set(b)) / 15 for a, b in zip(nn_in, nn_out)]))
    Ys = m.transform(X[:300])
    d1 = NearestNeighbors(n_neighbors=2).fit(m.embedding_).kneighbors(m.embedding_)[0][:, 1]
    selfd = float(np.median(np.linalg.norm(Ys - m.embedding_[:300], axis=1)) / np.median(d1))
    print(f"{loss} sgd transform: knn_acc {acc:.3f} nn15 {nn15:.3f} self {selfd:.2f}")
    assert acc > 0.95
