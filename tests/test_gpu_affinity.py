"""The affinity builders upstream of the t-SNE pass on the device (mg_perplexity_rows, mg_symmetrize_knn, the fuzzy
graph of the fit) against the oracle (oracle/affinity_oracle.py), and the whole device P builder against the oracle's
chain on the same neighbour lists."""
import numpy as np
import pytest

from magnet import capi
import affinity_oracle as AO

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


@pytest.mark.parametrize("k", [1, 7, 16, 64, 90, 128, 200, 512])
def test_perplexity_rows_match_the_oracle(k):
    """Each row's bisection on the device against the oracle's (the same steps; sums in another order): the same
    precision beta for every row but rows whose entropy sits within rounding of a branch, p within 1e-12 relative
    elsewhere; every row's entropy within tol of log(perplexity) where that is reachable."""
    rng = np.random.default_rng(k)
    n = 3000
    d2 = np.sort(rng.gamma(2.0, 3.0, (n, k)), axis=1).astype(np.float32)
    d2[0] = 0.0
    perp = min(30.0, max(1.5, k / 3.0))
    P, beta = capi.perplexity_rows(dev(d2), perp, 1e-5, 100, with_beta=True)
    P, beta = host(P), host(beta)
    Po, bo = AO.perplexity_rows(d2, perp, 1e-5, 100)
    same = beta == bo
    pos = same[:, None] & (Po > 0)
    print(f"k={k}: beta equal on {same.mean():.5f} of the rows; max p rel diff {np.abs(P[pos] / Po[pos] - 1).max():.2e}")
    assert same.mean() > 0.999
    assert np.allclose(P[same], Po[same], rtol=1e-12, atol=0)
    assert np.allclose(P.sum(1), 1.0, atol=1e-12)
    if k > 1:
        H = -(P * np.log(np.maximum(P, 1e-300))).sum(1)
        assert np.abs(H[1:] - np.log(perp)).max() < 1e-5


def test_symmetrize_equals_the_oracle_bit_for_bit():
    """mg_symmetrize_knn, both modes, on lists with a repeated neighbour, a self entry and zero values: indptr,
    indices and values equal the oracle's; tiny graphs (n = 2, k = 1) and a capacity too small raise."""
    import torch
    rng = np.random.default_rng(7)
    for n, k in ((2, 1), (50, 3), (4000, 15), (20000, 90)):
        idx = np.array([rng.choice(n, k, replace=False) for _ in range(n)], np.int32) if n < 5000 else \
            rng.integers(0, n, (n, k)).astype(np.int32)
        val = rng.random((n, k))
        if n >= 50:
            val[3, 1] = 0.0
            idx[9, 1] = idx[9, 0]
            idx[10, 0] = 10
        for mode, param in ((0, None), (1, 1.0), (1, 0.25)):
            ip, ix, dv = capi.symmetrize_knn(dev(idx), dev(val), mode, param)
            S = AO.symmetrize_knn(idx, val, mode, param)
            assert np.array_equal(host(ip), S.indptr) and np.array_equal(host(ix), S.indices), (n, k, mode)
            assert np.array_equal(host(dv), S.data), (n, k, mode)
    bad = np.zeros((10, 2), np.int32)
    bad[3, 1] = 10
    with pytest.raises(RuntimeError):
        capi.symmetrize_knn(dev(bad), dev(np.ones((10, 2))), 0)


def test_device_p_builder_equals_the_oracle_chain():
    """The device P builder against the oracle's calibration and symmetrisation of the same neighbour lists (the
    device's exact search and its NN-descent): the same sparsity; values within one float32 rounding (the
    calibrations' sums run in different orders) but for at most 0.1% of them (rows whose bisection took another
    branch within rounding), and all within 1e-5 relative; P sums to 1 and equals its transpose exactly;
    perplexity_affinities is the composition of the search and affinities_from_knn."""
    import torch
    from magnet.affinity import perplexity_affinities, affinities_from_knn
    rng = np.random.default_rng(3)
    for n, d, how in ((5000, 20, "exact"), (3000, 100, "exact"), (6000, 40, "nndescent")):
        X = (rng.normal(0, 4, (10, d))[rng.integers(0, 10, n)] + rng.normal(0, 1, (n, d))).astype(np.float32)
        P = perplexity_affinities(X, 30.0, knn=how)
        Xt = torch.as_tensor(X, device="cuda")
        if how == "exact":
            idx, d2 = capi.knn_exact_graph(Xt, 90)
        else:
            idx, d2 = capi.nn_descent(Xt, 90, 50, 10, 0.001, int(max(2, min(8, round(np.sqrt(n) / 40)))), 30, 0)
        P2 = affinities_from_knn(idx, d2, 30.0)
        assert np.array_equal(P.indptr, P2.indptr) and np.array_equal(P.indices, P2.indices) and np.array_equal(P.data, P2.data)
        Pc, _ = AO.perplexity_rows(host(d2), 30.0)
        S = AO.symmetrize_knn(host(idx), Pc, 0)
        assert np.array_equal(P.indptr, S.indptr) and np.array_equal(P.indices, S.indices)
        ulp = np.spacing(np.abs(S.data))
        print(f"n={n} d={d} {how}: nnz {P.nnz}, values equal {np.mean(P.data == S.data):.5f}, max |diff| / ulp "
              f"{(np.abs(P.data - S.data) / ulp).max():.1f}")
        assert np.mean(np.abs(P.data - S.data) > ulp) < 1e-3
        assert np.allclose(P.data, S.data, rtol=1e-5, atol=0)
        assert abs(float(P.data.astype(np.float64).sum()) - 1.0) < 1e-5
        assert (P - P.T).count_nonzero() == 0


def test_fuzzy_graph_on_the_device_is_the_fuzzy_union():
    """MAGNET's fit graph (magnet.graph.fuzzy_graph: device memberships, device union) against the oracle's union of
    its own memberships (exact) and of the oracle's memberships (to the memberships' bound, 1e-4 relative: their
    float32 bisections round differently)."""
    from magnet.graph import fuzzy_graph, memberships
    rng = np.random.default_rng(2)
    n, k = 3000, 14
    ind = np.array([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)])
    dist = np.sort(rng.uniform(0.5, 2.0, (n, k)).astype(np.float32), axis=1)
    P = fuzzy_graph(ind, dist, k + 1)
    v = memberships(dist, k + 1)
    S = AO.symmetrize_knn(ind, v.astype(np.float64), 1, 1.0)
    assert np.array_equal(P.indptr, S.indptr) and np.array_equal(P.indices, S.indices) and np.array_equal(P.data, S.data)
    vo = AO.smooth_knn(dist, k + 1)
    So = AO.symmetrize_knn(ind, vo.astype(np.float64), 1, 1.0)
    assert np.array_equal(P.indices, So.indices) and np.allclose(P.data, So.data, rtol=1e-4, atol=1e-6)
