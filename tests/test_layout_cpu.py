"""Host side of the sampled pass: the kernel fit, the sampling frequencies and the edge schedule, pinned by what
they must satisfy."""
import numpy as np
import scipy.sparse as sp

from magnet.layout import find_ab_params, sampling_frequencies, edge_schedule, reverse_pairs


def test_find_ab_params_matches_the_published_umap_fit():
    """min_dist 0.1, spread 1: the (a, b) UMAP publishes for its default curve (1.577, 0.8951); with b given, only a
    moves, and the fit reproduces the curve's value 1/2 where it should."""
    a, b = find_ab_params(1.0, 0.1)
    assert abs(a - 1.577) < 2e-3 and abs(b - 0.8951) < 1e-3
    a6, b6 = find_ab_params(1.0, 0.05, 0.6)
    assert b6 == 0.6 and 1.5 < a6 < 2.1
    x = np.linspace(0, 3, 300)
    target = np.where(x < 0.05, 1.0, np.exp(-(x - 0.05)))
    fit = 1 / (1 + a6 * x ** 1.2)
    assert np.sqrt(np.mean((fit - target) ** 2)) < 0.07   # the best b = 0.6 curve: RMS 0.056


def test_sampling_frequencies_sum_to_the_target():
    """min(1, s w) with the sum equal to the target, against a bisection on s; proportional below saturation."""
    rng = np.random.default_rng(0)
    for w, target in ((rng.random(1000), 300.0), (rng.random(2000) ** 6, 150.0), (np.ones(50), 20.0), (rng.random(10), 50.0)):
        f = sampling_frequencies(w, target)
        if len(w) <= target:
            assert np.all(f == 1)
            continue
        assert abs(f.sum() - target) < 1e-8 * target and f.max() <= 1.0
        lo, hi = 0.0, 1e12
        for _ in range(200):
            mid = (lo + hi) / 2
            lo, hi = (mid, hi) if np.minimum(1, mid * w).sum() < target else (lo, mid)
        assert np.allclose(f, np.minimum(1, lo * w), rtol=1e-9, atol=1e-12)


def test_edge_schedule_and_reverse_pairs():
    rng = np.random.default_rng(1)
    n = 400
    A = sp.random(n, n, density=0.05, random_state=1, format="csr")
    A = (A + A.T).tocsr()
    head, tail, eps = edge_schedule(A, 200, 5.0)
    assert np.all(np.diff(head) >= 0) and np.all(eps >= 1.0) and np.all(eps <= 200.0 + 1e-3)
    assert abs((1.0 / eps).sum() - 5.0 * n) / (5.0 * n) < 0.01   # samples per epoch (minus the dropped edges)
    pairs = rng.integers(0, n, (n, 3)).astype(np.int32)
    ip, ix = reverse_pairs(pairs, n)
    for u in (0, 17, 399):
        choosers = [i for i in range(n) for j in range(3) if pairs[i, j] == u]
        assert sorted(ix[ip[u]:ip[u + 1]].tolist()) == sorted(choosers)
