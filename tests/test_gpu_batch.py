"""The batched form (mg_tsne_optimize_batch): every stage launched once for many instances of uneven sizes (in
groups on several HIP streams and host threads) gives each instance exactly the result of its single-instance
loop (same block bodies, grids, block counts and order of operations), for any grouping."""
import numpy as np
import pytest

from magnet import capi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    capi.lib()
    capi.use_torch_stream()
    return torch


SIZES = (700, 2500, 1200, 4000, 900)
SCHED = dict(n_iter=80, early_iter=30)


def problems():
    from magnet.affinity import perplexity_affinities
    out = []
    for i, n in enumerate(SIZES):
        rng = np.random.default_rng(1000 + i)
        k = int(rng.integers(5, 21))
        c = rng.normal(0, 5, (k, 32))
        X = (c[rng.integers(0, k, n)] + rng.normal(0, 1, (n, 32))).astype(np.float32)
        out.append((perplexity_affinities(X, 30.0), rng.normal(0, 1e-4, (n, 2))))
    return out


_P = []


def cached():
    if not _P:
        _P.extend(problems())
    return _P


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("n_streams,n_threads", [(1, 1), (3, 1), (4, 2), (5, 5)])
def test_batch_equals_single_runs(precision, n_streams, n_threads):
    from magnet.batch import device_instances
    probs = cached()
    Ps, Ys = [p for p, _ in probs], [y for _, y in probs]
    batch = device_instances(Ps, Ys, precision)
    sq = capi.optimize_batch(batch, n_streams=n_streams, n_threads=n_threads, sumq=True, **SCHED)
    for b, single in enumerate(device_instances(Ps, Ys, precision)):
        ip, ix, dv, Y, U, G = single
        s1 = capi.optimize(ip, ix, dv, Y, U, G, 0, SCHED["n_iter"], **SCHED)
        for x, y in zip(batch[b][3:], (Y, U, G)):
            assert np.array_equal(host(x), host(y)), (b, SIZES[b])
        assert np.array_equal(sq[b], s1), b


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_equals_single_runs_with_options(precision):
    """The same equality with dof 0.6, five interpolation nodes, a 1.1 box and n_boxes clipped at 16 from spread
    starting layouts (every instance's spectra recomputed each iteration), in two groups."""
    from magnet.batch import device_instances
    probs = cached()
    Ps, Ys = [p for p, _ in probs], [y * 5e4 for _, y in probs]
    opts = dict(n_iter=40, early_iter=15, dof=0.6, n_interp=5, box_width=1.1, min_boxes=8, max_boxes=16)
    batch = device_instances(Ps, Ys, precision)
    sq = capi.optimize_batch(batch, n_streams=2, n_threads=1, sumq=True, **opts)
    for b, single in enumerate(device_instances(Ps, Ys, precision)):
        ip, ix, dv, Y, U, G = single
        s1 = capi.optimize(ip, ix, dv, Y, U, G, 0, opts["n_iter"], **opts)
        for x, y in zip(batch[b][3:], (Y, U, G)):
            assert np.array_equal(host(x), host(y)), (b, SIZES[b])
        assert np.array_equal(sq[b], s1), b


def test_optimize_tsne_batch_matches_optimize_tsne():
    from magnet.batch import optimize_tsne_batch
    from magnet.tsne import optimize_tsne
    probs = cached()[:3]
    out = optimize_tsne_batch([p for p, _ in probs], [y for _, y in probs], n_iter=60, early_iter=20, n_streams="auto")
    for (P, Y0), Yb in zip(probs, out):
        Ys = optimize_tsne(P, Y0.astype(np.float32), n_iter=60, early_iter=20)
        assert Yb.dtype == np.float32 and np.array_equal(Yb, Ys)


def test_batch_rejects_bad_instances():
    from magnet.batch import device_instances
    P, Y = cached()[0]
    inst = device_instances([P], [Y])
    ip, ix, dv, Yd, U, G = inst[0]
    with pytest.raises(ValueError):
        capi.optimize_batch([(ip, ix, dv, Yd, U[:-1], G)], **SCHED)
    with pytest.raises(capi.MgError):
        capi.optimize_batch(inst, n_streams=0, **SCHED)


def test_more_than_64_instances_and_mixed_fft_sizes():
    """70 small instances (block ranges split over several launches per stage), spans far apart so that several
    FFT sizes are in use at once; every instance equals its single run."""
    from magnet.affinity import perplexity_affinities
    from magnet.batch import device_instances
    Ps, Ys = [], []
    for i in range(70):
        rng = np.random.default_rng(50 + i)
        n = int(rng.integers(150, 420))
        X = (rng.normal(0, 6, (4, 8))[rng.integers(0, 4, n)] + rng.normal(0, 1, (n, 8))).astype(np.float32)
        Ps.append(perplexity_affinities(X, 10.0))
        Ys.append(rng.normal(0, 1e-4, (n, 2)) * (1 + 2e5 * (i % 3)))   # layouts of three spans
    kw = dict(n_iter=30, early_iter=10)
    batch = device_instances(Ps, Ys, "fp32")
    capi.optimize_batch(batch, n_streams=2, n_threads=1, **kw)
    for b, (ip, ix, dv, Y, U, G) in enumerate(device_instances(Ps, Ys, "fp32")):
        capi.optimize(ip, ix, dv, Y, U, G, 0, kw["n_iter"], **kw)
        assert np.array_equal(host(batch[b][3]), host(Y)), b


@pytest.mark.parametrize("graph,spread", [(1, False), (0, False), (1, True), (0, True)])
def test_device_driven_batch_with_stalls_and_replans(graph, spread):
    """The batch decides every grid on the device and replays its captured iteration (batch_graph 1) or issues the
    same launches one by one (0). Reading the states every 2 iterations, with windows that hold no smaller FFT size
    (every shrink stalls its instance until the next plan), many plans and stalls happen; each instance still equals
    its single run bit for bit, and the counters add up: every schedule iteration issued, stalls only as extra ones."""
    from magnet.batch import device_instances
    probs = cached()
    Ps = [p for p, _ in probs]
    # spread: starting layouts of span about 200 that contract under the exaggerated attraction, so grids shrink
    Ys = [np.random.default_rng(7 + b).normal(0, 30, y.shape) if spread else y for b, (_, y) in enumerate(probs)]
    old = {k: capi.tune(k) for k in ("batch_graph", "batch_chunk", "batch_window_down")}
    try:
        capi.tune("batch_graph", graph); capi.tune("batch_chunk", 2); capi.tune("batch_window_down", 0)
        batch = device_instances(Ps, Ys, "fp32")
        capi.optimize_batch(batch, n_streams=1, n_threads=1, **SCHED)
        st = capi.batch_stats()
    finally:
        for k, v in old.items():
            capi.tune(k, v)
    print(st)
    assert st["plans"] >= 2 and st["checks"] >= SCHED["n_iter"] // 2
    assert st["iterations"] >= SCHED["n_iter"] and st["replays"] == (st["iterations"] if graph else 0)
    assert st["launches"] > 0
    if spread:
        assert st["stalls"] > 0 and st["iterations"] > SCHED["n_iter"]
    for b, (ip, ix, dv, Y, U, G) in enumerate(device_instances(Ps, Ys, "fp32")):
        capi.optimize(ip, ix, dv, Y, U, G, 0, SCHED["n_iter"], **SCHED)
        for x, y in zip(batch[b][3:], (Y, U, G)):
            assert np.array_equal(host(x), host(y)), (b, SIZES[b])
