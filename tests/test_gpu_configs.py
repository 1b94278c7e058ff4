"""The five measured configurations under test, each at its full size:

- small (10,000 x 50, 500 iterations): 15-NN preservation of the GPU layouts (FP32 and FP64) within 0.005 of the
  layout the compiled reference computes (tests/golden/c1_layout.npz, make_c1_layout.py);
- MNIST-shaped (70,000 x 784, 750 iterations) and scale (1,000,000 x 50, 1,000 iterations): at states across the
  schedule -- the early phase, the switch, the clipped-grid stretch, the late phase -- one FP64 GPU iteration equals
  the oracle's from the same state (relative L2 <= 1e-10), and the FP32 gradient is within 1e-4 of the FP64 one on
  the same layout;
- batched (64 instances of 2,000-50,000 points, 500 iterations): every instance's batched result equals its single
  run bit for bit;
- serving (200,000-point model): a burst of requests of 1-4,096 rows gives each request its result alone.
"""
import os
import sys

import numpy as np
import pytest

from magnet import capi
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


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


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------------------------------------- small: 15-NN gate
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_small_config_15nn_preservation_matches_reference(prec):
    """benchmarks/metrics.py:53-59's 15-NN preservation of the final layout of the 500-iteration schedule on the
    10k mixture, from the same seeded start and the same affinities, within 0.005 of the reference's."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_c1_layout import c1_data, preservation15, N_ITER, EARLY
    g = np.load(os.path.join(ROOT, "tests", "golden", "c1_layout.npz"))
    X, lab, P, Y0 = c1_data()
    assert P.nnz == int(g["nnz"]) and abs(P.data.astype(np.float64).sum() / float(g["data_sum"]) - 1) < 1e-6
    dt = np.float32 if prec == "fp32" else np.float64
    data = (P.data / P.data.sum()).astype(dt)
    Y, U, G = dev(Y0.astype(dt)), dev(np.zeros_like(Y0, dt)), dev(np.ones_like(Y0, dt))
    capi.optimize(dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32)), dev(data), Y, U, G, 0, N_ITER,
                  n_iter=N_ITER, early_iter=EARLY)
    pres = preservation15(X, host(Y))
    ref = float(g["pres15"])
    print(f"{prec}: 15-NN preservation {pres:.4f}, reference {ref:.4f}")
    assert abs(pres - ref) <= 0.005


# ---------------------------------------------------------------------------- MNIST-shaped and scale: per iteration
def _trajectory(P, n_iter, early, every=25):
    """The FP32 schedule from the seeded start in runs of `every` iterations: {t: (Y, U, G) host copies} at every
    run start, and each start's grid box count."""
    import torch
    Y0 = np.random.default_rng(1).normal(0, 1e-4, (P.shape[0], 2))
    ip, ix, d = dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32)), dev(P.data.astype(np.float32))
    sess = capi.Session(ip, ix, d, n_iter=n_iter, early_iter=early)
    Y, U, G = dev(Y0.astype(np.float32)), dev(np.zeros_like(Y0, np.float32)), dev(np.ones_like(Y0, np.float32))
    snaps, boxes = {}, {}
    from magnet.tsne import grid_params
    for t in range(0, n_iter, every):
        snaps[t] = (host(Y).copy(), host(U).copy(), host(G).copy())
        boxes[t] = grid_params(capi.bbox(Y))[3]
        sess.run(Y, U, G, t, min(n_iter, t + every))
    sess.close()
    torch.cuda.synchronize()
    return snaps, boxes


def _pick(boxes, n_iter, early):
    """Checkpoints: start, middle of the early phase, its last iteration's run, the switch, the first and a middle
    clipped-grid run (1000 boxes; else the two widest grids), and the last run."""
    ts = sorted(boxes)
    clipped = [t for t in ts if boxes[t] >= 1000]
    if len(clipped) < 2:
        clipped = sorted(ts, key=lambda t: -boxes[t])[:2]
    mid_early = max(t for t in ts if t <= early // 2)
    last_early = max(t for t in ts if t < early)
    switch = min(t for t in ts if t >= early)
    return {"start": 0, "early": mid_early, "last_early": last_early, "switch": switch,
            "clipped_first": min(clipped), "clipped_mid": sorted(clipped)[len(clipped) // 2], "late": ts[-1]}


def _check_state(P, state, t, n_iter, early, label):
    """One FP64 GPU iteration from the (widened) state against the oracle's, and the FP32 gradient against FP64."""
    from magnet.tsne import grid_params
    Y, U, G = (a.astype(np.float64) for a in state)
    data64 = P.data.astype(np.float64)
    ip, ix = dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32))
    Yd, Ud, Gd = dev(Y), dev(U), dev(G)
    sq = capi.optimize(ip, ix, dev(data64), Yd, Ud, Gd, t, t + 1, n_iter=n_iter, early_iter=early)
    Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
    sqo = O.optimize(P.indptr, P.indices, data64, Yo, Uo, Go, t, t + 1, n_iter=n_iter, early_iter=early)
    Yg, Ug = host(Yd), host(Ud)
    e = max(rel(Yg - Y, Yo - Y), rel(Ug, Uo), abs(sq[0] - sqo[0]) / abs(sqo[0]))
    # FP32 gradient rho A - F / sumQ against FP64 on the same float32 layout
    Y32 = state[0].astype(np.float32)
    rho = 12.0 if t < early else 1.0
    xmin, ymin, width, nb = grid_params(capi.bbox(dev(Y32)))
    grads = []
    for dt, prec in ((np.float32, capi.MG_F32), (np.float64, capi.MG_F64)):
        Yx = dev(Y32.astype(dt))
        A = host(capi.attraction(ip, ix, dev(P.data.astype(dt)), Yx)).astype(np.float64)
        conv = capi.FFTConv(prec, 2 * nb * 3, 3)
        conv.set_tsne(width)
        s, F = conv.compute(Yx, xmin, ymin, width, nb)
        grads.append(rho * A - host(F).astype(np.float64) / s)
    eg = rel(grads[0], grads[1])
    print(f"{label} t={t} n_boxes={nb}: FP64 step vs oracle {e:.2e}, FP32 gradient vs FP64 {eg:.2e}", flush=True)
    assert e <= 1e-10, (label, t, e)
    assert eg <= 1e-4, (label, t, eg)


_CFG = {}


def _config(name):
    if name not in _CFG:
        from bench import make_data, WORKLOADS
        from magnet.affinity import perplexity_affinities
        W = WORKLOADS[name]
        X, _ = make_data(name)
        P = perplexity_affinities(X, 30.0)
        del X
        snaps, boxes = _trajectory(P, W["n_iter"], W["early"])
        _CFG[name] = (P, W, snaps, _pick(boxes, W["n_iter"], W["early"]), boxes)
    return _CFG[name]


ROLES = ["start", "early", "last_early", "switch", "clipped_first", "clipped_mid", "late"]


@pytest.mark.parametrize("role", ROLES)
def test_mnist_shaped_per_iteration(role):
    """70,000 x 784 (10 prototypes, quantised), perplexity 30, 750 iterations (250 early)."""
    P, W, snaps, picks, boxes = _config("mnist70k")
    t = picks[role]
    _check_state(P, snaps[t], t, W["n_iter"], W["early"], f"mnist70k {role}")


@pytest.mark.parametrize("role", ["start", "switch", "clipped_first", "late"])
def test_scale_1m_per_iteration(role):
    """1,000,000 x 50 (100 clusters, Dirichlet sizes), perplexity 30, 1,000 iterations (250 early)."""
    P, W, snaps, picks, boxes = _config("gmm1m")
    t = picks[role]
    _check_state(P, snaps[t], t, W["n_iter"], W["early"], f"gmm1m {role}")


# ------------------------------------------------------------------------------------------------ batched
def test_batched_64_instances_equal_single_runs():
    """The 64-instance configuration (n log-uniform in [2000, 50000], 500 iterations, 166 early), FP32: every
    instance's state after the batched run equals its own mg_tsne_optimize run bit for bit."""
    import torch
    from bench import batched_instances, BATCHED
    from magnet.batch import device_instances
    Ps, Ys = batched_instances()
    kw = dict(n_iter=BATCHED["n_iter"], early_iter=BATCHED["early"])
    single = device_instances(Ps, Ys, "fp32")
    for ip, ix, dv, Y, U, G in single:
        capi.optimize(ip, ix, dv, Y, U, G, 0, BATCHED["n_iter"], **kw)
    for n_streams in (1, 4):
        bat = device_instances(Ps, Ys, "fp32")
        capi.optimize_batch(bat, n_streams=n_streams, n_threads=1, **kw)
        torch.cuda.synchronize()
        bad = [i for i, (a, b) in enumerate(zip(single, bat))
               if not (torch.equal(a[3], b[3]) and torch.equal(a[4], b[4]) and torch.equal(a[5], b[5]))]
        print(f"batched, {n_streams} stream(s): {len(bat)} instances, {sum(P.shape[0] for P in Ps)} points; differing: {bad}")
        assert not bad


# ------------------------------------------------------------------------------------------------ serving
def test_serving_200k_burst_same_alone_and_in_the_mix():
    """The serving configuration's model (200,000 x 50, 50 clusters) resident; 40 requests of 1-4096 rows
    (log-uniform) submitted at once to the continuous-batching server with three workers: each result equals the
    request's placement alone (to 1e-5 of the layout scale)."""
    from bench import SERVING
    from magnet import MAGNET, TransformServer
    S = SERVING
    rng = np.random.default_rng(0)
    k, d = S["clusters"], S["d"]
    ntot = S["n_train"] + S["pool"]
    lab = rng.integers(0, k, ntot)
    X = (rng.normal(0, 4, (k, d))[lab] + rng.normal(0, 1, (ntot, d))).astype(np.float32)
    Xtr, pool = X[:S["n_train"]], X[S["n_train"]:]
    m = MAGNET(loss="tsne", random_state=0, transform_epochs=S["n_iter"]).fit(Xtr)
    r = np.random.default_rng(5)
    sizes = [int(round(np.exp(r.uniform(0, np.log(4096))))) for _ in range(40)]
    sizes[:2] = [1, 4096]
    reqs_q = [pool[r.integers(0, len(pool), s)] for s in sizes]
    alone = [m.transform(q) for q in reqs_q]
    srv = TransformServer(m, n_workers=3).start()
    try:
        reqs = [srv.submit(q) for q in reqs_q]
        outs = [q.wait(300) for q in reqs]
    finally:
        srv.stop()
    scale = float(np.abs(m.embedding_).max())
    worst = max(float(np.abs(a - o).max()) for a, o in zip(alone, outs))
    print(f"serving burst: {len(sizes)} requests, {sum(sizes)} rows, {len(srv.batches)} batches; worst difference "
          f"{worst:.2e} of layout scale {scale:.1f}")
    for a, o in zip(alone, outs):
        assert o.shape == a.shape and np.allclose(a, o, rtol=0, atol=1e-5 * scale)
