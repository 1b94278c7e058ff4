"""GPU parity of the whole optimiser loop (mg_tsne_optimize) against the oracle.

FP64: per iteration, from the same state, the GPU step equals the oracle step to relative L2 <= 1e-10
(the stated bar of the plain FP64 path). FP32: the gradient of the FP32 path against the FP64 one on
the same layouts (relative L2 <= 1e-4), determinism, resume, callbacks and the reference's quality test."""
import ctypes as C

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


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixture(n, d, k, seed, sep=5.0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, sep, (k, d))
    lab = rng.integers(0, k, n)
    return (c[lab] + rng.normal(0, 1, (n, d))).astype(np.float32), lab


_cache = {}


def problem(n=10_000, d=50, k=10, seed=0, perplexity=30.0):
    key = (n, d, k, seed, perplexity)
    if key not in _cache:
        from magnet.affinity import perplexity_affinities
        X, lab = mixture(n, d, k, seed)
        P = perplexity_affinities(X, perplexity)
        Y0 = np.random.default_rng(seed + 1).normal(0, 1e-4, (n, 2))
        _cache[key] = (P, Y0, lab)
    return _cache[key]


class State:
    def __init__(self, P, Y0, dtype):
        self.ip, self.ix = dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32))
        self.data = dev(P.data.astype(dtype))
        self.Y, self.U, self.G = dev(Y0.astype(dtype)), dev(np.zeros_like(Y0, dtype)), dev(np.ones_like(Y0, dtype))

    def run(self, it0, it1, **kw):
        return capi.optimize(self.ip, self.ix, self.data, self.Y, self.U, self.G, it0, it1, **kw)

    def snapshot(self):
        return host(self.Y).copy(), host(self.U).copy(), host(self.G).copy()


SCHED = dict(n_iter=500, early_iter=166)


def test_fp64_per_iteration_parity():
    """config C1's schedule (10k x 50, 10 clusters, perplexity 30, 500 iterations, early 166), FP64:
    at checkpoints across both phases the GPU iteration from state t equals the oracle's (1e-10)."""
    P, Y0, _ = problem()
    data = P.data.astype(np.float64)
    st = State(P, Y0, np.float64)
    it = 0
    worst = 0.0
    for t in (0, 1, 5, 60, 165, 166, 300, 499):
        if t > it:
            st.run(it, t, **SCHED)
        Y, U, G = st.snapshot()
        sq = st.run(t, t + 1, **SCHED)
        it = t + 1
        Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
        sqo = O.optimize(P.indptr, P.indices, data, Yo, Uo, Go, t, t + 1, **SCHED)
        Yg, Ug, Gg = st.snapshot()
        e = max(rel(Yg - Y, Yo - Y), rel(Ug, Uo), abs(sq[0] - sqo[0]) / sqo[0])
        worst = max(worst, e)
        assert e < 1e-10, (t, e)
        assert np.mean(Gg == Go) > 0.9999, t
    print("worst per-iteration relative error", worst)


def test_fp64_short_trajectory_matches_oracle():
    P, Y0, _ = problem()
    data = P.data.astype(np.float64)
    st = State(P, Y0, np.float64)
    sq = st.run(0, 25, **SCHED)
    Yo, Uo, Go = Y0.copy(), np.zeros_like(Y0), np.ones_like(Y0)
    sqo = O.optimize(P.indptr, P.indices, data, Yo, Uo, Go, 0, 25, **SCHED)
    assert rel(host(st.Y), Yo) < 1e-9 and np.abs(sq / sqo - 1).max() < 1e-10


def test_fp64_nonfused_path_with_mid_pairs():
    P, Y0, _ = problem(4000, 20, 5, 3)
    n = P.shape[0]
    pairs = np.random.default_rng(4).integers(0, n, (n, 5)).astype(np.int32)
    data = P.data.astype(np.float64)
    st = State(P, Y0, np.float64)
    kw = dict(n_iter=200, early_iter=66, pairs=dev(pairs), mid_weight=0.3, mid_phase=0.5)
    st.run(0, 40, **kw)
    Y, U, G = st.snapshot()
    st.run(40, 41, **kw)
    Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
    O.optimize(P.indptr, P.indices, data, Yo, Uo, Go, 40, 41, n_iter=200, early_iter=66, pairs=pairs, mid_weight=0.3,
               mid_phase=0.5)
    assert rel(host(st.Y) - Y, Yo - Y) < 1e-10


def fp32_vs_fp64_gradient(st, P, Y, rho, dof=1.0, ni=3):
    """relative L2 of the FP32 path's gradient against the FP64 path's on the same (float32) layout"""
    Y = Y.astype(np.float32)
    grads = {}
    for dt in (np.float64, np.float32):
        Yd = dev(Y.astype(dt))
        A = host(capi.attraction(st.ip, st.ix, dev(P.data.astype(dt)), Yd, dof))
        xmin, ymin, width, nb = O.grid(Y)
        conv = capi.FFTConv(capi.MG_F64 if dt == np.float64 else capi.MG_F32, nb * max(6, ni), ni)
        conv.set_tsne(width, dof)
        s, F = conv.compute(Yd, xmin, ymin, width, nb)
        grads[dt] = rho * A.astype(np.float64) - host(F).astype(np.float64) / s
    return rel(grads[np.float32], grads[np.float64])


def test_fp32_gradient_against_fp64():
    """FP32 path (bulk FP32, FP64 Z) against FP64 on identical inputs, on the layouts of an FP64 run
    through both phases: relative L2 <= 1e-4 (late in the run the gradient is a small difference of
    attraction and repulsion, which magnifies every rounding)."""
    P, Y0, _ = problem()
    st = State(P, Y0, np.float64)
    it = 0
    for t in (50, 100, 165, 200, 300, 400, 499):
        st.run(it, t, **SCHED); it = t
        e = fp32_vs_fp64_gradient(st, P, host(st.Y), 12.0 if t < 166 else 1.0)
        print(f"iteration {t}: FP32 gradient relative L2 error {e:.2e}")
        assert e < 1e-4, (t, e)


@pytest.mark.parametrize("dof,ni", [(0.6, 3), (2.0, 3), (1.0, 5)])
def test_fp32_gradient_against_fp64_options(dof, ni):
    """The same bound with a heavier and a lighter tailed kernel and with five interpolation nodes per box."""
    P, Y0, _ = problem(2000, 10, 5, seed=3)
    st = State(P, Y0, np.float64)
    kw = dict(n_iter=300, early_iter=100, dof=dof, n_interp=ni, max_boxes=min(1000, 4096 // ni // 4 * 4))
    it = 0
    for t in (60, 150, 299):
        st.run(it, t, **kw); it = t
        e = fp32_vs_fp64_gradient(st, P, host(st.Y), 12.0 if t < 100 else 1.0, dof, ni)
        assert e < 1e-4, (t, e)


def test_resume_and_determinism_fp32():
    P, Y0, _ = problem()
    a = State(P, Y0, np.float32)
    a.run(0, 40, **SCHED)
    b = State(P, Y0, np.float32)
    b.run(0, 17, **SCHED)
    b.run(17, 40, **SCHED)
    assert np.array_equal(host(a.Y), host(b.Y)) and np.array_equal(host(a.G), host(b.G))


def test_callback_cadence_and_stop():
    P, Y0, _ = problem(3000, 10, 5, 5)
    st = State(P, Y0, np.float32)
    seen = []

    def cb(it, d_Y, n, user):
        seen.append(it)
        return 1 if it >= 100 else 0

    st.run(0, 300, n_iter=300, early_iter=100, callback=cb)
    assert seen == [0, 50, 100]


def test_separates_blobs_like_the_reference():
    """tests/test_magnet.py:234-240 of the reference: 10-NN label purity > 0.95 on 5 blobs, d=10."""
    X, lab = mixture(2000, 10, 5, 5, sep=4.0)
    from magnet.affinity import perplexity_affinities
    P = perplexity_affinities(X, 30.0)
    st = State(P, np.random.default_rng(0).normal(0, 1e-4, (2000, 2)), np.float32)
    st.run(0, 300, n_iter=300, early_iter=100)
    Y = host(st.Y)
    assert np.isfinite(Y).all()
    from sklearn.neighbors import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=11).fit(Y).kneighbors(Y, return_distance=False)[:, 1:]
    assert (lab[nn] == lab[:, None]).mean() > 0.95


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_relabelling_changes_no_result(dtype):
    """The locality relabelling (every iteration, every 7, the default, or off) leaves the state, sumQ and the
    layouts handed to the callback bit-identical."""
    P, Y0, _ = problem(6000, 20, 6, 11)
    out = {}
    for every in (-1, 1, 7, 0):
        st = State(P, Y0, dtype)
        seen = {}
        def cb(it, y, n, user):
            import torch
            seen[it] = torch.empty((n, 2), dtype=torch.float32 if dtype == np.float32 else torch.float64, device="cuda")
            capi.lib().mg_memcpy_dtod(C.c_void_p(seen[it].data_ptr()), C.c_void_p(y),
                                      C.c_size_t(seen[it].numel() * seen[it].element_size()))
            return 0
        sq = capi.optimize(st.ip, st.ix, st.data, st.Y, st.U, st.G, 0, 120, n_iter=120, early_iter=40, callback=cb,
                           relabel_every=every)
        out[every] = (st.snapshot(), sq, {k: host(v) for k, v in seen.items()})
    ref = out[-1]
    for every in (1, 7, 0):
        (Y, U, G), sq, seen = out[every]
        assert np.array_equal(Y, ref[0][0]) and np.array_equal(U, ref[0][1]) and np.array_equal(G, ref[0][2]), every
        assert np.array_equal(sq, ref[1]), every
        assert sorted(seen) == sorted(ref[2]) and all(np.array_equal(seen[k], ref[2][k]) for k in seen), every


SCHED_OPTS = [dict(dof=0.6), dict(dof=2.0), dict(n_interp=2), dict(n_interp=5, max_boxes=816),
              dict(exaggeration=4.0, learning_rate=150.0, momentum=(0.3, 0.9)),
              dict(box_width=1.1, min_boxes=8, max_boxes=12)]


@pytest.mark.parametrize("opts", SCHED_OPTS, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()))
def test_fp64_loop_options_match_oracle(opts):
    """The whole FP64 loop with a non-default kernel exponent, interpolation order, schedule or grid limits
    (the last clips n_boxes at 12 with a 1.1 box) against the oracle over 20 iterations from a spread layout:
    every iteration's sumQ to 1e-10 and the final layout to 1e-9."""
    P, _, _ = problem(2000, 10, 5, seed=3)
    Y0 = np.random.default_rng(4).normal(0, 3.0, (2000, 2))
    data = P.data.astype(np.float64)
    st = State(P, Y0, np.float64)
    kw = dict(n_iter=40, early_iter=10, **opts)
    sq = st.run(0, 20, **kw)
    Yo, Uo, Go = Y0.copy(), np.zeros_like(Y0), np.ones_like(Y0)
    sqo = O.optimize(P.indptr, P.indices, data, Yo, Uo, Go, 0, 20, **kw)
    assert np.abs(sq / sqo - 1).max() < 1e-10
    assert rel(host(st.Y), Yo) < 1e-9


@pytest.mark.parametrize("n", [2, 3, 7])
def test_fp64_loop_tiny_layouts_match_oracle(n):
    """A handful of points (a complete graph P, default learning rate 200): the first iterations fling them from a
    span of 0.01 to about 100, with sumQ = sum pot0 - n of order 0.1 for n of order 1, so whole trajectories part
    by rounding alone; each iteration from the same state matches the oracle to 1e-10."""
    rng = np.random.default_rng(n)
    A = sp.csr_matrix(np.ones((n, n)) - np.eye(n))
    A = (A / A.sum()).tocsr()
    A.sort_indices()
    data = A.data.astype(np.float64)
    st = State(A, rng.normal(0, 1e-2, (n, 2)), np.float64)
    kw = dict(n_iter=30, early_iter=10)
    it = 0
    for t in (0, 1, 4):
        if t > it:
            st.run(it, t, **kw)
        Y, U, G = st.snapshot()
        sq = st.run(t, t + 1, **kw)
        it = t + 1
        Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
        sqo = O.optimize(A.indptr, A.indices, data, Yo, Uo, Go, t, t + 1, **kw)
        Yg, Ug, _ = st.snapshot()
        assert abs(sq[0] / sqo[0] - 1) < 1e-10, (t, sq[0], sqo[0])
        assert max(rel(Yg - Y, Yo - Y), rel(Ug, Uo)) < 1e-10, t


def test_fp64_loop_with_isolated_points_matches_oracle():
    """CSR rows without entries (points with no affinity): they feel only the repulsion."""
    P, Y0, _ = problem(2000, 10, 5, seed=3)
    P = P.tolil()
    for i in range(0, 2000, 97):
        P.rows[i] = []
        P.data[i] = []
    P = P.tocsr()
    P = P.multiply(1.0 / P.sum()).tocsr()
    P.sort_indices()
    assert np.any(np.diff(P.indptr) == 0)
    data = P.data.astype(np.float64)
    st = State(P, Y0, np.float64)
    sq = st.run(0, 25, **SCHED)
    Yo, Uo, Go = Y0.copy(), np.zeros_like(Y0), np.ones_like(Y0)
    sqo = O.optimize(P.indptr, P.indices, data, Yo, Uo, Go, 0, 25, **SCHED)
    assert np.abs(sq / sqo - 1).max() < 1e-10
    assert rel(host(st.Y), Yo) < 1e-9


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_session_runs_in_pieces_equal_one_call(dtype):
    """mg_tsne_session: a schedule run in pieces ([0, 17), [17, 60), [60, 90)) on one session gives the state and
    sumQ of one mg_tsne_optimize call over [0, 90), bit for bit; a second state from the same start on the same
    session (its buffers and spectra cache warm) gives them again."""
    import torch
    P, Y0, _ = problem(4000, 20, 5, 3)
    kw = dict(n_iter=90, early_iter=30)
    one = State(P, Y0, dtype)
    sq_one = one.run(0, 90, **kw)
    st = State(P, Y0, dtype)
    sess = capi.Session(st.ip, st.ix, st.data, **kw)
    for rep in range(2):
        Y, U, G = dev(Y0.astype(dtype)), dev(np.zeros_like(Y0, dtype)), dev(np.ones_like(Y0, dtype))
        sq = np.concatenate([sess.run(Y, U, G, a, b) for a, b in ((0, 17), (17, 60), (60, 90))])
        torch.cuda.synchronize()
        assert torch.equal(Y, one.Y) and torch.equal(U, one.U) and torch.equal(G, one.G), rep
        assert np.array_equal(sq, sq_one), rep
    sess.close()


def test_trace_kl_and_gradient_norm_match_oracle():
    """Per-iteration diagnostics of a session (KL divergence and gradient norm, FP64 with compensated sums) on the
    10k configuration: FP64 against the oracle's on the same layouts (1e-10 relative) across both phases; FP32 with
    and without relabelling equal bit for bit, and within 1e-5 (KL, absolute) / 1e-4 (norm, relative) of FP64 over
    the first iterations (later the two precisions' trajectories part)."""
    import torch
    P, Y0, _ = problem()
    n = P.shape[0]
    kw = dict(n_iter=500, early_iter=166)
    out = {}
    for dt, relabel in ((np.float64, 0), (np.float32, 0), (np.float32, -1)):
        st = State(P, Y0, dt)
        sess = capi.Session(st.ip, st.ix, st.data, relabel_every=relabel, **kw)
        states, sqs, kls, gns = [], [], [], []
        out_it = 0
        for a, b in ((0, 3), (160, 170), (300, 303)):
            if a > 0:
                sess.run(st.Y, st.U, st.G, out_it, a)
            states.append(host(st.Y).copy())
            sq, kl, gn = sess.run(st.Y, st.U, st.G, a, b, trace=True)
            sqs.append(sq); kls.append(kl); gns.append(gn)
            out_it = b
            if dt == np.float64:
                # the oracle on the layout each iteration starts from (the first of each run)
                Yt = states[-1]
                g, sqo = O.gradient(P.indptr, P.indices, P.data.astype(np.float64), Yt, a, **kw)
                klo = O.kl_divergence(P.indptr, P.indices, P.data.astype(np.float64), Yt, sqo)
                print(f"t={a}: KL {kl[0]:.12g} oracle {klo:.12g}; |g| {gn[0]:.12g} oracle {np.linalg.norm(g):.12g}")
                assert abs(kl[0] - klo) <= 1e-10 * max(1.0, abs(klo)) and abs(gn[0] / np.linalg.norm(g) - 1) < 1e-10
        sess.close()
        out[(dt, relabel)] = (np.concatenate(kls), np.concatenate(gns))
    a, b = out[(np.float32, 0)], out[(np.float32, -1)]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    k64, g64 = out[(np.float64, 0)]
    # the first run's iterations, before the two precisions' trajectories part
    print("FP32 - FP64 KL", np.abs(a[0][:3] - k64[:3]).max(), "norm rel", np.abs(a[1][:3] / g64[:3] - 1).max())
    assert np.abs(a[0][:3] - k64[:3]).max() < 1e-5 and np.abs(a[1][:3] / g64[:3] - 1).max() < 1e-4


def test_tunables_change_no_result():
    """Every launch shape the autotuner may pick (the sorts' keys per thread and digit width, the column pass in the
    row layout or between folded row passes or transposes, the attraction's LDS window on or off, its rows, window and
    threads per block) gives the same state bit for bit, FP32 and FP64; the autotuner leaves one of them set."""
    import torch
    from magnet.tune import autotune_loop, SPACE
    P, Y0, _ = problem()
    kw = dict(n_iter=500, early_iter=166)
    settings = [dict(sort_items=0), dict(sort_items=1), dict(sort_items=8), dict(attr_window=1, attr_rows=64, attr_win_kb=64,
                attr_tpb=512), dict(attr_window=1, attr_rows=512, attr_win_kb=128, attr_tpb=1024), dict(attr_window=0), dict(fft_strided_max=0), dict(fft_strided_max=0, fft_fold=0),
                dict(fft_strided_max=0, fft_fold=2), dict(sort_digit=10), dict(spec_issue=0), dict(spec_issue=2)]
    try:
        for dt in (np.float32, np.float64):
            outs = []
            for setting in settings:
                for name, values in SPACE:
                    capi.tune(name, values[0])
                for name, v in setting.items():
                    capi.tune(name, v)
                st = State(P, Y0, dt)
                capi.optimize(st.ip, st.ix, st.data, st.Y, st.U, st.G, 0, 40, **kw)
                outs.append(host(st.Y).copy())
            assert all(np.array_equal(outs[0], o) for o in outs[1:]), dt
        st = State(P, Y0, np.float32)
        sess = capi.Session(st.ip, st.ix, st.data, **kw)
        chosen, log = autotune_loop(sess, (st.Y, st.U, st.G), 0, iters=5)
        sess.close()
        print("autotuner chose", chosen)
        assert set(chosen) == {name for name, _ in SPACE} and all(capi.tune(k) == v for k, v in chosen.items())
    finally:
        for name, values in SPACE:
            capi.tune(name, values[0])
