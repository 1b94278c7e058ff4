"""Host side of MAGNET(loss="tsne"): parameter checks as the reference's (tests/test_magnet.py:425-428), and the
data preparation and the server's admission logic (the memberships, the fuzzy union, the mid-near pairs and the k-NN
graph run on the device and are tested there)."""
import numpy as np
import pytest

from magnet import MAGNET
from magnet.graph import prepare_data


def test_parameter_errors():
    X = np.random.default_rng(0).normal(size=(50, 4)).astype(np.float32)
    for bad in (dict(loss="foo"), dict(loss="tsne", n_components=3), dict(loss="tsne", metric="manhattan"),
                dict(loss="tsne", n_components=0), dict(loss="tsne", precision="fp16")):
        with pytest.raises(ValueError):
            MAGNET(**bad).fit(X)
    for other in (dict(loss="infonce"), dict(loss="umap_exact"), dict(loss="umap", kernel="gaussian"),
                  dict(loss="hybrid", kernel_b="consistent")):
        with pytest.raises(NotImplementedError):
            m = MAGNET(**other)
            m._validate()
            m._kernel_params(50)
    with pytest.raises(ValueError):
        MAGNET(loss="umap", transform_mode="beam")._validate()


def test_prepare_data_cosine_rows():
    X = np.array([[3.0, 4.0], [0.0, 0.0]], np.float32)
    Q = prepare_data(X, "cosine")
    assert np.allclose(Q, [[0.6, 0.8, 0.0], [0.0, 0.0, 1.0]])


def _bare_server(budget):
    """A TransformServer without its device state or worker threads: the admission logic alone."""
    import collections
    import threading
    from magnet.serving import TransformServer
    srv = TransformServer.__new__(TransformServer)
    srv._pending = collections.deque()
    srv._cv = threading.Condition()
    srv._ids = 0
    srv._stop = False
    srv._workers = []
    srv.max_points = budget
    srv.n_workers, srv.hbm_fraction, srv.knn_bytes, srv.per_point = 1, 1.0, 0, 1
    srv._mem_info = lambda: 1 << 40
    srv._warm = set()
    srv.batches = []
    srv.budgets = []
    return srv


def test_server_batches_in_arrival_order_within_the_budget():
    """A free worker takes the requests that have arrived, in order, while their rows fit the point budget; a
    request larger than the budget still forms a batch of its own; nothing is reordered or dropped."""
    srv = _bare_server(100)
    sizes = [30, 50, 40, 250, 10, 90, 5]
    reqs = [srv.submit(np.zeros((s, 3), np.float32)) for s in sizes]
    batches = []
    while srv._pending:
        batches.append([r.id for r in srv._next_batch()])
    assert batches == [[0, 1], [2], [3], [4, 5], [6]]
    assert [r.id for r in reqs] == list(range(len(sizes)))


def test_server_stop_fails_pending_requests_and_refuses_new_ones():
    srv = _bare_server(100)
    r = srv.submit(np.zeros((4, 3), np.float32))
    srv.stop()
    with pytest.raises(RuntimeError):
        r.wait(1)
    with pytest.raises(RuntimeError):
        srv.submit(np.zeros((4, 3), np.float32))
    assert srv._next_batch() is None


def test_server_admits_each_batch_by_the_hbm_free_at_that_moment():
    """The point budget is taken from the free HBM for every batch: as free memory shrinks, batches shrink; a
    worker's first batch also sets aside its k-NN scratch; with no room at all a batch still takes one request."""
    srv = _bare_server(10_000)
    srv.n_workers, srv.hbm_fraction, srv.knn_bytes, srv.per_point = 2, 0.5, 4000, 10
    free = iter([1_000_000, 1_000_000, 200_000, 40_000, 1000])
    srv._mem_info = lambda: next(free)
    for s in [100] * 300:
        srv.submit(np.zeros((s, 3), np.float32))
    sizes = []
    for worker in ("a", "a", "a", "a", "b"):
        sizes.append(sum(len(r.X) for r in srv._next_batch(worker)))
        srv._warm.add(worker)
    # budgets: (0.5 * free / 2 - scratch on a cold worker) / 10 bytes per point, capped at 10,000 points
    assert srv.budgets == [10_000, 10_000, 5_000, 1_000, 1]
    assert sizes == [10_000, 10_000, 5_000, 1_000, 100]
