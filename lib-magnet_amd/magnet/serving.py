"""A transform() service with continuous batching (the serving form of MAGNET.transform, magnet/__init__.py:876-930).

The fitted model's training rows, layout and potential field stay resident in HBM. Requests, each an array of
new rows, enter a queue. A worker thread forms device batches from them as they arrive:

- Several batches are in flight at once: n_workers worker threads, each with its own HIP stream, take the
  requests that have arrived whenever they are free, so requests that arrive while a batch is on the device
  start on another stream at once and their work overlaps with it on the device.
- A request is admitted into a batch while the batch's point budget allows. The budget is taken again for every
  batch, from the HBM free at that moment: the smaller of max_batch_points and the points that this worker's share
  of hbm_fraction of the free memory holds, after the k-NN scratch a worker takes on its first batch. When memory
  is short a batch still takes one request, so every request is served.
- A batch runs as one k-NN search and one placement launch; the requests in it share those launches.

New points do not interact, every per-request quantity is computed over the request's own rows, and the k-NN
search goes through in fixed-size query tiles, so a request's result does not depend on which other requests
share its batch, nor on what runs beside it. Latency is measured from submission to completion, per request.
"""
import collections
import threading
import time

import numpy as np

__all__ = ["TransformServer", "Request", "poisson_load", "latency_report"]


class Request:
    def __init__(self, X, rid):
        self.X = X
        self.id = rid
        self.t_submit = time.perf_counter()
        self.t_done = None
        self.result = None
        self.error = None
        self.batch = None
        self._done = threading.Event()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("request not served in time")
        if self.error is not None:
            raise self.error
        return self.result

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


class TransformServer:
    def __init__(self, model, max_batch_points=16384, hbm_fraction=0.5, n_iter=None, n_workers=3, mem_info=None):
        if not hasattr(model, "embedding_"):
            raise ValueError("the model must be fitted")
        if n_workers < 1:
            raise ValueError("n_workers must be >= 1")
        self.model = model
        self.n_iter = n_iter
        self.n_workers = int(n_workers)
        model._serving_state()   # training rows, layout and field to the device once
        n_train = len(model.embedding_)
        d = model._prepared_.shape[1]
        # a worker's first batch takes the library's k-NN lists (2 x 2^24 entries of 4 bytes) or, for rows of more
        # than 64 features, one GEMM distance tile of 256 rows and its top-k scratch
        self.knn_bytes = max(2 * (1 << 24) * 4, 256 * n_train * 4 * 3)
        # per new row: its features, k-NN indices / distances / weights (k <= 32), position, and torch temporaries
        self.per_point = (d + 32 * 3 + 8) * 4 * 2
        self.hbm_fraction = float(hbm_fraction)
        self.max_points = int(max_batch_points)
        self._mem_info = mem_info or _device_mem_info
        self._warm = set()   # workers whose k-NN scratch exists
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._ids = 0
        self._stop = False
        self._workers = [threading.Thread(target=self._run, daemon=True) for _ in range(self.n_workers)]
        self.batches = []   # (n_requests, n_points, seconds) of every batch served, in completion order
        self.budgets = []   # the point budget each batch was admitted under, in admission order

    def start(self):
        for w in self._workers:
            w.start()
        return self

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        with self._cv:
            left = list(self._pending)
            self._pending.clear()
        for r in left:
            r.error = RuntimeError("server stopped")
            r._done.set()

    def submit(self, X):
        with self._cv:
            if self._stop:
                raise RuntimeError("server stopped")
            r = Request(np.asarray(X), self._ids)
            self._ids += 1
            self._pending.append(r)
            self._cv.notify()
        return r

    def budget(self, worker=None):
        """Points a batch may take now: this worker's share of hbm_fraction of the HBM free at this moment, less the
        k-NN scratch its first batch allocates, capped by max_batch_points (at least 1)."""
        free = self._mem_info()
        share = self.hbm_fraction * free / self.n_workers - (0 if worker in self._warm else self.knn_bytes)
        return int(max(1, min(self.max_points, share // self.per_point)))

    def _next_batch(self, worker=None):
        """Everything that has arrived, up to the point budget taken now (at least one request); None when
        stopping."""
        with self._cv:
            while not self._pending and not self._stop:
                self._cv.wait()
            if self._stop:
                return None
            cap = self.budget(worker)
            batch, points = [], 0
            while self._pending and (not batch or points + len(self._pending[0].X) <= cap):
                r = self._pending.popleft()
                batch.append(r)
                points += len(r.X)
            self.budgets.append(cap)
            if self._pending:
                self._cv.notify()   # what did not fit goes to the next free worker
            return batch

    def _run(self):
        import torch
        from . import capi
        m = self.model
        stream = torch.cuda.Stream()
        me = threading.get_ident()
        with torch.cuda.stream(stream):
            capi.use_thread_stream(stream)
            while True:
                batch = self._next_batch(me)
                if batch is None:
                    break
                self._serve(m, batch)
                self._warm.add(me)

    def _serve(self, m, batch):
        import torch
        t0 = time.perf_counter()
        points = sum(len(r.X) for r in batch)
        try:
            Q = np.concatenate([m._prepare_query(r.X) for r in batch]) if points else np.zeros((0, m._prepared_.shape[1]), np.float32)
            bounds = np.cumsum([0] + [len(r.X) for r in batch])
            segs = list(zip(bounds[:-1].tolist(), bounds[1:].tolist()))
            Y = m._place(torch.as_tensor(Q, device="cuda"), self.n_iter, segs) if points else None
            out = Y.cpu().numpy().astype(np.float32) if points else np.zeros((0, 2), np.float32)
            err = None
        except Exception as e:   # noqa: BLE001 -- reported to every request of the batch
            out, err = None, e
        t_done = time.perf_counter()
        with self._cv:
            bid = len(self.batches)
            self.batches.append((len(batch), points, t_done - t0))
        o = 0
        for r in batch:
            r.batch = bid
            if err is None:
                r.result = out[o:o + len(r.X)]
                o += len(r.X)
            else:
                r.error = err
            r.t_done = t_done
            r._done.set()


def _device_mem_info():
    """Free HBM bytes on the current device (mg_mem_info)."""
    from . import capi
    return capi.mem_info()[0]


def poisson_load(server, sampler, qps, duration, seed=0):
    """Submits requests sampler(rng) at Poisson arrival times of rate qps for `duration` seconds; waits for all;
    returns the requests."""
    rng = np.random.default_rng(seed)
    reqs = []
    t_end = time.perf_counter() + duration
    t_next = time.perf_counter()
    while True:
        t_next += rng.exponential(1.0 / qps)
        if t_next > t_end:
            break
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        reqs.append(server.submit(sampler(rng)))
    for r in reqs:
        r.wait(600)
    return reqs


def latency_report(reqs, wall):
    lat = np.array([r.latency for r in reqs]) * 1e3
    pts = sum(len(r.X) for r in reqs)
    return {"requests": len(reqs), "points": int(pts), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
            "points_per_s": round(pts / wall, 1)}
