"""A transform() service with continuous batching (the serving form of MAGNET.transform, magnet/__init__.py:876-930).

The fitted model's training rows, layout and potential field stay resident in HBM. Requests, each an array of
new rows, enter a queue. A worker thread forms device batches from them as they arrive:

- A request is admitted into the next batch while the batch's point budget allows.
- The budget is the smaller of max_batch_points and what a fraction of the free HBM holds for the k-NN
  distance tiles.
- Requests that arrive while a batch is on the device wait for the next one.
- A batch runs as one k-NN search and one placement launch; the requests in it share those launches.

New points do not interact, and the k-NN search goes through in fixed-size query tiles, so a request's result
does not depend on which other requests share its batch. Latency is measured from submission to completion,
per request.
"""
import queue
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
    def __init__(self, model, max_batch_points=16384, hbm_fraction=0.5, n_iter=None):
        import torch
        if not hasattr(model, "embedding_"):
            raise ValueError("the model must be fitted")
        self.model = model
        self.n_iter = n_iter
        model._serving_state()   # training rows, layout and field to the device once
        n_train = len(model.embedding_)
        free, _ = torch.cuda.mem_get_info()
        tile_bytes = 256 * n_train * 4 * 3       # one k-NN distance tile and its top-k scratch
        per_point = 64 * 4 + 2 * 8               # neighbours, weights, position
        self.max_batch_points = int(max(1, min(max_batch_points, (hbm_fraction * free - tile_bytes) // per_point)))
        self._q = queue.Queue()
        self._ids = 0
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._worker = threading.Thread(target=self._run, daemon=True)
        self.batches = []   # (n_requests, n_points, seconds) of every batch served

    def start(self):
        self._worker.start()
        return self

    def stop(self):
        self._stop.set()
        self._q.put(None)
        self._worker.join()

    def submit(self, X):
        with self._lock:
            rid = self._ids
            self._ids += 1
        r = Request(np.asarray(X), rid)
        self._q.put(r)
        return r

    def _run(self):
        import torch
        m = self.model
        pending = []
        while not self._stop.is_set():
            if not pending:
                item = self._q.get()
                if item is None:
                    break
                pending.append(item)
            while True:   # everything that has arrived so far
                try:
                    item = self._q.get_nowait()
                except queue.Empty:
                    break
                if item is None:
                    self._stop.set()
                    break
                pending.append(item)
            batch, points = [], 0
            while pending and (not batch or points + len(pending[0].X) <= self.max_batch_points):
                r = pending.pop(0)
                batch.append(r)
                points += len(r.X)
            t0 = time.perf_counter()
            try:
                Q = np.concatenate([m._prepare_query(r.X) for r in batch]) if points else np.zeros((0, m._prepared_.shape[1]), np.float32)
                bounds = np.cumsum([0] + [len(r.X) for r in batch])
                segs = list(zip(bounds[:-1].tolist(), bounds[1:].tolist()))
                Y = m._place(torch.as_tensor(Q, device="cuda"), self.n_iter, segs) if points else None
                out = Y.cpu().numpy().astype(np.float32) if points else np.zeros((0, 2), np.float32)
                o = 0
                t_done = time.perf_counter()
                for r in batch:
                    r.result = out[o:o + len(r.X)]
                    o += len(r.X)
                    r.t_done = t_done
                    r.batch = len(self.batches)
                    r._done.set()
            except Exception as e:   # noqa: BLE001 -- reported to every request of the batch
                for r in batch:
                    r.error = e
                    r.t_done = time.perf_counter()
                    r._done.set()
            self.batches.append((len(batch), points, time.perf_counter() - t0))
        for r in pending:
            r.error = RuntimeError("server stopped")
            r._done.set()


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
