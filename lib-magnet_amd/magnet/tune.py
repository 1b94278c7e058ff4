"""Autotuner of the launch shapes that change no result (mg_tune_set): the radix sorts' keys per thread and digit
width, and the 2-D attraction's LDS-window sweep (on or off, rows per block, window size, threads per block).

autotune(run, space) times `run` (a short stretch of real work on copies of the state) for each candidate value of
each tunable in turn, keeps the fastest, and leaves it set: a coordinate sweep, one tunable at a time, in the order
given. Every candidate gives the same result bit for bit, so the choice only moves time.
"""
import time

__all__ = ["SPACE", "autotune", "autotune_loop"]

# candidates per tunable, in sweep order (the first value of each is its default)
SPACE = (
    ("sort_items", (0, 1, 2, 4, 8)),
    ("sort_digit", (8, 10)),
    ("fft_strided_max", (1024, 0, 2048)),
    ("fft_fold", (4, 2, 0)),
    ("fft_fold_persist", (3, 0)),
    ("spec_issue", (1, 0, 2)),
    ("attr_window", (-1, 0, 1)),
    ("attr_rows", (256, 64, 128, 512)),
    ("attr_win_kb", (64, 128, 0)),
    ("attr_tpb", (1024, 512, 256)),
)


def autotune(run, space=SPACE, repeats=2):
    """Sets each tunable of `space` to the value whose `run()` is fastest (the best of `repeats` timings), one tunable
    after another; returns {name: value} and the timings [(name, value, seconds)]."""
    import torch
    from . import capi
    chosen, log = {}, []
    run()   # warm-up: first-use allocations and caches are not charged to the first candidate
    for name, values in space:
        best, best_t = None, float("inf")
        for v in values:
            capi.tune(name, v)
            t = float("inf")
            for _ in range(repeats):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                run()
                torch.cuda.synchronize()
                t = min(t, time.perf_counter() - t0)
            log.append((name, v, round(t, 6)))
            if t < best_t:
                best, best_t = v, t
        capi.tune(name, best)
        chosen[name] = best
    return chosen, log


def autotune_loop(session, state, it0, iters=10, space=SPACE, repeats=2):
    """autotune over `iters` iterations [it0, it0 + iters) of a capi.Session from copies of `state` (Y, U, G device
    tensors); the state itself is not moved. `state` may also be a list of (state, it0) pairs: each candidate is
    then timed over `iters` iterations from every one of them (phases whose grids differ weigh by their cost)."""
    pairs = state if isinstance(state, list) else [(state, it0)]

    def run():
        for st, t0 in pairs:
            s = tuple(t.clone() for t in st)
            session.run(*s, t0, t0 + iters)
    run()   # warm the session's buffers and spectra for these stretches
    return autotune(run, space, repeats=repeats)
