"""mg_knn_noalloc, the allocation-free k-NN search, against mg_knn through the C-ABI: the same indices and squared
distances bit for bit, on two calls in a row (the second smaller), with the fixed scratch forcing fewer training
chunks than mg_knn uses; queries beyond the scratch's limit are refused before any launch."""
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


def _pair(T, Q, k, exclude):
    import torch
    ref_i, ref_d = capi.knn(T, Q, k, exclude)
    idx = torch.full((Q.shape[0], k), -7, dtype=torch.int32, device="cuda")
    d2 = torch.full((Q.shape[0], k), -7.0, dtype=torch.float32, device="cuda")
    capi.knn_noalloc(T, Q, idx, d2, k, exclude)
    torch.cuda.synchronize()
    return ref_i, ref_d, idx, d2


@pytest.mark.parametrize("d,k,exclude", [(24, 14, False), (9, 29, True), (50, 5, False)])
def test_noalloc_knn_equals_knn_on_two_calls(d, k, exclude):
    """m = 6000 then 700 queries against 60,000 rows: mg_knn splits the rows into 64 chunks, the fixed scratch
    takes 8 (k <= 16) or 4 chunks, and 64 and 512 chunks for the 700; integer-valued features give exact ties, broken by index in both."""
    import torch
    rng = np.random.default_rng(d * 31 + k)
    nt = 60_000
    T = rng.integers(0, 5, (nt, d)).astype(np.float32)
    T[: nt // 2] += rng.normal(0, 0.25, (nt // 2, d)).astype(np.float32)
    Tt = torch.as_tensor(T, device="cuda")
    for m in (6000, 700):
        print(f"noalloc knn: d={d} k={k} exclude={exclude} m={m} nt={nt}", flush=True)
        Q = T[:m].copy() if exclude else rng.integers(0, 5, (m, d)).astype(np.float32)
        ref_i, ref_d, idx, d2 = _pair(Tt, torch.as_tensor(Q, device="cuda"), k, exclude)
        assert torch.equal(idx, ref_i) and torch.equal(d2, ref_d), m


def test_noalloc_knn_refuses_more_queries_than_its_scratch():
    import torch
    T = torch.zeros((100, 4), dtype=torch.float32, device="cuda")
    Q = torch.zeros((32769, 4), dtype=torch.float32, device="cuda")
    idx = torch.empty((32769, 20), dtype=torch.int32, device="cuda")
    d2 = torch.empty((32769, 20), dtype=torch.float32, device="cuda")
    with pytest.raises(capi.MgError, match="at most 32768 query rows"):
        capi.knn_noalloc(T, Q, idx, d2, 20)
