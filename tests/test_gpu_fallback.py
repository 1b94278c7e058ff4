"""The reference's fallback kernels (src/fftrep.cpp:51-125, 294-411) on the GPU, through the C-ABI: float32 equal
to the reference's own outputs (tests/golden: the sums follow its single-thread order without fused multiply-adds),
float64 against the oracle to rounding (1e-12), and the pybind module's names against the same vectors."""
import numpy as np
import pytest

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


def shim_args(g):
    x0, y0, w, nb = g["shim_grid"]
    return float(np.float32(x0)), float(np.float32(y0)), float(np.float32(w)), int(nb)


def test_fallback_fp32_equal_to_reference(golden):
    g = golden
    x0, y0, w, nb = shim_args(g)
    Y = dev(g["shim_Y"])
    assert np.array_equal(host(capi.interp_scatter(Y, dev(g["shim_q"]), x0, y0, w, nb, 3)), g["shim_scatter"])
    assert np.array_equal(host(capi.interp_gather(Y, dev(g["shim_G"]), x0, y0, w, nb, 3)), g["shim_gather"])
    assert np.array_equal(host(capi.scatter_padded_2d(Y, x0, y0, w, nb, 3, int(g["shim_M"]))), g["shim_padded"])
    assert np.array_equal(host(capi.spectral_products(dev(g["shim_K1"]), dev(g["shim_K2"]), dev(g["shim_Wh"]))),
                          g["shim_products"])
    s, F = capi.gather_forces_2d(Y, dev(g["shim_phi"]), x0, y0, w, nb, 3)
    assert np.array_equal(host(F), g["shim_F"])
    assert abs(s - float(g["shim_sumQ"])) <= 1e-6 * abs(float(g["shim_sumQ"]))   # double sum, other order


@pytest.mark.parametrize("n,ni,nb,C", [(5000, 3, 40, 3), (20_000, 5, 64, 11), (1, 2, 20, 2), (0, 3, 20, 1)])
def test_fallback_fp64_against_oracle(n, ni, nb, C):
    """crowded boxes, many charge columns (more than one pass of the node kernel), a single point, no points"""
    rng = np.random.default_rng(n + ni)
    span = nb * 0.7
    Y = rng.normal(span / 2, span / 6, (n, 2)).clip(-1.0, span + 1.0)   # some points outside the grid: clamped
    q = rng.normal(0, 1, (n, C))
    args = (0.0, 0.0, 0.7, nb, ni)
    N = nb * ni
    Wo = O.interp_scatter(Y, q, *args)
    W = host(capi.interp_scatter(dev(Y), dev(q), *args))
    assert np.abs(W - Wo).max() <= 1e-12 * max(1.0, np.abs(Wo).max())
    G = rng.normal(0, 1, (C, N, N))
    if n:
        assert rel(host(capi.interp_gather(dev(Y), dev(G), *args)), O.interp_gather(Y, G, *args)) < 1e-13
    M = N + 6
    P = host(capi.scatter_padded_2d(dev(Y), *args, M))
    assert np.abs(P - O.scatter_padded_2d(Y, *args, M)).max() <= 1e-12 * max(1.0, np.abs(P).max())
    phi = rng.normal(0, 1, (4, M, M))
    s, F = capi.gather_forces_2d(dev(Y), dev(phi), *args)
    so, Fo = O.gather_forces_2d(Y, phi, *args)
    assert abs(s - so) <= 1e-12 * max(1.0, abs(so))
    if n:
        assert rel(host(F), Fo) < 1e-13


def test_fallback_pybind_names(golden):
    """the `_magnet` functions (host arrays in and out) give the reference's outputs"""
    import _magnet
    g = golden
    x0, y0, w, nb = shim_args(g)
    Y = g["shim_Y"]
    assert np.array_equal(np.asarray(_magnet.interp_scatter(Y, g["shim_q"], x0, y0, w, nb, 3)), g["shim_scatter"])
    assert np.array_equal(np.asarray(_magnet.interp_gather(Y, g["shim_G"], x0, y0, w, nb, 3, 1)), g["shim_gather"])
    buf = np.full((3, int(g["shim_M"]), int(g["shim_M"])), 7.0, np.float32)
    out = _magnet.scatter_padded_2d(Y, x0, y0, w, nb, 3, int(g["shim_M"]), -1, buf)
    assert np.array_equal(np.asarray(out), g["shim_padded"]) and np.array_equal(buf, g["shim_padded"])
    assert np.array_equal(np.asarray(_magnet.spectral_products(g["shim_K1"], g["shim_K2"], g["shim_Wh"])), g["shim_products"])
    s, F = _magnet.gather_forces_2d(Y, g["shim_phi"], x0, y0, w, nb, 3)
    assert np.array_equal(np.asarray(F), g["shim_F"]) and s == np.float32(g["shim_sumQ"])
