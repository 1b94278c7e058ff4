"""The reference-facing layer with values: the `_magnet` pybind module (host arrays in and out, the reference's
names and signatures, src/bindings.cpp:8-83 and src/fftconv.cpp:239-244) and magnet.tsne.FFTRepulsion2D over it,
on the reference's own acceptance test and on the golden vectors the compiled reference wrote."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _spectra(M, width, n_interp, dof=1.0):
    """The reference's real float32 kernel spectra (magnet/tsne.py:55-70) at padded size M."""
    import scipy.fft as sfft
    off = np.arange(M)
    off = np.where(off <= M // 2, off, off - M).astype(np.float64) * (width / n_interp)
    r2 = off[:, None] ** 2 + off[None, :] ** 2
    base = 1.0 + r2 / dof
    K1 = (base ** (-dof)).astype(np.float32)
    K2 = (base ** (-dof - 1.0)).astype(np.float32)
    return (np.ascontiguousarray(sfft.rfft2(K1).real, dtype=np.float32),
            np.ascontiguousarray(sfft.rfft2(K2).real, dtype=np.float32))


def test_fft_repulsion2d_matches_brute_force():
    """tests/test_magnet.py:221-231 of the reference, through magnet.tsne.FFTRepulsion2D (n_interp 5): sumQ to
    1e-3 and F to 2e-2 of the exact O(n^2) sums."""
    from magnet.tsne import FFTRepulsion2D
    rng = np.random.default_rng(0)
    Y = np.ascontiguousarray(rng.normal(0, 8, (1500, 2)).astype(np.float32))
    Yd = Y.astype(np.float64)
    D2 = ((Yd[:, None, :] - Yd[None, :, :]) ** 2).sum(-1)
    Q = 1.0 / (1.0 + D2)
    np.fill_diagonal(Q, 0)
    F_exact = ((Q ** 2)[:, :, None] * (Yd[:, None, :] - Yd[None, :, :])).sum(1)
    rep = FFTRepulsion2D(n_interp=5)
    sumQ, F = rep(Y)
    print(f"FFTRepulsion2D n_interp=5: sumQ rel {abs(sumQ - Q.sum()) / Q.sum():.2e}, F rel {rel(F, F_exact):.2e}")
    assert abs(sumQ - Q.sum()) / Q.sum() < 1e-3
    assert rel(F, F_exact) < 2e-2
    s2, F2 = rep(Y)                                   # the cached object gives the same result again
    assert s2 == sumQ and np.array_equal(F2, F)


@pytest.mark.parametrize("case,ni", [("rep3", 3), ("repc", 3), ("rep5", 5)])
def test_fftconv2d_set_kernels_compute_against_reference(golden, case, ni):
    """_magnet.FFTConv2D(M, n_interp).set_kernels(K1h, K2h).compute(...) at the reference's own padded size
    (scipy.fft.next_fast_len(2N)), on the inputs the compiled reference answered: the binding turns the real
    spectra back into kernel values, the device transforms at its own size. sumQ to 1e-5, F to 2e-4 (float32
    pocketfft against the device float32 FFT)."""
    import _magnet
    Y = np.ascontiguousarray(golden[case + "_Y"])
    xmin, ymin, width, nb, M = golden[case + "_grid"]
    nb, M = int(nb), int(M)
    conv = _magnet.FFTConv2D(M, ni)
    conv.set_kernels(*_spectra(M, width, ni))
    sumQ, F = conv.compute(Y, float(xmin), float(ymin), float(width), nb, -1)
    ref_s, ref_F = float(golden[case + "_sumQ"]), golden[case + "_F"]
    print(f"{case}: M={M} sumQ rel {abs(sumQ - ref_s) / ref_s:.2e}, F rel {rel(F, ref_F):.2e}")
    assert abs(sumQ - ref_s) / ref_s < 1e-5
    assert rel(F, ref_F) < 2e-4


def test_per_call_functions_against_reference(golden):
    """tsne_attraction (2-D, and 3-D with dof 0.7), tsne_step, tsne_update, mid_spring and bbox_2d through the
    pybind module on the reference's golden inputs: the update and step in place, as the reference does them."""
    import _magnet
    g = golden
    ip, ix, d = g["csr_indptr"], g["csr_indices"], g["csr_data"]
    out = np.asarray(_magnet.tsne_attraction(ip, ix, d, g["att_Y"], 1.0, -1))
    assert rel(out, g["att_out"]) < 1e-6
    out3 = np.asarray(_magnet.tsne_attraction(ip, ix, d, np.ascontiguousarray(g["att_Y3"]), 0.7, -1))
    assert rel(out3, g["att_out3_dof07"]) < 1e-6
    Y, U, G = g["att_Y"].copy(), g["st_U"].copy(), g["st_G"].copy()
    _magnet.tsne_step(Y, U, G, g["att_out"], g["st_F"], 12.0, float(np.float32(1 / 3.7)), 250.0, 0.5, -1)
    assert rel(Y, g["step_Y"]) < 1e-6 and rel(U, g["step_U"]) < 1e-6 and np.array_equal(G, g["step_G"])
    Y, U, G = g["att_Y"].copy(), g["st_U"].copy(), g["st_G"].copy()
    _magnet.tsne_update(Y, U, G, g["upd_grad"], 250.0, 0.8, -1)
    assert rel(Y, g["upd_Y"]) < 1e-6 and rel(U, g["upd_U"]) < 1e-6 and np.array_equal(G, g["upd_G"])
    gr = g["upd_grad"].copy()
    _magnet.mid_spring(g["att_Y"], g["spr_pairs"], gr, 0.01, -1)
    assert rel(gr, g["spr_grad"]) < 1e-6
    bb = _magnet.bbox_2d(g["att_Y"], -1)
    assert np.array_equal(np.array(bb, np.float32), g["bbox"])
