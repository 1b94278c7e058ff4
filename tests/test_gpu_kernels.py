"""GPU parity of every kernel of the pass, through the C-ABI, against the CPU oracle on the same inputs.

Tolerances (stated per test): FP64 kernels match the FP64 oracle to rounding (relative <= 1e-12; the
only differences are summation order and fused multiply-adds). FP32 kernels match the reference's own
float32 outputs (tests/golden) to float32 rounding of a different summation order / FFT."""
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


def dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def layout(n, seed, scale=4.0, k=6, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, scale, (k, 2))
    return np.ascontiguousarray(c[rng.integers(0, k, n)] + rng.normal(0, spread, (n, 2)))


# ------------------------------------------------------------------------------ point kernels
def test_bbox_exact():
    Y = layout(100_003, 1)
    assert capi.bbox(dev(Y)) == [Y[:, 0].min(), Y[:, 1].min(), Y[:, 0].max(), Y[:, 1].max()]
    Yf = Y.astype(np.float32)
    assert capi.bbox(dev(Yf)) == [float(v) for v in (Yf[:, 0].min(), Yf[:, 1].min(), Yf[:, 0].max(), Yf[:, 1].max())]


def test_attraction_fp64_and_fp32(golden):
    g = golden
    ip, ix = dev(g["csr_indptr"]), dev(g["csr_indices"])
    # FP64 vs FP64 oracle: summation order only (relative 1e-13)
    Y = g["att_Y"].astype(np.float64)
    ref = O.attraction(g["csr_indptr"], g["csr_indices"], g["csr_data"].astype(np.float64), Y, 1.0)
    out = host(capi.attraction(ip, ix, dev(g["csr_data"].astype(np.float64)), dev(Y), 1.0))
    assert rel(out, ref) < 1e-13
    # FP32 vs the reference's float32 output (relative 1e-6)
    out = host(capi.attraction(ip, ix, dev(g["csr_data"]), dev(g["att_Y"]), 1.0))
    assert rel(out, g["att_out"]) < 1e-6
    out3 = host(capi.attraction(ip, ix, dev(g["csr_data"]), dev(g["att_Y3"]), 0.7))
    assert rel(out3, g["att_out3_dof07"]) < 1e-6


@pytest.mark.parametrize("deg", [1, 3, 20, 150])
def test_attraction_row_widths(deg):
    # rows of every lane-group width (1..64 lanes per row), including empty rows
    rng = np.random.default_rng(deg)
    n = 5000
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, n, n * deg)
    keep = rng.random(n) < 0.9
    import scipy.sparse as sp
    P = sp.csr_matrix((rng.random(n * deg), (rows, cols)), shape=(n, n))
    P = sp.csr_matrix(sp.diags(keep.astype(float)) @ P)
    P.sort_indices()
    Y = rng.normal(0, 3, (n, 2))
    ref = O.attraction(P.indptr, P.indices, P.data, Y, 1.3)
    out = host(capi.attraction(dev(P.indptr.astype(np.int32)), dev(P.indices.astype(np.int32)), dev(P.data), dev(Y), 1.3))
    assert rel(out, ref) < 1e-13


def test_step_update_spring_fp64():
    rng = np.random.default_rng(3)
    n = 70_001
    Y, U = rng.normal(0, 1, (n, 2)), rng.normal(0, 0.1, (n, 2))
    G, A, F = rng.uniform(0.01, 2, (n, 2)), rng.normal(0, 1, (n, 2)), rng.normal(0, 1, (n, 2))
    Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
    O.step(Yo, Uo, Go, A, F, 12.0, 1 / 3.7, 5833.3, 0.5)
    dY, dU, dG = dev(Y), dev(U), dev(G)
    capi.step(dY, dU, dG, dev(A), dev(F), 12.0, 1 / 3.7, 5833.3, 0.5)
    assert rel(host(dY), Yo) < 1e-14 and rel(host(dU), Uo) < 1e-14 and np.array_equal(host(dG), Go)
    Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
    O.update(Yo, Uo, Go, A, 200.0, 0.8)
    dY, dU, dG = dev(Y), dev(U), dev(G)
    capi.update(dY, dU, dG, dev(A), 200.0, 0.8)
    assert rel(host(dY), Yo) < 1e-14 and rel(host(dU), Uo) < 1e-14 and np.array_equal(host(dG), Go)
    pairs = rng.integers(0, n, (n, 5)).astype(np.int32)
    go = A.copy(); O.mid_spring(Y, pairs, go, 0.003)
    dg = dev(A); capi.mid_spring(dev(Y), dev(pairs), dg, 0.003)
    assert rel(host(dg), go) < 1e-15


def test_step_update_fp32_against_reference(golden):
    g = golden
    dY, dU, dG = dev(g["att_Y"]), dev(g["st_U"]), dev(g["st_G"])
    capi.step(dY, dU, dG, dev(g["att_out"]), dev(g["st_F"]), 12.0, float(np.float32(1 / 3.7)), 250.0, 0.5)
    assert rel(host(dY), g["step_Y"]) < 1e-6 and rel(host(dU), g["step_U"]) < 1e-6
    assert np.mean(host(dG) == g["step_G"]) > 0.999   # gains: the sign test may flip only where g ~ 0
    dY, dU, dG = dev(g["att_Y"]), dev(g["st_U"]), dev(g["st_G"])
    capi.update(dY, dU, dG, dev(g["upd_grad"]), 250.0, 0.8)
    assert rel(host(dY), g["upd_Y"]) < 1e-6 and np.array_equal(host(dG), g["upd_G"])
    gr = dev(g["upd_grad"]); capi.mid_spring(dev(g["att_Y"]), dev(g["spr_pairs"]), gr, 0.01)
    assert rel(host(gr), g["spr_grad"]) < 1e-6


def test_update_any_dim():
    rng = np.random.default_rng(9)
    n, d = 3001, 5
    Y, U, G, D = rng.normal(0, 1, (n, d)), rng.normal(0, 0.1, (n, d)), rng.uniform(0.1, 2, (n, d)), rng.normal(0, 1, (n, d))
    Yo, Uo, Go = Y.copy(), U.copy(), G.copy()
    O.update(Yo, Uo, Go, D, 100.0, 0.8)
    dY, dU, dG = dev(Y), dev(U), dev(G)
    capi.update(dY, dU, dG, dev(D), 100.0, 0.8)
    assert rel(host(dY), Yo) < 1e-13 and np.array_equal(host(dG), Go)


# ------------------------------------------------------------------------------ repulsion
def rep_gpu(Y, ni=3, dof=1.0, grids=False):
    import torch
    prec = capi.MG_F64 if Y.dtype == np.float64 else capi.MG_F32
    xmin, ymin, width, nb = O.grid(Y)
    conv = capi.FFTConv(prec, nb * ni * 2, ni)
    conv.set_tsne(width, dof)
    dY = dev(Y)
    s, F = conv.compute(dY, xmin, ymin, width, nb)
    out = (s, host(F))
    if grids:
        W, phi = conv.grids(nb * ni, torch.float64 if prec == capi.MG_F64 else torch.float32)
        out = out + (host(W), host(phi))
    return out


@pytest.mark.parametrize("n,seed,scale", [(1500, 0, 8.0), (20_000, 1, 10.0), (70_000, 2, 25.0)])
def test_repulsion_fp64_matches_oracle(n, seed, scale):
    Y = layout(n, seed, scale, k=10)
    s, F, W, phi = rep_gpu(Y, grids=True)
    so, Fo, Wo, phio = O.repulsion(Y, *O.grid(Y), n_interp=3, grids=True)
    assert rel(W, Wo) < 1e-13                                   # spread: summation order only
    assert rel(phi.transpose(2, 0, 1), phio) < 1e-12            # FFT convolution
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-11     # gather, forces, Z


@pytest.mark.parametrize("ni", [1, 2, 4, 5, 6, 8])
def test_repulsion_fp64_interp_orders(ni):
    Y = layout(3000, 5, 3.0)
    xmin, ymin, width, nb = O.grid(Y)
    s, F = rep_gpu(Y, ni)
    so, Fo = O.repulsion(Y, xmin, ymin, width, nb, ni)
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-11


@pytest.mark.parametrize("dof", [0.5, 2.0])
def test_repulsion_fp64_dof(dof):
    Y = layout(4000, 6, 3.0)
    s, F = rep_gpu(Y, 3, dof)
    so, Fo = O.repulsion(Y, *O.grid(Y), n_interp=3, dof=dof)
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-11


def test_repulsion_fp32_against_reference(golden):
    for case in ("rep3", "repc"):
        Y = golden[case + "_Y"]
        s, F = rep_gpu(Y)
        # float32 spread / FFT / gather against the reference's float32 pocketfft path
        assert abs(s - float(golden[case + "_sumQ"])) / float(golden[case + "_sumQ"]) < 1e-5
        assert rel(F, golden[case + "_F"]) < 2e-4
    Y = golden["rep5_Y"]
    s, F = rep_gpu(Y, 5)
    assert rel(F, golden["rep5_F"]) < 2e-4


@pytest.mark.parametrize("n,scale", [(20_000, 25.0), (70_000, 60.0)])
def test_repulsion_fp32_against_fp64_oracle(n, scale):
    """FP32 path (node-relative charges, FP64 Z) against the FP64 oracle on the same float32 layout:
    sumQ to 2e-6 and F to 2e-6 relative (the reference's float32 path is at ~1e-5 on such layouts)."""
    Y = layout(n, 11, scale, k=12, spread=0.6).astype(np.float32)
    s, F = rep_gpu(Y)
    # the FP32 path works on the float-rounded grid (origin, box width): give the oracle the same grid
    xmin, ymin, width, nb = O.grid(Y)
    g32 = [float(np.float32(v)) for v in (xmin, ymin, width)]
    so, Fo = O.repulsion(Y.astype(np.float64), *g32, nb, n_interp=3)
    assert abs(s - so) / so < 2e-6 and rel(F, Fo) < 2e-6, (abs(s - so) / so, rel(F, Fo))


def test_repulsion_matches_exact_sum_reference_tolerance():
    # tests/test_magnet.py:221-231 of the reference, on the GPU: n_interp=5, sumQ 1e-3, F 2e-2
    Y = np.random.default_rng(0).normal(0, 8, (1500, 2))
    Z, Fe = O.exact_repulsion(Y)
    for dt in (np.float64, np.float32):
        s, F = rep_gpu(Y.astype(dt), 5)
        assert abs(s - Z) / Z < 1e-3 and rel(F, Fe) < 2e-2


def test_repulsion_fp64_clipped_grid_two_sort_passes():
    """a layout wider than 1000 boxes of 0.7: n_boxes clips at 1000 and the width grows to span / 1000
    (magnet/tsne.py:40-45); 10^6 box keys take two 10-bit radix passes. n_interp = 1 keeps N = 1000."""
    rng = np.random.default_rng(12)
    c = rng.uniform(-450, 450, (40, 2))
    Y = c[rng.integers(0, 40, 50_000)] + rng.normal(0, 2.0, (50_000, 2))
    xmin, ymin, width, nb = O.grid(Y)
    assert nb == 1000 and width > 0.7
    s, F = rep_gpu(Y, 1)
    so, Fo = O.repulsion(Y, xmin, ymin, width, nb, 1)
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-10


def test_repulsion_fp64_more_than_1024_boxes():
    """2048 boxes per side through the FFTConv2D call (n_interp 1): 2^22 box keys take the 8-bit digit sort
    (three passes)"""
    rng = np.random.default_rng(21)
    nb = 2048
    Y = rng.uniform(0.5, nb * 0.7 - 0.5, (20_000, 2))
    conv = capi.FFTConv(capi.MG_F64, 2 * nb, 1)
    conv.set_tsne(0.7)
    s, F = conv.compute(dev(Y), 0.0, 0.0, 0.7, nb)
    so, Fo = O.repulsion(Y, 0.0, 0.0, 0.7, nb, 1)
    assert abs(s - so) / so < 1e-12 and rel(host(F), Fo) < 1e-11


# one grid per compiled transform size M (2^s and 3 * 2^s, the smallest >= 2N): (n_interp, n_boxes)
FFT_SIZE_GRIDS = [(128, 3, 20), (192, 3, 32), (256, 3, 40), (384, 3, 64), (512, 3, 84), (768, 3, 128),
                  (1024, 3, 168), (1536, 3, 256), (2048, 3, 340), (3072, 3, 512), (4096, 3, 680),
                  (6144, 3, 1000)]


@pytest.mark.parametrize("M,ni,nb", FFT_SIZE_GRIDS)
def test_repulsion_every_fft_size(M, ni, nb):
    """each plan (first radix 3 .. 32, then radix 16) in all five FFT kernels, FP64 and FP32, against the
    oracle on a grid of that size: points spread over the whole grid, width 0.7."""
    N = ni * nb
    assert capi.fft_size(N) == M
    rng = np.random.default_rng(M)
    span = nb * 0.7
    Y = rng.uniform(0.05, span - 0.05, (4000, 2))
    Y[:2000] = rng.normal(span / 2, span / 10, (2000, 2)).clip(0.05, span - 0.05)
    xmin = ymin = 0.0
    out = {}
    for prec, dt in ((capi.MG_F64, np.float64), (capi.MG_F32, np.float32)):
        conv = capi.FFTConv(prec, 2 * N, ni)
        conv.set_tsne(0.7)
        s, F = conv.compute(dev(Y.astype(dt)), xmin, ymin, 0.7, nb)
        out[dt] = (s, host(F))
    so, Fo = O.repulsion(Y, xmin, ymin, 0.7, nb, ni)
    s, F = out[np.float64]
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-11, (abs(s - so) / so, rel(F, Fo))
    s, F = out[np.float32]
    so32, Fo32 = O.repulsion(Y.astype(np.float32).astype(np.float64), xmin, ymin, float(np.float32(0.7)), nb, ni)
    assert abs(s - so32) / so32 < 2e-6 and rel(F, Fo32) < 2e-6, (abs(s - so32) / so32, rel(F, Fo32))


@pytest.mark.parametrize("M,ni,nb", [g for g in FFT_SIZE_GRIDS if g[0] in (512, 768, 1024, 1536, 3072, 6144)] + [(8192, 4, 1000)]
                         + [(1536, 3, 255), (3072, 3, 511), (6144, 3, 999)])
def test_repulsion_fft_layouts_identical(M, ni, nb):
    """The three layouts of the 2-D FFT's middle -- the column pass in the row layout (strided), row passes that write
    and read the column layout themselves (folded, 4 or 2 rows per block, its column stride padded to 4 rows; the
    inverse pass persistent or one block per row group), row passes and transposes -- give the same sumQ and forces bit
    for bit, FP32 and FP64, at the sizes where each applies (grid sides N that are and are not multiples of 4)."""
    N = ni * nb
    assert capi.fft_size(N) == M
    rng = np.random.default_rng(M + 1)
    span = nb * 0.7
    Y = rng.uniform(0.05, span - 0.05, (3000, 2))
    layouts = [dict(fft_strided_max=8192, fft_fold=4), dict(fft_strided_max=0, fft_fold=0),
               dict(fft_strided_max=0, fft_fold=4), dict(fft_strided_max=0, fft_fold=2),
               dict(fft_strided_max=0, fft_fold=4, fft_fold_persist=0), dict(fft_strided_max=0, fft_fold=2, fft_fold_persist=0),
               dict(fft_strided_max=0, fft_fold=4, fft_fold_persist=1), dict(fft_strided_max=0, fft_fold=4, fft_fold_persist=2)]
    saved = {k: capi.tune(k) for k in ("fft_strided_max", "fft_fold", "fft_fold_persist")}
    try:
        for prec, dt in ((capi.MG_F64, np.float64), (capi.MG_F32, np.float32)):
            outs = []
            for lay in layouts:
                for k, v in lay.items():
                    capi.tune(k, v)
                conv = capi.FFTConv(prec, 2 * N, ni)
                conv.set_tsne(0.7)
                s, F = conv.compute(dev(Y.astype(dt)), 0.0, 0.0, 0.7, nb)
                outs.append((s, host(F)))
            for s, F in outs[1:]:
                assert s == outs[0][0] and np.array_equal(F, outs[0][1]), (M, dt)
    finally:
        for k, v in saved.items():
            capi.tune(k, v)


def test_repulsion_largest_fft_size():
    """M = 8192 (n_interp 4, 1000 boxes): FP64 against the exact O(n^2) sum at the interpolation error level
    of tests/test_magnet.py:221-231, and FP32 against FP64 on the same float32 layout."""
    rng = np.random.default_rng(15)
    nb, ni = 1000, 4
    assert capi.fft_size(nb * ni) == 8192
    span = nb * 0.7
    Yf = rng.uniform(1.0, span - 1.0, (3000, 2)).astype(np.float32)
    out = {}
    for prec, dt in ((capi.MG_F64, np.float64), (capi.MG_F32, np.float32)):
        conv = capi.FFTConv(prec, 2 * nb * ni, ni)
        conv.set_tsne(0.7)
        s, F = conv.compute(dev(Yf.astype(dt)), 0.0, 0.0, 0.7, nb)
        out[dt] = (s, host(F))
    Z, Fe = O.exact_repulsion(Yf.astype(np.float64))
    s64, F64 = out[np.float64]
    assert abs(s64 - Z) / Z < 1e-3 and rel(F64, Fe) < 2e-2
    s32, F32 = out[np.float32]
    assert abs(s32 - s64) / s64 < 2e-6 and rel(F32, F64) < 5e-6


def test_repulsion_largest_grid_fp32_against_fp64():
    """N = 3000 (n_boxes clipped at 1000, n_interp 3), M = 6144: FP32 (centred products) against FP64
    (reference products) on one float32 layout."""
    rng = np.random.default_rng(14)
    c = rng.uniform(-500, 500, (60, 2))
    Yf = (c[rng.integers(0, 60, 70_000)] + rng.normal(0, 3.0, (70_000, 2))).astype(np.float32)
    xmin, ymin, width, nb = O.grid(Yf)
    assert nb == 1000
    g = [float(np.float32(v)) for v in (xmin, ymin, width)]
    out = {}
    for prec, dt in ((capi.MG_F64, np.float64), (capi.MG_F32, np.float32)):
        conv = capi.FFTConv(prec, 6000, 3)
        conv.set_tsne(g[2])
        out[dt] = conv.compute(dev(Yf.astype(dt)), g[0], g[1], g[2], nb)
    s64, F64 = out[np.float64]
    s32, F32 = out[np.float32]
    assert abs(s32 - s64) / s64 < 2e-6 and rel(host(F32), host(F64)) < 5e-6


def test_repulsion_boxes_spanning_many_chunks():
    # all points in a few boxes (the early-exaggeration layout): boxes cross thousands of spread chunks
    rng = np.random.default_rng(7)
    Y = rng.normal(0, 1e-3, (300_000, 2))
    Y[:1000] += 0.9
    s, F = rep_gpu(Y)
    so, Fo = O.repulsion(Y, *O.grid(Y), n_interp=3)
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-10


def test_repulsion_edge_cases():
    # one point; points exactly on box boundaries and grid nodes; duplicates
    s, F = rep_gpu(np.array([[0.3, -0.2]]))
    so, Fo = O.repulsion(np.array([[0.3, -0.2]]), *O.grid(np.array([[0.3, -0.2]])))
    assert abs(s - so) <= 1e-12 * max(1, abs(so)) and np.allclose(F, Fo, atol=1e-12)
    g = np.arange(-10, 11) * 0.7
    Y = np.array([(a, b) for a in g for b in g[::2]], dtype=np.float64)
    Y = np.concatenate([Y, Y[:50]])
    s, F = rep_gpu(Y)
    so, Fo = O.repulsion(Y, *O.grid(Y))
    assert abs(s - so) / so < 1e-12 and rel(F, Fo) < 1e-11


def test_repulsion_is_deterministic():
    Y = layout(200_000, 8, 20.0, k=30)
    s1, F1 = rep_gpu(Y.astype(np.float32))
    s2, F2 = rep_gpu(Y.astype(np.float32))
    assert s1 == s2 and np.array_equal(F1, F2)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_custom_kernel_values_equal_tsne_kernels(prec):
    """A kernel given by its values (the form FFTConv2D.set_kernels hands over) against the oracle's t-SNE kernels:
    FP64 to rounding; FP32 to 1e-5, which needs the node spacing of each compute's grid for the centred product
    set's odd tables (a value table does not carry it); two layouts of different grid spacing."""
    for scale in (3.0, 9.0):
        Y = layout(5000, 9, scale)
        xmin, ymin, width, nb = O.grid(Y)
        N = nb * 3
        M = 2 * N + 6   # a 5-smooth-like size, resampled onto the device torus
        off = np.arange(M); off = np.where(off <= M // 2, off, off - M) * (width / 3)
        r2 = off[:, None] ** 2 + off[None, :] ** 2
        conv = capi.FFTConv(capi.MG_F64 if prec == "fp64" else capi.MG_F32, M, 3)
        conv.set_values(1 / (1 + r2), 1 / (1 + r2) ** 2)
        Yd = dev(Y if prec == "fp64" else Y.astype(np.float32))
        s, F = conv.compute(Yd, xmin, ymin, width, nb)
        so, Fo = O.repulsion(Y if prec == "fp64" else Y.astype(np.float32).astype(np.float64), xmin, ymin, width, nb, 3)
        tol_s, tol_F = (1e-12, 1e-11) if prec == "fp64" else (1e-5, 1e-5)
        assert abs(s - so) / so < tol_s and rel(host(F), Fo) < tol_F, (scale, s, so, rel(host(F), Fo))


def test_errors_are_reported():
    import torch
    ip = dev(np.array([0, 2, 1], np.int32)); ix = dev(np.array([0, 1], np.int32))
    with pytest.raises(capi.MgError, match="non-decreasing"):
        capi.attraction(ip, ix, dev(np.ones(2)), dev(np.zeros((2, 2))))
    ip = dev(np.array([0, 1, 2], np.int32)); ix = dev(np.array([0, 5], np.int32))
    with pytest.raises(capi.MgError, match="column index out of range"):
        capi.attraction(ip, ix, dev(np.ones(2)), dev(np.zeros((2, 2))))
    conv = capi.FFTConv(capi.MG_F64, 60, 3)
    conv.set_tsne(0.7)
    with pytest.raises(capi.MgError, match="grid larger than the buffer"):
        conv.compute(dev(np.zeros((4, 2))), 0.0, 0.0, 0.7, 40)
    with pytest.raises(capi.MgError):
        capi.FFTConv(capi.MG_F64, 60, 9)
    torch.cuda.synchronize()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_windowed_attraction_equals_the_plain_sweep(dtype):
    """The LDS-window sweep (every window size, block shape and rows per block, and rows only) gives the plain
    sweep's result bit for bit: rows of 0-300 edges, neighbours near the row (inside the window) and anywhere."""
    import torch
    rng = np.random.default_rng(11)
    n = 300_000
    lens = rng.integers(0, 300, n)
    lens[rng.integers(0, n, 2000)] = 0
    ip = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=ip[1:])
    rows = np.repeat(np.arange(n), lens)
    near = rows + rng.integers(-3000, 3000, rows.size)
    far = rng.integers(0, n, rows.size)
    ix = np.where(rng.random(rows.size) < 0.7, near, far).clip(0, n - 1).astype(np.int32)
    data = rng.random(rows.size).astype(dtype)
    Y = dev(rng.normal(0, 20, (n, 2)).astype(dtype))
    dip, dix, dd = dev(ip.astype(np.int32)), dev(ix), dev(data)
    saved = {k: capi.tune(k) for k in ("attr_window", "attr_rows", "attr_win_kb", "attr_tpb")}
    try:
        capi.tune("attr_window", 0)
        ref = capi.attraction(dip, dix, dd, Y, 0.8)
        for kb, tpb, rows_pb in ((64, 1024, 256), (128, 1024, 1024), (64, 512, 64), (128, 512, 300), (0, 256, 64)):
            capi.tune("attr_window", 1); capi.tune("attr_win_kb", kb); capi.tune("attr_tpb", tpb); capi.tune("attr_rows", rows_pb)
            out = capi.attraction(dip, dix, dd, Y, 0.8)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (kb, tpb, rows_pb)
    finally:
        for k, v in saved.items():
            capi.tune(k, v)
