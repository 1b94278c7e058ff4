"""Host-side checks that need no GPU: the C-ABI library loads and exports every symbol the header
declares, the pybind module keeps the reference's names, and the host logic (grid decision, affinities)
follows the reference."""
import ctypes
import os

import numpy as np
import pytest

from magnet import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    syms = capi.header_symbols()
    assert len(syms) >= 20
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_pybind_module_keeps_reference_names():
    import _magnet
    for name in ("FFTConv2D", "tsne_attraction", "tsne_step", "tsne_update", "mid_spring", "bbox_2d", "tsne_optimize",
                 "interp_scatter", "interp_gather", "scatter_padded_2d", "spectral_products", "gather_forces_2d"):
        assert hasattr(_magnet, name), name
    assert hasattr(_magnet.FFTConv2D, "set_kernels") and hasattr(_magnet.FFTConv2D, "compute")


def test_shape_errors_raise_before_any_device_call():
    import _magnet
    with pytest.raises(RuntimeError, match="indptr must have n \\+ 1 entries"):
        _magnet.tsne_attraction(np.zeros(3, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros((4, 2), np.float32))
    with pytest.raises(RuntimeError, match="dof must be positive"):
        _magnet.tsne_attraction(np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros((4, 2), np.float32), 0.0)
    with pytest.raises(RuntimeError, match="arrays must match Y"):
        _magnet.tsne_update(np.zeros((4, 2), np.float32), np.zeros((4, 2), np.float32), np.zeros((3, 2), np.float32),
                            np.zeros((4, 2), np.float32), 1.0, 0.5)
    with pytest.raises(RuntimeError, match="Y must be \\(n, 2\\)"):
        _magnet.bbox_2d(np.zeros((0, 2), np.float32))
    with pytest.raises(RuntimeError, match="charges must be \\(n, T\\)"):
        _magnet.interp_scatter(np.zeros((4, 2), np.float32), np.zeros((3, 2), np.float32), 0.0, 0.0, 0.7, 20, 3)
    with pytest.raises(RuntimeError, match="grid must be \\(T, N, N\\)"):
        _magnet.interp_gather(np.zeros((4, 2), np.float32), np.zeros((2, 59, 60), np.float32), 0.0, 0.0, 0.7, 20, 3)
    with pytest.raises(RuntimeError, match="bad grid parameters"):
        _magnet.scatter_padded_2d(np.zeros((4, 2), np.float32), 0.0, 0.0, 0.7, 20, 3, 50)
    with pytest.raises(RuntimeError, match="bad shapes"):
        _magnet.spectral_products(np.zeros((8, 5, 2), np.float32), np.zeros((8, 4, 2), np.float32), np.zeros((3, 8, 5, 2), np.float32))


def test_grid_decision_matches_oracle():
    import oracle as O
    from magnet.tsne import grid_params
    rng = np.random.default_rng(0)
    for scale in (1e-4, 1.0, 30.0, 400.0):
        Y = (rng.normal(0, scale, (500, 2))).astype(np.float32)
        bb = (Y[:, 0].min(), Y[:, 1].min(), Y[:, 0].max(), Y[:, 1].max())
        assert grid_params(bb) == O.grid(Y)


def test_fft_size_choice():
    """smallest 2^s or 3 * 2^s >= 2N, from 128 up to 8192 (no device needed)"""
    from magnet import capi
    sizes = [128 * 2 ** k for k in range(7)] + [192 * 2 ** k for k in range(6)]
    for N in [1, 60, 64, 65, 96, 97, 500, 1020, 1536, 2049, 3000, 3072, 3073, 4096]:
        want = min(m for m in sizes if m >= 2 * N)
        assert capi.fft_size(N) == want, N
    assert capi.fft_size(4097) == 0


def test_batch_entry_rejects_empty_and_bad_stream_counts():
    L = capi.lib()
    p = capi.Params(10, 3, 12.0, 1.0, 0.0, 0.5, 0.8, 3, 20, 1000, 0.7, 1.0, 0.0, 0.5)
    arr = (capi.Instance * 1)()
    assert L.mg_tsne_optimize_batch(capi.MG_F32, ctypes.byref(p), 0, arr, 1, 1) != 0
    assert b"no instances" in L.mg_last_error()
    assert L.mg_tsne_optimize_batch(capi.MG_F32, ctypes.byref(p), 1, arr, 0, 1) != 0
    assert b"n_streams" in L.mg_last_error()
    assert L.mg_tsne_optimize_batch(capi.MG_F32, ctypes.byref(p), 1, arr, 2, 3) != 0
    assert b"n_threads" in L.mg_last_error()
    assert L.mg_tsne_optimize_batch(7, ctypes.byref(p), 1, arr, 1, 1) != 0


def test_knn_entries_reject_bad_arguments():
    L = capi.lib()
    i64 = ctypes.c_int64
    assert L.mg_knn(None, i64(10), None, i64(3), 4, 2, i64(-1), None, None) != 0     # null arrays
    assert b"bad k-NN arguments" in L.mg_last_error()
    assert L.mg_knn(None, i64(10), None, i64(0), 4, 2, i64(-1), None, None) == 0     # no queries: nothing to do
    assert L.mg_topk_rows(None, i64(0), i64(5), 9, None, None) != 0                  # k above the row length
    assert b"k must be" in L.mg_last_error()


def test_optimize_rejects_a_grid_beyond_the_device_fft_up_front():
    """n_interp * max_boxes above 4096 nodes per side is refused before any device work (the reference's default of
    1000 boxes with 5 nodes would otherwise fail only when the layout first spans 820 boxes)."""
    L = capi.lib()
    i64 = ctypes.c_int64
    fake = ctypes.c_void_p(16)
    for ni, mb in ((5, 1000), (3, 1366), (8, 513)):
        p = capi.Params(10, 3, 12.0, 1.0, 0.0, 0.5, 0.8, ni, 20, mb, 0.7, 1.0, 0.0, 0.5)
        rc = L.mg_tsne_optimize(capi.MG_F32, ctypes.byref(p), fake, fake, fake, i64(1), i64(1), fake, fake, fake, None, 0, 0, 10,
                                None, capi.CALLBACK(), None)
        assert rc == 4 and b"exceeds 4096" in L.mg_last_error(), (ni, mb)
        arr = (capi.Instance * 1)()
        assert L.mg_tsne_optimize_batch(capi.MG_F32, ctypes.byref(p), 1, arr, 1, 1) == 4
