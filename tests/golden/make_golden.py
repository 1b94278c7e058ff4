"""Writes tests/golden/reference_f32.npz: input/output vectors of the reference's own native kernels.

The reference module is the reference's `_magnet` compiled from its C++ sources into oracle/_ref
(oracle/Makefile target `ref`); nothing of the reference's Python package is imported. The few lines
of Python glue the reference wraps around those kernels -- the grid decision (magnet/tsne.py:30-46),
the kernel spectra (magnet/tsne.py:48-71) and the fused optimiser loop (magnet/tsne.py:167-184) -- are
restated below. Every call uses n_threads=1, where the reference is deterministic.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_f32.npz")


def grid(R, Y, box_width=0.7, min_boxes=20, max_boxes=1000):
    """magnet/tsne.py:30-46 (with the native bbox_2d)."""
    xmin, ymin, xmax, ymax = R.bbox_2d(Y, 1)
    lo = np.array([xmin, ymin]) - 1e-3
    hi = np.array([xmax, ymax]) + 1e-3
    span = float((hi - lo).max())
    n_boxes = int(np.clip(np.ceil(span / box_width), min_boxes, max_boxes))
    n_boxes = min(max_boxes, -(-n_boxes // 4) * 4)
    width = box_width if n_boxes * box_width >= span else max(span / n_boxes, 1e-12)
    return float(lo[0]), float(lo[1]), width, n_boxes


def spectra(n_boxes, width, n_interp, dof=1.0):
    """magnet/tsne.py:55-70: K1, K2 on the padded torus, float32, and their rfft2 as complex64."""
    N = n_boxes * n_interp
    M = sfft.next_fast_len(2 * N, real=True)
    delta = width / n_interp
    off = np.arange(M)
    off = np.where(off <= M // 2, off, off - M).astype(np.float64) * delta
    r2 = off[:, None] ** 2 + off[None, :] ** 2
    base = 1.0 + r2 / dof
    K1 = (base ** (-dof)).astype(np.float32)
    K2 = (base ** (-dof - 1.0)).astype(np.float32)
    return M, np.ascontiguousarray(sfft.rfft2(K1), dtype=np.complex64), np.ascontiguousarray(sfft.rfft2(K2), dtype=np.complex64)


def repulsion(R, Y, n_interp=3, dof=1.0):
    xmin, ymin, width, nb = grid(R, Y)
    M, K1h, K2h = spectra(nb, width, n_interp, dof)
    conv = R.FFTConv2D(int(M), n_interp)
    conv.set_kernels(np.ascontiguousarray(K1h.real), np.ascontiguousarray(K2h.real))
    sumQ, F = conv.compute(Y, xmin, ymin, width, nb, 1)
    return (xmin, ymin, width, nb, M), float(sumQ), np.asarray(F)


def knn_csr(X, k, seed):
    """symmetric kNN affinities normalised to sum 1 (float32), exact neighbours by brute force"""
    n = len(X)
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(D, np.inf)
    nn = np.argsort(D, axis=1)[:, :k]
    w = np.exp(-np.take_along_axis(D, nn, 1) / np.median(D[np.arange(n)[:, None], nn]))
    A = sp.csr_matrix((w.ravel().astype(np.float32), (np.repeat(np.arange(n), k), nn.ravel())), shape=(n, n))
    P = (A + A.T).tocsr()
    P.sort_indices()
    P = sp.csr_matrix(P / P.sum(), dtype=np.float32)
    return P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data.astype(np.float32)


def main():
    R = oracle.ref_module()
    if R is None:
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    rng = np.random.default_rng(20240607)
    out = {}
    # --- repulsion: the reference test's layout (tests/test_magnet.py:221-231) and a clustered one
    Yr = np.ascontiguousarray(rng.normal(0, 8, (1500, 2)).astype(np.float32))
    for ni in (3, 5):
        g, s, F = repulsion(R, Yr, ni)
        out[f"rep{ni}_Y"] = Yr; out[f"rep{ni}_grid"] = np.array(g, np.float64); out[f"rep{ni}_sumQ"] = s; out[f"rep{ni}_F"] = F
    centres = rng.normal(0, 6, (8, 2))
    Yc = np.ascontiguousarray((centres[rng.integers(0, 8, 2500)] + rng.normal(0, 0.8, (2500, 2))).astype(np.float32))
    g, s, F = repulsion(R, Yc, 3)
    out["repc_Y"] = Yc; out["repc_grid"] = np.array(g, np.float64); out["repc_sumQ"] = s; out["repc_F"] = F
    # --- attraction (2-D and 3-D), step, update, spring
    n = 2000
    X = np.concatenate([rng.normal(c, 1.0, (n // 4, 10)) for c in rng.normal(0, 4, (4, 10))]).astype(np.float32)
    ip, ix, d = knn_csr(X, 12, 1)
    out["csr_indptr"], out["csr_indices"], out["csr_data"] = ip, ix, d
    Y = np.ascontiguousarray(rng.normal(0, 3, (n, 2)).astype(np.float32))
    Y3 = np.ascontiguousarray(rng.normal(0, 3, (n, 3)).astype(np.float32))
    out["att_Y"], out["att_Y3"] = Y, Y3
    out["att_out"] = np.asarray(R.tsne_attraction(ip, ix, d, Y, 1.0, 1))
    out["att_out3_dof07"] = np.asarray(R.tsne_attraction(ip, ix, d, Y3, 0.7, 1))
    F = rng.normal(0, 1, (n, 2)).astype(np.float32)
    U = rng.normal(0, 0.1, (n, 2)).astype(np.float32)
    G = rng.uniform(0.01, 2.0, (n, 2)).astype(np.float32)
    out["st_F"], out["st_U"], out["st_G"] = F, U, G
    Y1, U1, G1 = Y.copy(), U.copy(), G.copy()
    R.tsne_step(Y1, U1, G1, out["att_out"], F, 12.0, float(np.float32(1 / 3.7)), 250.0, 0.5, 1)
    out["step_Y"], out["step_U"], out["step_G"] = Y1, U1, G1
    grad = rng.normal(0, 1, (n, 2)).astype(np.float32)
    out["upd_grad"] = grad
    Y1, U1, G1 = Y.copy(), U.copy(), G.copy()
    R.tsne_update(Y1, U1, G1, grad, 250.0, 0.8, 1)
    out["upd_Y"], out["upd_U"], out["upd_G"] = Y1, U1, G1
    pairs = rng.integers(0, n, (n, 5)).astype(np.int32)
    g1 = grad.copy()
    R.mid_spring(Y, pairs, g1, 0.01, 1)
    out["spr_pairs"], out["spr_grad"] = pairs, g1
    out["bbox"] = np.array(R.bbox_2d(Y, 1), np.float32)
    # --- 40 iterations of the fused optimiser loop (magnet/tsne.py:148-184) from a tiny seeded init
    Y0 = np.ascontiguousarray(rng.normal(0, 1e-4, (n, 2)).astype(np.float32))
    out["traj_Y0"] = Y0
    Yt, Ut, Gt = Y0.copy(), np.zeros_like(Y0), np.ones_like(Y0)
    lr = max(200.0, n / 12.0)
    sq = []
    for it in range(40):
        early = it < 13
        rho, mom = (12.0, 0.5) if early else (1.0, 0.8)
        attr = np.asarray(R.tsne_attraction(ip, ix, d, Yt, 1.0, 1))
        _, sumQ, Fr = repulsion(R, Yt, 3)
        sq.append(sumQ)
        R.tsne_step(Yt, Ut, Gt, attr, Fr, float(rho), float(1.0 / max(sumQ, 1e-12)), float(lr), float(mom), 1)
        if it in (0, 9, 39):
            out[f"traj_Y{it + 1}"] = Yt.copy()
    out["traj_sumQ"] = np.array(sq)

    # fallback kernels (src/fftrep.cpp:51-125, 294-411), one thread
    rng = np.random.default_rng(31)
    Ys = np.ascontiguousarray(rng.normal(0, 2.0, (1500, 2)).astype(np.float32))
    xmin, ymin, width, nb = grid(R, Ys)
    q = np.ascontiguousarray(np.column_stack([np.ones(1500), Ys[:, 0], rng.normal(0, 1, 1500)]).astype(np.float32))
    out["shim_Y"], out["shim_q"], out["shim_grid"] = Ys, q, np.array([xmin, ymin, width, nb], np.float64)
    out["shim_scatter"] = np.asarray(R.interp_scatter(Ys, q, xmin, ymin, width, nb, 3, 1))
    Gg = np.ascontiguousarray(rng.normal(0, 1, out["shim_scatter"].shape).astype(np.float32))
    out["shim_G"], out["shim_gather"] = Gg, np.asarray(R.interp_gather(Ys, Gg, xmin, ymin, width, nb, 3, 1))
    Mp = 3 * nb + 4
    out["shim_M"] = np.array(Mp)
    out["shim_padded"] = np.asarray(R.scatter_padded_2d(Ys, xmin, ymin, width, nb, 3, Mp, 1))
    phi = np.ascontiguousarray(rng.normal(0, 1, (4, Mp, Mp)).astype(np.float32))
    sQ, Fs = R.gather_forces_2d(Ys, phi, xmin, ymin, width, nb, 3, 1)
    out["shim_phi"], out["shim_sumQ"], out["shim_F"] = phi, np.float64(sQ), np.asarray(Fs)
    Ms, Msh = 24, 13
    K1 = np.ascontiguousarray(rng.normal(0, 1, (Ms, Msh, 2)).astype(np.float32))
    K2 = np.ascontiguousarray(rng.normal(0, 1, (Ms, Msh, 2)).astype(np.float32))
    Wh = np.ascontiguousarray(rng.normal(0, 1, (3, Ms, Msh, 2)).astype(np.float32))
    out["shim_K1"], out["shim_K2"], out["shim_Wh"] = K1, K2, Wh
    out["shim_products"] = np.asarray(R.spectral_products(K1, K2, Wh, 1))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
