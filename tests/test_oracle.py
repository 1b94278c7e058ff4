"""The CPU oracle, pinned: against vectors the reference's own kernels wrote (tests/golden), against the
compiled reference itself when oracle/_ref is built, and against what the mathematics fixes (closed
forms, conservation laws, brute force, library FFTs). Every check here would fail for a plausible
slip in the restatement (a dropped term, a wrong sign or index, a transposed operand)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-300))


# ------------------------------------------------------------------------------ golden vectors
def test_attraction_bit_exact_to_reference(golden):
    g = golden
    out = O.attraction(g["csr_indptr"], g["csr_indices"], g["csr_data"], g["att_Y"], np.float32(1.0))
    assert np.array_equal(out, g["att_out"])
    out3 = O.attraction(g["csr_indptr"], g["csr_indices"], g["csr_data"], g["att_Y3"], np.float32(0.7))
    assert np.array_equal(out3, g["att_out3_dof07"])


def test_step_update_spring_bit_exact_to_reference(golden):
    g = golden
    Y, U, G = g["att_Y"].copy(), g["st_U"].copy(), g["st_G"].copy()
    O.step(Y, U, G, g["att_out"], g["st_F"], 12.0, np.float32(1 / 3.7), 250.0, 0.5)
    assert np.array_equal(Y, g["step_Y"]) and np.array_equal(U, g["step_U"]) and np.array_equal(G, g["step_G"])
    Y, U, G = g["att_Y"].copy(), g["st_U"].copy(), g["st_G"].copy()
    O.update(Y, U, G, g["upd_grad"], 250.0, 0.8)
    assert np.array_equal(Y, g["upd_Y"]) and np.array_equal(U, g["upd_U"]) and np.array_equal(G, g["upd_G"])
    gr = g["upd_grad"].copy()
    O.mid_spring(g["att_Y"], g["spr_pairs"], gr, np.float32(0.01))
    assert np.array_equal(gr, g["spr_grad"])


@pytest.mark.parametrize("case,ni", [("rep3", 3), ("rep5", 5), ("repc", 3)])
def test_repulsion_matches_reference_float32(golden, case, ni):
    Y = golden[case + "_Y"]
    gx = golden[case + "_grid"]
    # the same grid decision as the reference's (float32 bbox, double arithmetic)
    xmin, ymin, width, nb = O.grid(Y)
    assert (xmin, ymin, width, nb) == (gx[0], gx[1], gx[2], int(gx[3]))
    s, F = O.repulsion(Y, np.float32(xmin), np.float32(ymin), np.float32(width), nb, ni)
    # the reference transforms in float32 (pocketfft) on a 5-smooth padding; the oracle in double on
    # a power of two: the two agree to float32 rounding of the FFT
    assert abs(s - float(golden[case + "_sumQ"])) / float(golden[case + "_sumQ"]) < 2e-6
    assert rel(F, golden[case + "_F"]) < 5e-5


def test_trajectory_matches_reference_float32(golden):
    g = golden
    n = len(g["traj_Y0"])
    Y = g["traj_Y0"].copy()
    U, G = np.zeros_like(Y), np.ones_like(Y)
    sq = O.optimize(g["csr_indptr"], g["csr_indices"], g["csr_data"], Y, U, G, 0, 1, n_iter=40, early_iter=13,
                    learning_rate=max(200.0, n / 12.0))
    assert rel(Y, g["traj_Y1"]) < 1e-5
    O.optimize(g["csr_indptr"], g["csr_indices"], g["csr_data"], Y, U, G, 1, 10, n_iter=40, early_iter=13,
               learning_rate=max(200.0, n / 12.0))
    assert rel(Y, g["traj_Y10"]) < 1e-4
    sq2 = O.optimize(g["csr_indptr"], g["csr_indices"], g["csr_data"], Y, U, G, 10, 40, n_iter=40, early_iter=13,
                     learning_rate=max(200.0, n / 12.0))
    assert rel(Y, g["traj_Y40"]) < 1e-3
    assert abs(sq[0] - g["traj_sumQ"][0]) / g["traj_sumQ"][0] < 1e-5
    assert np.abs(sq2 / g["traj_sumQ"][10:] - 1).max() < 1e-3


def test_fallback_kernels_bit_exact_to_reference(golden):
    """interp_scatter / interp_gather / scatter_padded_2d / spectral_products / gather_forces_2d
    (src/fftrep.cpp:51-125, 294-411) against the reference's own outputs at n_threads = 1"""
    g = golden
    Y, q = g["shim_Y"], g["shim_q"]
    xmin, ymin, width, nb = [float(v) for v in g["shim_grid"][:3]] + [int(g["shim_grid"][3])]
    x0, y0, w = np.float32(xmin), np.float32(ymin), np.float32(width)
    assert np.array_equal(O.interp_scatter(Y, q, x0, y0, w, nb, 3), g["shim_scatter"])
    assert np.array_equal(O.interp_gather(Y, g["shim_G"], x0, y0, w, nb, 3), g["shim_gather"])
    assert np.array_equal(O.scatter_padded_2d(Y, x0, y0, w, nb, 3, int(g["shim_M"])), g["shim_padded"])
    assert np.array_equal(O.spectral_products(g["shim_K1"], g["shim_K2"], g["shim_Wh"]), g["shim_products"])
    s, F = O.gather_forces_2d(Y, g["shim_phi"], x0, y0, w, nb, 3)
    assert np.array_equal(F, g["shim_F"]) and np.float32(s) == np.float32(g["shim_sumQ"])


def test_fallback_kernels_math():
    """scatter: conservation of each charge column; gather is its adjoint; the padded scatter is the
    (1, x, y) scatter placed in the corner; products are complex products"""
    rng = np.random.default_rng(3)
    Y = rng.normal(0, 2, (700, 2))
    q = rng.normal(0, 1, (700, 4))
    args = (-7.0, -7.5, 0.7, 22, 3)
    Wg = O.interp_scatter(Y, q, *args)
    assert np.allclose(Wg.reshape(4, -1).sum(1), q.sum(0), rtol=1e-12, atol=1e-10)
    G = rng.normal(0, 1, Wg.shape)
    assert abs((Wg * G).sum() - (q * O.interp_gather(Y, G, *args)).sum()) < 1e-9 * np.abs(Wg * G).sum()
    P = O.scatter_padded_2d(Y, *args, 70)
    W3 = O.interp_scatter(Y, np.column_stack([np.ones(700), Y]), *args)
    assert np.allclose(P[:, :66, :66], W3, rtol=1e-12, atol=1e-12) and not P[:, 66:, :].any() and not P[:, :, 66:].any()
    K1, K2, Wh = rng.normal(0, 1, (10, 6, 2)), rng.normal(0, 1, (10, 6, 2)), rng.normal(0, 1, (3, 10, 6, 2))
    c = lambda a: a[..., 0] + 1j * a[..., 1]  # noqa: E731
    out = c(O.spectral_products(K1, K2, Wh))
    assert np.allclose(out[0], c(K1) * c(Wh)[0]) and np.allclose(out[1:], c(K2)[None] * c(Wh))


_REF_CHECK = r"""
import sys, numpy as np, scipy.sparse as sp
sys.path.insert(0, sys.argv[1])
import oracle as O
R = O.ref_module()
rng = np.random.default_rng(5)
Y = np.ascontiguousarray(rng.normal(0, 5, (3000, 2)).astype(np.float32))
assert np.allclose(R.bbox_2d(Y, 1), (Y[:, 0].min(), Y[:, 1].min(), Y[:, 0].max(), Y[:, 1].max()))
A = sp.random(3000, 3000, density=0.004, random_state=2, format="csr", dtype=np.float32)
A = (A + A.T).tocsr(); A.sort_indices(); A = sp.csr_matrix(A / A.sum(), dtype=np.float32)
ip, ix, d = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float32)
assert np.array_equal(np.asarray(R.tsne_attraction(ip, ix, d, Y, 1.0, 1)), O.attraction(ip, ix, d, Y, np.float32(1.0)))
F = rng.normal(0, 1, (3000, 2)).astype(np.float32); U = rng.normal(0, .1, (3000, 2)).astype(np.float32)
G = rng.uniform(.01, 2, (3000, 2)).astype(np.float32); A2 = rng.normal(0, 1, (3000, 2)).astype(np.float32)
Y1, U1, G1 = Y.copy(), U.copy(), G.copy(); R.tsne_step(Y1, U1, G1, A2, F, 4.0, 0.01, 300.0, 0.8, 1)
Y2, U2, G2 = Y.copy(), U.copy(), G.copy(); O.step(Y2, U2, G2, A2, F, 4.0, np.float32(0.01), 300.0, 0.8)
assert np.array_equal(Y1, Y2) and np.array_equal(U1, U2) and np.array_equal(G1, G2)
print("ok")
"""


def test_against_compiled_reference_when_built():
    import glob
    import os
    import subprocess
    import sys
    if not glob.glob(os.path.join(os.path.dirname(O.__file__), "_ref", "_magnet*.so")):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    r = subprocess.run([sys.executable, "-c", _REF_CHECK, os.path.dirname(O.__file__)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]


# ------------------------------------------------------------------------------ mathematics
@pytest.mark.parametrize("ni", [1, 2, 3, 4, 5, 8])
def test_lagrange_partition_reproduction_and_nodes(ni):
    nodes = (np.arange(ni) + 0.5) / ni
    for u in np.linspace(0, 1, 17):
        w = np.zeros(8)
        O.lib().or_lagrange_f64(u, ni, w)
        w = w[:ni]
        for p in range(ni):   # reproduces polynomials up to degree ni - 1 (p = 0: partition of unity)
            assert abs((w * nodes ** p).sum() - u ** p) < 1e-12
    for j, u in enumerate(nodes):
        w = np.zeros(8)
        O.lib().or_lagrange_f64(u, ni, w)
        assert np.allclose(w[:ni], np.eye(ni)[j], atol=1e-12)


def _layout(n=800, seed=0, scale=4.0):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, scale, (5, 2))
    return np.ascontiguousarray(c[rng.integers(0, 5, n)] + rng.normal(0, 1.0, (n, 2)))


def test_spread_conserves_charges_and_gather_is_its_adjoint():
    Y = _layout()
    xmin, ymin, width, nb = O.grid(Y)
    s, F, W, phi = O.repulsion(Y, xmin, ymin, width, nb, 3, grids=True)
    assert abs(W[0].sum() - len(Y)) < 1e-9
    assert abs(W[1].sum() - Y[:, 0].sum()) < 1e-9 * np.abs(Y).sum()
    assert abs(W[2].sum() - Y[:, 1].sum()) < 1e-9 * np.abs(Y).sum()
    # Z = sum_i sum_g w_i(g) phi0(g) = <W0, phi0>
    assert abs((W[0] * phi[0]).sum() - (s + len(Y))) < 1e-9 * (s + len(Y))


def test_point_on_a_node_charges_one_cell():
    nb, ni, width = 20, 3, 0.7
    xmin = ymin = 0.0
    Y = np.array([[xmin + width * (7 + 0.5 / 3), ymin + width * (11 + 2.5 / 3)]])
    _, _, W, _ = O.repulsion(Y, xmin, ymin, width, nb, ni, grids=True)
    nz = np.argwhere(np.abs(W[0]) > 1e-12)
    assert nz.tolist() == [[7 * 3 + 0, 11 * 3 + 2]]
    assert abs(W[0][21, 35] - 1) < 1e-12 and abs(W[1][21, 35] - Y[0, 0]) < 1e-12


def test_fft_convolution_equals_direct_sum_and_ignores_padding():
    Y = _layout(300, 1, 1.5)
    xmin, ymin, width, nb = O.grid(Y)
    assert nb == 20
    N = nb * 3
    s_d, F_d = O.repulsion(Y, xmin, ymin, width, nb, 3, direct=True)
    for M in (2 * N, 128, 256, 2 * N + 7):   # 5-smooth, powers of two, odd size: the same convolution
        s, F = O.repulsion(Y, xmin, ymin, width, nb, 3, M=M)
        assert abs(s - s_d) / s_d < 1e-12 and rel(F, F_d) < 1e-11


def test_kernel_spectra_match_numpy_and_are_real():
    M, delta = 64, 0.7 / 3
    K1, K2 = O.kernel_spectra(M, delta, 1.0)
    off = np.arange(M); off = np.where(off <= M // 2, off, off - M) * delta
    r2 = off[:, None] ** 2 + off[None, :] ** 2
    for Kh, K in ((K1, 1 / (1 + r2)), (K2, 1 / (1 + r2) ** 2)):
        full = np.fft.rfft2(K)
        assert np.abs(full.imag).max() < 1e-12 * np.abs(full.real).max()
        assert np.allclose(Kh, full.real, rtol=0, atol=1e-12 * np.abs(full.real).max())
    K1s, _ = O.kernel_spectra(M, delta, 0.6)   # heavier tail, dof != 1
    full = np.fft.rfft2((1 + r2 / 0.6) ** -0.6)
    assert np.allclose(K1s, full.real, atol=1e-11 * np.abs(full.real).max())


@pytest.mark.parametrize("ni,tz,tf", [(3, 1e-3, 2e-2), (5, 1e-4, 2e-3)])
def test_repulsion_matches_exact_sum(ni, tz, tf):
    # the reference's own acceptance test (tests/test_magnet.py:221-231: n_interp=5, 1e-3 / 2e-2)
    Y = np.random.default_rng(0).normal(0, 8, (1500, 2))
    Z, Fe = O.exact_repulsion(Y)
    s, F = O.repulsion(Y, *O.grid(Y), n_interp=ni)
    assert abs(s - Z) / Z < tz and rel(F, Fe) < tf


def test_repulsion_symmetries():
    Y = _layout(600, 3)
    s, F = O.repulsion(Y, *O.grid(Y))
    assert np.abs(F.sum(0)).max() < 1e-3 * np.abs(F).sum()   # Newton's third law, to the interpolation error
    Ys = np.ascontiguousarray(Y[:, ::-1])
    s2, F2 = O.repulsion(Ys, *O.grid(Ys))
    assert abs(s2 - s) / s < 1e-12 and rel(F2[:, ::-1], F) < 1e-11
    Yt = Y + np.array([0.5, -0.25])   # exact shift: the grid moves with the points
    s3, F3 = O.repulsion(Yt, *O.grid(Yt))
    assert abs(s3 - s) / s < 1e-9 and rel(F3, F) < 1e-9


def test_dof_kernels_match_exact_sum():
    Y = _layout(500, 4, 2.0)
    for dof in (0.5, 1.0, 2.0):
        Z, Fe = O.exact_repulsion(Y, dof)
        s, F = O.repulsion(Y, *O.grid(Y), n_interp=5, dof=dof)
        assert abs(s - Z) / Z < 1e-3 and rel(F, Fe) < 5e-3


def _dense_P(n, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.02, random_state=seed, format="csr")
    A = (A + A.T).tocsr(); A.setdiag(0); A.eliminate_zeros(); A.sort_indices()
    A = A / A.sum()
    return sp.csr_matrix(A), rng


def test_attraction_brute_force_and_zero_sum():
    P, rng = _dense_P(400, 7)
    Y = rng.normal(0, 2, (400, 2))
    out = O.attraction(P.indptr, P.indices, P.data, Y, 1.0)
    D = Y[:, None, :] - Y[None, :, :]
    Q = 1.0 / (1.0 + (D ** 2).sum(-1))
    ref = ((P.toarray() * Q)[:, :, None] * D).sum(1)
    assert rel(out, ref) < 1e-13
    assert np.abs(out.sum(0)).max() < 1e-12 * np.abs(out).sum()


def test_step_closed_form_and_recentring():
    rng = np.random.default_rng(3)
    n = 3000
    Y, U = rng.normal(0, 1, (n, 2)), rng.normal(0, 0.1, (n, 2))
    G, A, F = rng.uniform(0.01, 2, (n, 2)), rng.normal(0, 1, (n, 2)), rng.normal(0, 1, (n, 2))
    rho, inv, lr, mom = 4.0, 0.01, 200.0, 0.8
    g = rho * A - F * inv
    Gn = np.maximum(np.where((g > 0) == (U > 0), G * 0.8, G + 0.2), 0.01)
    Un = mom * U - lr * Gn * g
    Yn = Y + Un
    Yn -= Yn.mean(0)
    Y2, U2, G2 = Y.copy(), U.copy(), G.copy()
    O.step(Y2, U2, G2, A, F, rho, inv, lr, mom)
    assert np.allclose(G2, Gn, rtol=0, atol=1e-15) and np.allclose(U2, Un, rtol=1e-14) and np.allclose(Y2, Yn, atol=1e-12)
    assert np.abs(Y2.mean(0)).max() < 1e-13


def test_one_iteration_follows_the_exact_gradient():
    """Whole-loop restatement: one early-exaggeration step from a layout equals the closed-form update
    with the exact t-SNE gradient, to the interpolation error of the repulsion."""
    P, rng = _dense_P(500, 11)
    Y = rng.normal(0, 3, (500, 2))
    Y1 = Y.copy(); U = np.zeros_like(Y); G = np.ones_like(Y)
    O.optimize(P.indptr, P.indices, P.data, Y1, U, G, 0, 1, n_iter=10, early_iter=5, n_interp=5)
    D = Y[:, None, :] - Y[None, :, :]
    Q = 1.0 / (1.0 + (D ** 2).sum(-1)); np.fill_diagonal(Q, 0)
    Z = Q.sum()
    grad = 12.0 * ((P.toarray() * Q)[:, :, None] * D).sum(1) - ((Q ** 2)[:, :, None] * D).sum(1) / Z
    lr = max(200.0, 500 / 12.0)
    Gn = np.where(grad > 0, 1.2, 0.8)   # U = 0: (g > 0) == (U > 0) holds where g <= 0 -> 0.8, else 1 + 0.2
    Yn = Y - lr * Gn * grad
    Yn -= Yn.mean(0)
    assert rel(Y1 - Y, Yn - Y) < 1e-3


def test_kl_divergence_closed_form_and_gradient_of_the_oracle():
    """kl_divergence against the definition sum p log(p / q) on a tiny complete graph with the exact Z (dof 1 and
    0.7); gradient against the exact O(n^2) t-SNE gradient on a small layout at the interpolation error level."""
    rng = np.random.default_rng(5)
    n = 6
    Y = rng.normal(0, 2, (n, 2))
    W = rng.random((n, n)); W = W + W.T; np.fill_diagonal(W, 0); W /= W.sum()
    import scipy.sparse as sp
    P = sp.csr_matrix(W)
    for dof in (1.0, 0.7):
        D2 = ((Y[:, None] - Y[None]) ** 2).sum(-1)
        K = (1 + D2 / dof) ** (-dof); np.fill_diagonal(K, 0)
        Z = K.sum()
        m = W > 0
        exact = float((W[m] * np.log(W[m] / (K[m] / Z))).sum())
        assert abs(O.kl_divergence(P.indptr, P.indices, P.data, Y, Z, dof) - exact) < 1e-13
    # the gradient of the fused loop against exact sums (rho A - F / Z), away from the grid's interpolation error
    n = 400
    Y = rng.normal(0, 5, (n, 2))
    W = rng.random((n, n)) * (rng.random((n, n)) < 0.05); W = W + W.T; np.fill_diagonal(W, 0); W /= W.sum()
    P = sp.csr_matrix(W)
    g, sq = O.gradient(P.indptr, P.indices, P.data, Y, it=3, n_iter=10, early_iter=5)
    D2 = ((Y[:, None] - Y[None]) ** 2).sum(-1)
    Q = 1 / (1 + D2); np.fill_diagonal(Q, 0)
    A = ((W * Q)[:, :, None] * (Y[:, None] - Y[None])).sum(1)
    F = ((Q ** 2)[:, :, None] * (Y[:, None] - Y[None])).sum(1)
    exact = 12.0 * A - F / Q.sum()
    assert np.linalg.norm(g - exact) / np.linalg.norm(exact) < 1e-3 and abs(sq / Q.sum() - 1) < 1e-3


def _run_reference_check(name):
    """runs tests/ref_layout_checks.py `name` in a process of its own (the reference registers the module name
    `_magnet`, like the product)"""
    import glob
    import os
    import subprocess
    import sys
    if not glob.glob(os.path.join(os.path.dirname(O.__file__), "_ref", "_magnet*.so")):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_layout_checks.py")
    r = subprocess.run([sys.executable, script, name], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout[-2000:], r.stderr[-2000:])


def test_layout_oracle_equals_the_reference_owner_mode_when_nothing_is_sampled():
    """With no negative samples (negative_sample_rate 1e-6 puts the first one past the run) and no mid-range pairs
    the owner-mode epoch uses no random numbers: the oracle's layout equals the compiled reference's
    (optimize_layout, parallel=1, n_threads=1) bit for bit after 30 epochs."""
    _run_reference_check("exact")


def test_layout_oracle_matches_the_reference_in_law():
    """With negative samples and mid-range pairs (random streams per node here, per chunk in the reference) both
    layouts after 200 epochs separate the blobs (10-NN accuracy > 0.95, the reference's test_owner_parallel_mode) and
    have the same spread within 15%."""
    _run_reference_check("law")


# ---------------------------------------------------------------------------------------------- upstream rows
def test_perplexity_rows_hit_the_entropy_and_a_root_finders_beta():
    """perplexity_rows: every row's entropy is log(perplexity) within the tolerance, its p sums to 1, and its beta is
    the root of H(beta) = log(perplexity) that scipy's brentq finds (an independent method) to the tolerance's width."""
    from scipy.optimize import brentq
    from affinity_oracle import perplexity_rows
    rng = np.random.default_rng(1)
    d2 = np.sort(rng.gamma(2.0, 3.0, (200, 90)), axis=1).astype(np.float32)
    d2[5] = 0.0            # all-zero row: uniform p (any beta)
    d2[6, 1:] = 1e3        # a row whose mass can only sit on one neighbour
    for perp in (5.0, 30.0):
        Pc, beta = perplexity_rows(d2, perp, tol=1e-10, n_iter=200)
        assert np.allclose(Pc.sum(1), 1.0, rtol=0, atol=1e-12)
        H = -(Pc * np.log(np.maximum(Pc, 1e-300))).sum(1)
        ok = np.ones(200, bool); ok[[5, 6]] = False
        assert np.abs(H[ok] - np.log(perp)).max() < 1e-8
        assert np.allclose(Pc[5], 1 / 90)
        for i in (0, 17, 100, 199):
            d = d2[i].astype(np.float64) - d2[i].min()
            def f(b):
                w = np.exp(-d * b)
                return np.log(w.sum()) + b * (d * w).sum() / w.sum() - np.log(perp)
            root = brentq(f, 1e-8, 1e3, xtol=1e-14)
            assert abs(beta[i] / root - 1) < 1e-6, (i, beta[i], root)


def test_symmetrize_sum_is_p_plus_p_transpose_over_2n():
    """symmetrize_knn mode 0 against the dense (P + P^T) / 2n on the same values (ties, a value of zero, a repeated
    neighbour and a self entry included): the same entries and values (float32 of the float64 sum)."""
    from affinity_oracle import symmetrize_knn
    rng = np.random.default_rng(4)
    n, k = 120, 7
    idx = np.array([rng.choice(n, k, replace=False) for _ in range(n)])
    val = rng.random((n, k))
    val[3, 2] = 0.0
    idx[9, 1] = idx[9, 0]          # a neighbour listed twice: its values add up
    idx[10, 4] = 10                 # a self entry: dropped
    S = symmetrize_knn(idx, val, 0)
    D = np.zeros((n, n))
    np.add.at(D, (np.repeat(np.arange(n), k), idx.ravel()), val.ravel())
    D = (D + D.T) * (1.0 / (2.0 * n))
    np.fill_diagonal(D, 0.0)
    assert np.array_equal(S.toarray(), D.astype(np.float32))
    assert S.nnz == int((D != 0).sum()) and S.has_sorted_indices


def test_fuzzy_union_matches_the_dense_formula():
    """symmetrize_knn mode 1 against A + A^T - A * A^T (set_op_mix_ratio 1) and the mix with the intersection
    (0.3): symmetric, no diagonal, zeros dropped (magnet/knn.py:215-230)."""
    from affinity_oracle import symmetrize_knn
    rng = np.random.default_rng(2)
    n, k = 300, 10
    ind = np.array([rng.choice(np.delete(np.arange(n), i), k, replace=False) for i in range(n)])
    v = rng.random((n, k)).astype(np.float32)
    v[0, :3] = 0.0
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), k), ind.ravel()] = v.ravel().astype(np.float64)
    for r in (1.0, 0.3):
        P = symmetrize_knn(ind, v.astype(np.float64), 1, r).toarray()
        want = (r * (A + A.T - A * A.T) + (1 - r) * A * A.T).astype(np.float32)
        assert np.array_equal(P, want) and np.array_equal(P, P.T) and not np.diag(P).any()


def test_smooth_knn_hits_the_target_sum():
    """smooth_knn: the nearest non-zero neighbour has membership 1 and each row sums to log2(k) (magnet/knn.py:165-198)."""
    from affinity_oracle import smooth_knn
    rng = np.random.default_rng(1)
    d = np.sort(rng.uniform(0.1, 3.0, (500, 14)).astype(np.float32), axis=1)
    v = smooth_knn(d, 15)
    assert v.shape == d.shape and (v > 0).all() and (v <= 1).all()
    assert np.allclose(v[:, 0], 1.0)
    assert np.abs(v.sum(1) - np.log2(15)).max() < 1e-3


def test_mid_near_rule_and_its_hash():
    """mid_near_pairs: the draws come from splitmix64's finaliser (its published first output from state 0 is
    0xE220A8397B1DCDAF); every partner is a drawn row other than the row itself, and at most one drawn row is
    strictly nearer (the second nearest, src/nndescent.cpp:730-759)."""
    from affinity_oracle import _mix64, hash4, mid_near_pairs
    assert int(_mix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, 4)).astype(np.float32)
    out = mid_near_pairs(X, 3, 6, seed=11)
    for i in (0, 7, 299):
        for p in range(3):
            j = (hash4(11, i, p, np.arange(6)) % np.uint64(300)).astype(np.int64)
            j = np.where(j == i, (j + 1) % 300, j)
            d = ((X[j].astype(np.float64) - X[i]) ** 2).sum(1)
            partner = out[i, p]
            assert partner != i and partner in j
            dp = ((X[partner].astype(np.float64) - X[i]) ** 2).sum()
            assert (d < dp).sum() <= 1
