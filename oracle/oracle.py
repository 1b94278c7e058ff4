"""numpy front end of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only there, as the
checker. Each wrapper names the reference code it restates (kerighan/lib-magnet v3.0.2):
grid magnet/tsne.py:30-46; spectra magnet/tsne.py:48-71; repulsion src/fftconv.cpp:58-230;
attraction src/fftrep.cpp:129-177; step/update src/fftrep.cpp:181-266; spring src/fftrep.cpp:270-287;
optimiser magnet/tsne.py:132-214.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("n_iter", C.c_int32), ("early_iter", C.c_int32),
                ("early_exaggeration", C.c_double), ("exaggeration", C.c_double),
                ("learning_rate", C.c_double), ("momentum0", C.c_double), ("momentum1", C.c_double),
                ("n_interp", C.c_int32), ("min_boxes", C.c_int32), ("max_boxes", C.c_int32),
                ("box_width", C.c_double), ("dof", C.c_double), ("mid_weight", C.c_double),
                ("mid_phase", C.c_double)]


def build():
    """Compile liboracle.so (and the reference module into _ref/ when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def ref_module():
    """The reference's own _magnet compiled from its sources into oracle/_ref, or None.

    pybind11 keeps one module object per module name in a process, so this must run in a process that
    has not imported the product's `_magnet` (tests run it in a subprocess)."""
    import importlib.machinery
    import importlib.util
    import sys
    paths = glob.glob(os.path.join(HERE, "_ref", "_magnet*.so"))
    if not paths:
        return None
    if "_magnet" in sys.modules:
        raise RuntimeError("the product's _magnet is loaded in this process; load the reference in another process")
    loader = importlib.machinery.ExtensionFileLoader("_magnet", paths[0])
    spec = importlib.util.spec_from_file_location("_magnet", paths[0], loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


class LayoutParams(C.Structure):
    _fields_ = [("epoch_start", C.c_int32), ("epoch_end", C.c_int32), ("alpha0", C.c_float), ("alpha1", C.c_float),
                ("mid_w0", C.c_float), ("mid_w1", C.c_float), ("a", C.c_float), ("b", C.c_float), ("rep_a", C.c_float),
                ("rep_b", C.c_float), ("negative_sample_rate", C.c_float), ("exaggeration", C.c_float),
                ("repulsion_strength", C.c_float), ("mid_rate", C.c_float), ("clip_mode", C.c_int32),
                ("n_fixed", C.c_int64), ("seed", C.c_uint64), ("attraction_scale", C.c_float)]


def layout_epochs(Y, head, tail, eps, eons, eonns, epoch_start, epoch_end, alpha0, alpha1, mid_w0=0.0, mid_w1=0.0, a=1.0,
                  b=1.0, rep_a=0.0, rep_b=0.0, negative_sample_rate=5.0, exaggeration=1.0, repulsion_strength=1.0,
                  mid_rate=0.0, clip_mode=0, n_fixed=0, seed=0, pairs=None, rev_indptr=None, rev_indices=None,
                    attraction_scale=1.5):
    """Owner-mode sampled-layout epochs (src/layout.cpp:129-260) on float32 Y (n, 2), eons / eonns in place."""
    p = LayoutParams(int(epoch_start), int(epoch_end), alpha0, alpha1, mid_w0, mid_w1, a, b, rep_a, rep_b,
                     negative_sample_rate, exaggeration, repulsion_strength, mid_rate, int(clip_mode), int(n_fixed),
                     int(seed) & ((1 << 64) - 1), attraction_scale)
    keep = []
    def cp(x):
        if x is None:
            return None
        x = np.ascontiguousarray(x, np.int32)
        keep.append(x)
        return x.ctypes.data_as(C.c_void_p)
    lib().or_layout_epochs(C.byref(p), Y, len(Y), np.ascontiguousarray(head, np.int32), np.ascontiguousarray(tail, np.int32),
                           np.ascontiguousarray(eps, np.float32), eons, eonns, len(head), cp(pairs),
                           pairs.shape[1] if pairs is not None else 0, cp(rev_indptr), cp(rev_indices))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_grid_f64.argtypes = [_f64p, C.c_int64, C.c_double, C.c_int, C.c_int, _f64p]
        L.or_grid_f32.argtypes = [_f32p, C.c_int64, C.c_double, C.c_int, C.c_int, _f64p]
        L.or_fft_size.argtypes = [C.c_int]; L.or_fft_size.restype = C.c_int
        L.or_kernel_spectra.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _f64p, _f64p]
        for suf, fp, T in (("f64", _f64p, C.c_double), ("f32", _f32p, C.c_float)):
            getattr(L, "or_repulsion_" + suf).argtypes = [fp, C.c_int64, T, T, T, C.c_int, C.c_int, C.c_int, C.c_double,
                                                         C.c_void_p, C.c_void_p, fp, C.c_void_p, C.c_void_p, C.c_int]
            getattr(L, "or_repulsion_" + suf).restype = C.c_double
            getattr(L, "or_attraction_" + suf).argtypes = [_i32p, _i32p, fp, fp, C.c_int64, C.c_int, T, fp]
            getattr(L, "or_step_" + suf).argtypes = [fp, fp, fp, fp, fp, C.c_int64, C.c_int, T, T, T, T]
            getattr(L, "or_update_" + suf).argtypes = [fp, fp, fp, fp, C.c_int64, C.c_int, T, T]
            getattr(L, "or_mid_spring_" + suf).argtypes = [fp, _i32p, fp, C.c_int64, C.c_int, C.c_int, T]
            getattr(L, "or_optimize_" + suf).argtypes = [C.POINTER(Params), _i32p, _i32p, fp, C.c_int64, fp, fp, fp,
                                                        C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        for suf, fp, T in (("f64", _f64p, C.c_double), ("f32", _f32p, C.c_float)):
            getattr(L, "or_interp_scatter_" + suf).argtypes = [fp, C.c_int64, fp, C.c_int, T, T, T, C.c_int, C.c_int, fp]
            getattr(L, "or_interp_gather_" + suf).argtypes = [fp, C.c_int64, fp, C.c_int, T, T, T, C.c_int, C.c_int, fp]
            getattr(L, "or_scatter_padded_" + suf).argtypes = [fp, C.c_int64, T, T, T, C.c_int, C.c_int, C.c_int, fp]
            getattr(L, "or_spectral_products_" + suf).argtypes = [fp, fp, fp, C.c_int64, C.c_int64, fp]
            getattr(L, "or_gather_forces_" + suf).argtypes = [fp, C.c_int64, fp, C.c_int, T, T, T, C.c_int, C.c_int, fp]
            getattr(L, "or_gather_forces_" + suf).restype = C.c_double
        L.or_layout_epochs.argtypes = [C.POINTER(LayoutParams), _f32p, C.c_int64, _i32p, _i32p, _f32p, _f32p, _f32p,
                                       C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.or_exact_repulsion.argtypes = [_f64p, C.c_int64, C.c_double, _f64p]
        L.or_exact_repulsion.restype = C.c_double
        L.or_lagrange_f64.argtypes = [C.c_double, C.c_int, _f64p]
        _LIB = L
    return _LIB


def _suffix(a):
    return "f64" if a.dtype == np.float64 else "f32"


def grid(Y, box_width=0.7, min_boxes=20, max_boxes=1000):
    """(xmin, ymin, width, n_boxes) of FFTRepulsion2D._grid; bbox in Y's dtype, the rest in double."""
    out = np.zeros(4)
    getattr(lib(), "or_grid_" + _suffix(Y))(np.ascontiguousarray(Y), len(Y), box_width, min_boxes, max_boxes, out)
    return float(out[0]), float(out[1]), float(out[2]), int(out[3])


def fft_size(N):
    return lib().or_fft_size(int(N))


def kernel_spectra(M, delta, dof=1.0, round_f32=False):
    Mh = M // 2 + 1
    K1 = np.zeros((M, Mh)); K2 = np.zeros((M, Mh))
    lib().or_kernel_spectra(int(M), float(delta), float(dof), int(round_f32), K1, K2)
    return K1, K2


def repulsion(Y, xmin, ymin, width, n_boxes, n_interp=3, M=0, dof=1.0, spectra=None, grids=False, direct=False):
    """(sumQ, F) -- and (W (3,N,N), phi (4,N,N)) when grids=True -- in Y's precision."""
    Y = np.ascontiguousarray(Y)
    n = len(Y)
    F = np.zeros_like(Y)
    N = n_boxes * n_interp
    W = np.zeros((3, N, N), Y.dtype) if grids else None
    phi = np.zeros((4, N, N), Y.dtype) if grids else None
    k1 = k2 = None
    if spectra is not None:
        k1 = np.ascontiguousarray(spectra[0], dtype=np.float64); k2 = np.ascontiguousarray(spectra[1], dtype=np.float64)
        M = k1.shape[0]
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
    s = getattr(lib(), "or_repulsion_" + _suffix(Y))(Y, n, xmin, ymin, width, int(n_boxes), int(n_interp), int(M),
                                                      float(dof), ptr(k1), ptr(k2), F, ptr(W), ptr(phi), int(direct))
    return (s, F, W, phi) if grids else (s, F)


def exact_repulsion(Y, dof=1.0):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    F = np.zeros_like(Y)
    Z = lib().or_exact_repulsion(Y, len(Y), float(dof), F)
    return Z, F


def attraction(indptr, indices, data, Y, dof=1.0):
    Y = np.ascontiguousarray(Y)
    out = np.zeros_like(Y)
    getattr(lib(), "or_attraction_" + _suffix(Y))(np.ascontiguousarray(indptr, np.int32), np.ascontiguousarray(indices, np.int32),
                                                   np.ascontiguousarray(data, Y.dtype), Y, Y.shape[0], Y.shape[1], dof, out)
    return out


def step(Y, U, G, attr, F, rho, inv_sumQ, lr, mom):
    """In place on Y, U, G (tsne_step)."""
    getattr(lib(), "or_step_" + _suffix(Y))(Y, U, G, np.ascontiguousarray(attr, Y.dtype), np.ascontiguousarray(F, Y.dtype),
                                             Y.shape[0], Y.shape[1], rho, inv_sumQ, lr, mom)


def update(Y, U, G, grad, lr, mom):
    """In place on Y, U, G (tsne_update)."""
    getattr(lib(), "or_update_" + _suffix(Y))(Y, U, G, np.ascontiguousarray(grad, Y.dtype), Y.shape[0], Y.shape[1], lr, mom)


def mid_spring(Y, pairs, grad, w):
    """In place on grad."""
    getattr(lib(), "or_mid_spring_" + _suffix(Y))(Y, np.ascontiguousarray(pairs, np.int32), grad, Y.shape[0], Y.shape[1],
                                                   pairs.shape[1], w)


def optimize(indptr, indices, data, Y, U=None, G=None, it0=0, it1=None, n_iter=750, early_iter=250,
             early_exaggeration=12.0, exaggeration=1.0, learning_rate=None, momentum=(0.5, 0.8), n_interp=3,
             box_width=0.7, dof=1.0, pairs=None, mid_weight=0.0, mid_phase=0.5, min_boxes=20, max_boxes=1000):
    """optimize_tsne on iterations [it0, it1) with data already normalised; returns per-iteration sumQ."""
    p = Params(n_iter, early_iter, early_exaggeration, exaggeration, learning_rate or 0.0, momentum[0], momentum[1],
               n_interp, min_boxes, max_boxes, box_width, dof, mid_weight, mid_phase)
    it1 = n_iter if it1 is None else it1
    if U is None:
        U = np.zeros_like(Y)
    if G is None:
        G = np.ones_like(Y)
    sq = np.zeros(max(0, it1 - it0))
    pr = np.ascontiguousarray(pairs, np.int32) if pairs is not None else None
    getattr(lib(), "or_optimize_" + _suffix(Y))(C.byref(p), np.ascontiguousarray(indptr, np.int32),
                                                 np.ascontiguousarray(indices, np.int32), np.ascontiguousarray(data, Y.dtype),
                                                 len(Y), Y, U, G, pr.ctypes.data_as(C.c_void_p) if pr is not None else None,
                                                 pr.shape[1] if pr is not None else 0, it0, it1, sq.ctypes.data_as(C.c_void_p))
    return sq


# ---- per-iteration diagnostics (no counterpart in the reference; the KL divergence t-SNE minimises, and the norm of
# the gradient the loop applies), in float64 numpy
def kl_divergence(indptr, indices, data, Y, Z, dof=1.0):
    """KL(P || Q) with q_ij = (1 + |y_i - y_j|^2 / dof)^-dof / Z over the CSR entries of P (P normalised, zeros
    skipped): sum p log p - sum p log q = sum p log p + dof sum p log1p(d^2 / dof) + (sum p) log Z."""
    ip = np.asarray(indptr, np.int64)
    rows = np.repeat(np.arange(len(ip) - 1), np.diff(ip))
    p = np.asarray(data, np.float64)
    Yd = np.asarray(Y, np.float64)
    d2 = ((Yd[rows] - Yd[np.asarray(indices, np.int64)]) ** 2).sum(1)
    nz = p > 0
    return float((p[nz] * np.log(p[nz])).sum() + dof * (p * np.log1p(d2 / dof)).sum() + p.sum() * np.log(Z))


def gradient(indptr, indices, data, Y, it, n_iter=750, early_iter=250, early_exaggeration=12.0, exaggeration=1.0,
             n_interp=3, box_width=0.7, dof=1.0, min_boxes=20, max_boxes=1000):
    """(gradient rho A - F / sumQ of the fused loop at iteration it (magnet/tsne.py:167-184), sumQ) for layout Y."""
    Yd = np.ascontiguousarray(Y, np.float64)
    A = attraction(indptr, indices, np.asarray(data, np.float64), Yd, dof)
    sq, F = repulsion(Yd, *grid(Yd, box_width, min_boxes, max_boxes), n_interp=n_interp, dof=dof)
    rho = early_exaggeration if it < early_iter else exaggeration
    return rho * A - F / max(sq, 1e-12), sq


# ---- fallback kernels (src/fftrep.cpp:24-125, 294-411); arrays in Y's dtype
def interp_scatter(Y, charges, xmin, ymin, width, n_boxes, n_interp):
    """(C, N, N) charge grids, points summed in index order."""
    Y = np.ascontiguousarray(Y); q = np.ascontiguousarray(charges, Y.dtype)
    N = n_boxes * n_interp
    out = np.zeros((q.shape[1], N, N), Y.dtype)
    getattr(lib(), "or_interp_scatter_" + _suffix(Y))(Y, len(Y), q, q.shape[1], xmin, ymin, width, n_boxes, n_interp, out)
    return out


def interp_gather(Y, grid_, xmin, ymin, width, n_boxes, n_interp):
    Y = np.ascontiguousarray(Y); G = np.ascontiguousarray(grid_, Y.dtype)
    out = np.zeros((len(Y), G.shape[0]), Y.dtype)
    getattr(lib(), "or_interp_gather_" + _suffix(Y))(Y, len(Y), G, G.shape[0], xmin, ymin, width, n_boxes, n_interp, out)
    return out


def scatter_padded_2d(Y, xmin, ymin, width, n_boxes, n_interp, M):
    Y = np.ascontiguousarray(Y)
    out = np.zeros((3, M, M), Y.dtype)
    getattr(lib(), "or_scatter_padded_" + _suffix(Y))(Y, len(Y), xmin, ymin, width, n_boxes, n_interp, M, out)
    return out


def spectral_products(K1h, K2h, Wh):
    """complex arrays as float pairs: K (M, Mh, 2), Wh (3, M, Mh, 2) -> (4, M, Mh, 2)"""
    K1h = np.ascontiguousarray(K1h); K2h = np.ascontiguousarray(K2h, K1h.dtype); Wh = np.ascontiguousarray(Wh, K1h.dtype)
    M, Mh = K1h.shape[0], K1h.shape[1]
    out = np.zeros((4, M, Mh, 2), K1h.dtype)
    getattr(lib(), "or_spectral_products_" + _suffix(K1h))(K1h, K2h, Wh, M, Mh, out)
    return out


def gather_forces_2d(Y, phi, xmin, ymin, width, n_boxes, n_interp):
    """(sumQ, F) from the padded potentials (4, M, M)."""
    Y = np.ascontiguousarray(Y); phi = np.ascontiguousarray(phi, Y.dtype)
    F = np.zeros_like(Y)
    s = getattr(lib(), "or_gather_forces_" + _suffix(Y))(Y, len(Y), phi, phi.shape[1], xmin, ymin, width, n_boxes, n_interp, F)
    return s, F
