"""ctypes binding of libmagnet_hip.so, the C-ABI of include/magnet_hip.h, for device-resident use.

The pybind module `_magnet` copies host arrays in and out on every call, like the reference's module;
callers whose data already sits in HBM (the benchmark, serving, the parity tests) call the C-ABI
directly through this binding. Device memory comes from torch tensors (allocation and streams only):
every compute call is a C-ABI call, and the library runs on torch's current stream.
"""
import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libmagnet_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "magnet_hip.h")

MG_F32, MG_F64 = 0, 1
_lib = None


class Params(C.Structure):
    _fields_ = [("n_iter", C.c_int32), ("early_iter", C.c_int32), ("early_exaggeration", C.c_double),
                ("exaggeration", C.c_double), ("learning_rate", C.c_double), ("momentum0", C.c_double),
                ("momentum1", C.c_double), ("n_interp", C.c_int32), ("min_boxes", C.c_int32), ("max_boxes", C.c_int32),
                ("box_width", C.c_double), ("dof", C.c_double), ("mid_weight", C.c_double), ("mid_phase", C.c_double),
                ("relabel_every", C.c_int32)]


class LayoutParams(C.Structure):
    _fields_ = [("epoch_start", C.c_int32), ("epoch_end", C.c_int32), ("alpha0", C.c_float), ("alpha1", C.c_float),
                ("mid_w0", C.c_float), ("mid_w1", C.c_float), ("a", C.c_float), ("b", C.c_float), ("rep_a", C.c_float),
                ("rep_b", C.c_float), ("negative_sample_rate", C.c_float), ("exaggeration", C.c_float),
                ("repulsion_strength", C.c_float), ("mid_rate", C.c_float), ("clip_mode", C.c_int32),
                ("n_fixed", C.c_int64), ("seed", C.c_uint64), ("attraction_scale", C.c_float)]


class Instance(C.Structure):
    _fields_ = [("indptr", C.c_void_p), ("indices", C.c_void_p), ("data", C.c_void_p), ("nnz", C.c_int64),
                ("n", C.c_int64), ("Y", C.c_void_p), ("U", C.c_void_p), ("G", C.c_void_p),
                ("h_sumQ", C.POINTER(C.c_double))]


CALLBACK = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p)


def header_symbols():
    """Names of the functions include/magnet_hip.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mg_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmagnet_hip.so is not built (make -C lib-magnet_amd)")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
        sig = {
            "mg_last_error": ([], C.c_char_p), "mg_version": ([], i32), "mg_set_stream": ([vp], i32), "mg_set_thread_stream": ([vp], i32),
            "mg_synchronize": ([], i32), "mg_mem_info": ([C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], i32),
            "mg_bbox_2d": ([i32, vp, i64, C.POINTER(C.c_double)], i32),
            "mg_profile_enable": ([i32], i32),
            "mg_tune_set": ([C.c_char_p, i32], i32), "mg_tune_get": ([C.c_char_p, C.POINTER(C.c_int)], i32), "mg_fft_size": ([i32], i32),
            "mg_profile_grid": ([C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)], i32), "mg_profile_reset": ([], i32), "mg_profile_num_stages": ([], i32),
            "mg_profile_sized": ([i32, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_double)], i32),
            "mg_profile_get": ([i32, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                C.POINTER(C.c_double)], i32),
            "mg_tsne_attraction": ([i32, vp, vp, vp, i64, vp, i64, i32, d, vp], i32),
            "mg_tsne_step": ([i32, vp, vp, vp, vp, vp, i64, i32, d, d, d, d], i32),
            "mg_tsne_update": ([i32, vp, vp, vp, vp, i64, i32, d, d], i32),
            "mg_mid_spring": ([i32, vp, vp, vp, i64, i32, i32, d], i32),
            "mg_fftconv_create": ([C.POINTER(vp), i32, i32, i32], i32), "mg_fftconv_destroy": ([vp], i32),
            "mg_fftconv_set_kernel_values": ([vp, i32, vp, vp], i32),
            "mg_fftconv_set_kernels_tsne": ([vp, d, d], i32),
            "mg_fftconv_compute": ([vp, vp, i64, d, d, d, i32, vp, C.POINTER(C.c_double)], i32),
            "mg_fftconv_grids": ([vp, vp, vp], i32),
            "mg_tsne_transform": ([vp, vp, i64, vp, vp, i32, vp, i64, i32, d, d, d, i32, d, d], i32),
            "mg_interp_scatter": ([i32, vp, i64, vp, i32, d, d, d, i32, i32, vp], i32),
            "mg_interp_gather": ([i32, vp, i64, vp, i32, d, d, d, i32, i32, vp], i32),
            "mg_scatter_padded_2d": ([i32, vp, i64, d, d, d, i32, i32, i32, vp], i32),
            "mg_spectral_products": ([i32, vp, vp, vp, i64, i64, vp], i32),
            "mg_gather_forces_2d": ([i32, vp, i64, vp, i32, d, d, d, i32, i32, vp, C.POINTER(C.c_double)], i32),
            "mg_tsne_optimize": ([i32, C.POINTER(Params), vp, vp, vp, i64, i64, vp, vp, vp, vp, i32, i32, i32,
                                  C.POINTER(C.c_double), CALLBACK, vp], i32),
            "mg_tsne_session_create": ([C.POINTER(vp), i32, C.POINTER(Params), vp, vp, vp, i64, i64, vp, i32], i32),
            "mg_tsne_session_run": ([vp, vp, vp, vp, i32, i32, C.POINTER(C.c_double), CALLBACK, vp], i32),
            "mg_tsne_session_destroy": ([vp], i32),
            "mg_tsne_session_trace": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], i32),
            "mg_optimize_layout": ([C.POINTER(LayoutParams), vp, i64, i32, vp, vp, vp, vp, vp, i64, vp, i32, vp, vp], i32),
            "mg_tsne_optimize_batch": ([i32, C.POINTER(Params), i32, C.POINTER(Instance), i32, i32], i32),
            "mg_tsne_batch_stats": ([C.POINTER(C.c_int64)] * 6 + [C.POINTER(C.c_double)], i32),
            "mg_topk_rows": ([vp, i64, i64, i32, vp, vp], i32),
            "mg_knn": ([vp, i64, vp, i64, i32, i32, i64, vp, vp], i32),
            "mg_knn_noalloc": ([vp, i64, vp, i64, i32, i32, i64, vp, vp], i32),
            "mg_smooth_knn": ([vp, i64, i32, i32, vp, vp], i32),
            "mg_perplexity_rows": ([vp, i64, i32, C.c_double, C.c_double, i32, vp, vp], i32),
            "mg_nn_descent": ([vp, i64, i32, i32, i32, i32, C.c_float, i32, i32, C.c_uint64, i32, vp, vp], i32),
            "mg_mid_near_pairs": ([vp, i64, i32, i32, i32, C.c_uint64, vp], i32),
            "mg_knn_exact_graph": ([vp, i64, i32, i32, vp, vp], i32),
            "mg_symmetrize_knn": ([i32, vp, vp, i64, i32, C.c_double, vp, vp, vp, i64, C.POINTER(C.c_int64)], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class MgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise MgError(rc, lib().mg_last_error().decode())


def use_torch_stream():
    import torch
    check(lib().mg_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream)))


def use_thread_stream(stream):
    """The calling thread's library calls go to this torch stream (the other threads keep theirs)."""
    check(lib().mg_set_thread_stream(C.c_void_p(stream.cuda_stream)))


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def prec_of(t):
    import torch
    return MG_F64 if t.dtype == torch.float64 else MG_F32


# ---- thin wrappers on torch device tensors
def bbox(Y):
    out = (C.c_double * 4)()
    check(lib().mg_bbox_2d(prec_of(Y), ptr(Y), Y.shape[0], out))
    return list(out)


def attraction(indptr, indices, data, Y, dof=1.0):
    import torch
    out = torch.empty_like(Y)
    check(lib().mg_tsne_attraction(prec_of(Y), ptr(indptr), ptr(indices), ptr(data), indices.numel(), ptr(Y),
                                   Y.shape[0], Y.shape[1], dof, ptr(out)))
    return out


def step(Y, U, G, A, F, rho, inv, lr, mom):
    check(lib().mg_tsne_step(prec_of(Y), ptr(Y), ptr(U), ptr(G), ptr(A), ptr(F), Y.shape[0], Y.shape[1], rho, inv, lr, mom))


def update(Y, U, G, grad, lr, mom):
    check(lib().mg_tsne_update(prec_of(Y), ptr(Y), ptr(U), ptr(G), ptr(grad), Y.shape[0], Y.shape[1], lr, mom))


def mid_spring(Y, pairs, grad, w):
    check(lib().mg_mid_spring(prec_of(Y), ptr(Y), ptr(pairs), ptr(grad), Y.shape[0], Y.shape[1], pairs.shape[1], w))


class FFTConv:
    def __init__(self, prec, M, n_interp=3):
        self.h = C.c_void_p()
        self.prec = prec
        check(lib().mg_fftconv_create(C.byref(self.h), prec, M, n_interp))
        self._destroy = lib().mg_fftconv_destroy   # kept: module globals may be gone when the interpreter exits

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "_destroy", None):
            self._destroy(self.h)
            self.h = None

    def set_tsne(self, width, dof=1.0):
        check(lib().mg_fftconv_set_kernels_tsne(self.h, width, dof))

    def set_values(self, K1, K2):
        import numpy as np
        K1 = np.ascontiguousarray(K1, np.float64); K2 = np.ascontiguousarray(K2, np.float64)
        check(lib().mg_fftconv_set_kernel_values(self.h, K1.shape[0], K1.ctypes.data_as(C.c_void_p),
                                                  K2.ctypes.data_as(C.c_void_p)))

    def compute(self, Y, xmin, ymin, width, n_boxes):
        import torch
        F = torch.empty_like(Y)
        s = C.c_double()
        check(lib().mg_fftconv_compute(self.h, ptr(Y), Y.shape[0], xmin, ymin, width, n_boxes, ptr(F), C.byref(s)))
        return s.value, F

    def transform(self, Y_train, nbr, w, Y_new, n_iter=100, learning_rate=200.0, momentum=(0.5, 0.8), momentum_switch=25,
                  exaggeration=1.0, dof=1.0):
        """In place on Y_new (m, 2): mg_tsne_transform against the field of the last compute (on Y_train)."""
        check(lib().mg_tsne_transform(self.h, ptr(Y_train), Y_train.shape[0], ptr(nbr), ptr(w), nbr.shape[1], ptr(Y_new),
                                      Y_new.shape[0], n_iter, learning_rate, momentum[0], momentum[1], momentum_switch,
                                      exaggeration, dof))

    def grids(self, N, dtype):
        import torch
        W = torch.empty((3, N, N), dtype=dtype, device="cuda")
        phi = torch.empty((N, N, 4), dtype=dtype, device="cuda")
        check(lib().mg_fftconv_grids(self.h, ptr(W), ptr(phi)))
        return W, phi


def optimize(indptr, indices, data, Y, U, G, it0=0, it1=None, n_iter=750, early_iter=250, early_exaggeration=12.0,
             exaggeration=1.0, learning_rate=0.0, momentum=(0.5, 0.8), n_interp=3, box_width=0.7, dof=1.0, pairs=None,
             mid_weight=0.0, mid_phase=0.5, min_boxes=20, max_boxes=1000, callback=None, relabel_every=0):
    """Runs mg_tsne_optimize on device tensors; returns per-iteration sumQ (numpy)."""
    import numpy as np
    it1 = n_iter if it1 is None else it1
    p = Params(n_iter, early_iter, early_exaggeration, exaggeration, learning_rate, momentum[0], momentum[1], n_interp,
               min_boxes, max_boxes, box_width, dof, mid_weight, mid_phase, relabel_every)
    sq = np.zeros(max(1, it1 - it0))
    cb = CALLBACK(callback) if callback else CALLBACK(0)
    check(lib().mg_tsne_optimize(prec_of(Y), C.byref(p), ptr(indptr), ptr(indices), ptr(data), indices.numel(), Y.shape[0],
                                 ptr(Y), ptr(U), ptr(G), ptr(pairs), pairs.shape[1] if pairs is not None else 0, it0, it1,
                                 sq.ctypes.data_as(C.POINTER(C.c_double)), cb, None))
    return sq[: max(0, it1 - it0)]


class Session:
    """mg_tsne_session: one CSR P (device tensors kept alive by the object) and one schedule; run(Y, U, G, it0, it1)
    advances a state in place and returns the per-iteration sumQ; buffers, streams and spectra persist between runs."""

    def __init__(self, indptr, indices, data, prec=None, n_iter=750, early_iter=250, early_exaggeration=12.0,
                 exaggeration=1.0, learning_rate=0.0, momentum=(0.5, 0.8), n_interp=3, box_width=0.7, dof=1.0, pairs=None,
                 mid_weight=0.0, mid_phase=0.5, min_boxes=20, max_boxes=1000, relabel_every=0):
        self.h = C.c_void_p()
        self._keep = (indptr, indices, data, pairs)
        self.n_iter = n_iter
        self.p = Params(n_iter, early_iter, early_exaggeration, exaggeration, learning_rate, momentum[0], momentum[1],
                        n_interp, min_boxes, max_boxes, box_width, dof, mid_weight, mid_phase, relabel_every)
        prec = prec_of(data) if prec is None else prec
        check(lib().mg_tsne_session_create(C.byref(self.h), prec, C.byref(self.p), ptr(indptr), ptr(indices), ptr(data),
                                           indices.numel(), indptr.numel() - 1, ptr(pairs),
                                           pairs.shape[1] if pairs is not None else 0))

    def run(self, Y, U, G, it0=0, it1=None, callback=None, trace=False):
        """Iterations [it0, it1) on the state; returns sumQ per iteration, or (sumQ, KL, gradient norm) with trace."""
        import numpy as np
        it1 = self.n_iter if it1 is None else it1
        m = max(1, it1 - it0)
        sq = np.zeros(m)
        cb = CALLBACK(callback) if callback else CALLBACK(0)
        dp = C.POINTER(C.c_double)
        kl, gn = (np.zeros(m), np.zeros(m)) if trace else (None, None)
        check(lib().mg_tsne_session_trace(self.h, kl.ctypes.data_as(dp) if trace else None,
                                          gn.ctypes.data_as(dp) if trace else None))
        try:
            check(lib().mg_tsne_session_run(self.h, ptr(Y), ptr(U), ptr(G), int(it0), int(it1), sq.ctypes.data_as(dp), cb, None))
        finally:
            if trace:
                check(lib().mg_tsne_session_trace(self.h, None, None))
        k = max(0, it1 - it0)
        return (sq[:k], kl[:k], gn[:k]) if trace else sq[:k]

    def close(self):
        if self.h:
            lib().mg_tsne_session_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def optimize_batch(instances, n_iter=750, early_iter=250, early_exaggeration=12.0, exaggeration=1.0, learning_rate=0.0,
                   momentum=(0.5, 0.8), n_interp=3, box_width=0.7, dof=1.0, min_boxes=20, max_boxes=1000, n_streams=4,
                   n_threads=1, sumq=False):
    """Runs mg_tsne_optimize_batch on a list of (indptr, indices, data, Y, U, G) device-tensor tuples, all of one
    dtype; the states are updated in place. Returns the per-instance sumQ arrays when sumq is set."""
    import numpy as np
    if not instances:
        raise ValueError("no instances")
    prec = prec_of(instances[0][3])
    if any(prec_of(t[3]) != prec for t in instances):
        raise ValueError("all instances must have one dtype")
    p = Params(n_iter, early_iter, early_exaggeration, exaggeration, learning_rate, momentum[0], momentum[1], n_interp,
               min_boxes, max_boxes, box_width, dof, 0.0, 0.5)
    arr = (Instance * len(instances))()
    sq = [np.zeros(max(1, n_iter)) for _ in instances] if sumq else None
    for b, (ip, ix, dv, Y, U, G) in enumerate(instances):
        if Y.dim() != 2 or Y.shape[1] != 2 or U.shape != Y.shape or G.shape != Y.shape:
            raise ValueError("Y, U, G must be (n, 2)")
        if ip.numel() != Y.shape[0] + 1 or ix.numel() != dv.numel():
            raise ValueError("CSR arrays do not match the layout")
        arr[b] = Instance(ip.data_ptr(), ix.data_ptr(), dv.data_ptr(), ix.numel(), Y.shape[0], Y.data_ptr(),
                          U.data_ptr(), G.data_ptr(), sq[b].ctypes.data_as(C.POINTER(C.c_double)) if sumq else None)
    check(lib().mg_tsne_optimize_batch(prec, C.byref(p), len(instances), arr, int(n_streams), int(n_threads)))
    return [s[:n_iter] for s in sq] if sumq else None


def batch_stats():
    """{plans, iterations, replays, launches, checks, stalls, plan_s} of the last optimize_batch call
    (mg_tsne_batch_stats)."""
    v = [C.c_int64() for _ in range(6)]
    ps = C.c_double()
    check(lib().mg_tsne_batch_stats(*[C.byref(x) for x in v], C.byref(ps)))
    names = ("plans", "iterations", "replays", "launches", "checks", "stalls")
    return dict({k: x.value for k, x in zip(names, v)}, plan_s=ps.value)


def optimize_layout(Y, head, tail, eps, eons, eonns, epoch_start, epoch_end, alpha0, alpha1, mid_w0=0.0, mid_w1=0.0,
                    a=1.0, b=1.0, rep_a=0.0, rep_b=0.0, negative_sample_rate=5.0, exaggeration=1.0, repulsion_strength=1.0,
                    mid_rate=0.0, clip_mode=0, n_fixed=0, seed=0, pairs=None, rev_indptr=None, rev_indices=None,
                    attraction_scale=1.5):
    """mg_optimize_layout on device tensors: Y (n, 2) float32, edges sorted by head (int32), eps / eons / eonns (m)
    float32 (the last two updated in place), pairs (n, m) int32 or None, reverse pairs CSR or None."""
    p = LayoutParams(int(epoch_start), int(epoch_end), alpha0, alpha1, mid_w0, mid_w1, a, b, rep_a, rep_b,
                     negative_sample_rate, exaggeration, repulsion_strength, mid_rate, int(clip_mode), int(n_fixed),
                     int(seed) & ((1 << 64) - 1), attraction_scale)
    check(lib().mg_optimize_layout(C.byref(p), ptr(Y), Y.shape[0], Y.shape[1], ptr(head), ptr(tail), ptr(eps), ptr(eons),
                                   ptr(eonns), head.numel(), ptr(pairs), pairs.shape[1] if pairs is not None else 0,
                                   ptr(rev_indptr), ptr(rev_indices)))


def topk_rows(D, k):
    """(indices (rows, k) int32, values (rows, k) float32): the k smallest entries of each row of the float32
    device tile D, sorted by (value, column)."""
    import torch
    if D.dtype != torch.float32 or D.dim() != 2 or not D.is_contiguous():
        raise ValueError("D must be a contiguous (rows, cols) float32 tensor")
    idx = torch.empty((D.shape[0], k), dtype=torch.int32, device=D.device)
    val = torch.empty((D.shape[0], k), dtype=torch.float32, device=D.device)
    check(lib().mg_topk_rows(ptr(D), D.shape[0], D.shape[1], int(k), ptr(idx), ptr(val)))
    return idx, val


def knn(T, Q, k, exclude_self=False):
    """(indices (m, k) int32, squared distances (m, k) float32) of the k nearest rows of T (nt, d) to each row of
    Q (m, d), float32 device tensors; exclude_self (Q is T) leaves each row out of its own list."""
    import torch
    for a in (T, Q):
        if a.dtype != torch.float32 or a.dim() != 2 or not a.is_contiguous():
            raise ValueError("T and Q must be contiguous 2-d float32 tensors")
    if T.shape[1] != Q.shape[1]:
        raise ValueError("T and Q must have the same number of features")
    idx = torch.empty((Q.shape[0], k), dtype=torch.int32, device=Q.device)
    d2 = torch.empty((Q.shape[0], k), dtype=torch.float32, device=Q.device)
    check(lib().mg_knn(ptr(T), T.shape[0], ptr(Q), Q.shape[0], T.shape[1], int(k), 0 if exclude_self else -1, ptr(idx), ptr(d2)))
    return idx, d2


def knn_noalloc(T, Q, idx, d2, k, exclude_self=False):
    """mg_knn_noalloc: the search of knn() into the caller's idx (m, k) int32 / d2 (m, k) float32 device tensors,
    with no device allocation and no synchronisation (the results are ready once the library stream reaches them)."""
    import torch
    for a in (T, Q, d2):
        if a.dtype != torch.float32 or a.dim() != 2 or not a.is_contiguous():
            raise ValueError("T, Q and d2 must be contiguous 2-d float32 tensors")
    if idx.dtype != torch.int32 or tuple(idx.shape) != (Q.shape[0], k) or tuple(d2.shape) != (Q.shape[0], k):
        raise ValueError("idx / d2 must be (m, k) int32 / float32 tensors")
    if T.shape[1] != Q.shape[1]:
        raise ValueError("T and Q must have the same number of features")
    check(lib().mg_knn_noalloc(ptr(T), T.shape[0], ptr(Q), Q.shape[0], T.shape[1], int(k), 0 if exclude_self else -1,
                               ptr(idx), ptr(d2)))


def smooth_knn(dist, n_neighbors, mean_all):
    """Smooth k-NN memberships (m, k) of the sorted neighbour distances dist (m, k) on the device (float32);
    mean_all (m,): each row's fallback mean distance."""
    import torch
    if dist.dtype != torch.float32 or dist.dim() != 2 or not dist.is_contiguous():
        raise ValueError("dist must be a contiguous (m, k) float32 tensor")
    mean_all = mean_all.to(torch.float32).contiguous()
    vals = torch.empty_like(dist)
    check(lib().mg_smooth_knn(ptr(dist), dist.shape[0], dist.shape[1], int(n_neighbors), ptr(mean_all), ptr(vals)))
    return vals


def perplexity_rows(d2, perplexity=30.0, tol=1e-5, n_iter=100, with_beta=False):
    """Conditional affinities p_j|i (n, k) float64 of the squared neighbour distances d2 (n, k) float32 (device), the
    entropy of each row log(perplexity) within tol (bisection on the Gaussian precision, at most n_iter steps);
    with_beta: also the precisions (n,)."""
    import torch
    if d2.dtype != torch.float32 or d2.dim() != 2 or not d2.is_contiguous():
        raise ValueError("d2 must be a contiguous (n, k) float32 tensor")
    P = torch.empty(d2.shape, dtype=torch.float64, device=d2.device)
    beta = torch.empty(d2.shape[0], dtype=torch.float64, device=d2.device) if with_beta else None
    check(lib().mg_perplexity_rows(ptr(d2), d2.shape[0], d2.shape[1], float(perplexity), float(tol), int(n_iter), ptr(P),
                                   ptr(beta) if beta is not None else None))
    return (P, beta) if with_beta else P


def symmetrize_knn(idx, val, mode=0, param=None):
    """Symmetric CSR (indptr (n+1) int32, indices int32, data float32; device tensors) of the directed graph
    idx (n, k) int32 with values val (n, k) float64: mode 0 gives each pair (a + b) * param (param defaults to 1 / 2n:
    the perplexity affinities (P + P^T) / 2n), mode 1 param (a + b - a b) + (1 - param) a b (the fuzzy union,
    param = set_op_mix_ratio, default 1)."""
    import torch
    if idx.dtype != torch.int32 or idx.dim() != 2 or not idx.is_contiguous():
        raise ValueError("idx must be a contiguous (n, k) int32 tensor")
    val = val.to(torch.float64).contiguous()
    if val.shape != idx.shape:
        raise ValueError("val must have the shape of idx")
    n, k = idx.shape
    if param is None:
        param = 1.0 / (2.0 * n) if mode == 0 else 1.0
    cap = 2 * n * k
    indptr = torch.empty(n + 1, dtype=torch.int32, device=idx.device)
    indices = torch.empty(max(cap, 1), dtype=torch.int32, device=idx.device)
    data = torch.empty(max(cap, 1), dtype=torch.float32, device=idx.device)
    nnz = C.c_int64(0)
    check(lib().mg_symmetrize_knn(int(mode), ptr(idx), ptr(val), n, k, float(param), ptr(indptr), ptr(indices), ptr(data),
                                  cap, C.byref(nnz)))
    return indptr, indices[: nnz.value].clone(), data[: nnz.value].clone()


def nn_descent(X, k, max_candidates=50, n_iters=10, delta=0.001, n_trees=4, leaf_size=30, seed=0, verbose=False):
    """(indices (n, k) int32, squared distances (n, k) float32) of the NN-descent graph of the rows of X (contiguous
    (n, d) float32 device tensor), rows sorted by (distance, index)."""
    import torch
    if X.dtype != torch.float32 or X.dim() != 2 or not X.is_contiguous():
        raise ValueError("X must be a contiguous (n, d) float32 tensor")
    n = X.shape[0]
    idx = torch.empty((n, k), dtype=torch.int32, device=X.device)
    d2 = torch.empty((n, k), dtype=torch.float32, device=X.device)
    check(lib().mg_nn_descent(ptr(X), n, X.shape[1], int(k), int(max_candidates), int(n_iters), float(delta), int(n_trees),
                              int(leaf_size), int(seed) & 0xFFFFFFFFFFFFFFFF, 1 if verbose else 0, ptr(idx), ptr(d2)))
    return idx, d2


def knn_exact_graph(X, k):
    """(indices (n, k) int32, squared distances (n, k) float32): the exact k nearest other rows of every row of X
    (contiguous (n, d) float32 device tensor, k <= 128), sorted by (distance, index)."""
    import torch
    if X.dtype != torch.float32 or X.dim() != 2 or not X.is_contiguous():
        raise ValueError("X must be a contiguous (n, d) float32 tensor")
    n = X.shape[0]
    idx = torch.empty((n, k), dtype=torch.int32, device=X.device)
    d2 = torch.empty((n, k), dtype=torch.float32, device=X.device)
    check(lib().mg_knn_exact_graph(ptr(X), n, X.shape[1], int(k), ptr(idx), ptr(d2)))
    return idx, d2


def mid_near_pairs(X, n_pairs=5, n_cand=6, seed=0):
    """(n, n_pairs) int32 device tensor: the second nearest of n_cand hashed draws per pair (mg_mid_near_pairs)."""
    import torch
    if X.dtype != torch.float32 or X.dim() != 2 or not X.is_contiguous():
        raise ValueError("X must be a contiguous (n, d) float32 tensor")
    out = torch.empty((X.shape[0], n_pairs), dtype=torch.int32, device=X.device)
    check(lib().mg_mid_near_pairs(ptr(X), X.shape[0], X.shape[1], int(n_pairs), int(n_cand), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                  ptr(out)))
    return out


def mem_info():
    """(free, total) HBM bytes of the current device."""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().mg_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def tune(name, value=None):
    """Sets (value given) or reads a launch-shape tunable (mg_tune_set / mg_tune_get); returns the value."""
    if value is not None:
        check(lib().mg_tune_set(name.encode(), int(value)))
    v = C.c_int()
    check(lib().mg_tune_get(name.encode(), C.byref(v)))
    return v.value


def profile(on=True):
    check(lib().mg_profile_enable(1 if on else 0))
    check(lib().mg_profile_reset())


def profile_read():
    """{stage: (total_ms, launches, algorithmic_bytes)} accumulated by mg_tsne_optimize since profile()."""
    out = {}
    for s in range(lib().mg_profile_num_stages()):
        name, tot, cnt, by = C.c_char_p(), C.c_double(), C.c_int64(), C.c_double()
        check(lib().mg_profile_get(s, C.byref(name), C.byref(tot), C.byref(cnt), C.byref(by)))
        out[name.value.decode()] = (tot.value, cnt.value, by.value)
    return out


def profile_sized():
    """{(stage, M): (total_ms, launches, algorithmic_bytes)} per padded FFT size M (0 for size-independent stages)."""
    out = {}
    for s in range(lib().mg_profile_num_stages()):
        name, tot, cnt, by = C.c_char_p(), C.c_double(), C.c_int64(), C.c_double()
        check(lib().mg_profile_get(s, C.byref(name), C.byref(tot), C.byref(cnt), C.byref(by)))
        k = C.c_int()
        check(lib().mg_profile_sized(s, 0, C.byref(k), None, None, None, None))
        if k.value == 0:
            continue
        Ms, T, N, B = (C.c_int * k.value)(), (C.c_double * k.value)(), (C.c_int64 * k.value)(), (C.c_double * k.value)()
        check(lib().mg_profile_sized(s, k.value, C.byref(k), Ms, T, N, B))
        for j in range(k.value):
            out[(name.value.decode(), Ms[j])] = (T[j], N[j], B[j])
    return out


def fft_size(N):
    """Padded FFT size the device uses for a live grid side N (0: unsupported)."""
    return int(lib().mg_fft_size(int(N)))


def profile_grid():
    """(mean N, max N, max M) of the profiled iterations."""
    a, b, c = C.c_double(), C.c_int(), C.c_int()
    check(lib().mg_profile_grid(C.byref(a), C.byref(b), C.byref(c)))
    return a.value, b.value, c.value


# ---- fallback kernels (src/fftrep.cpp:51-125, 294-411) on device tensors
def interp_scatter(Y, charges, xmin, ymin, width, n_boxes, n_interp):
    import torch
    N = n_boxes * n_interp
    out = torch.empty((charges.shape[1], N, N), dtype=Y.dtype, device=Y.device)
    check(lib().mg_interp_scatter(prec_of(Y), ptr(Y), Y.shape[0], ptr(charges), charges.shape[1], xmin, ymin, width,
                                  n_boxes, n_interp, ptr(out)))
    return out


def interp_gather(Y, grid, xmin, ymin, width, n_boxes, n_interp):
    import torch
    out = torch.empty((Y.shape[0], grid.shape[0]), dtype=Y.dtype, device=Y.device)
    check(lib().mg_interp_gather(prec_of(Y), ptr(Y), Y.shape[0], ptr(grid), grid.shape[0], xmin, ymin, width, n_boxes,
                                 n_interp, ptr(out)))
    return out


def scatter_padded_2d(Y, xmin, ymin, width, n_boxes, n_interp, M):
    import torch
    out = torch.empty((3, M, M), dtype=Y.dtype, device=Y.device)
    check(lib().mg_scatter_padded_2d(prec_of(Y), ptr(Y), Y.shape[0], xmin, ymin, width, n_boxes, n_interp, M, ptr(out)))
    return out


def spectral_products(K1h, K2h, Wh):
    """float pairs: K (M, Mh, 2), Wh (3, M, Mh, 2) -> (4, M, Mh, 2)"""
    import torch
    M, Mh = K1h.shape[0], K1h.shape[1]
    out = torch.empty((4, M, Mh, 2), dtype=K1h.dtype, device=K1h.device)
    check(lib().mg_spectral_products(prec_of(K1h), ptr(K1h), ptr(K2h), ptr(Wh), M, Mh, ptr(out)))
    return out


def gather_forces_2d(Y, phi, xmin, ymin, width, n_boxes, n_interp):
    import torch
    F = torch.empty_like(Y)
    s = C.c_double()
    check(lib().mg_gather_forces_2d(prec_of(Y), ptr(Y), Y.shape[0], ptr(phi), phi.shape[1], xmin, ymin, width, n_boxes,
                                    n_interp, ptr(F), C.byref(s)))
    return s.value, F
