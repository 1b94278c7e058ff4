"""Full-batch t-SNE with FFT-interpolated repulsion on MI355X: the interface of the reference's
magnet/tsne.py (FFTRepulsion2D, optimize_tsne; kerighan/lib-magnet v3.0.2, magnet/tsne.py:11-214)
over the HIP implementation in `_magnet` / libmagnet_hip.

optimize_tsne runs the whole loop on the device (one call; the state stays in HBM). precision="fp32"
(default, float32 state and arithmetic with FP64 Z and reductions, like the reference's float32 path)
or "fp64" (the plain FP64 path). FFTRepulsion2D keeps the reference's per-call interface.
"""
import warnings

import numpy as np
import scipy.sparse as sp

__all__ = ["FFTRepulsion2D", "optimize_tsne", "grid_params"]


def _native():
    try:
        import _magnet
    except ImportError as e:   # the HIP path is the product: no silent CPU fallback
        raise ImportError("the native module _magnet (libmagnet_hip) is not built or not loadable: "
                          "build it with __graft_entry__.build()") from e
    return _magnet


def grid_params(bbox, box_width=0.7, min_boxes=20, max_boxes=1000):
    """Grid decision of FFTRepulsion2D._grid (magnet/tsne.py:30-46) from (xmin, ymin, xmax, ymax)."""
    xmin, ymin, xmax, ymax = bbox
    lo = np.array([xmin, ymin], dtype=np.float64) - 1e-3
    hi = np.array([xmax, ymax], dtype=np.float64) + 1e-3
    span = float((hi - lo).max())
    n_boxes = int(np.clip(np.ceil(span / box_width), min_boxes, max_boxes))
    n_boxes = min(max_boxes, -(-n_boxes // 4) * 4)
    width = box_width if n_boxes * box_width >= span else max(span / n_boxes, 1e-12)
    return float(lo[0]), float(lo[1]), width, n_boxes


class FFTRepulsion2D:
    """Potentials sum_j K(|y_i - y_j|) c_j for K1 = (1 + r^2/dof)^-dof and K2 = (1 + r^2/dof)^(-dof-1) by
    Lagrange interpolation on boxes of fixed width and FFT convolution on the nodes (magnet/tsne.py:11-90).
    Each call returns (sumQ, F (n, 2)) with F_i = sum_j K2(r_ij) (y_i - y_j). `kernels`, as in the
    reference, is an optional callable r2 -> (K1, K2) evaluated on the padded torus."""

    def __init__(self, n_interp=3, box_width=0.7, min_boxes=20, max_boxes=1000, n_threads=-1, dof=1.0, kernels=None):
        self.dof = float(dof)
        self.kernels = kernels
        self.n_interp = n_interp
        self.box_width = box_width
        self.min_boxes = min_boxes
        self.max_boxes = max_boxes
        self.n_threads = n_threads
        self._convs = {}

    def _grid(self, Y):
        bbox = _native().bbox_2d(np.ascontiguousarray(Y, dtype=np.float32), self.n_threads)
        return grid_params(bbox, self.box_width, self.min_boxes, self.max_boxes)

    def _conv(self, n_boxes, width):
        """One device object per (M, width), kernels set once (magnet/tsne.py:82-88)."""
        import scipy.fft as fft
        N = n_boxes * self.n_interp
        M = int(fft.next_fast_len(2 * N, real=True))
        key = (M, round(width, 6))
        conv = self._convs.get(key)
        if conv is None:
            _m = _native()
            conv = _m.FFTConv2D(M, self.n_interp)
            delta = width / self.n_interp
            off = np.arange(M)
            off = np.where(off <= M // 2, off, off - M).astype(np.float64) * delta
            r2 = off[:, None] ** 2 + off[None, :] ** 2
            if self.kernels is not None:
                K1, K2 = self.kernels(r2)
            else:
                base = 1.0 + r2 / self.dof
                K1, K2 = base ** (-self.dof), base ** (-self.dof - 1.0)
            # the reference hands real float32 spectra to FFTConv2D.set_kernels; so does this binding
            K1h = np.ascontiguousarray(fft.rfft2(np.asarray(K1, dtype=np.float32)).real, dtype=np.float32)
            K2h = np.ascontiguousarray(fft.rfft2(np.asarray(K2, dtype=np.float32)).real, dtype=np.float32)
            conv.set_kernels(K1h, K2h)
            if len(self._convs) > 4:
                self._convs.pop(next(iter(self._convs)))
            self._convs[key] = conv
        return conv

    def __call__(self, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        xmin, ymin, width, n_boxes = self._grid(Y)
        sumQ, F = self._conv(n_boxes, width).compute(Y, xmin, ymin, width, n_boxes, self.n_threads)
        return float(sumQ), np.asarray(F)


def optimize_tsne(P, Y, n_iter=750, early_iter=250, early_exaggeration=12.0, exaggeration=1.0,
                  learning_rate=None, momentum=(0.5, 0.8), n_interp=3, box_width=0.7, n_threads=-1,
                  mid_pairs=None, mid_weight=0.0, mid_phase=0.5, dof=1.0, verbose=False, callback=None,
                  far_weight=0.0, far_r0=0.1, far_a=1.74, mid_form="mean", mid_c=0.3, precision="fp32",
                  state=None, it0=0, it1=None, return_sumq=False):
    """Full-batch t-SNE gradient descent (magnet/tsne.py:132-214). P: symmetric CSR affinities
    (normalised to sum 1 here), Y: (n, 2) initial layout. Returns the final layout (float32 for
    precision="fp32", float64 for "fp64"); the caller's array is updated in place when it already has
    that dtype and is C-contiguous.

    Extensions: precision ("fp32" / "fp64"); state=(update, gains) arrays to resume a run exactly (the
    reference restarts them at 0 / 1); it0 / it1 run part of the n_iter schedule; return_sumq returns
    (Y, sumQ per iteration). far_weight > 0 and mid_form != "mean" are research options of the
    reference that this path does not implement."""
    if far_weight > 0:
        raise NotImplementedError("far_weight (research far-field repulsion) is not implemented on the MI355X path")
    if mid_form != "mean" and mid_pairs is not None and mid_weight > 0:
        raise NotImplementedError("only mid_form='mean' is implemented on the MI355X path")
    if precision not in ("fp32", "fp64"):
        raise ValueError("precision must be 'fp32' or 'fp64'")
    dt = np.float32 if precision == "fp32" else np.float64
    n = P.shape[0]
    P = sp.csr_matrix(P, dtype=dt)
    P = P / P.sum()
    P = sp.csr_matrix(P)
    indptr, indices, data = P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data.astype(dt)
    if learning_rate is None:
        learning_rate = max(200.0, n / early_exaggeration)
    if not (isinstance(Y, np.ndarray) and Y.dtype == dt and Y.flags.c_contiguous):
        Y = np.ascontiguousarray(Y, dtype=dt)
    if Y.shape != (n, 2):
        raise ValueError("Y must be (n, 2) with n = P.shape[0]")
    if state is None:
        update, gains = np.zeros_like(Y), np.ones_like(Y)
    else:
        update, gains = state
    if mid_pairs is not None:
        mid_pairs = np.ascontiguousarray(mid_pairs, dtype=np.int32)
    it1 = n_iter if it1 is None else it1
    cb = None
    if callback is not None or verbose:
        def cb(it, Yh):
            if verbose and (it % 100 == 0 or it == n_iter - 1):
                rho = early_exaggeration if it < early_iter else exaggeration
                print(f"  t-SNE iteration {it:4d}  exaggeration={rho:g}")
            if callback is not None:
                return callback(it, Yh)
            return None
    max_boxes = 1000
    if n_interp * max_boxes > 4096:
        # the largest device FFT holds grids of 4096 nodes per side; the reference's limit of 1000 boxes would pass
        # that with more than 4 nodes per box, so the box count is capped up front instead of failing mid-run
        max_boxes = 4096 // n_interp
        warnings.warn(f"n_interp={n_interp}: the grid is capped at {max_boxes} boxes per side (the reference allows 1000); "
                      "results differ from the reference only once the layout spans more boxes", RuntimeWarning)
    sumq = _native().tsne_optimize(indptr, indices, data, Y, update, gains, int(n_iter), int(early_iter),
                                   float(early_exaggeration), float(exaggeration), float(learning_rate),
                                   float(momentum[0]), float(momentum[1]), int(n_interp), float(box_width), float(dof),
                                   mid_pairs, float(mid_weight), float(mid_phase), int(it0), int(it1), 20, max_boxes, cb)
    return (Y, np.asarray(sumq)) if return_sumq else Y
