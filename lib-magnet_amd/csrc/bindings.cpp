// `_magnet`: the reference's native-module names for the t-SNE path (src/bindings.cpp:8-83 and
// src/fftconv.cpp:239-244), implemented over the C-ABI of libmagnet_hip (include/magnet_hip.h).
// Host numpy arrays are copied to HBM, the HIP path runs, results are copied back; in-place functions
// write into the caller's arrays as the reference does (arr<T> = py::array_t<T, c_style | forcecast>,
// src/layout.hpp:14-15). Errors raise RuntimeError with the library's message (src/layout.hpp:34).
// The whole-loop entry `tsne_optimize` keeps the state on the device for all iterations; it is what
// magnet.optimize_tsne uses, in FP32 or FP64.
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/functional.h>
#include "magnet_hip.h"
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace py = pybind11;
template <typename T> using arr = py::array_t<T, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc) {
    if (rc != MG_OK) throw std::runtime_error(mg_last_error());
}
void require(bool ok, const char* msg) {
    if (!ok) throw std::runtime_error(msg);
}

// device copy of a host buffer, freed on scope exit
struct Dev {
    void* p = nullptr;
    size_t bytes = 0;
    Dev() = default;
    explicit Dev(size_t b) : bytes(b) { check(mg_malloc(&p, b ? b : 1)); }
    Dev(const void* h, size_t b) : Dev(b) { if (b) check(mg_memcpy_htod(p, h, b)); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    Dev(Dev&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    Dev& operator=(Dev&& o) noexcept {
        if (this != &o) { if (p) mg_free(p); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~Dev() { if (p) mg_free(p); }
    void to_host(void* h) const { if (bytes) check(mg_memcpy_dtoh(h, p, bytes)); }
};
template <class T> Dev up(const arr<T>& a) { return Dev(a.data(), (size_t)a.size() * sizeof(T)); }

void validate_csr_host(const arr<int32_t>& indptr, const arr<int32_t>& indices, int64_t data_size, int64_t n) {
    require(indptr.ndim() == 1 && indptr.shape(0) == n + 1, "indptr must have n + 1 entries");
    const int32_t* ip = indptr.data();
    require(ip[0] == 0, "indptr[0] must be 0");
    const int64_t nnz = ip[n];
    require(indices.ndim() == 1 && indices.shape(0) == nnz, "indices must have indptr[n] entries");
    require(data_size == nnz, "data must have indptr[n] entries");
}

// ---------------------------------------------------------------------------------- per-call kernels
// nn_descent (src/nndescent.cpp:636-645): (indices (n, k) int32, squared distances (n, k) float32), rows sorted
py::tuple nn_descent(arr<float> X, int k, int max_candidates, int n_iters, float delta, int n_trees, int leaf_size,
                     uint64_t seed, int, bool verbose) {
    require(X.ndim() == 2, "X must be 2-d");
    require(X.shape(1) >= 1, "X must have at least one column");
    const int64_t n = X.shape(0);
    require(k >= 1 && k < n, "k must be between 1 and n - 1");
    arr<int32_t> idx({n, (int64_t)k});
    arr<float> dist({n, (int64_t)k});
    Dev dx = up(X), di((size_t)n * k * sizeof(int32_t)), dd((size_t)n * k * sizeof(float));
    {
        py::gil_scoped_release rel;
        check(mg_nn_descent((const float*)dx.p, n, (int)X.shape(1), k, max_candidates, n_iters, delta, n_trees, leaf_size, seed,
                            verbose ? 1 : 0, (int32_t*)di.p, (float*)dd.p));
        di.to_host(idx.mutable_data());
        dd.to_host(dist.mutable_data());
    }
    return py::make_tuple(idx, dist);
}

// mid_near_pairs_dense (src/nndescent.cpp:761-767): (n, n_pairs) int32
arr<int32_t> mid_near_pairs_dense(arr<float> X, int n_pairs, int n_cand, uint64_t seed, int) {
    require(X.ndim() == 2 && X.shape(1) >= 1, "X must be 2-d");
    const int64_t n = X.shape(0);
    arr<int32_t> out({n, (int64_t)n_pairs});
    Dev dx = up(X), dout((size_t)n * (n_pairs > 0 ? n_pairs : 1) * sizeof(int32_t));
    {
        py::gil_scoped_release rel;
        check(mg_mid_near_pairs((const float*)dx.p, n, (int)X.shape(1), n_pairs, n_cand, seed, (int32_t*)dout.p));
        dout.to_host(out.mutable_data());
    }
    return out;
}

arr<float> tsne_attraction(arr<int32_t> indptr, arr<int32_t> indices, arr<float> data, arr<float> Y, float dof, int) {
    require(dof > 0.f, "dof must be positive");
    require(Y.ndim() == 2, "Y must be 2-d");
    const int64_t n = Y.shape(0);
    const int dim = (int)Y.shape(1);
    require(dim >= 1 && dim <= 64, "dim out of range");
    validate_csr_host(indptr, indices, data.ndim() == 1 ? data.shape(0) : -1, n);
    arr<float> out({n, (int64_t)dim});
    if (n == 0) return out;
    Dev dip = up(indptr), dix = up(indices), dd = up(data), dy = up(Y), dout((size_t)n * dim * sizeof(float));
    {
        py::gil_scoped_release rel;
        check(mg_tsne_attraction(MG_F32, (const int32_t*)dip.p, (const int32_t*)dix.p, dd.p, (int64_t)indices.shape(0), dy.p,
                                 n, dim, dof, dout.p));
        dout.to_host(out.mutable_data());
    }
    return out;
}

void step_common(int mode, arr<float>& Y, arr<float>& U, arr<float>& G, const arr<float>& A, const arr<float>* F, float rho,
                 float inv, float lr, float mom) {
    require(Y.ndim() == 2, "Y must be 2-d");
    const int64_t n = Y.shape(0);
    const int dim = (int)Y.shape(1);
    const int64_t total = n * dim;
    require(U.size() == total && G.size() == total && A.size() == total && (!F || F->size() == total), "arrays must match Y");
    if (n == 0) return;
    Dev dy = up(Y), du = up(U), dg = up(G), da = up(A);
    Dev df = F ? up(*F) : Dev();
    {
        py::gil_scoped_release rel;
        if (mode == 0) check(mg_tsne_step(MG_F32, dy.p, du.p, dg.p, da.p, df.p, n, dim, rho, inv, lr, mom));
        else check(mg_tsne_update(MG_F32, dy.p, du.p, dg.p, da.p, n, dim, lr, mom));
        dy.to_host(Y.mutable_data()); du.to_host(U.mutable_data()); dg.to_host(G.mutable_data());
    }
}

void tsne_step(arr<float> Y, arr<float> U, arr<float> G, arr<float> A, arr<float> F, float rho, float inv_sumQ, float lr,
               float momentum, int) {
    step_common(0, Y, U, G, A, &F, rho, inv_sumQ, lr, momentum);
}

void tsne_update(arr<float> Y, arr<float> U, arr<float> G, arr<float> grad, float lr, float momentum, int) {
    step_common(1, Y, U, G, grad, nullptr, 0.f, 0.f, lr, momentum);
}

void mid_spring(arr<float> Y, arr<int32_t> pairs, arr<float> grad, float w, int) {
    require(Y.ndim() == 2 && grad.ndim() == 2 && pairs.ndim() == 2, "bad shapes");
    const int64_t n = Y.shape(0);
    const int dim = (int)Y.shape(1), m = (int)pairs.shape(1);
    require(pairs.shape(0) == n && grad.shape(0) == n && grad.shape(1) == dim && m >= 1, "bad shapes");
    if (n == 0) return;
    const int32_t* pp = pairs.data();
    for (int64_t k = 0; k < n * m; ++k) require(pp[k] >= 0 && pp[k] < n, "pair index out of range");
    Dev dy = up(Y), dp = up(pairs), dg = up(grad);
    py::gil_scoped_release rel;
    check(mg_mid_spring(MG_F32, dy.p, (const int32_t*)dp.p, dg.p, n, dim, m, w));
    dg.to_host(grad.mutable_data());
}

std::tuple<float, float, float, float> bbox_2d(arr<float> Y, int) {
    require(Y.ndim() == 2 && Y.shape(1) == 2 && Y.shape(0) >= 1, "Y must be (n, 2)");
    Dev dy = up(Y);
    double out[4];
    {
        py::gil_scoped_release rel;
        check(mg_bbox_2d(MG_F32, dy.p, Y.shape(0), out));
    }
    return std::make_tuple((float)out[0], (float)out[1], (float)out[2], (float)out[3]);
}

// --------------------------------------------------------------------------- fallback kernels
// (src/fftrep.cpp:51-125, 294-411), float32 like the reference's, run on the device
arr<float> interp_scatter(arr<float> Y, arr<float> charges, float xmin, float ymin, float width, int n_boxes, int n_interp, int) {
    require(Y.ndim() == 2 && Y.shape(1) == 2, "Y must be (n, 2)");
    const int64_t n = Y.shape(0);
    require(charges.ndim() == 2 && charges.shape(0) == n, "charges must be (n, T)");
    require(n_boxes >= 1 && n_interp >= 1 && n_interp <= 8 && width > 0, "bad grid parameters");
    const int C = (int)charges.shape(1);
    const int64_t N = (int64_t)n_boxes * n_interp;
    arr<float> out({(int64_t)C, N, N});
    Dev dy = up(Y), dq = up(charges), dout((size_t)C * N * N * sizeof(float));
    py::gil_scoped_release rel;
    check(mg_interp_scatter(MG_F32, dy.p, n, dq.p, C, xmin, ymin, width, n_boxes, n_interp, dout.p));
    dout.to_host(out.mutable_data());
    return out;
}

arr<float> interp_gather(arr<float> Y, arr<float> grid, float xmin, float ymin, float width, int n_boxes, int n_interp, int) {
    require(Y.ndim() == 2 && Y.shape(1) == 2, "Y must be (n, 2)");
    const int64_t n = Y.shape(0);
    require(grid.ndim() == 3 && grid.shape(1) == (int64_t)n_boxes * n_interp && grid.shape(2) == (int64_t)n_boxes * n_interp,
            "grid must be (T, N, N)");
    const int C = (int)grid.shape(0);
    arr<float> out({n, (int64_t)C});
    if (n == 0) return out;
    Dev dy = up(Y), dg = up(grid), dout((size_t)n * C * sizeof(float));
    py::gil_scoped_release rel;
    check(mg_interp_gather(MG_F32, dy.p, n, dg.p, C, xmin, ymin, width, n_boxes, n_interp, dout.p));
    dout.to_host(out.mutable_data());
    return out;
}

arr<float> scatter_padded_2d(arr<float> Y, float xmin, float ymin, float width, int n_boxes, int n_interp, int M, int,
                             arr<float> out) {
    require(Y.ndim() == 2 && Y.shape(1) == 2, "Y must be (n, 2)");
    const int64_t n = Y.shape(0);
    require(n_interp >= 1 && n_interp <= 8 && M >= (int64_t)n_boxes * n_interp && width > 0, "bad grid parameters");
    if (out.size() == 0) out = arr<float>({(int64_t)3, (int64_t)M, (int64_t)M});
    require(out.ndim() == 3 && out.shape(0) == 3 && out.shape(1) == M && out.shape(2) == M, "out must be (3, M, M)");
    Dev dy = up(Y), dout((size_t)3 * M * M * sizeof(float));
    {
        py::gil_scoped_release rel;
        check(mg_scatter_padded_2d(MG_F32, dy.p, n, xmin, ymin, width, n_boxes, n_interp, M, dout.p));
        dout.to_host(out.mutable_data());
    }
    return out;
}

arr<float> spectral_products(arr<float> K1h, arr<float> K2h, arr<float> Wh, int) {
    require(K1h.ndim() == 3 && K1h.shape(2) == 2 && Wh.ndim() == 4 && Wh.shape(0) == 3 && Wh.shape(3) == 2, "bad shapes");
    const int64_t M = K1h.shape(0), Mh = K1h.shape(1);
    require(K2h.ndim() == 3 && K2h.shape(0) == M && K2h.shape(1) == Mh && K2h.shape(2) == 2 && Wh.shape(1) == M &&
            Wh.shape(2) == Mh, "bad shapes");
    arr<float> out({(int64_t)4, M, Mh, (int64_t)2});
    Dev d1 = up(K1h), d2 = up(K2h), dw = up(Wh), dout((size_t)8 * M * Mh * sizeof(float));
    py::gil_scoped_release rel;
    check(mg_spectral_products(MG_F32, d1.p, d2.p, dw.p, M, Mh, dout.p));
    dout.to_host(out.mutable_data());
    return out;
}

std::tuple<float, arr<float>> gather_forces_2d(arr<float> Y, arr<float> phi, float xmin, float ymin, float width, int n_boxes,
                                               int n_interp, int) {
    require(Y.ndim() == 2 && Y.shape(1) == 2 && phi.ndim() == 3 && phi.shape(0) == 4, "bad shapes");
    const int64_t n = Y.shape(0), M = phi.shape(1);
    require(phi.shape(2) == M && M >= (int64_t)n_boxes * n_interp, "phi must be (4, M, M) with M >= N");
    arr<float> F({n, (int64_t)2});
    double sq = -(double)n;
    if (n > 0) {
        Dev dy = up(Y), dp = up(phi), df((size_t)n * 2 * sizeof(float));
        py::gil_scoped_release rel;
        check(mg_gather_forces_2d(MG_F32, dy.p, n, dp.p, (int)M, xmin, ymin, width, n_boxes, n_interp, df.p, &sq));
        df.to_host(F.mutable_data());
    }
    return std::make_tuple((float)sq, F);
}

// ------------------------------------------------------------------------------------- FFTConv2D
class FFTConv2D {
public:
    FFTConv2D(int M, int n_interp) : M_(M), ni_(n_interp) {
        require(M >= 1 && n_interp >= 1 && n_interp <= 8, "bad grid parameters");
        check(mg_fftconv_create(&c_, MG_F32, M, n_interp));
    }
    ~FFTConv2D() { if (c_) mg_fftconv_destroy(c_); }

    // spectra (M, M/2+1), real: the kernel values are their inverse real 2-D transform
    void set_kernels(arr<float> K1h, arr<float> K2h) {
        const int Mh = M_ / 2 + 1;
        require(K1h.ndim() == 2 && K1h.shape(0) == M_ && K1h.shape(1) == Mh, "K1h must be (M, M/2+1) real");
        require(K2h.ndim() == 2 && K2h.shape(0) == M_ && K2h.shape(1) == Mh, "K2h must be (M, M/2+1) real");
        py::module_ fft = py::module_::import("numpy.fft");
        auto values = [&](const arr<float>& K) {
            py::object v = fft.attr("irfft2")(K.attr("astype")("complex128"), py::make_tuple(M_, M_));
            return arr<double>(v);
        };
        arr<double> v1 = values(K1h), v2 = values(K2h);
        check(mg_fftconv_set_kernel_values(c_, M_, v1.data(), v2.data()));
    }

    std::tuple<float, arr<float>> compute(arr<float> Y, float xmin, float ymin, float width, int n_boxes, int) {
        require(Y.ndim() == 2 && Y.shape(1) == 2, "Y must be (n, 2)");
        require(n_boxes * ni_ <= M_ && width > 0, "grid larger than the buffer");
        const int64_t n = Y.shape(0);
        arr<float> F({n, (int64_t)2});
        if (n == 0) return std::make_tuple(0.f, F);
        Dev dy = up(Y), df((size_t)n * 2 * sizeof(float));
        double sumq = 0;
        {
            py::gil_scoped_release rel;
            check(mg_fftconv_compute(c_, dy.p, n, xmin, ymin, width, n_boxes, df.p, &sumq));
            df.to_host(F.mutable_data());
        }
        return std::make_tuple((float)sumq, F);
    }

private:
    int M_, ni_;
    mg_fftconv* c_ = nullptr;
};

// ------------------------------------------------------------------------------------ whole loop
template <class T>
py::array_t<double> optimize_impl(int prec, arr<int32_t> indptr, arr<int32_t> indices, arr<T> data, arr<T> Y, arr<T> U,
                                  arr<T> G, const mg_tsne_params& p, py::object pairs_obj, int it0, int it1,
                                  py::object callback) {
    require(Y.ndim() == 2 && Y.shape(1) == 2, "Y must be (n, 2)");
    const int64_t n = Y.shape(0);
    require(n >= 1, "Y must be (n, 2) with n >= 1");
    require(U.size() == n * 2 && G.size() == n * 2, "arrays must match Y");
    validate_csr_host(indptr, indices, data.ndim() == 1 ? data.shape(0) : -1, n);
    const int64_t nnz = indices.shape(0);
    Dev dip = up(indptr), dix = up(indices), dd = up(data), dy = up(Y), du = up(U), dg = up(G);
    Dev dp;
    int m = 0;
    if (!pairs_obj.is_none()) {
        arr<int32_t> pairs(pairs_obj);
        require(pairs.ndim() == 2 && pairs.shape(0) == n && pairs.shape(1) >= 1, "pairs must be (n, m)");
        m = (int)pairs.shape(1);
        const int32_t* pp = pairs.data();
        for (int64_t k = 0; k < n * m; ++k) require(pp[k] >= 0 && pp[k] < n, "pair index out of range");
        dp = up(pairs);
    }
    const int span = it1 - it0;
    py::array_t<double> sumq(span > 0 ? span : 0);
    struct Ctx { py::object cb; int64_t n; size_t ts; std::string err; } ctx{callback, n, sizeof(T), {}};
    auto tramp = [](int it, const void* d_Y, int64_t nn, void* user) -> int {
        Ctx* c = static_cast<Ctx*>(user);
        py::gil_scoped_acquire acq;
        try {
            arr<T> Yh({nn, (int64_t)2});
            if (mg_memcpy_dtoh(Yh.mutable_data(), d_Y, (size_t)nn * 2 * sizeof(T)) != MG_OK) { c->err = mg_last_error(); return 1; }
            py::object r = c->cb(it, Yh);
            return (!r.is_none() && py::bool_(r)) ? 1 : 0;   // a truthy return value stops the run
        } catch (py::error_already_set& e) {
            c->err = e.what();
            return 1;
        }
    };
    int rc;
    {
        py::gil_scoped_release rel;
        rc = mg_tsne_optimize(prec, &p, (const int32_t*)dip.p, (const int32_t*)dix.p, dd.p, nnz, n, dy.p, du.p, dg.p,
                              m ? (const int32_t*)dp.p : nullptr, m, it0, it1, span > 0 ? sumq.mutable_data() : nullptr,
                              callback.is_none() ? nullptr : +tramp, &ctx);
    }
    if (!ctx.err.empty()) throw std::runtime_error("callback failed: " + ctx.err);
    check(rc);
    dy.to_host(Y.mutable_data()); du.to_host(U.mutable_data()); dg.to_host(G.mutable_data());
    return sumq;
}

py::array_t<double> tsne_optimize(py::array indptr, py::array indices, py::array data, py::array Y, py::array U, py::array G,
                                  int n_iter, int early_iter, double early_exaggeration, double exaggeration,
                                  double learning_rate, double momentum0, double momentum1, int n_interp, double box_width,
                                  double dof, py::object pairs, double mid_weight, double mid_phase, int it0, int it1,
                                  int min_boxes, int max_boxes, py::object callback) {
    mg_tsne_params p;
    p.n_iter = n_iter; p.early_iter = early_iter; p.early_exaggeration = early_exaggeration; p.exaggeration = exaggeration;
    p.learning_rate = learning_rate; p.momentum0 = momentum0; p.momentum1 = momentum1; p.n_interp = n_interp;
    p.min_boxes = min_boxes; p.max_boxes = max_boxes; p.box_width = box_width; p.dof = dof;
    p.mid_weight = mid_weight; p.mid_phase = mid_phase; p.relabel_every = 0;
    if (it1 < 0) it1 = n_iter;
    const bool f64 = py::dtype::of<double>().is(Y.dtype());
    // in-place state must already have the working dtype and layout (forcecast would copy it)
    auto inplace_ok = [&](const py::array& a) {
        return (f64 ? py::dtype::of<double>().is(a.dtype()) : py::dtype::of<float>().is(a.dtype())) &&
               (a.flags() & py::array::c_style);
    };
    require(inplace_ok(Y) && inplace_ok(U) && inplace_ok(G), "Y, update and gains must be C-contiguous float32 or float64 arrays of one dtype");
    if (f64)
        return optimize_impl<double>(MG_F64, arr<int32_t>(indptr), arr<int32_t>(indices), arr<double>(data), arr<double>(Y),
                                     arr<double>(U), arr<double>(G), p, pairs, it0, it1, callback);
    return optimize_impl<float>(MG_F32, arr<int32_t>(indptr), arr<int32_t>(indices), arr<float>(data), arr<float>(Y),
                                arr<float>(U), arr<float>(G), p, pairs, it0, it1, callback);
}

// ------------------------------------------------------------------------------------ sampled layout
// optimize_layout with the reference's arguments (src/bindings.cpp:11-23, src/layout.cpp:343-472) over
// mg_optimize_layout: loss 0 (umap), kernel 0 (power), optimizer 0 (sgd), 2-D, mid-range pairs or none; the options
// this path does not implement (InfoNCE, other kernels, Adam, per-node exponents, random-walk or two-layer partners)
// raise. Every parallel mode runs as the reference's owner mode (deterministic); the thread count has no effect.
// For parallel 0 (hogwild) with fixed nodes, the attraction factor is 1 (the new points' edges have no mirror).
// Y, epoch_of_next_sample and epoch_of_next_negative_sample are updated in place.
void optimize_layout(py::array Y_, arr<int32_t> head, arr<int32_t> tail, arr<float> eps, py::array eons_, py::array eonns_,
                     arr<int32_t> indptr, arr<int32_t> indices, int epoch_start, int epoch_end, float alpha0, float alpha1,
                     float mid_w0, float mid_w1, float a, float b, float rep_a, float rep_b, float negative_sample_rate, int loss,
                     float exaggeration, float repulsion_strength, int walk_len, float mid_rate, int optimizer, int clip_mode,
                     int64_t n_fixed, py::array adam_m, py::array adam_v, py::array adam_t, float beta1, float beta2,
                     arr<int32_t> walk_lens, arr<int32_t> mid_pairs, arr<int32_t> jump_indptr, arr<int32_t> jump_indices,
                     float mid_jump, int kernel, arr<float> node_b, uint64_t seed, int n_threads, int parallel,
                     arr<int32_t> rev_indptr, arr<int32_t> rev_indices) {
    (void)indptr; (void)indices; (void)adam_m; (void)adam_v; (void)adam_t; (void)beta1; (void)beta2; (void)n_threads;
    auto f32_inplace = [](const py::array& x) { return py::dtype::of<float>().is(x.dtype()) && (x.flags() & py::array::c_style); };
    require(f32_inplace(Y_) && f32_inplace(eons_) && f32_inplace(eonns_),
            "Y and the epoch arrays must be C-contiguous float32 (they are updated in place)");
    require(Y_.ndim() == 2, "Y must be 2-dimensional");
    const int64_t n = Y_.shape(0);
    const int dim = (int)Y_.shape(1);
    require(head.ndim() == 1, "head must be 1-d");
    const int64_t m = head.shape(0);
    require(tail.ndim() == 1 && tail.shape(0) == m && eps.shape(0) == m && eons_.shape(0) == m && eonns_.shape(0) == m,
            "head, tail and epoch arrays must have the same length");
    require(parallel == 0 || parallel == 1, "parallel must be 0 (hogwild) or 1 (owner)");
    require(loss == 0, "this path implements loss 0 (umap) only");
    require(kernel == 0, "this path implements kernel 0 (power) only");
    require(optimizer == 0, "this path implements optimizer 0 (sgd) only");
    require(node_b.size() == 0, "per-node kernel exponents are not implemented on this path");
    require(jump_indptr.size() == 0 || mid_jump <= 0.f, "two-layer random-walk partners are not implemented on this path");
    const bool use_mid = (mid_w0 > 0.f || mid_w1 > 0.f) && mid_rate > 0.f;
    require(!use_mid || mid_pairs.size() > 0 || walk_len <= 0, "random-walk partners are not implemented on this path");
    (void)walk_lens;
    const int n_mid = (use_mid && mid_pairs.size() > 0) ? (int)mid_pairs.shape(1) : 0;
    require(n_mid == 0 || (mid_pairs.ndim() == 2 && mid_pairs.shape(0) == n), "mid_pairs must be (n, m)");
    mg_layout_params p;
    p.epoch_start = epoch_start; p.epoch_end = epoch_end; p.alpha0 = alpha0; p.alpha1 = alpha1; p.mid_w0 = mid_w0;
    p.mid_w1 = mid_w1; p.a = a; p.b = b; p.rep_a = rep_a; p.rep_b = rep_b; p.negative_sample_rate = negative_sample_rate;
    p.exaggeration = exaggeration; p.repulsion_strength = repulsion_strength; p.mid_rate = mid_rate; p.clip_mode = clip_mode;
    p.n_fixed = n_fixed; p.seed = seed;
    // the owner mode's factor 1.5 stands for the pull of the mirrored edge; the hogwild placement of new points
    // (n_fixed > 0, edges from the new points only) has no mirrored edges and pulls with factor 1
    p.attraction_scale = (parallel == 0 && n_fixed > 0) ? 1.f : 1.5f;
    float* Y = static_cast<float*>(Y_.mutable_data());
    float* eons = static_cast<float*>(eons_.mutable_data());
    float* eonns = static_cast<float*>(eonns_.mutable_data());
    Dev dY(Y, (size_t)n * dim * sizeof(float)), dh = up(head), dt = up(tail), de = up(eps);
    Dev don(eons, (size_t)m * sizeof(float)), donn(eonns, (size_t)m * sizeof(float));
    Dev dp = n_mid ? up(mid_pairs) : Dev();
    const bool rev = parallel == 1 && n_mid > 0 && rev_indptr.size() > 0;
    Dev drp = rev ? up(rev_indptr) : Dev(), dri = rev ? up(rev_indices) : Dev();
    {
        py::gil_scoped_release rel;
        check(mg_optimize_layout(&p, (float*)dY.p, n, dim, (const int32_t*)dh.p, (const int32_t*)dt.p, (const float*)de.p,
                                 (float*)don.p, (float*)donn.p, m, n_mid ? (const int32_t*)dp.p : nullptr, n_mid,
                                 rev ? (const int32_t*)drp.p : nullptr, rev ? (const int32_t*)dri.p : nullptr));
        dY.to_host(Y);
        don.to_host(eons);
        donn.to_host(eonns);
    }
}

}  // namespace

PYBIND11_MODULE(_magnet, m) {
    m.doc() = "MAGNET t-SNE path on MI355X (HIP, gfx950): the reference's _magnet names over libmagnet_hip";
    py::class_<FFTConv2D>(m, "FFTConv2D", "FFT repulsion with persistent device buffers: sort, spread, FFTs, products, gather")
        .def(py::init<int, int>(), py::arg("M"), py::arg("n_interp"))
        .def("set_kernels", &FFTConv2D::set_kernels, py::arg("K1h"), py::arg("K2h"))
        .def("compute", &FFTConv2D::compute, py::arg("Y"), py::arg("xmin"), py::arg("ymin"), py::arg("width"), py::arg("n_boxes"),
             py::arg("n_threads") = -1);
    m.def("tsne_attraction", &tsne_attraction, "sum_j p_ij (1 + d^2/dof)^-1 (y_i - y_j) over a CSR graph", py::arg("indptr"),
          py::arg("indices"), py::arg("data"), py::arg("Y"), py::arg("dof") = 1.0f, py::arg("n_threads") = -1);
    m.def("tsne_step", &tsne_step, "Fused t-SNE iteration: gradient combination, gains, momentum, update, recentring",
          py::arg("Y"), py::arg("update"), py::arg("gains"), py::arg("attr"), py::arg("F"), py::arg("rho"), py::arg("inv_sumQ"),
          py::arg("lr"), py::arg("momentum"), py::arg("n_threads") = -1);
    m.def("tsne_update", &tsne_update, "Fused gains / momentum / position update / recentring", py::arg("Y"), py::arg("update"),
          py::arg("gains"), py::arg("grad"), py::arg("lr"), py::arg("momentum"), py::arg("n_threads") = -1);
    m.def("mid_spring", &mid_spring, "grad += w (Y_i - mean of partner positions)", py::arg("Y"), py::arg("pairs"),
          py::arg("grad"), py::arg("w"), py::arg("n_threads") = -1);
    m.def("bbox_2d", &bbox_2d, "Bounding box of a (n, 2) array", py::arg("Y"), py::arg("n_threads") = -1);
    m.def("interp_scatter", &interp_scatter, "Spread point charges on an interpolation grid (2-D)", py::arg("Y"),
          py::arg("charges"), py::arg("xmin"), py::arg("ymin"), py::arg("width"), py::arg("n_boxes"), py::arg("n_interp"),
          py::arg("n_threads") = -1);
    m.def("interp_gather", &interp_gather, "Read grid potentials back at the points (2-D)", py::arg("Y"), py::arg("grid"),
          py::arg("xmin"), py::arg("ymin"), py::arg("width"), py::arg("n_boxes"), py::arg("n_interp"), py::arg("n_threads") = -1);
    m.def("scatter_padded_2d", &scatter_padded_2d, "Charges (1, x, y) on a zero-padded M x M grid", py::arg("Y"),
          py::arg("xmin"), py::arg("ymin"), py::arg("width"), py::arg("n_boxes"), py::arg("n_interp"), py::arg("M"),
          py::arg("n_threads") = -1, py::arg("out") = arr<float>(0));
    m.def("spectral_products", &spectral_products, "Kernel-spectrum products of the FFT repulsion", py::arg("K1h"),
          py::arg("K2h"), py::arg("Wh"), py::arg("n_threads") = -1);
    m.def("gather_forces_2d", &gather_forces_2d, "Potentials at the points and repulsive forces: (sumQ, F)", py::arg("Y"),
          py::arg("phi"), py::arg("xmin"), py::arg("ymin"), py::arg("width"), py::arg("n_boxes"), py::arg("n_interp"),
          py::arg("n_threads") = -1);
    m.def("tsne_optimize", &tsne_optimize,
          "Whole t-SNE loop on the device (float32 or float64 state, in place); returns sumQ per iteration",
          py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("Y"), py::arg("update"), py::arg("gains"),
          py::arg("n_iter"), py::arg("early_iter"), py::arg("early_exaggeration"), py::arg("exaggeration"),
          py::arg("learning_rate"), py::arg("momentum0"), py::arg("momentum1"), py::arg("n_interp"), py::arg("box_width"),
          py::arg("dof"), py::arg("pairs") = py::none(), py::arg("mid_weight") = 0.0, py::arg("mid_phase") = 0.5,
          py::arg("it0") = 0, py::arg("it1") = -1, py::arg("min_boxes") = 20, py::arg("max_boxes") = 1000,
          py::arg("callback") = py::none());
    m.def("mid_near_pairs_dense", &mid_near_pairs_dense, "PaCMAP-style mid-near partners on the device", py::arg("X"),
          py::arg("n_pairs"), py::arg("n_cand"), py::arg("seed"), py::arg("n_threads") = -1);
    m.def("nn_descent", &nn_descent, "NN-descent k-NN graph on the device: (indices, squared distances), rows sorted",
          py::arg("X"), py::arg("k"), py::arg("max_candidates") = 50, py::arg("n_iters") = 10, py::arg("delta") = 0.001f,
          py::arg("n_trees") = 4, py::arg("leaf_size") = 30, py::arg("seed") = 0, py::arg("n_threads") = -1,
          py::arg("verbose") = false);
    m.def("optimize_layout", &optimize_layout, "Sampled SGD layout optimisation on the device (in place)",
          py::arg("Y"), py::arg("head"), py::arg("tail"), py::arg("epochs_per_sample"),
          py::arg("epoch_of_next_sample"), py::arg("epoch_of_next_negative_sample"),
          py::arg("indptr"), py::arg("indices"),
          py::arg("epoch_start"), py::arg("epoch_end"), py::arg("alpha0"), py::arg("alpha1"),
          py::arg("mid_w0"), py::arg("mid_w1"),
          py::arg("a"), py::arg("b"), py::arg("rep_a"), py::arg("rep_b"), py::arg("negative_sample_rate"), py::arg("loss"),
          py::arg("exaggeration"), py::arg("repulsion_strength"), py::arg("walk_len"), py::arg("mid_rate"),
          py::arg("optimizer"), py::arg("clip_mode"), py::arg("n_fixed"),
          py::arg("adam_m"), py::arg("adam_v"), py::arg("adam_t"), py::arg("beta1"), py::arg("beta2"),
          py::arg("walk_lens"), py::arg("mid_pairs"), py::arg("jump_indptr"), py::arg("jump_indices"),
          py::arg("mid_jump"), py::arg("kernel"), py::arg("node_b"), py::arg("seed"), py::arg("n_threads"), py::arg("parallel") = 0,
          py::arg("rev_indptr") = arr<int32_t>(0), py::arg("rev_indices") = arr<int32_t>(0));
    m.attr("__version__") = "3.0.2+mi355x";
}
