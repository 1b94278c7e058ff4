// Shared device helpers and the internal launcher interface of libmagnet_hip (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <stdexcept>
#include <cmath>

namespace mg {

typedef int64_t i64;

// ---- errors: every launcher throws mg::Error; the C-ABI turns it into a status code + message
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void hip_check(hipError_t e, const char* what);
#define MG_HIP(call) ::mg::hip_check((call), #call)
inline void require(bool ok, const std::string& msg, int code = 1) { if (!ok) throw Error(code, msg); }
constexpr int ERR_UNSUP = 4;   // the C-ABI's MG_ERR_UNSUP

constexpr int WAVE = 64;
constexpr int MAX_M = 8192;          // largest device FFT size (n_interp * n_boxes <= 4096)

template <class T> struct cplx { T x, y; };
template <class T> __host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class T> __host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __host__ __device__ inline cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }

template <class T> struct vec2;
template <> struct vec2<float> { typedef float2 type; };
template <> struct vec2<double> { typedef double2 type; };

__device__ inline int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Lagrange weights of n_interp nodes at (k + 1/2) / n_interp, local coordinate u in [0, 1]
// (closed form for three nodes as in src/fftconv.cpp:24-29, product form otherwise).
template <class T, int NI> __device__ inline void lagrange(T u, T* w) {
    if constexpr (NI == 3) {
        const T a = u - T(1) / T(6), b = u - T(0.5), c = u - T(5) / T(6);
        w[0] = b * c * T(4.5); w[1] = -a * c * T(9); w[2] = a * b * T(4.5);
    } else {
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const T uk = (k + T(0.5)) / NI;
            T num = 1, den = 1;
#pragma unroll
            for (int m = 0; m < NI; ++m) {
                if (m == k) continue;
                const T um = (m + T(0.5)) / NI;
                num *= (u - um); den *= (uk - um);
            }
            w[k] = num / den;
        }
    }
}

// Grid of one repulsion call, in the arithmetic type T (float for the FP32 path, as the reference
// passes float xmin / ymin / width to FFTConv2D::compute), and in double. The box of a point and its
// coordinates within the box use the double values, which are not rounded to T: the FP32 path then
// places every point exactly as the FP64 path does (with a grid origin rounded to float, a point within
// a rounding of a box edge would change box, and carry the other box's interpolation error).
template <class T> struct GridT {
    T xmin, ymin, width, inv_w;
    double xmin_d, ymin_d, width_d, inv_w_d;
    int nb, N;
};

// Grid decision of FFTRepulsion2D._grid (magnet/tsne.py:30-46) on a bounding box, in double. The same function
// runs on the host (the single loop) and on the device (the device-driven batch), so both decide the same grid.
struct GridDecision { double xmin, ymin, width; int nb; };
__host__ __device__ inline GridDecision decide_grid(const double* bb, double bw, int minb, int maxb) {
    const double lox = bb[0] - 1e-3, loy = bb[1] - 1e-3, hix = bb[2] + 1e-3, hiy = bb[3] + 1e-3;
    const double sx = hix - lox, sy = hiy - loy;
    const double span = sx > sy ? sx : sy;
    double nbf = ceil(span / bw);
    nbf = nbf < (double)minb ? (double)minb : nbf;
    nbf = nbf > (double)maxb ? (double)maxb : nbf;
    int nb = (int)nbf;
    nb = ((nb + 3) / 4) * 4;
    nb = nb < maxb ? nb : maxb;
    const double q = span / nb;
    const double width = (nb * bw >= span) ? bw : (q > 1e-12 ? q : 1e-12);
    return {lox, loy, width, nb};
}

template <class T> __host__ __device__ inline GridT<T> make_grid(double xmin, double ymin, double width, int nb, int ni) {
    GridT<T> g;
    g.xmin = (T)xmin; g.ymin = (T)ymin; g.width = (T)width;
    g.inv_w = T(1) / g.width;
    g.xmin_d = xmin; g.ymin_d = ymin; g.width_d = width;   // unrounded: boxes and local coordinates of both paths agree
    g.inv_w_d = 1.0 / g.width_d;
    g.nb = nb; g.N = nb * ni;
    return g;
}

// The padded FFT sizes with compiled plans: 2^s and 3 * 2^s from 128 to MAX_M; the smallest M >= 2N (0 when none).
__host__ __device__ inline int fft_size(int N) {
    for (int M = 128; M <= MAX_M; M *= 2) {
        if (3 * (M / 4) >= 2 * N && 3 * (M / 4) >= 128) return 3 * (M / 4);
        if (M >= 2 * N) return M;
    }
    return 0;
}
// the supported size after / before M (0 when none)
__host__ __device__ inline int fft_size_next(int M) { return fft_size(M / 2 + 1); }
__host__ __device__ inline int fft_size_prev(int M) {
    int best = 0;
    for (int P = 128; P <= MAX_M; P *= 2) {
        const int a = 3 * (P / 4);
        if (a >= 128 && a < M && a > best) best = a;
        if (P < M && P > best) best = P;
    }
    return best;
}

// Box of a point: x-box by multiplication with 1/width (the counting-sort key of src/fftconv.cpp:88),
// y-box by division (:106-110). Key = bx * nb + by. Evaluated in double for both paths: the FP32 path
// then puts every point in the box the FP64 path would on the same grid (float evaluation moves
// points that lie within a rounding of a box edge, and each such point carries the interpolation
// error of the neighbouring box).
template <class T> __device__ inline int box_key(T x, T y, const GridT<T>& g) {
    const int bx = min(g.nb - 1, max(0, (int)(((double)x - g.xmin_d) * g.inv_w_d)));
    const double uy = ((double)y - g.ymin_d) / g.width_d;
    const int by = min(g.nb - 1, max(0, (int)uy));
    return bx * g.nb + by;
}

// ---- compensated FP64 sums: a running sum s and the rounding error c it has dropped (Neumaier's variant of Kahan
// summation); two such pairs combine through an error-free TwoSum of their sums. The value is s + c.
struct CSum { double s, c; };
__device__ __host__ inline void csum_add(CSum& a, double x) {
    const double t = a.s + x;
    a.c += (fabs(a.s) >= fabs(x)) ? (a.s - t) + x : (x - t) + a.s;
    a.s = t;
}
__device__ __host__ inline CSum csum_merge(CSum a, CSum b) {
    const double t = a.s + b.s, bp = t - a.s;
    const double e = (a.s - (t - bp)) + (b.s - bp);
    return {t, a.c + b.c + e};
}

// ---- launch-shape choices that change no result (set through mg_tune_set; the defaults are measured choices)
struct Tunables {
    int attr_window = -1;   // LDS window sweep of the 2-D attraction: 1 on, 0 off, -1 from 262,144 points on
    int attr_rows = 256;    // rows per block of the window sweep
    int attr_win_kb = 64;   // LDS window of the sweep: 128 or 64 KiB of positions (0: no window, rows only)
    int attr_tpb = 1024;    // threads per block of the sweep: 1024, 512 or 256
    int sort_items = 0;     // keys per thread of a radix-sort pass up to 2^20 keys: 1, 2, 4 or 8 (0: 2 up to 262,144 keys, else 8)
    int fft_strided_max = 1024;   // padded FFT sizes up to which the column pass works in the row layout (no transposes)
    int fft_fold = 4;       // above that size, rows per block of the row passes that write / read the column layout
                            // themselves (4 or 2; FP64 takes 2), or 0: row passes and transposes
    int fft_fold_persist = 3;   // folded row passes as one block per CU over its XCD's row groups, with the next group's
                                // loads in flight: bit 0 the inverse pass, bit 1 the forward pass (else one block per
                                // row group); used where the per-group form holds one block per CU
    int spec_issue = 1;     // recomputed spectra (second stream) issued first (0), after the box sort (1) or the row pass (2)
    int sort_digit = 8;     // digit bits of the radix sorts of keys of 9-16 bits: 8 (two passes either way), or 10
    int batch_graph = 1;    // the batched loop replays its iteration as a captured hipGraph (0: issues the launches)
    int batch_chunk = 16;   // iterations the batched loop queues between two looks at the instances' states
    int batch_window_down = 1;   // the batched plan's window of FFT sizes holds the size below the current one
    int batch_window_up = 2;     // and this many sizes above it
    int batch_wl_grid = 2048;    // most blocks of a batched grid-stage launch over its work list
};
Tunables& tunables();

// Raises a kernel's dynamic LDS limit to `bytes` (when above the 64 KiB default), once per kernel and size for the
// process: the HIP runtime's function attributes are shared by every host thread, so the calls are serialised
// under one lock and none is made again on later launches (nor while a stream is being captured).
int device_cus();   // compute units of the current device (read once per device)
void kernel_lds_limit(const void* f, int bytes);

// ---- scratch of one repulsion engine (device pointers), sized by the host
template <class T> struct RepBuffers {
    int *keysA, *valsA, *keysB, *valsB;   // radix-sort ping-pong (box key, point index)
    int* hist;                            // digit rows (up to 1024 x nblocks) + digit totals
    int* item_key;                        // 2 per spread chunk
    int* item_valid;
    T* item_val;                          // 2 per chunk x NI x 3NI
    T* W;                                 // (3, N, N) charges
    cplx<T>* Wh;                          // (3, Mh, N) row spectra, column-major for the column pass
    cplx<T>* Ph;                          // (4, Mh, N) products after the column inverse
    cplx<T>* Tr;                          // (4, N, Mh) row-major staging of the transposes
    T* phi;                               // (4, N, N) potentials, planar
    double* zpart;                        // per-block Z partials
    double* zout;                         // [0] sumQ, [1] 1/max(sumQ, 1e-12)
    T* inv_out;                           // 1/max(sumQ, 1e-12) rounded to T
};

// ---- launchers (hipcc-compiled translation units), all on `st`
template <class T> void launch_bbox(const T* Y, i64 n, T* part, unsigned* counter, T* out, hipStream_t st);
template <class T> void launch_attraction(const int32_t* ip, const int32_t* ix, const T* P, const T* Y, i64 n,
                                          int dim, T inv_dof, T* out, i64 nnz, hipStream_t st);
// mode 0: grad = rho * attr - F * inv (inv from d_inv when non-null, else inv_h); mode 1: grad given in A.
// pairs != null adds the mid-range spring w * (Y_i - mean partners) to the gradient.
template <class T> void launch_update(int mode, T* Y, T* U, T* G, const T* A, const T* F, const T* d_inv, T inv_h,
                                      T rho, T lr, T mom, const int32_t* pairs, int m, T w_mid, i64 n, int dim,
                                      double* part, T* mean_bbox, const int* inv, hipStream_t st, T* host_mb = nullptr,
                                      int* host_tag = nullptr, int tag = 0);
// host_mb / host_tag (pinned host memory, 2-D only): mean_bbox's six values are also written there, then tag into
// *host_tag (system scope, after the values), so the host can wait for them without a copy or an event
// inv != null: the layout is relabelled (position inv[i] holds original point i); the mean and bbox are then
// summed in original point order
template <class T> void launch_mid_spring(const T* Y, const int32_t* pairs, T* grad, i64 n, int dim, int m, T w,
                                          hipStream_t st);
// gradient of the non-fused path: combine ? rho * A - F * inv : grad, plus the spring when pairs != null
template <class T> void launch_grad(const T* A, const T* F, const T* d_inv, T inv_h, T rho, const T* Y, const int32_t* pairs,
                                    int m, T w, i64 n, int dim, int combine, T* grad, hipStream_t st);
template <class T> void launch_fill(T* p, T v, i64 count, hipStream_t st);

// box sort: returns which buffer holds the result (0 = A, 1 = B); inv != null: the layout is relabelled (position
// inv[i] holds original point i) and the points are taken in original order, values their positions
template <class T> int launch_box_sort(const T* Y, i64 n, const GridT<T>& g, RepBuffers<T>& b, hipStream_t st,
                                       const int* inv = nullptr);
int sort_hist_blocks(i64 n);
// ints of a radix sort's digit-row scratch for n keys, enough for any sort_items setting
i64 sort_hist_ints(i64 n);
int spread_chunks(i64 n);
// centred: charges (1, x - x_node, y - y_node) and forces from the centred potentials (FP32 path)
// fill: clear the grid's 3 N^2 charges first (false: the caller keeps the grid zero between calls)
template <class T> void launch_spread(const int* keys, const int* vals, const T* Y, i64 n, const GridT<T>& g, int ni,
                                      bool centred, RepBuffers<T>& b, hipStream_t st, bool fill = true);
// Wclr != null: the gather also sets the charge nodes of every occupied box of W = Wclr back to zero
template <class T> void launch_gather(const int* keys, const int* vals, const T* Y, i64 n, const GridT<T>& g, int ni,
                                      bool centred, const T* phi, T* F, RepBuffers<T>& b, double* sumq_hist, hipStream_t st,
                                      T* Wclr = nullptr);
int gather_blocks(i64 n);
// out-of-sample placement of m new points against a fixed training layout and its potential field phi
template <class T> void launch_transform(const T* Ytr, i64 n_train, const int32_t* nbr, const T* w, int k, T* Yn, i64 m,
                                         const GridT<T>& g, int ni, bool centred, const T* phi, double inv_Z, int n_iter,
                                         double lr, double mom0, double mom1, int mom_switch, double rho, double dof,
                                         const double* centroid, hipStream_t st);
// the mean of a layout (n, 2) in double into c[0], c[1]
template <class T> void launch_centroid(const T* Y, i64 n, double* c, hipStream_t st);

// the k (<= 32) smallest entries of each row of a (rows, cols) float tile, sorted by (value, column) (k_knn.hip)
void launch_topk_rows(const float* D, i64 rows, i64 cols, int k, int32_t* oi, float* ov, hipStream_t st);

// fused exact k-NN search (features <= 64, k <= 32): squared distances and indices sorted by (distance, index)
int knn_chunks(i64 m, i64 nt);
void launch_knn(const float* T, i64 nt, const float* Q, i64 m, int d, int k, i64 self_offset, float* part_d, int* part_i,
                int32_t* oi, float* ov, hipStream_t st);
// the same search with its scratch in a fixed device array (no allocation); at most knn_noalloc_max_queries(k) queries
i64 knn_noalloc_max_queries(int k);
void launch_knn_noalloc(const float* T, i64 nt, const float* Q, i64 m, int d, int k, i64 self_offset, int32_t* oi, float* ov,
                        hipStream_t st);

// smooth k-NN memberships of sorted neighbour distances (m, k), per row; mean_all (m): the fallback mean distance
void launch_smooth_knn(const float* dist, i64 m, int k, int n_neighbors, const float* mean_all, float* vals, hipStream_t st);
// perplexity calibration of k squared neighbour distances per row (P (n, k) double, beta (n) double or null)
void launch_perplexity(const float* d2, i64 n, int k, double perplexity, double tol, int n_iter, double* P, double* beta,
                       hipStream_t st);
// symmetrisation of a directed k-NN graph into CSR (mode 0: (a + b) * param; 1: fuzzy union mixed by param); scratch of
// symmetrize_scratch_ints(n, k) ints; returns nnz, or -nnz when nnz exceeds cap (nothing written but the row pointers)
i64 symmetrize_scratch_ints(i64 n, int k);
// exclusive scan of n ints into out[0 .. n] (out[n] the total), with scan_tmp_ints(n) ints of scratch
i64 scan_tmp_ints(i64 n);
void launch_exscan(const int* a, i64 n, int* out, int* tmp, hipStream_t st);
// NN-descent k-NN graph of the rows of X (n, d): out_idx / out_dist (n, K), squared distances, sorted; synchronises st
void nn_descent(const float* X, i64 n, int d, int K, int max_candidates, int n_iters, float delta, int n_trees, int leaf_size,
                uint64_t seed, int32_t* out_idx, float* out_dist, int verbose, hipStream_t st);
void mid_near_pairs(const float* X, i64 n, int d, int n_pairs, int n_cand, uint64_t seed, int32_t* out, hipStream_t st);
// exact k nearest other rows of every row of X (k <= 128), NN-descent's lists as the bound; synchronises st
void knn_self_exact(const float* X, i64 n, int d, int K, int32_t* out_idx, float* out_dist, hipStream_t st);
i64 symmetrize_knn(int mode, const int32_t* idx, const double* val, i64 n, int k, double param, int32_t* indptr,
                   int32_t* out_idx, float* out_val, i64 cap, int* scratch, hipStream_t st);

// locality relabelling (k_order.hip)
void launch_identity_labels(int* perm, int* inv, i64 n, hipStream_t st);
template <class T> void launch_relabel(const int* vals, const int* perm, i64 n, const int32_t* ip, const int32_t* ix, const T* P,
                                       int* order, int* inv_new, int* len, int* tiles, int32_t* ipn, int32_t* ixn, T* Pn,
                                       hipStream_t st);
template <class T> void launch_perm_state(const T* Y, const T* U, const T* G, const int* inv_old, const int* order, i64 n,
                                          T* Yn, T* Un, T* Gn, hipStream_t st);
// sampled layout optimisation (k_layout.hip): epochs [epoch_start, epoch_end) of the owner-mode SGD on Y (n, 2) in
// place; scratch of layout_scratch_bytes(n)
size_t layout_scratch_bytes(i64 n);
void validate_csr(const int32_t* ip, const int32_t* ix, i64 n, i64 nnz, hipStream_t st);

// per-iteration diagnostics (KL divergence, gradient norm) in FP64 with compensated sums; scratch of
// trace_scratch_doubles() doubles, whose last part launch_trace_plogp fills once per run
int trace_scratch_doubles();
template <class T> void launch_trace_plogp(const T* P, i64 nnz, double* scratch, hipStream_t st);
template <class T> void launch_trace(const int32_t* ip, const int32_t* ix, const T* P, const T* Y, const int* inv, i64 n,
                                     double dof, const T* A, const T* F, const T* d_inv, T rho, const double* zout,
                                     double* scratch, double* out, hipStream_t st);
// relabelling sort input: the Morton key of each point's box (original order), value its current position
template <class T> void launch_morton_keys(const T* Y, const int* inv, i64 n, const GridT<T>& g, int* keys, int* vals,
                                           hipStream_t st);
template <class T> void launch_unperm(const T* src, const int* inv, i64 n, T* dst, hipStream_t st);
template <class T> int launch_key_sort(int* keysA, int* valsA, int* keysB, int* valsB, int* hist, i64 n, int bits, hipStream_t st);

// FFT stages. M = 2^s or 3 * 2^s, 128 <= M <= MAX_M, N <= M / 2.
void fft_init_twiddles();   // uploads the twiddle tables once per device
template <class T> void launch_rows_forward(const T* W, int N, int M, cplx<T>* out, hipStream_t st);   // out (3, N, Mh)
template <class T> void launch_transpose(const cplx<T>* src, cplx<T>* dst, int batch, int R, int C, hipStream_t st);
// K: spectra tables K1, K2 (and x K2, y K2 when centred), each (M/2+1, M/2+1), [ky][kx <= M/2]. Wh (3, Mh, N),
// Ph (4, Mh, N). centred must be true for float and false for double.
template <class T> void launch_columns(const cplx<T>* Wh, int N, int M, const T* const* K, bool centred, cplx<T>* Ph,
                                       hipStream_t st, bool strided = false, int LD = 0);
// strided: Wh is the row pass's (3, N, Mh) output and Ph receives (4, N, Mh) for the inverse row pass, with no
// transposes (the same values); otherwise column ky of channel c starts at (c Mh + ky) LD (LD 0: N)
template <class T> void launch_rows_inverse(const cplx<T>* X, int N, int M, T* phi, hipStream_t st);   // X (4, N, Mh)
// the row passes with the transposes folded in: W (3, N, N) -> Wh (3, Mh, LD) and Ph (4, Mh, LD) -> phi (4, N, N), the
// same values as a row pass and a transpose; `rows` rows per block (4 or 2; FP64 takes 2); padded sizes 512-8192. The
// column layout's stride LD = fold_stride_n(N) rounds N up to 4 rows, so every 32-byte segment is aligned (pass the
// same LD to launch_columns; LD <= M/2 holds, M/2 being a multiple of 4)
inline int fold_stride_n(int N) { return (N + 3) & ~3; }
template <class T> bool fold_supported(int M);
template <class T> void launch_rows_forward_fold(const T* W, int N, int M, cplx<T>* Wh, int rows, hipStream_t st);
template <class T> void launch_rows_inverse_fold(const cplx<T>* Ph, int N, int M, T* phi, int rows, hipStream_t st);
// sym (ntab 4): the kernel is symmetric in x <-> y (K(a, b) = K(b, a)); y K2's table is then x K2's transposed
// (fewer transforms). kernel spectra on the M x M torus: values from (delta, dof) or from a host-built table (double, (2, M, M), even
// in each offset); ntab = 2 (K1, K2) or 4 (K1, K2, x K2, y K2); outputs (Mh, Mh) real, [ky][kx]
// (round_kernel: kernel values rounded to float before the transform). tmp: 4 Mh^2 complex values.
template <class T> void launch_spectra(int M, double delta, double dof, const double* table, bool round_kernel, int ntab,
                                       cplx<T>* tmp, T* const* K, hipStream_t st, bool sym = false);
template <class T> void launch_phi_interleave(const T* planar, T* out, i64 NN, hipStream_t st);   // (4,N,N) -> (N,N,4)

}  // namespace mg
