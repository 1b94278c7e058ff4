/* C-ABI of the MI355X (gfx950) implementation of MAGNET's full-batch loss="tsne" gradient pass.
 *
 * Every entry point is a drop-in for one function of the reference's native module `_magnet`
 * (kerighan/lib-magnet v3.0.2, src/bindings.cpp:8-83); the reference interface each one replaces is
 * cited beside it. Conventions shared by all entry points:
 *   - Return value: MG_OK (0) or an MG_ERR_* code; mg_last_error() gives the message of the calling
 *     thread's last failure (the reference's `require` -> std::runtime_error, src/layout.hpp:34).
 *   - `prec` selects the arithmetic: MG_F64 (double arrays) or MG_F32 (float arrays). Index arrays are
 *     int32 in both. Point arrays are (n, dim) row-major, 2-D paths use dim == 2.
 *   - Pointers named d_* are device (HBM) pointers; h_* are host pointers. Work is issued on the
 *     library stream and is ordered on it: the calling thread's own stream if it set one
 *     (mg_set_thread_stream), else the process-wide one (mg_set_stream, default the null stream).
 *   - Entry points that return a host scalar (sumQ, bbox) synchronise the stream before returning.
 */
#ifndef MAGNET_HIP_H
#define MAGNET_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MG_OK 0
#define MG_ERR_ARG 1     /* invalid argument or shape (reference: require() -> RuntimeError) */
#define MG_ERR_HIP 2     /* HIP runtime error */
#define MG_ERR_NOMEM 3   /* device allocation failed */
#define MG_ERR_UNSUP 4   /* valid in the reference, not supported on this path (message says why) */

#define MG_F32 0
#define MG_F64 1

/* ---- runtime ------------------------------------------------------------------------------------ */
const char* mg_last_error(void);
int mg_version(void);                                   /* 100 * major + minor */
int mg_set_stream(void* hip_stream);                    /* hipStream_t; NULL = null stream */
/* The calling thread's stream, in place of the process-wide one, for the rest of the thread's life. No
 * counterpart in the reference (its calls run on host threads). For callers that drive the device from
 * several host threads at once, e.g. a server with more than one batch in flight: entry points called from
 * different threads on different streams overlap on the device. mg_knn keeps its scratch per thread; the
 * other entry points share no scratch between threads, except those on one mg_fftconv object, which is
 * not re-entrant (mg_tsne_transform only reads the object's field and may run from several threads). */
int mg_set_thread_stream(void* hip_stream);
int mg_synchronize(void);
int mg_malloc(void** d_ptr, size_t bytes);
int mg_free(void* d_ptr);
int mg_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int mg_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int mg_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
/* HBM bytes free / total on the current device (serving admission). */
int mg_mem_info(size_t* free_bytes, size_t* total_bytes);

/* Launch-shape choices that change no result, process-wide (no counterpart in the reference; for tuning sweeps):
 *   "attr_window"  the 2-D attraction's LDS-window sweep: 1 on, 0 off, -1 (default) from 262,144 points on;
 *   "attr_rows"    rows per block of that sweep (default 256);
 *   "attr_win_kb"  its window of positions in LDS: 64 (default) or 128 KiB, or 0 for none (rows only);
 *   "attr_tpb"     its threads per block: 1024 (default), 512 or 256;
 *   "sort_items"   keys per thread of the radix sorts of up to 2^20 keys: 1, 2, 4 or 8, or 0 (default: 2 up to
 *                  262,144 keys, 8 above);
 *   "sort_digit"   digit bits of the radix sorts of keys of 9 to 16 bits (two passes either way): 8 (default) or 10;
 *   "fft_strided_max" padded FFT sizes up to which the column pass of the repulsion's 2-D FFT works in the row
 *                  pass's layout, with no transposes (default 1024; 0: never);
 *   "fft_fold"     above that size, rows per block of the row passes that write and read the column pass's layout
 *                  themselves: 4 (default) or 2 (FP64 takes 2), or 0 for row passes and transposes;
 *   "fft_fold_persist" bit 0: the folded inverse row pass runs one block per CU over its XCD's row groups with the
 *                  next group's loads in flight; bit 1: the folded forward pass does; 3 (default) both, 0 neither
 *                  (one block per group). Used where the per-group form holds one block per CU and the persistent
 *                  form's buffers and twiddles fit in LDS;
 *   "spec_issue"   when the kernel spectra are recomputed beside the loop's repulsion: issued first (0), after the
 *                  box sort (1, default) or after the row pass (2);
 *   "batch_graph"  1 (default): mg_tsne_optimize_batch replays its captured iteration (a hipGraph); 0: it issues the
 *                  same launches one by one;
 *   "batch_chunk"  iterations mg_tsne_optimize_batch queues between two reads of the instances' states (default 16);
 *   "batch_window_down", "batch_window_up"  FFT sizes below (0 or 1, default 1) and above (0 to 2, default 2) the
 *                  current one in each instance's window (a grid that leaves its window stalls its instance until the
 *                  next plan);
 *   "batch_wl_grid" most blocks of a batched grid-stage launch over its work list (default 2048).
 * MG_ERR_ARG for an unknown name or a value out of range. magnet.tune.autotune picks them by timing. */
int mg_tune_set(const char* name, int value);
int mg_tune_get(const char* name, int* value);

/* Per-stage timing of mg_tsne_optimize with HIP events on the working stream (the reference prints
 * per-stage laps of FFTConv2D::compute under MAGNET_DEBUG, src/fftconv.cpp:69-70; setting MAGNET_DEBUG
 * here enables this and prints every stage of every iteration to stderr). Stages: attraction,
 * box_sort, spread, rows_forward, columns, rows_inverse, gather, update, spectra. algo_bytes sums the
 * algorithmic HBM bytes of the timed launches (each input read once, each output written once). */
int mg_profile_enable(int on);
int mg_profile_reset(void);
int mg_profile_num_stages(void);
int mg_profile_get(int stage, const char** name, double* total_ms, int64_t* count, double* algo_bytes);
/* The same per padded FFT size M (the FFT stages' kernels are compiled per size; other stages report M = 0): the
 * number of sizes seen in *n_sizes, and up to max_sizes of them in M / total_ms / count / algo_bytes. */
int mg_profile_sized(int stage, int max_sizes, int* n_sizes, int* M, double* total_ms, int64_t* count, double* algo_bytes);
/* Grid sizes of the profiled iterations: mean and largest live side N = n_interp * n_boxes, largest M. */
int mg_profile_grid(double* mean_N, int* max_N, int* max_M);

/* Padded FFT size the device uses for a live grid side N = n_interp * n_boxes: the smallest of 2^s and
 * 3 * 2^s (128 <= M <= 8192) that is >= 2N, or 0 when N > 4096. The reference pads to
 * scipy.fft.next_fast_len(2N) (magnet/tsne.py:50); the results do not depend on the padding beyond
 * rounding, since any M >= 2N - 1 holds the linear convolution. */
int mg_fft_size(int N);

/* ---- per-call kernels (device pointers) ----------------------------------------------------------- */

/* bbox_2d (src/fftrep.cpp:415-434): h_out = {xmin, ymin, xmax, ymax} of d_Y (n, 2), in double. */
int mg_bbox_2d(int prec, const void* d_Y, int64_t n, double* h_out);

/* tsne_attraction (src/fftrep.cpp:129-177): out_i = sum_j p_ij (1 + |y_i - y_j|^2 / dof)^-1 (y_i - y_j).
 * CSR (n + 1, nnz, nnz) with data in `prec`; dim in [1, 64]. The CSR is validated on the device
 * (validate_csr, src/layout.hpp:21-32). The order of every sum is fixed, so the result is deterministic. */
int mg_tsne_attraction(int prec, const int32_t* d_indptr, const int32_t* d_indices, const void* d_data,
                       int64_t nnz, const void* d_Y, int64_t n, int dim, double dof, void* d_out);

/* tsne_step (src/fftrep.cpp:225-266): grad = rho * attr - F * inv_sumQ, delta-bar-delta gains,
 * momentum, Y += U, Y -= mean(Y); Y, U, G updated in place. */
int mg_tsne_step(int prec, void* d_Y, void* d_U, void* d_G, const void* d_attr, const void* d_F, int64_t n,
                 int dim, double rho, double inv_sumQ, double lr, double momentum);

/* tsne_update (src/fftrep.cpp:181-220): the same update from a given gradient. */
int mg_tsne_update(int prec, void* d_Y, void* d_U, void* d_G, const void* d_grad, int64_t n, int dim,
                   double lr, double momentum);

/* mid_spring (src/fftrep.cpp:270-287): grad_i += w (Y_i - mean_k Y[pairs[i, k]]), pairs (n, m). */
int mg_mid_spring(int prec, const void* d_Y, const int32_t* d_pairs, void* d_grad, int64_t n, int dim, int m,
                  double w);

/* ---- FFT repulsion object: FFTConv2D (src/fftconv.cpp:43-237) ------------------------------------- */
typedef struct mg_fftconv mg_fftconv;

/* FFTConv2D(M, n_interp) (src/fftconv.cpp:45-47). M is the padded size the caller's spectra use; the
 * device transforms run on mg_fft_size(N), the smallest of 2^s and 3 * 2^s that is >= 2N, whatever M is
 * (any padded size >= 2N - 1 gives the same convolution). n_interp in [1, 8]. */
int mg_fftconv_create(mg_fftconv** out, int prec, int M, int n_interp);
int mg_fftconv_destroy(mg_fftconv* conv);

/* set_kernels (src/fftconv.cpp:51-55) in kernel-value form: h_K1, h_K2 are (M_in, M_in) kernel tables
 * on the padded torus, sampled at offsets (m or m - M_in) * delta (magnet/tsne.py:57-67), i.e. the
 * inverse transforms of the spectra FFTConv2D.set_kernels takes (the binding converts one into the
 * other). Offsets up to min(M_in, M_dev) / 2 are used; the device re-transforms at its own size. */
int mg_fftconv_set_kernel_values(mg_fftconv* conv, int M_in, const double* h_K1, const double* h_K2);

/* The t-SNE kernels of FFTRepulsion2D._kernels (magnet/tsne.py:48-71): K1 = (1 + r^2/dof)^-dof,
 * K2 = (1 + r^2/dof)^(-dof-1) at node spacing width / n_interp. Spectra are computed on the device
 * (FP32 objects round the kernel values to float first, like the reference) for each padded size a
 * compute call needs, and cached. */
int mg_fftconv_set_kernels_tsne(mg_fftconv* conv, double width, double dof);

/* compute (src/fftconv.cpp:58-230): F (n, 2) and h_sumQ = sum_i pot0_i - n for the grid
 * (xmin, ymin, width, n_boxes). */
int mg_fftconv_compute(mg_fftconv* conv, const void* d_Y, int64_t n, double xmin, double ymin, double width,
                       int n_boxes, void* d_F, double* h_sumQ);

/* Intermediate grids of the last compute, for tests: W (3, N, N) charges, phi (N, N, 4) potentials
 * (interleaved, after the 1/M^2 scale). Either pointer may be NULL. */
int mg_fftconv_grids(mg_fftconv* conv, void* d_W, void* d_phi);

/* Out-of-sample placement (MAGNET.transform, magnet/__init__.py:876-930, in the full-batch t-SNE form): m new points
 * Y_new (m, 2, initialised by the caller) move for n_iter iterations under the attraction of their k training
 * neighbours (nbr (m, k) int32 indices into Y_train, weights w (m, k) with rows summing to one, scaled by 1/n_train
 * like the rows of the training P) and the repulsion of the training layout, read from the potential field `conv`
 * holds from its last mg_fftconv_compute (call it on Y_train first); gradient rho A - F/Z with Z the training
 * layout's, gains and momentum as in the training update (momentum0 before iteration momentum_switch, then
 * momentum1). A point outside the field's grid takes the far-field repulsion of the training layout as n_train
 * points at its centroid (the centroid of the layout of the last mg_fftconv_compute). The training layout is not
 * moved, and every new point is independent of the others: a point's result does not depend on which points share
 * the call. */
int mg_tsne_transform(mg_fftconv* conv, const void* d_Ytrain, int64_t n_train, const int32_t* d_nbr, const void* d_w,
                      int k, void* d_Ynew, int64_t m, int n_iter, double learning_rate, double momentum0, double momentum1,
                      int momentum_switch, double exaggeration, double dof);

/* Selection step of the serving k-NN search (the reference's transform() queries its k-NN index,
 * magnet/__init__.py:876-900): for each row of the (rows, cols) float32 device tile D, the k (1..32) smallest
 * entries sorted by (value, column): d_idx / d_val (rows, k). */
int mg_topk_rows(const float* d_D, int64_t rows, int64_t cols, int k, int32_t* d_idx, float* d_val);
/* The whole serving k-NN search for rows of d float32 features (the reference's k-NN query, magnet/knn.py; rows of
 * more than 64 features go through the features in chunks of 64, with the same sums):
 * for each of the m query rows d_Q, the k (1..32) nearest of the nt training rows d_T, exact squared euclidean
 * distances summed over the features in order, sorted by (distance, index): d_idx / d_dist2 (m, k).
 * self_offset >= 0 leaves out training row q + self_offset for query q (the graph of the training rows). A query's
 * result does not depend on the other queries of the call. Calls are serialised (one scratch per process). */
int mg_knn(const float* d_T, int64_t nt, const float* d_Q, int64_t m, int d, int k, int64_t self_offset, int32_t* d_idx,
           float* d_dist2);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops): the search of mg_knn, with the same results bit for bit, that allocates nothing and does
 * not synchronise (it returns once its launches are queued on the library stream). Its scratch is a fixed 8 MiB
 * device array that each device holds until the process ends, so m is limited to 65536 query rows for k <= 16 and
 * 32768 for k <= 32 (MG_ERR_ARG beyond). Calls that overlap on one device share that array and race. */
int mg_knn_noalloc(const float* d_T, int64_t nt, const float* d_Q, int64_t m, int d, int k, int64_t self_offset,
                   int32_t* d_idx, float* d_dist2);
/* smooth_knn (magnet/knn.py:165-198) on the device, float32: memberships d_vals (m, k) of the sorted neighbour
 * distances d_dist (m, k), one row per thread; d_mean_all (m) is each row's fallback mean distance (the mean over
 * its call, used when the row has no non-zero distance). */
int mg_smooth_knn(const float* d_dist, int64_t m, int k, int n_neighbors, const float* d_mean_all, float* d_vals);
/* Perplexity calibration on the device: for each row of k squared neighbour distances d_dist2 (n, k) float32, the
 * Gaussian precision beta whose conditional distribution p_j|i = exp(-beta (d_ij - min_j d_ij)) / sum has entropy
 * log(perplexity) (natural log) within tol, by bisection from beta = 1 (doubling / halving while a bound is open),
 * at most n_iter steps, each row on its own; d_P (n, k) float64 receives p_j|i, d_beta (n, may be NULL) beta. No
 * counterpart in the reference library, which takes these affinities from openTSNE in its benchmark scripts
 * (benchmarks/opentsne_audit.py:18-23); the configurations it reports on are built from them. */
int mg_perplexity_rows(const float* d_dist2, int64_t n, int k, double perplexity, double tol, int n_iter, double* d_P,
                       double* d_beta);
/* Symmetric CSR of a directed k-NN graph d_idx (n, k) int32 (indices in [0, n)) with values d_val (n, k) float64:
 * each unordered pair {i, j} gets a = its i -> j value(s) and b = its j -> i value(s), and the entries (i, j) and
 * (j, i) the value
 *   mode 0: (a + b) * param                                   ((P + P^T) / 2n for param = 1 / 2n),
 *   mode 1: param (a + b - a b) + (1 - param) a b             (the fuzzy union mixed with the intersection by
 *                                                              set_op_mix_ratio = param: magnet/knn.py:215-230),
 * in float64, stored as float32. Diagonal entries are dropped, and zeros (mode 0: in float64, mode 1: in float32).
 * Rows sorted by column. d_indptr (n + 1), d_indices / d_data of capacity cap (2 n k always suffices); *h_nnz gets the
 * entry count (MG_ERR_ARG when it exceeds cap). Deterministic; allocates its scratch (about 20 n k bytes). */
int mg_symmetrize_knn(int mode, const int32_t* d_idx, const double* d_val, int64_t n, int k, double param,
                      int32_t* d_indptr, int32_t* d_indices, float* d_data, int64_t cap, int64_t* h_nnz);
/* nn_descent (src/nndescent.cpp:390-583, the binding of :636-645) on the device, dense rows: the k (<= 128) nearest
 * other rows of each row of d_X (n, d) float32, as d_idx (n, k) int32 and squared euclidean distances d_dist2 (n, k),
 * each row sorted by (distance, index). Random-projection trees (n_trees, leaf_size) seed the lists, random rows fill
 * what they leave open, then up to n_iters rounds of candidate sampling (at most max_candidates new and old per node,
 * capped at 64) and local joins, stopping when a round changes at most delta * n * k entries. Deterministic for a
 * seed; agrees with the reference's graph in law (recall against exact search), not in value. Allocates its working
 * memory (about 30 n k bytes plus the join's proposals); synchronises the library stream. verbose prints a line per
 * tree and round to stderr. */
int mg_nn_descent(const float* d_X, int64_t n, int d, int k, int max_candidates, int n_iters, float delta, int n_trees,
                  int leaf_size, uint64_t seed, int verbose, int32_t* d_idx, float* d_dist2);
/* mid_near_pairs_dense (src/nndescent.cpp:730-767) on the device: d_out (n, n_pairs) int32, for pair p of row i the
 * second nearest (squared euclidean, ties by index) of n_cand rows drawn by a hash of (seed, i, p, c) (a draw equal to
 * i moves to i + 1). The reference draws from per-thread generators: the partners agree in law, not in value. */
int mg_mid_near_pairs(const float* d_X, int64_t n, int d, int n_pairs, int n_cand, uint64_t seed, int32_t* d_out);
/* The exact k-NN graph of the rows of d_X (n, d) float32, k <= 128: for each row its k nearest other rows by squared
 * euclidean distance (one fused multiply-add per feature, in order), sorted by (distance, index), as d_idx (n, k)
 * and d_dist2 (n, k). NN-descent's lists bound the search (every true neighbour lies within the k-th distance of the
 * approximate list), and a scan of all pairs keeps the rows within that bound. No counterpart in the reference,
 * whose exact search is brute force in Python (benchmarks); the graph of the perplexity affinities the
 * configurations use. Allocates its scratch (about 16 k n bytes and NN-descent's); synchronises the library stream. */
int mg_knn_exact_graph(const float* d_X, int64_t n, int d, int k, int32_t* d_idx, float* d_dist2);

/* ---- sampled layout: optimize_layout (src/layout.cpp:343-472, src/bindings.cpp) ----------------------------------
 * Epochs [epoch_start, epoch_end) of MAGNET's sampled pass on the device, for loss "umap" with the power kernel
 * q = 1 / (1 + a d^2b), plain SGD (optimizer "sgd"), in the reference's owner parallel mode: every edge (head, tail)
 * is sampled every epochs_per_sample[e] epochs (state in d_eons / d_eonns, as the reference keeps it), moves its
 * head by the attraction (times 1.5), negative_sample_rate random nodes per sample repel it, and with mid-range
 * pairs (n, n_mid) a partner pulls it at rate mid_rate; the partners' share comes through the reverse pairs
 * (CSR rev_indptr / rev_indices: who chose each node; may be NULL). alpha and the mid-range weight move linearly
 * from (alpha0, mid_w0) to (alpha1, mid_w1) over the epochs. Nodes below n_fixed do not move. Edges must be sorted
 * by head. Y (n, 2) float32 in place. Every node is moved by one thread reading the others' positions at the start
 * of the epoch, so the result is deterministic; its random streams differ from the reference's (per node and epoch
 * instead of per thread chunk), so it agrees with the reference in law, not in value. The reference's other losses,
 * kernels, Adam, node exponents and random-walk partners are not on this path (MG_ERR_UNSUP: dim != 2). */
typedef struct {
    int32_t epoch_start, epoch_end;
    float alpha0, alpha1, mid_w0, mid_w1;
    float a, b, rep_a, rep_b;          /* rep_b <= 0: the repulsion uses (a, b) */
    float negative_sample_rate, exaggeration, repulsion_strength, mid_rate;
    int32_t clip_mode;                 /* 0: per coordinate ("coord"), 1: in norm ("norm") */
    int64_t n_fixed;
    uint64_t seed;
    float attraction_scale;            /* factor on the attraction of a head: 1.5 (the owner mode's, for graphs whose
                                          edges come in both directions; <= 0 also means 1.5) or 1 (edges without
                                          their reverse, as in out-of-sample placement) */
} mg_layout_params;
int mg_optimize_layout(const mg_layout_params* p, float* d_Y, int64_t n, int dim, const int32_t* d_head, const int32_t* d_tail,
                       const float* d_eps, float* d_eons, float* d_eonns, int64_t n_edges, const int32_t* d_pairs, int n_mid,
                       const int32_t* d_rev_indptr, const int32_t* d_rev_indices);

/* ---- fallback kernels (src/fftrep.cpp:51-125, 294-411; magnet/tsne.py:92-129 calls them when FFTConv2D is not
 * used). Box of a point: u = (x - xmin) / width clamped to [0, n_boxes - 1], product-form Lagrange weights;
 * N = n_boxes * n_interp. Sums run in the reference's single-thread order without fused multiply-adds, so the
 * float32 results are the reference's. Device pointers; complex arrays are interleaved (re, im) pairs. */
/* interp_scatter (src/fftrep.cpp:51-92): Y (n, 2), charges (n, C) -> out (C, N, N) */
int mg_interp_scatter(int prec, const void* d_Y, int64_t n, const void* d_charges, int C, double xmin, double ymin,
                      double width, int n_boxes, int n_interp, void* d_out);
/* interp_gather (src/fftrep.cpp:94-125): grid (C, N, N) -> out (n, C) */
int mg_interp_gather(int prec, const void* d_Y, int64_t n, const void* d_grid, int C, double xmin, double ymin,
                     double width, int n_boxes, int n_interp, void* d_out);
/* scatter_padded_2d (src/fftrep.cpp:294-343): charges (1, x, y) on the zero-padded grid, out (3, M, M), M >= N */
int mg_scatter_padded_2d(int prec, const void* d_Y, int64_t n, double xmin, double ymin, double width, int n_boxes,
                         int n_interp, int M, void* d_out);
/* spectral_products (src/fftrep.cpp:346-371): K1h, K2h (M, Mh) complex, Wh (3, M, Mh) complex ->
 * out (4, M, Mh) complex: K1h Wh[0], K2h Wh[0..2] */
int mg_spectral_products(int prec, const void* d_K1h, const void* d_K2h, const void* d_Wh, int64_t M, int64_t Mh, void* d_out);
/* gather_forces_2d (src/fftrep.cpp:373-411): phi (4, M, M) -> F (n, 2), *h_sumQ = sum pot0 - n */
int mg_gather_forces_2d(int prec, const void* d_Y, int64_t n, const void* d_phi, int M, double xmin, double ymin,
                        double width, int n_boxes, int n_interp, void* d_F, double* h_sumQ);

/* ---- whole loop: optimize_tsne (magnet/tsne.py:132-214) ------------------------------------------ */
typedef struct {
    int32_t n_iter;                /* schedule length (tsne.py:132 n_iter) */
    int32_t early_iter;            /* iterations with early exaggeration */
    double early_exaggeration;     /* rho during the early phase (12) */
    double exaggeration;           /* rho afterwards (1) */
    double learning_rate;          /* <= 0: max(200, n / early_exaggeration) (tsne.py:152-153) */
    double momentum0, momentum1;   /* (0.5, 0.8) */
    int32_t n_interp;              /* Lagrange nodes per box (3) */
    int32_t min_boxes, max_boxes;  /* (20, 1000) (tsne.py:18) */
    double box_width;              /* 0.7 */
    double dof;                    /* Student-t degrees of freedom (1) */
    double mid_weight, mid_phase;  /* mid-range spring (tsne.py:194-198); pairs may be NULL */
    int32_t relabel_every;         /* locality relabelling of the points every this many iterations (0: 25; < 0: off).
                                      It never changes a result (no counterpart in the reference) */
} mg_tsne_params;

/* Called after iteration `it` when it % 50 == 0 or it == n_iter - 1 (tsne.py:180-181) with the device
 * layout; return non-zero to stop. */
typedef int (*mg_tsne_callback)(int it, const void* d_Y, int64_t n, void* user);

/* Runs iterations [it0, it1) of the n_iter schedule on the device state (Y, U, G), all (n, 2) in `prec`
 * (U = 0 and G = 1 at it0 = 0 for a fresh run; keeping them makes a resumed run identical to an
 * uninterrupted one). The CSR data must be normalised to sum 1 (tsne.py:149-151). d_pairs (n, m_pairs)
 * may be NULL. h_sumQ, when non-NULL, receives sumQ of each iteration (it1 - it0 doubles).
 * Besides the library stream, the loop runs the next iteration's attraction and any spectra recomputation on
 * two streams of its own (default priority), ordered against the library stream by events; it returns once
 * all three have finished. */
int mg_tsne_optimize(int prec, const mg_tsne_params* params, const int32_t* d_indptr, const int32_t* d_indices,
                     const void* d_data, int64_t nnz, int64_t n, void* d_Y, void* d_U, void* d_G, const int32_t* d_pairs,
                     int m_pairs, int it0, int it1, double* h_sumQ, mg_tsne_callback cb, void* user);

/* A reusable optimize_tsne loop over one CSR P (device arrays that must outlive the session): the CSR is validated
 * once at creation, and the loop's device buffers, streams and kernel-spectra cache persist between runs, so a run
 * costs only its iterations. mg_tsne_session_run(s, Y, U, G, it0, it1, ...) is mg_tsne_optimize with the session's
 * P, schedule and pairs; consecutive runs of [a, b) and [b, c) give the state of one run of [a, c). No counterpart in
 * the reference (its loop is a Python function over per-call kernels); for callers that run a schedule in pieces
 * (checkpointing, timing, interleaving several layouts). A session is not re-entrant. */
typedef struct mg_tsne_session mg_tsne_session;
int mg_tsne_session_create(mg_tsne_session** out, int prec, const mg_tsne_params* params, const int32_t* d_indptr,
                           const int32_t* d_indices, const void* d_data, int64_t nnz, int64_t n, const int32_t* d_pairs,
                           int m_pairs);
int mg_tsne_session_run(mg_tsne_session* s, void* d_Y, void* d_U, void* d_G, int it0, int it1, double* h_sumQ,
                        mg_tsne_callback cb, void* user);
int mg_tsne_session_destroy(mg_tsne_session* s);
/* Per-iteration diagnostics of the session's next runs (no counterpart in the reference): for every iteration of
 * [it0, it1), h_kl receives KL(P || Q) = sum p log p + dof sum p log(1 + d^2 / dof) + (sum p) log Z on the layout
 * the iteration starts from (Z its sumQ), and h_grad_norm the L2 norm of its gradient (rho A - F / Z, plus the
 * mid-range spring when pairs are set). Both in FP64 with compensated sums over the points and edges in original
 * order, for either precision; an extra CSR sweep per iteration. Host arrays of it1 - it0 doubles; either may be
 * NULL, both NULL turns the diagnostics off. */
int mg_tsne_session_trace(mg_tsne_session* s, double* h_kl, double* h_grad_norm);

/* Batched form: `count` (<= 1024) independent embeddings (CSR P, state Y / U / G of their own sizes, device
 * arrays) on one schedule p (learning_rate 0: per instance max(200, n / early_exaggeration); no mid-range pairs;
 * max_boxes <= 1024). Every stage of an iteration is one launch for all instances of a group (one per FFT size for
 * the FFT stages). The iteration is device-driven: a kernel decides every instance's grid from the bbox of its last
 * update, and the launches depend only on a plan of each instance's window of FFT sizes (the one its grid needs and
 * the sizes either side), so the iteration is captured once in a hipGraph and replayed; the host reads the
 * instances' states every few iterations and plans again when one nears the edge of its window (an instance whose
 * grid leaves its window waits, its state unchanged, for the next plan). The instances are split into n_streams
 * groups of about equal point counts, each on its own HIP stream (and a second one for its attraction), driven by
 * n_threads host threads (1 <= n_threads <= n_streams).
 * Each instance's result is bit for bit the one mg_tsne_optimize gives it alone. h_sumQ (host, n_iter doubles)
 * may be null. The call returns when every instance is done. No counterpart in the reference (it fits one
 * embedding per call). */
typedef struct {
    const int32_t* indptr;
    const int32_t* indices;
    const void* data;
    int64_t nnz, n;
    void *Y, *U, *G;
    double* h_sumQ;
} mg_tsne_instance;
int mg_tsne_optimize_batch(int prec, const mg_tsne_params* p, int count, const mg_tsne_instance* inst, int n_streams,
                           int n_threads);
/* Counters of the last mg_tsne_optimize_batch call, summed over its groups (no counterpart in the reference): plans
 * made (a plan fixes each instance's window of FFT sizes, its buffers, the launches of an iteration and their
 * hipGraph), iterations issued (the schedule's length, plus those some instance spent waiting for a plan), graph
 * replays among them, the kernel launches they hold, reads of the instances' states, instance stalls seen at those
 * reads, and the host seconds spent planning. Any pointer may be NULL. */
int mg_tsne_batch_stats(int64_t* plans, int64_t* iterations, int64_t* replays, int64_t* launches, int64_t* checks,
                        int64_t* stalls, double* plan_seconds);

#ifdef __cplusplus
}
#endif
#endif /* MAGNET_HIP_H */
