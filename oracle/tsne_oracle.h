// CPU oracle for the full-batch t-SNE gradient pass -- TEST INFRASTRUCTURE ONLY.
//
// A restatement, in plain C++, of the reference algorithm (kerighan/lib-magnet v3.0.2):
//   grid decision        magnet/tsne.py:30-46, src/fftrep.cpp:415-434
//   kernel spectra       magnet/tsne.py:48-71
//   FFT repulsion        src/fftconv.cpp:24-230 (spread, r2c/c2c FFTs, products, inverse, gather, Z)
//   attraction           src/fftrep.cpp:129-177
//   update / step        src/fftrep.cpp:181-266
//   mid-range spring     src/fftrep.cpp:270-287
//   optimiser schedule   magnet/tsne.py:132-214
//   fallback kernels     src/fftrep.cpp:24-125, 294-411
//   sampled layout       src/layout.cpp:118-276, 343-472 (owner mode, loss 0, power kernel, SGD) (interp_scatter / interp_gather, scatter_padded_2d,
//                        spectral_products, gather_forces_2d)
// Every function exists in a double (_f64) and a float (_f32) form. The f64 form is the reference the
// FP64 GPU path is checked against; the f32 form follows the reference's float32 arithmetic order and
// is checked against vectors written by the compiled reference.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only
// as the checker. The product (lib-magnet_amd) never links, loads or calls it.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Grid decision of FFTRepulsion2D._grid on the bbox of Y (n, 2): out = {xmin, ymin, width, n_boxes}.
// The f32 form takes the bbox in float (bbox_2d) and the rest in double, like the reference.
void or_grid_f64(const double* Y, int64_t n, double box_width, int min_boxes, int max_boxes, double* out);
void or_grid_f32(const float* Y, int64_t n, double box_width, int min_boxes, int max_boxes, double* out);

// Smallest power of two >= 2N (any M >= 2N - 1 gives the same convolution).
int or_fft_size(int N);

// Real spectra (M, M/2+1) of the kernels K1 = (1 + r^2/dof)^-dof and K2 = (1 + r^2/dof)^(-dof-1)
// sampled at offsets (m or m - M) * width / n_interp on the M x M torus. round_f32 = 1 rounds the
// kernel tables to float before the transform (magnet/tsne.py:65-67).
void or_kernel_spectra(int M, double delta, double dof, int round_f32, double* K1h, double* K2h);

// Repulsion: returns sumQ = Z - n and fills F (n, 2). If K1h/K2h are null, the spectra are computed
// here (or_kernel_spectra with M). W_out (3, N, N) and phi_out (4, N, N) are optional intermediate
// grids (unscaled charges, potentials). direct = 1 replaces the FFT convolution by the O(N^4) sum.
double or_repulsion_f64(const double* Y, int64_t n, double xmin, double ymin, double width, int n_boxes,
                        int n_interp, int M, double dof, const double* K1h, const double* K2h, double* F,
                        double* W_out, double* phi_out, int direct);
double or_repulsion_f32(const float* Y, int64_t n, float xmin, float ymin, float width, int n_boxes,
                        int n_interp, int M, double dof, const double* K1h, const double* K2h, float* F,
                        float* W_out, float* phi_out, int direct);

// Exact O(n^2) repulsion: Z = sum_{i != j} K1(r_ij), F_i = sum_j K2(r_ij) (y_i - y_j).
double or_exact_repulsion(const double* Y, int64_t n, double dof, double* F);

// Attraction out_i = sum_j p_ij (1 + |y_i - y_j|^2 / dof)^-1 (y_i - y_j), any dim.
void or_attraction_f64(const int32_t* indptr, const int32_t* indices, const double* data, const double* Y,
                       int64_t n, int dim, double dof, double* out);
void or_attraction_f32(const int32_t* indptr, const int32_t* indices, const float* data, const float* Y,
                       int64_t n, int dim, float dof, float* out);

// Fused step (grad = rho*attr - F*inv_sumQ) and plain update (given grad); both recentre Y.
void or_step_f64(double* Y, double* U, double* G, const double* attr, const double* F, int64_t n, int dim,
                 double rho, double inv_sumQ, double lr, double mom);
void or_step_f32(float* Y, float* U, float* G, const float* attr, const float* F, int64_t n, int dim,
                 float rho, float inv_sumQ, float lr, float mom);
void or_update_f64(double* Y, double* U, double* G, const double* grad, int64_t n, int dim, double lr, double mom);
void or_update_f32(float* Y, float* U, float* G, const float* grad, int64_t n, int dim, float lr, float mom);
void or_mid_spring_f64(const double* Y, const int32_t* pairs, double* grad, int64_t n, int dim, int m, double w);
void or_mid_spring_f32(const float* Y, const int32_t* pairs, float* grad, int64_t n, int dim, int m, float w);

// Whole optimiser (magnet/tsne.py:132-214) for iterations [it0, it1) of an n_iter schedule, on the
// caller's state (Y, U, G). data must already be normalised to sum 1. pairs may be null.
// Per-iteration sumQ goes to sumQ_out[it - it0] when non-null.
typedef struct {
    int32_t n_iter, early_iter;
    double early_exaggeration, exaggeration, learning_rate, momentum0, momentum1;
    int32_t n_interp, min_boxes, max_boxes;
    double box_width, dof, mid_weight, mid_phase;
} or_params;
void or_optimize_f64(const or_params* p, const int32_t* indptr, const int32_t* indices, const double* data,
                     int64_t n, double* Y, double* U, double* G, const int32_t* pairs, int m_pairs,
                     int it0, int it1, double* sumQ_out);
void or_optimize_f32(const or_params* p, const int32_t* indptr, const int32_t* indices, const float* data,
                     int64_t n, float* Y, float* U, float* G, const int32_t* pairs, int m_pairs,
                     int it0, int it1, double* sumQ_out);

// Lagrange weights of n_interp nodes at (k + 1/2) / n_interp for local coordinate u (src/fftconv.cpp:24-40).
void or_lagrange_f64(double u, int n_interp, double* w);

// ---- fallback kernels (src/fftrep.cpp), in the order of the reference at n_threads = 1. Box of a point:
// u = (x - xmin) / width, clamped to [0, n_boxes - 1]; weights from the product-form Lagrange polynomials
// at the clamped local coordinate (src/fftrep.cpp:24-45).
// interp_scatter: Y (n, 2), charges (n, C) -> out (C, N, N), N = n_boxes * n_interp (zeroed here).
void or_interp_scatter_f64(const double* Y, int64_t n, const double* charges, int C, double xmin, double ymin, double width,
                           int n_boxes, int n_interp, double* out);
void or_interp_scatter_f32(const float* Y, int64_t n, const float* charges, int C, float xmin, float ymin, float width,
                           int n_boxes, int n_interp, float* out);
// interp_gather: grid (C, N, N) -> out (n, C)
void or_interp_gather_f64(const double* Y, int64_t n, const double* grid, int C, double xmin, double ymin, double width,
                          int n_boxes, int n_interp, double* out);
void or_interp_gather_f32(const float* Y, int64_t n, const float* grid, int C, float xmin, float ymin, float width,
                          int n_boxes, int n_interp, float* out);
// scatter_padded_2d: charges (1, x, y) on the zero-padded (3, M, M) grid (zeroed here)
void or_scatter_padded_f64(const double* Y, int64_t n, double xmin, double ymin, double width, int n_boxes, int n_interp,
                           int M, double* out);
void or_scatter_padded_f32(const float* Y, int64_t n, float xmin, float ymin, float width, int n_boxes, int n_interp,
                           int M, float* out);
// spectral_products: K1h, K2h (M, Mh) complex, Wh (3, M, Mh) complex -> out (4, M, Mh) complex (interleaved)
void or_spectral_products_f64(const double* K1h, const double* K2h, const double* Wh, int64_t M, int64_t Mh, double* out);
void or_spectral_products_f32(const float* K1h, const float* K2h, const float* Wh, int64_t M, int64_t Mh, float* out);
// gather_forces_2d: phi (4, M, M) -> F (n, 2); returns sumQ = sum pot0 - n (double accumulation)
double or_gather_forces_f64(const double* Y, int64_t n, const double* phi, int M, double xmin, double ymin, double width,
                            int n_boxes, int n_interp, double* F);
double or_gather_forces_f32(const float* Y, int64_t n, const float* phi, int M, float xmin, float ymin, float width,
                            int n_boxes, int n_interp, float* F);

// Sampled layout (src/layout.cpp:118-276, 343-472), owner mode, loss 0, power kernel, SGD: epochs
// [epoch_start, epoch_end) on Y (n, 2) in place, edges sorted by head; the fields as mg_layout_params.
typedef struct {
    int32_t epoch_start, epoch_end;
    float alpha0, alpha1, mid_w0, mid_w1;
    float a, b, rep_a, rep_b;
    float negative_sample_rate, exaggeration, repulsion_strength, mid_rate;
    int32_t clip_mode;
    int64_t n_fixed;
    uint64_t seed;
    float attraction_scale;
} or_layout_params;
void or_layout_epochs(const or_layout_params* p, float* Y, int64_t n, const int32_t* head, const int32_t* tail, const float* eps,
                      float* eons, float* eonns, int64_t m, const int32_t* pairs, int n_mid, const int32_t* rev_ptr,
                      const int32_t* rev_idx);

#ifdef __cplusplus
}
#endif
