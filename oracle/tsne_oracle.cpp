// CPU oracle for the full-batch t-SNE gradient pass -- TEST INFRASTRUCTURE ONLY (see tsne_oracle.h).
// Written for clarity (the FFT transforms alone split their independent rows and columns over host threads,
// which changes no result): every loop follows the reference file:line cited above it,
// in the same arithmetic order, so that the float form can be compared bit for bit with the compiled
// reference where the reference is itself deterministic (attraction, step, update, spring).
// Build: g++ -O3 -ffp-contract=off (no fused multiply-add, like the reference's x86-64 build).
#include "tsne_oracle.h"
#include <cmath>
#include <complex>
#include <vector>
#include <algorithm>
#include <cstring>
#include <thread>

namespace {

using cd = std::complex<double>;

// ---- Lagrange weights: closed form for 3 nodes (src/fftconv.cpp:24-29), generic product otherwise (:30-39)
template <class T> inline void lagrange(T u, int ni, T* w) {
    if (ni == 3) {
        const T a = u - T(1) / T(6), b = u - T(0.5), c = u - T(5) / T(6);
        w[0] = b * c * T(4.5); w[1] = -a * c * T(9); w[2] = a * b * T(4.5);
        return;
    }
    for (int k = 0; k < ni; ++k) {
        const T uk = (k + T(0.5)) / ni;
        T num = 1, den = 1;
        for (int m = 0; m < ni; ++m) {
            if (m == k) continue;
            const T um = (m + T(0.5)) / ni;
            num *= (u - um); den *= (uk - um);
        }
        w[k] = num / den;
    }
}

// ---- 1-D complex DFT, sign -1 forward, +1 inverse, unnormalised. Radix-2 when L is a power of two, with the
// twiddles exp(-2 pi i k / L) computed once per transform set (each from its own cos / sin; the inverse uses their
// conjugates); the O(L^2) definition otherwise (used only for small non-power-of-two M in padding tests).
struct Twiddles {
    int L = 0;
    std::vector<cd> w;   // exp(-2 pi i k / L), k < L / 2
    explicit Twiddles(int L_) : L(L_) {
        if ((L & (L - 1)) != 0) return;
        w.resize(L / 2);
        for (int k = 0; k < L / 2; ++k) { const double ang = -2.0 * M_PI * k / L; w[k] = cd(std::cos(ang), std::sin(ang)); }
    }
};

void dft(cd* a, const Twiddles& T, int sign) {
    const int L = T.L;
    if (L <= 1) return;
    if ((L & (L - 1)) == 0) {
        for (int i = 1, j = 0; i < L; ++i) {      // bit-reversal permutation
            int bit = L >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (int len = 2; len <= L; len <<= 1) {
            const int step = L / len;
            for (int i = 0; i < L; i += len) {
                for (int k = 0; k < len / 2; ++k) {
                    const cd t = T.w[(size_t)k * step];
                    const double tr = t.real(), ti = sign < 0 ? t.imag() : -t.imag();
                    const cd x = a[i + k + len / 2];
                    // the complex product written out: the finite-value result of operator*, without its
                    // library call for infinities
                    const cd u = a[i + k], v(x.real() * tr - x.imag() * ti, x.real() * ti + x.imag() * tr);
                    a[i + k] = u + v; a[i + k + len / 2] = u - v;
                }
            }
        }
        return;
    }
    std::vector<cd> out(L);
    for (int k = 0; k < L; ++k) {
        cd s = 0;
        for (int m = 0; m < L; ++m) {
            const double ang = sign * 2.0 * M_PI * (double)(((int64_t)k * m) % L) / L;
            s += a[m] * cd(std::cos(ang), std::sin(ang));
        }
        out[k] = s;
    }
    std::copy(out.begin(), out.end(), a);
}

// f(i) for i in [0, count), split over host threads by index (each i is independent, so the result does not
// depend on the thread count)
template <class F> void parallel_for(int count, F f) {
    const int nt = std::max(1, std::min({count, 16, (int)std::thread::hardware_concurrency()}));
    if (nt <= 1) { for (int i = 0; i < count; ++i) f(i); return; }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] { for (int i = t; i < count; i += nt) f(i); });
    for (auto& th : pool) th.join();
}

// The column transforms of an M x M row-major grid, eight adjacent columns copied out together (whole cache
// lines) and transformed one by one.
void dft_columns(cd* g, int M, int sign, const Twiddles& T) {
    constexpr int CB = 8;
    parallel_for((M + CB - 1) / CB, [&](int blk) {
        const int c0 = blk * CB, nc = std::min(CB, M - c0);
        std::vector<cd> cols((size_t)CB * M);
        for (int r = 0; r < M; ++r)
            for (int j = 0; j < nc; ++j) cols[(size_t)j * M + r] = g[(size_t)r * M + c0 + j];
        for (int j = 0; j < nc; ++j) dft(cols.data() + (size_t)j * M, T, sign);
        for (int r = 0; r < M; ++r)
            for (int j = 0; j < nc; ++j) g[(size_t)r * M + c0 + j] = cols[(size_t)j * M + r];
    });
}

// 2-D transform of an M x M row-major grid: rows (last axis) then columns; only rows [0, live_rows) of the input
// may be non-zero (the zero rows transform to zero and are skipped).
void dft2(cd* g, int M, int sign, int live_rows) {
    const Twiddles T(M);
    parallel_for(live_rows, [&](int r) { dft(g + (size_t)r * M, T, sign); });
    dft_columns(g, M, sign, T);
}

// Inverse-direction 2-D transform of which only rows [0, out_rows) are needed: columns first, then those rows.
void dft2_rows_out(cd* g, int M, int sign, int out_rows) {
    const Twiddles T(M);
    dft_columns(g, M, sign, T);
    parallel_for(out_rows, [&](int r) { dft(g + (size_t)r * M, T, sign); });
}

inline double kern1(double r2, double dof) { return std::pow(1.0 + r2 / dof, -dof); }
inline double kern2(double r2, double dof) { return std::pow(1.0 + r2 / dof, -dof - 1.0); }

// Convolution of the (3, N, N) charges with K1 (channel 0) and K2 (channels 0..2) -> (4, N, N),
// cyclic on the M x M torus with the kernel wrapped at M/2 (magnet/tsne.py:57-59; src/fftconv.cpp:124-198).
void convolve_fft(const std::vector<double>& W, int N, int M, const double* K1h, const double* K2h,
                  std::vector<double>& phi) {
    const int Mh = M / 2 + 1;
    std::vector<cd> spec[3];
    for (int c = 0; c < 3; ++c) {
        spec[c].assign((size_t)M * M, cd(0, 0));
        for (int gx = 0; gx < N; ++gx)
            for (int gy = 0; gy < N; ++gy) spec[c][(size_t)gx * M + gy] = W[((size_t)c * N + gx) * N + gy];
        dft2(spec[c].data(), M, -1, N);
    }
    // real spectra given on (M, Mh); the kernels are even in both axes, so K(kx, ky) = K(kx, M - ky)
    auto kval = [&](const double* Kh, int kx, int ky) { return Kh[(size_t)kx * Mh + (ky < Mh ? ky : M - ky)]; };
    phi.assign((size_t)4 * N * N, 0.0);
    std::vector<cd> prod((size_t)M * M);
    const double scale = 1.0 / ((double)M * (double)M);
    for (int o = 0; o < 4; ++o) {
        const int src = o == 0 ? 0 : o - 1;
        const double* Kh = o == 0 ? K1h : K2h;
        parallel_for(M, [&](int kx) {
            for (int ky = 0; ky < M; ++ky) prod[(size_t)kx * M + ky] = spec[src][(size_t)kx * M + ky] * kval(Kh, kx, ky);
        });
        dft2_rows_out(prod.data(), M, +1, N);
        for (int gx = 0; gx < N; ++gx)
            for (int gy = 0; gy < N; ++gy) phi[((size_t)o * N + gx) * N + gy] = prod[(size_t)gx * M + gy].real() * scale;
    }
}

// The same convolution evaluated as a plain sum over grid nodes (no FFT, no padding).
void convolve_direct(const std::vector<double>& W, int N, double delta, double dof, std::vector<double>& phi) {
    phi.assign((size_t)4 * N * N, 0.0);
    for (int gx = 0; gx < N; ++gx)
        for (int gy = 0; gy < N; ++gy) {
            double s[4] = {0, 0, 0, 0};
            for (int hx = 0; hx < N; ++hx)
                for (int hy = 0; hy < N; ++hy) {
                    const double dx = (gx - hx) * delta, dy = (gy - hy) * delta, r2 = dx * dx + dy * dy;
                    const double k1 = kern1(r2, dof), k2 = kern2(r2, dof);
                    const double w0 = W[((size_t)0 * N + hx) * N + hy], w1 = W[((size_t)1 * N + hx) * N + hy],
                                 w2 = W[((size_t)2 * N + hx) * N + hy];
                    s[0] += k1 * w0; s[1] += k2 * w0; s[2] += k2 * w1; s[3] += k2 * w2;
                }
            for (int o = 0; o < 4; ++o) phi[((size_t)o * N + gx) * N + gy] = s[o];
        }
}

// ---- FFTConv2D::compute (src/fftconv.cpp:58-230), one thread
template <class T>
double repulsion(const T* Y, int64_t n, T xmin, T ymin, T width, int nb, int ni, int M, double dof,
                 const double* K1h_in, const double* K2h_in, T* F, T* W_out, T* phi_out, int direct) {
    const int N = nb * ni;
    if (M <= 0) M = or_fft_size(N);
    const T inv_w = T(1) / width;
    // x-box of every point (multiplication by 1/width, :85-88) and a stable counting sort (:90-96)
    std::vector<int32_t> bxs(n);
    std::vector<int64_t> count(nb + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int bx = std::min(nb - 1, std::max(0, (int)((Y[2 * i] - xmin) * inv_w)));
        bxs[i] = bx; ++count[bx + 1];
    }
    for (int b = 0; b < nb; ++b) count[b + 1] += count[b];
    std::vector<int32_t> order(n);
    { std::vector<int64_t> off(count.begin(), count.end() - 1); for (int64_t i = 0; i < n; ++i) order[off[bxs[i]]++] = (int32_t)i; }
    // spread in x-box order (:98-122), accumulated in T like the reference
    std::vector<T> Wt((size_t)3 * N * N, T(0));
    T wx[8], wy[8];
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = order[k];
        const T x = Y[2 * i], y = Y[2 * i + 1];
        const T ux = (x - xmin) / width, uy = (y - ymin) / width;
        const int bx = bxs[i], by = std::min(nb - 1, std::max(0, (int)uy));
        lagrange<T>(std::min(T(1), std::max(T(0), ux - bx)), ni, wx);
        lagrange<T>(std::min(T(1), std::max(T(0), uy - by)), ni, wy);
        for (int a = 0; a < ni; ++a) {
            const int gx = bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const int gy = by * ni + b; const T w = wx[a] * wy[b];
                Wt[((size_t)0 * N + gx) * N + gy] += w;
                Wt[((size_t)1 * N + gx) * N + gy] += w * x;
                Wt[((size_t)2 * N + gx) * N + gy] += w * y;
            }
        }
    }
    if (W_out) std::copy(Wt.begin(), Wt.end(), W_out);
    std::vector<double> Wd(Wt.begin(), Wt.end()), phid;
    if (direct) {
        convolve_direct(Wd, N, (double)width / ni, dof, phid);
    } else {
        std::vector<double> k1, k2;
        const double *K1h = K1h_in, *K2h = K2h_in;
        if (!K1h || !K2h) {
            const int Mh = M / 2 + 1;
            k1.resize((size_t)M * Mh); k2.resize((size_t)M * Mh);
            or_kernel_spectra(M, (double)width / ni, dof, sizeof(T) == 4, k1.data(), k2.data());
            K1h = k1.data(); K2h = k2.data();
        }
        convolve_fft(Wd, N, M, K1h, K2h, phid);
    }
    std::vector<T> phi(phid.begin(), phid.end());   // potentials rounded to T (float32 grid in the reference)
    if (phi_out) std::copy(phi.begin(), phi.end(), phi_out);
    // gather in x-box order, forces and Z (:201-229); Z accumulated in double
    double tot = 0.0;
    for (int64_t kk = 0; kk < n; ++kk) {
        const int64_t i = order[kk];
        const T x = Y[2 * i], y = Y[2 * i + 1];
        const T ux = (x - xmin) / width, uy = (y - ymin) / width;
        const int bx = bxs[i], by = std::min(nb - 1, std::max(0, (int)uy));
        lagrange<T>(std::min(T(1), std::max(T(0), ux - bx)), ni, wx);
        lagrange<T>(std::min(T(1), std::max(T(0), uy - by)), ni, wy);
        T pot[4] = {0, 0, 0, 0};
        for (int a = 0; a < ni; ++a) {
            const size_t gx = (size_t)bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const size_t gy = (size_t)by * ni + b; const T w = wx[a] * wy[b];
                for (int c = 0; c < 4; ++c) pot[c] += w * phi[((size_t)c * N + gx) * N + gy];
            }
        }
        tot += pot[0];
        F[2 * i] = pot[1] * x - pot[2];
        F[2 * i + 1] = pot[1] * y - pot[3];
    }
    if (sizeof(T) == 4) return (double)(float)(tot - (double)n);   // returned as float (:229)
    return tot - (double)n;
}

// ---- tsne_attraction (src/fftrep.cpp:129-177)
template <class T>
void attraction(const int32_t* ip, const int32_t* ix, const T* P, const T* Y, int64_t n, int dim, T dof, T* O) {
    const T inv_dof = T(1) / dof;
    if (dim == 2) {
        for (int64_t i = 0; i < n; ++i) {
            const T xi = Y[2 * i], yi = Y[2 * i + 1];
            T ax = 0, ay = 0;
            for (int32_t e = ip[i]; e < ip[i + 1]; ++e) {
                const int64_t j = ix[e];
                const T dx = xi - Y[2 * j], dy = yi - Y[2 * j + 1];
                const T w = P[e] / (T(1) + (dx * dx + dy * dy) * inv_dof);
                ax += w * dx; ay += w * dy;
            }
            O[2 * i] = ax; O[2 * i + 1] = ay;
        }
        return;
    }
    std::vector<T> acc(dim), diff(dim);
    for (int64_t i = 0; i < n; ++i) {
        std::fill(acc.begin(), acc.end(), T(0));
        const T* yi = Y + i * dim;
        for (int32_t e = ip[i]; e < ip[i + 1]; ++e) {
            const T* yj = Y + (int64_t)ix[e] * dim;
            T d2 = 0;
            for (int d = 0; d < dim; ++d) { diff[d] = yi[d] - yj[d]; d2 += diff[d] * diff[d]; }
            const T w = P[e] / (T(1) + d2 * inv_dof);
            for (int d = 0; d < dim; ++d) acc[d] += w * diff[d];
        }
        for (int d = 0; d < dim; ++d) O[i * dim + d] = acc[d];
    }
}

// ---- gains / momentum / position update and recentring (src/fftrep.cpp:181-266). The mean is
// accumulated as the reference does with one thread: T partial sums flushed to double every 1024 rows.
template <class T, class GradF>
void update_core(T* Y, T* U, T* G, int64_t n, int dim, T lr, T mom, GradF grad) {
    std::vector<double> s(dim, 0.0);
    std::vector<T> sl(dim, T(0));
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < dim; ++d) {
            const int64_t k = i * dim + d;
            const T g_ = grad(k);
            const bool same = (g_ > T(0)) == (U[k] > T(0));
            T g = same ? G[k] * T(0.8) : G[k] + T(0.2);
            if (g < T(0.01)) g = T(0.01);
            G[k] = g;
            const T u = mom * U[k] - lr * g * g_;
            U[k] = u;
            const T y = Y[k] + u;
            Y[k] = y;
            sl[d] += y;
        }
        if ((i & 1023) == 1023) for (int d = 0; d < dim; ++d) { s[d] += sl[d]; sl[d] = T(0); }
    }
    for (int d = 0; d < dim; ++d) s[d] += sl[d];
    std::vector<T> mean(dim);
    for (int d = 0; d < dim; ++d) mean[d] = (T)(s[d] / (double)n);
    for (int64_t i = 0; i < n; ++i) for (int d = 0; d < dim; ++d) Y[i * dim + d] -= mean[d];
}

// ---- mid_spring (src/fftrep.cpp:270-287)
template <class T>
void mid_spring(const T* Y, const int32_t* Pp, T* G, int64_t n, int dim, int m, T w) {
    std::vector<T> acc(dim);
    for (int64_t i = 0; i < n; ++i) {
        std::fill(acc.begin(), acc.end(), T(0));
        for (int k = 0; k < m; ++k) { const T* yp = Y + (int64_t)Pp[i * m + k] * dim; for (int d = 0; d < dim; ++d) acc[d] += yp[d]; }
        for (int d = 0; d < dim; ++d) G[i * dim + d] += w * (Y[i * dim + d] - acc[d] / m);
    }
}

// ---- grid decision (magnet/tsne.py:30-46); bbox values in T (bbox_2d returns floats), the rest in double
template <class T>
void grid(const T* Y, int64_t n, double bw, int minb, int maxb, double* out) {
    T x0 = Y[0], x1 = Y[0], y0 = Y[1], y1 = Y[1];
    for (int64_t i = 0; i < n; ++i) {
        x0 = std::min(x0, Y[2 * i]); x1 = std::max(x1, Y[2 * i]);
        y0 = std::min(y0, Y[2 * i + 1]); y1 = std::max(y1, Y[2 * i + 1]);
    }
    const double lox = (double)x0 - 1e-3, loy = (double)y0 - 1e-3, hix = (double)x1 + 1e-3, hiy = (double)y1 + 1e-3;
    const double span = std::max(hix - lox, hiy - loy);
    double nbf = std::ceil(span / bw);
    nbf = std::min((double)maxb, std::max((double)minb, nbf));
    int nb = (int)nbf;
    nb = std::min(maxb, ((nb + 3) / 4) * 4);
    const double width = (nb * bw >= span) ? bw : std::max(span / nb, 1e-12);
    out[0] = lox; out[1] = loy; out[2] = width; out[3] = nb;
}

// ---- optimize_tsne (magnet/tsne.py:132-214) on iterations [it0, it1)
template <class T>
void optimize(const or_params* p, const int32_t* ip, const int32_t* ix, const T* data, int64_t n, T* Y, T* U,
              T* G, const int32_t* pairs, int m, int it0, int it1, double* sumQ_out) {
    const double lr = p->learning_rate > 0 ? p->learning_rate : std::max(200.0, n / p->early_exaggeration);
    std::vector<T> attr(2 * n), F(2 * n), grad(2 * n);
    const bool fused = !(pairs && m > 0 && p->mid_weight > 0);
    for (int it = it0; it < it1; ++it) {
        const bool early = it < p->early_iter;
        const double rho = early ? p->early_exaggeration : p->exaggeration;
        const double mom = early ? p->momentum0 : p->momentum1;
        attraction<T>(ip, ix, data, Y, n, 2, (T)p->dof, attr.data());
        double g[4];
        grid<T>(Y, n, p->box_width, p->min_boxes, p->max_boxes, g);
        const int nb = (int)g[3];
        const double sumQ = repulsion<T>(Y, n, (T)g[0], (T)g[1], (T)g[2], nb, p->n_interp, 0, p->dof, nullptr,
                                         nullptr, F.data(), nullptr, nullptr, 0);
        if (sumQ_out) sumQ_out[it - it0] = sumQ;
        const T inv = (T)(1.0 / std::max(sumQ, 1e-12));
        const T rho_t = (T)rho;
        if (fused) {
            const T* A = attr.data(); const T* Fp = F.data();
            update_core<T>(Y, U, G, n, 2, (T)lr, (T)mom, [&](int64_t k) { return rho_t * A[k] - Fp[k] * inv; });
        } else {
            for (int64_t k = 0; k < 2 * n; ++k) grad[k] = rho_t * attr[k] - F[k] * inv;
            const double w_mid = p->mid_weight * std::max(0.0, 1.0 - it / std::max(1.0, p->mid_phase * p->n_iter)) / n;
            if (w_mid > 0) mid_spring<T>(Y, pairs, grad.data(), n, 2, m, (T)w_mid);
            const T* D = grad.data();
            update_core<T>(Y, U, G, n, 2, (T)lr, (T)mom, [&](int64_t k) { return D[k]; });
        }
    }
}

}  // namespace

extern "C" {

int or_fft_size(int N) { int M = 1; while (M < 2 * N) M <<= 1; return M; }

void or_lagrange_f64(double u, int ni, double* w) { lagrange<double>(u, ni, w); }

void or_kernel_spectra(int M, double delta, double dof, int round_f32, double* K1h, double* K2h) {
    const int Mh = M / 2 + 1;
    std::vector<cd> g1((size_t)M * M), g2((size_t)M * M);
    parallel_for(M, [&](int a) {
        const double ox = (a <= M / 2 ? a : a - M) * delta;
        for (int b = 0; b < M; ++b) {
            const double oy = (b <= M / 2 ? b : b - M) * delta;
            const double r2 = ox * ox + oy * oy;
            double k1 = kern1(r2, dof), k2 = kern2(r2, dof);
            if (round_f32) { k1 = (float)k1; k2 = (float)k2; }
            g1[(size_t)a * M + b] = k1; g2[(size_t)a * M + b] = k2;
        }
    });
    dft2(g1.data(), M, -1, M); dft2(g2.data(), M, -1, M);
    for (int a = 0; a < M; ++a)
        for (int b = 0; b < Mh; ++b) { K1h[(size_t)a * Mh + b] = g1[(size_t)a * M + b].real(); K2h[(size_t)a * Mh + b] = g2[(size_t)a * M + b].real(); }
}

void or_grid_f64(const double* Y, int64_t n, double bw, int minb, int maxb, double* out) { grid<double>(Y, n, bw, minb, maxb, out); }
void or_grid_f32(const float* Y, int64_t n, double bw, int minb, int maxb, double* out) { grid<float>(Y, n, bw, minb, maxb, out); }

double or_repulsion_f64(const double* Y, int64_t n, double xmin, double ymin, double width, int nb, int ni, int M,
                        double dof, const double* K1h, const double* K2h, double* F, double* W_out, double* phi_out, int direct) {
    return repulsion<double>(Y, n, xmin, ymin, width, nb, ni, M, dof, K1h, K2h, F, W_out, phi_out, direct);
}
double or_repulsion_f32(const float* Y, int64_t n, float xmin, float ymin, float width, int nb, int ni, int M,
                        double dof, const double* K1h, const double* K2h, float* F, float* W_out, float* phi_out, int direct) {
    return repulsion<float>(Y, n, xmin, ymin, width, nb, ni, M, dof, K1h, K2h, F, W_out, phi_out, direct);
}

double or_exact_repulsion(const double* Y, int64_t n, double dof, double* F) {
    double Z = 0;
    for (int64_t i = 0; i < n; ++i) {
        double fx = 0, fy = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = Y[2 * i] - Y[2 * j], dy = Y[2 * i + 1] - Y[2 * j + 1], r2 = dx * dx + dy * dy;
            Z += kern1(r2, dof);
            const double k2 = kern2(r2, dof);
            fx += k2 * dx; fy += k2 * dy;
        }
        F[2 * i] = fx; F[2 * i + 1] = fy;
    }
    return Z;
}

void or_attraction_f64(const int32_t* ip, const int32_t* ix, const double* P, const double* Y, int64_t n, int dim, double dof, double* O) { attraction<double>(ip, ix, P, Y, n, dim, dof, O); }
void or_attraction_f32(const int32_t* ip, const int32_t* ix, const float* P, const float* Y, int64_t n, int dim, float dof, float* O) { attraction<float>(ip, ix, P, Y, n, dim, dof, O); }

void or_step_f64(double* Y, double* U, double* G, const double* A, const double* F, int64_t n, int dim, double rho, double inv, double lr, double mom) {
    update_core<double>(Y, U, G, n, dim, lr, mom, [&](int64_t k) { return rho * A[k] - F[k] * inv; });
}
void or_step_f32(float* Y, float* U, float* G, const float* A, const float* F, int64_t n, int dim, float rho, float inv, float lr, float mom) {
    update_core<float>(Y, U, G, n, dim, lr, mom, [&](int64_t k) { return rho * A[k] - F[k] * inv; });
}
void or_update_f64(double* Y, double* U, double* G, const double* D, int64_t n, int dim, double lr, double mom) {
    update_core<double>(Y, U, G, n, dim, lr, mom, [&](int64_t k) { return D[k]; });
}
void or_update_f32(float* Y, float* U, float* G, const float* D, int64_t n, int dim, float lr, float mom) {
    update_core<float>(Y, U, G, n, dim, lr, mom, [&](int64_t k) { return D[k]; });
}
void or_mid_spring_f64(const double* Y, const int32_t* pairs, double* grad, int64_t n, int dim, int m, double w) { mid_spring<double>(Y, pairs, grad, n, dim, m, w); }
void or_mid_spring_f32(const float* Y, const int32_t* pairs, float* grad, int64_t n, int dim, int m, float w) { mid_spring<float>(Y, pairs, grad, n, dim, m, w); }

void or_optimize_f64(const or_params* p, const int32_t* ip, const int32_t* ix, const double* data, int64_t n, double* Y,
                     double* U, double* G, const int32_t* pairs, int m, int it0, int it1, double* sumQ_out) {
    optimize<double>(p, ip, ix, data, n, Y, U, G, pairs, m, it0, it1, sumQ_out);
}
void or_optimize_f32(const or_params* p, const int32_t* ip, const int32_t* ix, const float* data, int64_t n, float* Y,
                     float* U, float* G, const int32_t* pairs, int m, int it0, int it1, double* sumQ_out) {
    optimize<float>(p, ip, ix, data, n, Y, U, G, pairs, m, it0, it1, sumQ_out);
}

}  // extern "C"

// ============================================================== fallback kernels (src/fftrep.cpp)
namespace {

// product-form Lagrange weights of n_interp nodes at (k + 1/2) / n_interp (src/fftrep.cpp:24-35)
template <class T> void lagrange_prod(T u, int ni, T* w) {
    for (int k = 0; k < ni; ++k) {
        const T uk = (k + T(0.5)) / ni;
        T num = 1, den = 1;
        for (int m = 0; m < ni; ++m) {
            if (m == k) continue;
            const T um = (m + T(0.5)) / ni;
            num *= (u - um); den *= (uk - um);
        }
        w[k] = num / den;
    }
}

// Box::locate (src/fftrep.cpp:38-45)
template <class T> void locate(T x, T y, T xmin, T ymin, T width, int nb, int ni, int& bx, int& by, T* wx, T* wy) {
    const T ux = (x - xmin) / width, uy = (y - ymin) / width;
    bx = std::min(nb - 1, std::max(0, (int)ux));
    by = std::min(nb - 1, std::max(0, (int)uy));
    lagrange_prod<T>(std::min(T(1), std::max(T(0), ux - bx)), ni, wx);
    lagrange_prod<T>(std::min(T(1), std::max(T(0), uy - by)), ni, wy);
}

// interp_scatter (src/fftrep.cpp:51-92), one thread: points in index order
template <class T> void interp_scatter(const T* Y, int64_t n, const T* Cq, int C, T xmin, T ymin, T width, int nb, int ni, T* O) {
    const int N = nb * ni;
    std::fill(O, O + (size_t)C * N * N, T(0));
    T wx[8], wy[8]; int bx, by;
    for (int64_t i = 0; i < n; ++i) {
        locate<T>(Y[2 * i], Y[2 * i + 1], xmin, ymin, width, nb, ni, bx, by, wx, wy);
        for (int a = 0; a < ni; ++a) {
            const int gx = bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const int gy = by * ni + b;
                const T w = wx[a] * wy[b];
                for (int c = 0; c < C; ++c) O[((size_t)c * N + gx) * N + gy] += w * Cq[i * C + c];
            }
        }
    }
}

// interp_gather (src/fftrep.cpp:94-125)
template <class T> void interp_gather(const T* Y, int64_t n, const T* G, int C, T xmin, T ymin, T width, int nb, int ni, T* O) {
    const int N = nb * ni;
    T wx[8], wy[8]; int bx, by;
    for (int64_t i = 0; i < n; ++i) {
        locate<T>(Y[2 * i], Y[2 * i + 1], xmin, ymin, width, nb, ni, bx, by, wx, wy);
        for (int c = 0; c < C; ++c) O[i * C + c] = T(0);
        for (int a = 0; a < ni; ++a) {
            const int gx = bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const int gy = by * ni + b;
                const T w = wx[a] * wy[b];
                for (int c = 0; c < C; ++c) O[i * C + c] += w * G[((size_t)c * N + gx) * N + gy];
            }
        }
    }
}

// scatter_padded_2d (src/fftrep.cpp:294-343), one thread: points in x-box counting-sort order
template <class T> void scatter_padded(const T* Y, int64_t n, T xmin, T ymin, T width, int nb, int ni, int M, T* O) {
    std::fill(O, O + (size_t)3 * M * M, T(0));
    std::vector<int> bxs(n);
    std::vector<int64_t> cnt(nb + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        const T ux = (Y[2 * i] - xmin) / width;
        bxs[i] = std::min(nb - 1, std::max(0, (int)ux));
        ++cnt[bxs[i] + 1];
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    std::vector<int64_t> order(n), fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i) order[fill[bxs[i]]++] = i;
    T wx[8], wy[8]; int bx, by;
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = order[k];
        locate<T>(Y[2 * i], Y[2 * i + 1], xmin, ymin, width, nb, ni, bx, by, wx, wy);
        const T ch[3] = {T(1), Y[2 * i], Y[2 * i + 1]};
        for (int a = 0; a < ni; ++a) {
            const int gx = bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const int gy = by * ni + b;
                const T w = wx[a] * wy[b];
                for (int c = 0; c < 3; ++c) O[((size_t)c * M + gx) * M + gy] += w * ch[c];
            }
        }
    }
}

// spectral_products (src/fftrep.cpp:346-371)
template <class T> void spectral_products(const T* K1, const T* K2, const T* W, int64_t M, int64_t Mh, T* O) {
    const int64_t L = M * Mh;
    for (int64_t k = 0; k < L; ++k) {
        const T ar = K1[2 * k], ai = K1[2 * k + 1], br = W[2 * k], bi = W[2 * k + 1];
        O[2 * k] = ar * br - ai * bi; O[2 * k + 1] = ar * bi + ai * br;
        const T cr = K2[2 * k], ci = K2[2 * k + 1];
        for (int c = 0; c < 3; ++c) {
            const T wr = W[2 * (c * L + k)], wi = W[2 * (c * L + k) + 1];
            O[2 * ((c + 1) * L + k)] = cr * wr - ci * wi; O[2 * ((c + 1) * L + k) + 1] = cr * wi + ci * wr;
        }
    }
}

// gather_forces_2d (src/fftrep.cpp:373-411)
template <class T> double gather_forces(const T* Y, int64_t n, const T* phi, int M, T xmin, T ymin, T width, int nb, int ni, T* F) {
    T wx[8], wy[8]; int bx, by;
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        locate<T>(Y[2 * i], Y[2 * i + 1], xmin, ymin, width, nb, ni, bx, by, wx, wy);
        T pot[4] = {0, 0, 0, 0};
        for (int a = 0; a < ni; ++a) {
            const int64_t gx = bx * ni + a;
            for (int b = 0; b < ni; ++b) {
                const int64_t gy = by * ni + b;
                const T w = wx[a] * wy[b];
                for (int c = 0; c < 4; ++c) pot[c] += w * phi[((size_t)c * M + gx) * M + gy];
            }
        }
        s += pot[0];
        F[2 * i] = pot[1] * Y[2 * i] - pot[2];
        F[2 * i + 1] = pot[1] * Y[2 * i + 1] - pot[3];
    }
    return s - (double)n;
}

}  // namespace

extern "C" {
#define OR_SHIMS(T, suf) \
    void or_interp_scatter_##suf(const T* Y, int64_t n, const T* q, int C, T x0, T y0, T w, int nb, int ni, T* O) { interp_scatter<T>(Y, n, q, C, x0, y0, w, nb, ni, O); } \
    void or_interp_gather_##suf(const T* Y, int64_t n, const T* G, int C, T x0, T y0, T w, int nb, int ni, T* O) { interp_gather<T>(Y, n, G, C, x0, y0, w, nb, ni, O); } \
    void or_scatter_padded_##suf(const T* Y, int64_t n, T x0, T y0, T w, int nb, int ni, int M, T* O) { scatter_padded<T>(Y, n, x0, y0, w, nb, ni, M, O); } \
    void or_spectral_products_##suf(const T* K1, const T* K2, const T* W, int64_t M, int64_t Mh, T* O) { spectral_products<T>(K1, K2, W, M, Mh, O); } \
    double or_gather_forces_##suf(const T* Y, int64_t n, const T* phi, int M, T x0, T y0, T w, int nb, int ni, T* F) { return gather_forces<T>(Y, n, phi, M, x0, y0, w, nb, ni, F); }
OR_SHIMS(double, f64)
OR_SHIMS(float, f32)
#undef OR_SHIMS
}  // extern "C"

// ============================================================== sampled layout (src/layout.cpp:118-276, owner mode)
// The owner-mode epoch of the reference's optimize_layout (src/layout.cpp:129-260 with OWNER = true, loss 0, power
// kernel, SGD): node by node in index order, each node's due edges in order, every other node read from the layout
// at the start of the epoch; the random numbers from the generator of src/rng.hpp, one stream per node and epoch
// (the reference keeps one per chunk of nodes). Float arithmetic without fused multiply-adds, as the reference.
namespace {
struct LRng {
    uint64_t s;
    explicit LRng(uint64_t seed) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s = (z ^ (z >> 31)) | 1ULL;
    }
    uint64_t next() { s ^= s >> 12; s ^= s << 25; s ^= s >> 27; return s * 0x2545F4914F6CDD1DULL; }
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};
void lstep(float& x, float& y, float gx, float gy, float alpha, int clip_mode) {   // src/layout.cpp:80-116
    float sx = alpha * gx, sy = alpha * gy;
    const float cap = 4.0f * alpha;
    if (clip_mode == 1) {
        const float n2 = sx * sx + sy * sy;
        if (n2 > cap * cap) { const float f = cap / std::sqrt(n2); sx *= f; sy *= f; }
    } else {
        sx = sx > cap ? cap : (sx < -cap ? -cap : sx);
        sy = sy > cap ? cap : (sy < -cap ? -cap : sy);
    }
    x += sx; y += sy;
}
}  // namespace

extern "C" void or_layout_epochs(const or_layout_params* p, float* Y, int64_t n, const int32_t* head, const int32_t* tail,
                                 const float* eps, float* eons, float* eonns, int64_t m, const int32_t* pairs, int n_mid,
                                 const int32_t* rev_ptr, const int32_t* rev_idx) {
    const float OWNER_ATT = p->attraction_scale > 0.f ? p->attraction_scale : 1.5f;   // src/layout.cpp:31-38
    const float rep_a = p->rep_b > 0 ? p->rep_a : p->a, rep_b = p->rep_b > 0 ? p->rep_b : p->b;
    const float two_ab = 2.f * p->a * p->b, two_b = 2.f * rep_b * p->repulsion_strength;
    const bool b_one = std::fabs(p->b - 1.f) < 1e-6f, rb_one = std::fabs(rep_b - 1.f) < 1e-6f;
    std::vector<int64_t> eptr(n + 1, 0);
    for (int64_t e = 0; e < m; ++e) ++eptr[head[e] + 1];
    for (int64_t u = 0; u < n; ++u) eptr[u + 1] += eptr[u];
    std::vector<float> snap(Y, Y + 2 * n);
    const int span = std::max(1, p->epoch_end - p->epoch_start);
    for (int epoch = p->epoch_start; epoch < p->epoch_end; ++epoch) {
        const float frac = (float)(epoch - p->epoch_start) / (float)span;
        const float alpha = p->alpha0 + (p->alpha1 - p->alpha0) * frac;
        const float mid_w = p->mid_w0 + (p->mid_w1 - p->mid_w0) * frac;
        const float fepoch = (float)epoch;
        for (int64_t u = p->n_fixed; u < n; ++u) {
            float x = snap[2 * u], y = snap[2 * u + 1];
            LRng rng(p->seed * 0x9E3779B1ULL + (uint64_t)epoch * 0x85EBCA6BULL + (uint64_t)u * 0xC2B2AE35ULL);
            for (int64_t e = eptr[u]; e < eptr[u + 1]; ++e) {
                if (eons[e] > fepoch) continue;
                const int64_t v = tail[e];
                const float dx = x - snap[2 * v], dy = y - snap[2 * v + 1];
                const float d2 = dx * dx + dy * dy;
                float gx = 0.f, gy = 0.f;
                if (d2 > 0.f) {
                    const float pw = b_one ? d2 : std::pow(d2, p->b);
                    const float c = -OWNER_ATT * p->exaggeration * two_ab * pw / (d2 * (p->a * pw + 1.f));
                    gx += c * dx; gy += c * dy;
                }
                eons[e] += eps[e];
                const float epn = eps[e] / p->negative_sample_rate;
                const int n_neg = (int)((fepoch - eonns[e]) / epn);
                for (int q = 0; q < n_neg; ++q) {
                    const int64_t k = rng.below((uint32_t)n);
                    if (k == u) continue;
                    const float kx = x - snap[2 * k], ky = y - snap[2 * k + 1];
                    const float dd = kx * kx + ky * ky;
                    if (dd > 0.f) {
                        const float pw = rb_one ? dd : std::pow(dd, rep_b);
                        const float c = two_b / ((0.001f + dd) * (rep_a * pw + 1.f));
                        gx += c * kx; gy += c * ky;
                    } else {
                        const float sc = 1e-3f / std::max(alpha, 1e-6f);
                        gx += (rng.uniform() - 0.5f) * sc;
                        gy += (rng.uniform() - 0.5f) * sc;
                    }
                }
                eonns[e] += n_neg * epn;
                if (mid_w > 0.f && pairs && n_mid > 0 && rng.uniform() < p->mid_rate) {
                    const int64_t w = pairs[u * n_mid + rng.below((uint32_t)n_mid)];
                    if (w != u) { gx += -mid_w * (x - snap[2 * w]); gy += -mid_w * (y - snap[2 * w + 1]); }
                }
                lstep(x, y, gx, gy, alpha, p->clip_mode);
            }
            if (mid_w > 0.f && rev_ptr && n_mid > 0) {
                const int rs = rev_ptr[u], re = rev_ptr[u + 1];
                float pr = p->mid_rate * (float)(re - rs) / (float)n_mid;
                while (pr > 0.f && re > rs) {
                    if (rng.uniform() < pr) {
                        const int64_t c = rev_idx[rs + (int)rng.below((uint32_t)(re - rs))];
                        if (c != u) lstep(x, y, -mid_w * (x - snap[2 * c]), -mid_w * (y - snap[2 * c + 1]), alpha, p->clip_mode);
                    }
                    pr -= 1.f;
                }
            }
            Y[2 * u] = x; Y[2 * u + 1] = y;
        }
        std::copy(Y, Y + 2 * n, snap.begin());
    }
}
