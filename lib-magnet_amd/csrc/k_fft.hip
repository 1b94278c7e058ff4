// FFT convolution of the charge grids with the Cauchy kernels (src/fftconv.cpp:124-198), LDS resident
// on gfx950 and written for this path (no rocFFT). Transforms are the Stockham radix-16/8/4/2 passes of
// fft_lds.hpp.
//   rows forward : real rows of length M (N live values, zero padding) -> M/2 + 1 complex, through one
//                  complex transform of length M/2 (even / odd samples as real / imaginary parts); the
//                  output is stored transposed, (channel, ky, gx), so the column pass reads contiguously;
//   columns      : per ky, the three columns are transformed (rows >= N are zero), multiplied by the
//                  kernel spectra and transformed back; only the N live rows are stored. Forward,
//                  products and inverse stay in one workgroup;
//   rows inverse : Hermitian half rows -> real rows of length M through one complex transform of length
//                  M/2, scaled by 1/M^2, stored interleaved (gx, gy, 4) for the gather.
// Two product sets. The reference's, used by the FP64 path: phi0 = K1*W0, phi1..3 = K2*W0..2 with the
// charges (1, x, y). The centred one, used by the FP32 path, where the charges are (1, x - x_node,
// y - y_node): phi0 = K1*W0, phi1 = K2*W0, phi2 = (x K2)*W0 - K2*W1', phi3 = (y K2)*W0 - K2*W2'. The
// odd kernels x K2, y K2 have imaginary spectra (i Tx, i Ty). Both sets give the same forces in exact
// arithmetic; the centred one avoids the cancellation of pot1 * x - pot2 at large |x|.
// Kernel spectra are stored for kx <= M/2 only (they are even, or odd for Tx, in kx).
#include "common.hpp"
#include "fft_lds.hpp"
#include <cmath>
#include <mutex>
#include <vector>

namespace mg {

namespace {

constexpr int TPB = 256;

struct Twiddles { cplx<float>* f32 = nullptr; cplx<double>* f64 = nullptr; int lds_max = 65536; };
Twiddles g_tw[64];
std::mutex g_tw_mu;

template <class T> const cplx<T>* twiddles(int dev);
template <> const cplx<float>* twiddles<float>(int dev) { return g_tw[dev].f32; }
template <> const cplx<double>* twiddles<double>(int dev) { return g_tw[dev].f64; }

int current_device() { int d = 0; MG_HIP(hipGetDevice(&d)); return d; }

// X[k] (k = 0..L) of a real sequence from Z = DFT_L(even + i odd): E = (Z_k + conj Z_{L-k}) / 2,
// O = -i (Z_k - conj Z_{L-k}) / 2, X_k = E + exp(-2 pi i k / M) O
template <class T> __device__ inline cplx<T> r2c_post(const cplx<T>* z, int k, int L, int logM, const cplx<T>* tw) {
    const cplx<T> zk = z[k & (L - 1)], zc = cconj(z[(L - k) & (L - 1)]);
    const cplx<T> e = {(zk.x + zc.x) * T(0.5), (zk.y + zc.y) * T(0.5)};
    const cplx<T> d = csub(zk, zc);
    const cplx<T> o = {d.y * T(0.5), -d.x * T(0.5)};
    return cadd(e, cmul(tw[(k << (LOG_MAX_M - logM)) & (MAX_M - 1)], o));
}

// ---- rows forward: blockIdx.y = channel, R rows per workgroup
template <class T>
__global__ __launch_bounds__(TPB) void k_rows_fwd(const T* __restrict__ W, int N, int logM, int R,
                                                  cplx<T>* __restrict__ Wh, const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
    const int L = 1 << (logM - 1), logL = logM - 1, Mh = L + 1;
    const int c = blockIdx.y, gx0 = blockIdx.x * R;
    const int rows = min(R, N - gx0);
    for (int e = threadIdx.x; e < R * L; e += TPB) {
        const int r = e >> logL, j = e & (L - 1);
        cplx<T> z = {T(0), T(0)};
        if (r < rows) {
            const T* src = W + ((size_t)c * N + gx0 + r) * N;
            if (2 * j < N) z.x = src[2 * j];
            if (2 * j + 1 < N) z.y = src[2 * j + 1];
        }
        buf[e] = z;
    }
    __syncthreads();
    lds_fft_batch<T, -1>(buf, R, logL, tw);
    for (int e = threadIdx.x; e < R * Mh; e += TPB) {
        const int k = e / R, r = e - k * R;
        if (r < rows) Wh[((size_t)c * Mh + k) * N + gx0 + r] = r2c_post(buf + r * L, k, L, logM, tw);
    }
}

// kernel table value at column frequency kx (stored for kx <= M/2; `odd` flips the sign above M/2)
template <class T> __device__ inline T ktab(const T* row, int kx, int M, bool odd) {
    const int Mh = M / 2 + 1;
    if (kx < Mh) return row[kx];
    const T v = row[M - kx];
    return odd ? -v : v;
}

template <class T> __device__ inline void load_column(cplx<T>* A, const cplx<T>* __restrict__ src, int N, int M) {
    for (int i = threadIdx.x; i < M; i += TPB) A[i] = i < N ? src[i] : cplx<T>{T(0), T(0)};
}
template <class T> __device__ inline void store_column(cplx<T>* __restrict__ dst, const cplx<T>* A, int N) {
    for (int g = threadIdx.x; g < N; g += TPB) dst[g] = A[g];
}

// ---- columns, reference product set. NBUF 2: both channel-0 products held in LDS; 1: spectrum of
// channel 0 parked in `stash` (global) between its two products.
template <class T, int NBUF>
__global__ __launch_bounds__(TPB) void k_columns(const cplx<T>* __restrict__ Wh, int N, int logM,
                                                 const T* __restrict__ K1, const T* __restrict__ K2,
                                                 cplx<T>* __restrict__ Ph, cplx<T>* __restrict__ stash,
                                                 const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = 1 << logM, Mh = M / 2 + 1, ky = blockIdx.x;
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* B = A + M;
    const T* k1 = K1 + (size_t)ky * Mh;
    const T* k2 = K2 + (size_t)ky * Mh;
    auto col = [&](int ch) { return Wh + ((size_t)ch * Mh + ky) * N; };
    auto out = [&](int o) { return Ph + ((size_t)o * Mh + ky) * N; };
    // channel 0: phi0 = K1 W0, phi1 = K2 W0
    load_column(A, col(0), N, M);
    __syncthreads();
    lds_fft<T, -1>(A, 1, logM, tw);
    for (int k = threadIdx.x; k < M; k += TPB) {
        const cplx<T> a = A[k];
        const T c1 = ktab(k1, k, M, false), c2 = ktab(k2, k, M, false);
        if (NBUF == 2) B[k] = {a.x * c2, a.y * c2};
        else stash[(size_t)ky * M + k] = {a.x * c2, a.y * c2};
        A[k] = {a.x * c1, a.y * c1};
    }
    __syncthreads();
    if (NBUF == 2) {
        lds_fft_batch<T, +1>(A, 2, logM, tw);
        store_column(out(0), A, N);
        store_column(out(1), B, N);
        __syncthreads();
        // channels 1 and 2 together: phi2 = K2 W1, phi3 = K2 W2
        load_column(A, col(1), N, M);
        load_column(B, col(2), N, M);
        __syncthreads();
        lds_fft_batch<T, -1>(A, 2, logM, tw);
        for (int k = threadIdx.x; k < M; k += TPB) {
            const T c2 = ktab(k2, k, M, false);
            const cplx<T> a = A[k], b = B[k];
            A[k] = {a.x * c2, a.y * c2};
            B[k] = {b.x * c2, b.y * c2};
        }
        __syncthreads();
        lds_fft_batch<T, +1>(A, 2, logM, tw);
        store_column(out(2), A, N);
        store_column(out(3), B, N);
    } else {
        lds_fft<T, +1>(A, 1, logM, tw);
        store_column(out(0), A, N);
        __syncthreads();
        for (int k = threadIdx.x; k < M; k += TPB) A[k] = stash[(size_t)ky * M + k];
        __syncthreads();
        lds_fft<T, +1>(A, 1, logM, tw);
        store_column(out(1), A, N);
        for (int ch = 1; ch < 3; ++ch) {
            __syncthreads();
            load_column(A, col(ch), N, M);
            __syncthreads();
            lds_fft<T, -1>(A, 1, logM, tw);
            for (int k = threadIdx.x; k < M; k += TPB) {
                const T c2 = ktab(k2, k, M, false);
                const cplx<T> a = A[k];
                A[k] = {a.x * c2, a.y * c2};
            }
            __syncthreads();
            lds_fft<T, +1>(A, 1, logM, tw);
            store_column(out(ch + 1), A, N);
        }
    }
}

// ---- columns, centred product set. NBUF 3: the three spectra in LDS; 2: the channel-0 spectrum parked
// in `stash` while channels 1 and 2 are transformed one at a time.
template <class T, int NBUF>
__global__ __launch_bounds__(TPB) void k_columns_centred(const cplx<T>* __restrict__ Wh, int N, int logM,
                                                         const T* __restrict__ K1, const T* __restrict__ K2,
                                                         const T* __restrict__ TX, const T* __restrict__ TY,
                                                         cplx<T>* __restrict__ Ph, cplx<T>* __restrict__ stash,
                                                         const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int M = 1 << logM, Mh = M / 2 + 1, ky = blockIdx.x;
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* B = A + M;
    cplx<T>* C = B + M;
    const T* k1 = K1 + (size_t)ky * Mh;
    const T* k2 = K2 + (size_t)ky * Mh;
    const T* tx = TX + (size_t)ky * Mh;
    const T* ty = TY + (size_t)ky * Mh;
    auto col = [&](int ch) { return Wh + ((size_t)ch * Mh + ky) * N; };
    auto out = [&](int o) { return Ph + ((size_t)o * Mh + ky) * N; };
    // odd-kernel product i t s0 - k2 s1
    auto oddprod = [](cplx<T> s0, cplx<T> s1, T t, T c2) { return cplx<T>{-t * s0.y - c2 * s1.x, t * s0.x - c2 * s1.y}; };
    if (NBUF == 3) {
        load_column(A, col(0), N, M);
        load_column(B, col(1), N, M);
        load_column(C, col(2), N, M);
        __syncthreads();
        lds_fft_batch<T, -1>(A, 3, logM, tw);
        for (int k = threadIdx.x; k < M; k += TPB) {
            const T c2 = ktab(k2, k, M, false), tX = ktab(tx, k, M, true), tY = ktab(ty, k, M, false);
            const cplx<T> s0 = A[k];
            B[k] = oddprod(s0, B[k], tX, c2);
            C[k] = oddprod(s0, C[k], tY, c2);
        }
        __syncthreads();
        lds_fft_batch<T, +1>(B, 2, logM, tw);
        store_column(out(2), B, N);
        store_column(out(3), C, N);
        __syncthreads();
        for (int k = threadIdx.x; k < M; k += TPB) {
            const T c1 = ktab(k1, k, M, false), c2 = ktab(k2, k, M, false);
            const cplx<T> s0 = A[k];
            B[k] = {s0.x * c2, s0.y * c2};
            A[k] = {s0.x * c1, s0.y * c1};
        }
        __syncthreads();
        lds_fft_batch<T, +1>(A, 2, logM, tw);
        store_column(out(0), A, N);
        store_column(out(1), B, N);
    } else {
        cplx<T>* S0 = stash + (size_t)ky * M;
        load_column(A, col(0), N, M);
        __syncthreads();
        lds_fft<T, -1>(A, 1, logM, tw);
        for (int k = threadIdx.x; k < M; k += TPB) {
            const T c1 = ktab(k1, k, M, false), c2 = ktab(k2, k, M, false);
            const cplx<T> s0 = A[k];
            S0[k] = s0;
            B[k] = {s0.x * c2, s0.y * c2};
            A[k] = {s0.x * c1, s0.y * c1};
        }
        __syncthreads();
        lds_fft_batch<T, +1>(A, 2, logM, tw);
        store_column(out(0), A, N);
        store_column(out(1), B, N);
        for (int ch = 1; ch < 3; ++ch) {
            __syncthreads();
            load_column(B, col(ch), N, M);
            __syncthreads();
            lds_fft<T, -1>(B, 1, logM, tw);
            const T* tt = ch == 1 ? tx : ty;
            for (int k = threadIdx.x; k < M; k += TPB) {
                const T c2 = ktab(k2, k, M, false), t = ktab(tt, k, M, ch == 1);
                B[k] = oddprod(S0[k], B[k], t, c2);
            }
            __syncthreads();
            lds_fft<T, +1>(B, 1, logM, tw);
            store_column(out(ch + 1), B, N);
        }
    }
}

// ---- rows inverse: CH channels per workgroup (4: all, interleaved store; 1: blockIdx.y = channel)
template <class T> struct vec4;
template <> struct vec4<float> { typedef float4 type; };
template <> struct vec4<double> { typedef double4 type; };

template <class T, int CH>
__global__ __launch_bounds__(TPB) void k_rows_inv(const cplx<T>* __restrict__ Ph, int N, int logM, int R,
                                                  T* __restrict__ phi, T scale, const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
    const int L = 1 << (logM - 1), logL = logM - 1, Mh = L + 1;
    const int gx0 = blockIdx.x * R, o0 = CH == 4 ? 0 : blockIdx.y;
    const int rows = min(R, N - gx0);
    // Z_k = E + i O with E = X_k + conj X_{L-k}, O = (X_k - conj X_{L-k}) exp(+2 pi i k / M)
    for (int e = threadIdx.x; e < CH * R * L; e += TPB) {
        const int oc = e / (R * L), rem = e - oc * (R * L);
        const int k = rem / R, r = rem - k * R;
        cplx<T> z = {T(0), T(0)};
        if (r < rows) {
            const size_t base = (size_t)(o0 + oc) * Mh * N + gx0 + r;
            const cplx<T> xk = Ph[base + (size_t)k * N], xc = cconj(Ph[base + (size_t)(L - k) * N]);
            const cplx<T> ev = cadd(xk, xc);
            const cplx<T> od = cmul(csub(xk, xc), cconj(tw[(k << (LOG_MAX_M - logM)) & (MAX_M - 1)]));
            z = {ev.x - od.y, ev.y + od.x};
        }
        buf[(oc * R + r) * L + k] = z;
    }
    __syncthreads();
    lds_fft_batch<T, +1>(buf, CH * R, logL, tw);
    for (int e = threadIdx.x; e < R * N; e += TPB) {
        const int r = e / N, gy = e - r * N;
        if (r >= rows) continue;
        const int j = gy >> 1;
        const bool odd = gy & 1;
        if (CH == 4) {
            typename vec4<T>::type v;
            const cplx<T> a = buf[(0 * R + r) * L + j], b = buf[(1 * R + r) * L + j];
            const cplx<T> c = buf[(2 * R + r) * L + j], d = buf[(3 * R + r) * L + j];
            v.x = (odd ? a.y : a.x) * scale; v.y = (odd ? b.y : b.x) * scale;
            v.z = (odd ? c.y : c.x) * scale; v.w = (odd ? d.y : d.x) * scale;
            reinterpret_cast<typename vec4<T>::type*>(phi)[(size_t)(gx0 + r) * N + gy] = v;
        } else {
            const cplx<T> a = buf[r * L + j];
            phi[((size_t)(gx0 + r) * N + gy) * 4 + o0] = (odd ? a.y : a.x) * scale;
        }
    }
}

// ---- kernel spectra (double). Tables: 0 K1, 1 K2, 2 x K2, 3 y K2 (offsets in grid units * delta).
// Pass 1: per row a, the real transform along b (keep Re, or Im for the table odd in b: y K2).
// Pass 2: per column kb, the real transform along a (keep Re, or Im for the table odd in a: x K2);
// output (Mh, Mh) per table: [ky][kx], kx <= M/2.
__device__ inline double kernel_value(int a, int b, int M, double delta, double dof, const double* table, int which,
                                      bool round_kernel) {
    const double ox = (a <= M / 2 ? a : a - M) * delta, oy = (b <= M / 2 ? b : b - M) * delta;
    double k2;
    if (table) {
        if (which == 0) return table[(size_t)a * M + b];
        k2 = table[(size_t)M * M + (size_t)a * M + b];
    } else {
        const double base = 1.0 + (ox * ox + oy * oy) / dof;
        if (which == 0) { double v = pow(base, -dof); return round_kernel ? (double)(float)v : v; }
        k2 = pow(base, -dof - 1.0);
        if (round_kernel) k2 = (double)(float)k2;
    }
    // the odd tables drop the Nyquist offset M/2 (no negative twin on the torus), which keeps them exactly
    // odd -- so their spectra are imaginary -- and is never used: the live offsets stay below N <= M/2
    if (which == 2) return a == M / 2 ? 0.0 : ox * k2;
    if (which == 3) return b == M / 2 ? 0.0 : oy * k2;
    return k2;
}

__global__ __launch_bounds__(TPB) void k_spec_rows(int logM, int ntab, double delta, double dof,
                                                   const double* __restrict__ table, int round_kernel,
                                                   double* __restrict__ tmp, const cplx<double>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, a = blockIdx.x;
    for (int which = 0; which < ntab; ++which) {
        for (int j = threadIdx.x; j < L; j += TPB)
            buf[j] = {kernel_value(a, 2 * j, M, delta, dof, table, which, round_kernel),
                      kernel_value(a, 2 * j + 1, M, delta, dof, table, which, round_kernel)};
        __syncthreads();
        lds_fft<double, -1>(buf, 1, logL, tw);
        for (int k = threadIdx.x; k < Mh; k += TPB) {
            const cplx<double> X = r2c_post(buf, k, L, logM, tw);
            tmp[((size_t)which * M + a) * Mh + k] = which == 3 ? X.y : X.x;
        }
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(TPB) void k_spec_cols(int logM, int ntab, const double* __restrict__ tmp, T* __restrict__ K1,
                                                   T* __restrict__ K2, T* __restrict__ TX, T* __restrict__ TY,
                                                   const cplx<double>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, kb = blockIdx.x;
    for (int which = 0; which < ntab; ++which) {
        const double* src = tmp + (size_t)which * M * Mh + kb;
        for (int j = threadIdx.x; j < L; j += TPB) buf[j] = {src[(size_t)(2 * j) * Mh], src[(size_t)(2 * j + 1) * Mh]};
        __syncthreads();
        lds_fft<double, -1>(buf, 1, logL, tw);
        T* out = (which == 0 ? K1 : which == 1 ? K2 : which == 2 ? TX : TY) + (size_t)kb * Mh;
        for (int k = threadIdx.x; k < Mh; k += TPB) {
            const cplx<double> X = r2c_post(buf, k, L, logM, tw);
            out[k] = (T)(which == 2 ? X.y : X.x);
        }
        __syncthreads();
    }
}

template <class F> void set_lds(F* f, int bytes) {
    if (bytes > 65536) MG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
}

int log2i(int M) { int l = 0; while ((1 << l) < M) ++l; return l; }

}  // namespace

void fft_init_twiddles() {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    const int dev = current_device();
    require(dev >= 0 && dev < 64, "device index out of range");
    if (g_tw[dev].f64) return;
    std::vector<cplx<double>> t64(MAX_M);
    std::vector<cplx<float>> t32(MAX_M);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < MAX_M; ++k) {
        const long double ang = -2.0L * pi * (long double)k / (long double)MAX_M;
        t64[k] = {(double)cosl(ang), (double)sinl(ang)};
        t32[k] = {(float)t64[k].x, (float)t64[k].y};
    }
    MG_HIP(hipMalloc(&g_tw[dev].f64, MAX_M * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&g_tw[dev].f32, MAX_M * sizeof(cplx<float>)));
    MG_HIP(hipMemcpy(g_tw[dev].f64, t64.data(), MAX_M * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g_tw[dev].f32, t32.data(), MAX_M * sizeof(cplx<float>), hipMemcpyHostToDevice));
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
    g_tw[dev].lds_max = lds;
}

static int g_lds_cap = 0;   // test knob: LDS bytes the FFT planner may assume (0 = the device's)
int fft_lds_max() {
    const int dev = g_tw[current_device()].lds_max;
    return g_lds_cap > 0 && g_lds_cap < dev ? g_lds_cap : dev;
}
void fft_set_lds_cap(int bytes) { g_lds_cap = bytes; }

// LDS per workgroup for one stage, and whether the column pass needs the global stash
int column_buffers(int M, int elem, bool centred) {
    const int lds = fft_lds_max(), one = M * elem;
    if (centred) return 3 * one <= lds ? 3 : (2 * one <= lds ? 2 : 0);
    return 2 * one <= lds ? 2 : (one <= lds ? 1 : 0);
}

template <class T> void launch_rows_forward(const T* W, int N, int M, cplx<T>* Wh, hipStream_t st) {
    const int logM = log2i(M), L = M / 2;
    const int lds_max = fft_lds_max();
    const int per = L * (int)sizeof(cplx<T>);
    require(per <= lds_max, "FFT size too large for LDS", 4);
    const int budget = lds_max < 65536 ? lds_max : 65536;
    int R = budget / per; R = R < 1 ? 1 : (R > 8 ? 8 : R);
    const int bytes = R * per;
    set_lds(k_rows_fwd<T>, bytes);
    hipLaunchKernelGGL(k_rows_fwd<T>, dim3((N + R - 1) / R, 3), dim3(TPB), bytes, st, W, N, logM, R, Wh, twiddles<T>(current_device()));
    MG_HIP(hipGetLastError());
}

template <class T> void launch_columns(const cplx<T>* Wh, int N, int M, const T* const* K, bool centred, cplx<T>* Ph,
                                       cplx<T>* stash, hipStream_t st) {
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int one = M * (int)sizeof(cplx<T>);
    const int nb = column_buffers(M, (int)sizeof(cplx<T>), centred);
    require(nb > 0, "FFT size too large for LDS", 4);
    const cplx<T>* tw = twiddles<T>(current_device());
    if (centred) {
        if (nb == 3) {
            set_lds(k_columns_centred<T, 3>, 3 * one);
            hipLaunchKernelGGL((k_columns_centred<T, 3>), dim3(Mh), dim3(TPB), 3 * one, st, Wh, N, logM, K[0], K[1], K[2], K[3], Ph, stash, tw);
        } else {
            require(stash != nullptr, "column stash missing");
            set_lds(k_columns_centred<T, 2>, 2 * one);
            hipLaunchKernelGGL((k_columns_centred<T, 2>), dim3(Mh), dim3(TPB), 2 * one, st, Wh, N, logM, K[0], K[1], K[2], K[3], Ph, stash, tw);
        }
    } else {
        if (nb == 2) {
            set_lds(k_columns<T, 2>, 2 * one);
            hipLaunchKernelGGL((k_columns<T, 2>), dim3(Mh), dim3(TPB), 2 * one, st, Wh, N, logM, K[0], K[1], Ph, stash, tw);
        } else {
            require(stash != nullptr, "column stash missing");
            set_lds(k_columns<T, 1>, one);
            hipLaunchKernelGGL((k_columns<T, 1>), dim3(Mh), dim3(TPB), one, st, Wh, N, logM, K[0], K[1], Ph, stash, tw);
        }
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_inverse(const cplx<T>* Ph, int N, int M, T* phi, hipStream_t st) {
    const int logM = log2i(M), L = M / 2;
    const int lds_max = fft_lds_max();
    const int per = L * (int)sizeof(cplx<T>);
    const T scale = (T)(1.0 / ((double)M * (double)M));
    const int budget = lds_max < 65536 ? lds_max : 65536;
    if (4 * per <= lds_max) {
        int R = budget / (4 * per); R = R < 1 ? 1 : (R > 8 ? 8 : R);
        const int bytes = 4 * R * per;
        set_lds(k_rows_inv<T, 4>, bytes);
        hipLaunchKernelGGL((k_rows_inv<T, 4>), dim3((N + R - 1) / R), dim3(TPB), bytes, st, Ph, N, logM, R, phi, scale, twiddles<T>(current_device()));
    } else {
        require(per <= lds_max, "FFT size too large for LDS", 4);
        int R = budget / per; R = R < 1 ? 1 : (R > 8 ? 8 : R);
        const int bytes = R * per;
        set_lds(k_rows_inv<T, 1>, bytes);
        hipLaunchKernelGGL((k_rows_inv<T, 1>), dim3((N + R - 1) / R, 4), dim3(TPB), bytes, st, Ph, N, logM, R, phi, scale, twiddles<T>(current_device()));
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_spectra(int M, double delta, double dof, const double* table, bool round_kernel, int ntab,
                                       double* tmp, T* const* K, hipStream_t st) {
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int bytes = (M / 2) * (int)sizeof(cplx<double>);
    require(bytes <= fft_lds_max(), "FFT size too large for LDS", 4);
    const cplx<double>* tw = twiddles<double>(current_device());
    set_lds(k_spec_rows, bytes);
    hipLaunchKernelGGL(k_spec_rows, dim3(M), dim3(TPB), bytes, st, logM, ntab, delta, dof, table, (int)round_kernel, tmp, tw);
    MG_HIP(hipGetLastError());
    set_lds(k_spec_cols<T>, bytes);
    hipLaunchKernelGGL(k_spec_cols<T>, dim3(Mh), dim3(TPB), bytes, st, logM, ntab, tmp, K[0], K[1], K[2], K[3], tw);
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template void launch_rows_forward<T>(const T*, int, int, cplx<T>*, hipStream_t); \
    template void launch_columns<T>(const cplx<T>*, int, int, const T* const*, bool, cplx<T>*, cplx<T>*, hipStream_t); \
    template void launch_rows_inverse<T>(const cplx<T>*, int, int, T*, hipStream_t); \
    template void launch_spectra<T>(int, double, double, const double*, bool, int, double*, T* const*, hipStream_t);
MG_INST(float)
MG_INST(double)

}  // namespace mg
