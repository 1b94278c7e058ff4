// FFT convolution of the charge grids with the two Cauchy kernels (src/fftconv.cpp:124-198), LDS
// resident on gfx950, written for this path (no rocFFT):
//   rows forward : real rows of length M (N live values, zero padding) -> M/2 + 1 complex, through one
//                  complex transform of length M/2 (even / odd samples as real / imaginary parts); the
//                  output is stored transposed, (channel, ky, gx), so the column pass reads contiguously;
//   columns      : per ky, the three columns are transformed (rows >= N are zero), multiplied by the real
//                  kernel spectra (K1 for channel 0, K2 for channels 0, 1, 2) and transformed back; only
//                  the N live output rows are stored -- forward, products and inverse in one workgroup;
//   rows inverse : Hermitian half rows -> real rows of length M through one complex transform of length
//                  M/2, scaled by 1/M^2, stored interleaved (gx, gy, 4) for the gather.
// All transforms are radix-2 in place in LDS: decimation in time on bit-reversed input, decimation in
// frequency for the column inverse (natural input, bit-reversed output, read back through bitrev).
// Twiddles come from one table exp(-2 pi i k / MAX_M) per precision, computed on the host in long double.
#include "common.hpp"
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

__device__ inline int bitrev(int i, int logL) { return logL == 0 ? 0 : (int)(__brev((unsigned)i) >> (32 - logL)); }

// radix-2 DIT over `ntr` consecutive transforms of length L = 2^logL, bit-reversed input, natural output
template <class T, int SIGN> __device__ void lds_dit(cplx<T>* buf, int ntr, int logL, const cplx<T>* __restrict__ tw) {
    const int L = 1 << logL;
    const int half_total = ntr * (L >> 1);
    for (int s = 1; s <= logL; ++s) {
        const int half = 1 << (s - 1), tws = LOG_MAX_M - s;
        for (int b = threadIdx.x; b < half_total; b += blockDim.x) {
            const int tr = b >> (logL - 1), bb = b & ((L >> 1) - 1);
            const int k = bb & (half - 1);
            const int i0 = tr * L + ((bb >> (s - 1)) << s) + k, i1 = i0 + half;
            cplx<T> w = tw[k << tws];
            if (SIGN > 0) w.y = -w.y;
            const cplx<T> u = buf[i0], v = cmul(buf[i1], w);
            buf[i0] = cadd(u, v); buf[i1] = csub(u, v);
        }
        __syncthreads();
    }
}

// radix-2 DIF, natural input, bit-reversed output
template <class T, int SIGN> __device__ void lds_dif(cplx<T>* buf, int ntr, int logL, const cplx<T>* __restrict__ tw) {
    const int L = 1 << logL;
    const int half_total = ntr * (L >> 1);
    for (int s = logL; s >= 1; --s) {
        const int half = 1 << (s - 1), tws = LOG_MAX_M - s;
        for (int b = threadIdx.x; b < half_total; b += blockDim.x) {
            const int tr = b >> (logL - 1), bb = b & ((L >> 1) - 1);
            const int k = bb & (half - 1);
            const int i0 = tr * L + ((bb >> (s - 1)) << s) + k, i1 = i0 + half;
            cplx<T> w = tw[k << tws];
            if (SIGN > 0) w.y = -w.y;
            const cplx<T> u = buf[i0], v = buf[i1];
            buf[i0] = cadd(u, v); buf[i1] = cmul(csub(u, v), w);
        }
        __syncthreads();
    }
}

// X[k] (k = 0..L) of a real sequence from Z = DFT_L(even + i odd): E = (Z_k + conj Z_{L-k}) / 2,
// O = -i (Z_k - conj Z_{L-k}) / 2, X_k = E + exp(-2 pi i k / M) O
template <class T> __device__ inline cplx<T> r2c_post(const cplx<T>* z, int k, int L, int logM, const cplx<T>* tw) {
    const cplx<T> zk = z[k & (L - 1)], zc = cconj(z[(L - k) & (L - 1)]);
    const cplx<T> e = {(zk.x + zc.x) * T(0.5), (zk.y + zc.y) * T(0.5)};
    const cplx<T> d = csub(zk, zc);
    const cplx<T> o = {d.y * T(0.5), -d.x * T(0.5)};
    if (k == L) return {e.x - o.x, e.y - o.y};
    return cadd(e, cmul(tw[k << (LOG_MAX_M - logM)], o));
}

// ---- rows forward: blockIdx.y = channel, R rows per workgroup
template <class T>
__global__ __launch_bounds__(TPB) void k_rows_fwd(const T* __restrict__ W, int N, int logM, int R,
                                                  cplx<T>* __restrict__ Wh, const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1;
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
        buf[r * L + bitrev(j, logL)] = z;
    }
    __syncthreads();
    lds_dit<T, -1>(buf, R, logL, tw);
    for (int e = threadIdx.x; e < R * Mh; e += TPB) {
        const int k = e / R, r = e - k * R;
        if (r < rows) Wh[((size_t)c * Mh + k) * N + gx0 + r] = r2c_post(buf + r * L, k, L, logM, tw);
    }
}

// ---- columns: forward, kernel products, inverse; TWO = both channel-0 products held in LDS
template <class T, bool TWO>
__global__ __launch_bounds__(TPB) void k_columns(const cplx<T>* __restrict__ Wh, int N, int logM,
                                                 const T* __restrict__ K1, const T* __restrict__ K2,
                                                 cplx<T>* __restrict__ Ph, cplx<T>* __restrict__ stash,
                                                 const cplx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* B = A + (1 << logM);
    const int M = 1 << logM, Mh = M / 2 + 1, ky = blockIdx.x;
    const T* k1 = K1 + (size_t)ky * M;
    const T* k2 = K2 + (size_t)ky * M;
    for (int c = 0; c < 3; ++c) {
        const cplx<T>* src = Wh + ((size_t)c * Mh + ky) * N;
        for (int i = threadIdx.x; i < M; i += TPB) {
            const int j = bitrev(i, logM);
            A[i] = j < N ? src[j] : cplx<T>{T(0), T(0)};
        }
        __syncthreads();
        lds_dit<T, -1>(A, 1, logM, tw);
        int nout = 1;
        if (c == 0) {
            for (int k = threadIdx.x; k < M; k += TPB) {
                const cplx<T> a = A[k];
                if (TWO) B[k] = {a.x * k2[k], a.y * k2[k]};
                else stash[(size_t)ky * M + k] = a;
                A[k] = {a.x * k1[k], a.y * k1[k]};
            }
            nout = TWO ? 2 : 1;
        } else {
            for (int k = threadIdx.x; k < M; k += TPB) { const cplx<T> a = A[k]; A[k] = {a.x * k2[k], a.y * k2[k]}; }
        }
        __syncthreads();
        lds_dif<T, +1>(A, nout, logM, tw);
        const int o = c == 0 ? 0 : c + 1;
        for (int g = threadIdx.x; g < N; g += TPB) {
            const int br = bitrev(g, logM);
            Ph[((size_t)o * Mh + ky) * N + g] = A[br];
            if (c == 0 && TWO) Ph[((size_t)1 * Mh + ky) * N + g] = B[br];
        }
        __syncthreads();
        if (c == 0 && !TWO) {
            for (int k = threadIdx.x; k < M; k += TPB) {
                const cplx<T> a = stash[(size_t)ky * M + k];
                A[k] = {a.x * k2[k], a.y * k2[k]};
            }
            __syncthreads();
            lds_dif<T, +1>(A, 1, logM, tw);
            for (int g = threadIdx.x; g < N; g += TPB) Ph[((size_t)1 * Mh + ky) * N + g] = A[bitrev(g, logM)];
            __syncthreads();
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
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1;
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
            const cplx<T> od = cmul(csub(xk, xc), cconj(tw[k << (LOG_MAX_M - logM)]));
            z = {ev.x - od.y, ev.y + od.x};
        }
        buf[(oc * R + r) * L + bitrev(k, logL)] = z;
    }
    __syncthreads();
    lds_dit<T, +1>(buf, CH * R, logL, tw);
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

// ---- kernel spectra. Pass 1: per row a, the real transform along b of K(a, b) (double), kept real
// (the kernels are even, so the spectra are real); pass 2: per column kb the transform along a.
__device__ inline double kernel_value(int a, int b, int M, double delta, double dof, const double* table, int which,
                                      bool round_kernel) {
    double v;
    if (table) v = table[((size_t)which * M + a) * M + b];
    else {
        const double ox = (a <= M / 2 ? a : a - M) * delta, oy = (b <= M / 2 ? b : b - M) * delta;
        const double base = 1.0 + (ox * ox + oy * oy) / dof;
        v = which == 0 ? pow(base, -dof) : pow(base, -dof - 1.0);
    }
    if (round_kernel) v = (double)(float)v;
    return v;
}

__global__ __launch_bounds__(TPB) void k_spec_rows(int logM, double delta, double dof, const double* __restrict__ table,
                                                   int round_kernel, double* __restrict__ tmp,
                                                   const cplx<double>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, a = blockIdx.x;
    for (int which = 0; which < 2; ++which) {
        for (int j = threadIdx.x; j < L; j += TPB)
            buf[bitrev(j, logL)] = {kernel_value(a, 2 * j, M, delta, dof, table, which, round_kernel),
                                    kernel_value(a, 2 * j + 1, M, delta, dof, table, which, round_kernel)};
        __syncthreads();
        lds_dit<double, -1>(buf, 1, logL, tw);
        for (int k = threadIdx.x; k < Mh; k += TPB) tmp[((size_t)which * M + a) * Mh + k] = r2c_post(buf, k, L, logM, tw).x;
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(TPB) void k_spec_cols(int logM, const double* __restrict__ tmp, T* __restrict__ K1,
                                                   T* __restrict__ K2, const cplx<double>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<double>* buf = reinterpret_cast<cplx<double>*>(smem);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, kb = blockIdx.x;
    for (int which = 0; which < 2; ++which) {
        const double* src = tmp + (size_t)which * M * Mh + kb;
        for (int j = threadIdx.x; j < L; j += TPB) buf[bitrev(j, logL)] = {src[(size_t)(2 * j) * Mh], src[(size_t)(2 * j + 1) * Mh]};
        __syncthreads();
        lds_dit<double, -1>(buf, 1, logL, tw);
        T* out = (which == 0 ? K1 : K2) + (size_t)kb * M;
        for (int k = threadIdx.x; k < Mh; k += TPB) {
            const T v = (T)r2c_post(buf, k, L, logM, tw).x;
            out[k] = v;
            if (k > 0 && k < L) out[M - k] = v;
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
    const int H = MAX_M / 2;
    std::vector<cplx<double>> t64(H);
    std::vector<cplx<float>> t32(H);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < H; ++k) {
        const long double ang = -2.0L * pi * (long double)k / (long double)MAX_M;
        t64[k] = {(double)cosl(ang), (double)sinl(ang)};
        t32[k] = {(float)t64[k].x, (float)t64[k].y};
    }
    MG_HIP(hipMalloc(&g_tw[dev].f64, H * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&g_tw[dev].f32, H * sizeof(cplx<float>)));
    MG_HIP(hipMemcpy(g_tw[dev].f64, t64.data(), H * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g_tw[dev].f32, t32.data(), H * sizeof(cplx<float>), hipMemcpyHostToDevice));
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
    g_tw[dev].lds_max = lds;
}

int fft_lds_max() { return g_tw[current_device()].lds_max; }

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

template <class T> void launch_columns(const cplx<T>* Wh, int N, int M, const T* K1, const T* K2, cplx<T>* Ph,
                                       cplx<T>* stash, hipStream_t st) {
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int lds_max = fft_lds_max();
    const int one = M * (int)sizeof(cplx<T>);
    if (2 * one <= lds_max) {
        set_lds(k_columns<T, true>, 2 * one);
        hipLaunchKernelGGL((k_columns<T, true>), dim3(Mh), dim3(TPB), 2 * one, st, Wh, N, logM, K1, K2, Ph, stash, twiddles<T>(current_device()));
    } else {
        require(one <= lds_max && stash != nullptr, "FFT size too large for LDS", 4);
        set_lds(k_columns<T, false>, one);
        hipLaunchKernelGGL((k_columns<T, false>), dim3(Mh), dim3(TPB), one, st, Wh, N, logM, K1, K2, Ph, stash, twiddles<T>(current_device()));
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

template <class T> void launch_spectra(int M, double delta, double dof, const double* table, bool round_kernel,
                                       double* tmp, T* K1, T* K2, hipStream_t st) {
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int bytes = (M / 2) * (int)sizeof(cplx<double>);
    require(bytes <= fft_lds_max(), "FFT size too large for LDS", 4);
    const cplx<double>* tw = twiddles<double>(current_device());
    set_lds(k_spec_rows, bytes);
    hipLaunchKernelGGL(k_spec_rows, dim3(M), dim3(TPB), bytes, st, logM, delta, dof, table, (int)round_kernel, tmp, tw);
    MG_HIP(hipGetLastError());
    set_lds(k_spec_cols<T>, bytes);
    hipLaunchKernelGGL(k_spec_cols<T>, dim3(Mh), dim3(TPB), bytes, st, logM, tmp, K1, K2, tw);
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template void launch_rows_forward<T>(const T*, int, int, cplx<T>*, hipStream_t); \
    template void launch_columns<T>(const cplx<T>*, int, int, const T*, const T*, cplx<T>*, cplx<T>*, hipStream_t); \
    template void launch_rows_inverse<T>(const cplx<T>*, int, int, T*, hipStream_t); \
    template void launch_spectra<T>(int, double, double, const double*, bool, double*, T*, T*, hipStream_t);
MG_INST(float)
MG_INST(double)

}  // namespace mg
