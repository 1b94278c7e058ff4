// FFT convolution of the charge grids with the Cauchy kernels (src/fftconv.cpp:124-198), LDS resident
// on gfx950 and written for this path (no rocFFT). Transforms are the Stockham radix-16/8/4/2 passes of
// fft_lds.hpp.
//   rows forward : real rows of length M (N live values, zero padding) -> M/2 + 1 complex, through one
//                  complex transform of length M/2 (even / odd samples as real / imaginary parts), stored
//                  by rows; a tiled transpose then lays them out (channel, ky, gx) for the column pass;
//   columns      : per ky, the three columns are transformed (rows >= N are zero), multiplied by the
//                  kernel spectra and transformed back; only the N live rows are stored. Forward,
//                  products and inverse stay in one workgroup;
//   rows inverse : (after a transpose back to rows) Hermitian half rows -> real rows of length M through
//                  one complex transform of length M/2, scaled by 1/M^2, stored interleaved (gx, gy, 4).
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

struct Twiddles {
    cplx<float>* hi32 = nullptr; cplx<float>* lo32 = nullptr;
    cplx<double>* hi64 = nullptr; cplx<double>* lo64 = nullptr;
    int lds_max = 65536;
};
Twiddles g_tw[64];
std::mutex g_tw_mu;

template <class T> struct TwPtr { const cplx<T>* hi; const cplx<T>* lo; };
template <class T> TwPtr<T> twiddles(int dev);
template <> TwPtr<float> twiddles<float>(int dev) { return {g_tw[dev].hi32, g_tw[dev].lo32}; }
template <> TwPtr<double> twiddles<double>(int dev) { return {g_tw[dev].hi64, g_tw[dev].lo64}; }

// threads of the FFT workgroups; __launch_bounds__(FTPB, 1) lets the compiler use up to 256 VGPRs
constexpr int FTPB = 512;

int current_device() { int d = 0; MG_HIP(hipGetDevice(&d)); return d; }

// X[k] (k = 0..L) of a real sequence from Z = DFT_L(even + i odd): E = (Z_k + conj Z_{L-k}) / 2,
// O = -i (Z_k - conj Z_{L-k}) / 2, X_k = E + exp(-2 pi i k / M) O
template <class T> __device__ inline cplx<T> r2c_post(const cplx<T>* lds, int base, int k, int L, int logM,
                                                     const TwLDS<T>& tw) {
    const cplx<T> zk = lds[swz<T>(base + (k & (L - 1)))], zc = cconj(lds[swz<T>(base + ((L - k) & (L - 1)))]);
    const cplx<T> e = {(zk.x + zc.x) * T(0.5), (zk.y + zc.y) * T(0.5)};
    const cplx<T> d = csub(zk, zc);
    const cplx<T> o = {d.y * T(0.5), -d.x * T(0.5)};
    return cadd(e, cmul(twiddle<T, -1>(tw, (k << (LOG_MAX_M - logM)) & (MAX_M - 1)), o));
}

// ---- rows forward: blockIdx.y = channel, R rows per workgroup
template <class T>
__global__ __launch_bounds__(FTPB, 1) void k_rows_fwd(const T* __restrict__ W, int N, int logM, int R,
                                                   cplx<T>* __restrict__ Wh, TwPtr<T> twp) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TwLDS<T> tw;
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    fft_stage_twiddles(tw, twp.hi, twp.lo);
    const int L = 1 << (logM - 1), logL = logM - 1, Mh = L + 1;
    const int c = blockIdx.y, gx0 = blockIdx.x * R;
    const int rows = min(R, N - gx0);
    for (int e = threadIdx.x; e < R * L; e += FTPB) {
        const int r = e >> logL, j = e & (L - 1);
        cplx<T> z = {T(0), T(0)};
        if (r < rows) {
            const T* src = W + ((size_t)c * N + gx0 + r) * N;
            if (2 * j < N) z.x = src[2 * j];
            if (2 * j + 1 < N) z.y = src[2 * j + 1];
        }
        lds[swz<T>(e)] = z;
    }
    __syncthreads();
    lds_fft_batch<T, -1>(lds, 0, R, logL, tw);
    for (int e = threadIdx.x; e < R * Mh; e += FTPB) {   // row-major (channel, gx, ky): contiguous rows
        const int r = e / Mh, k = e - r * Mh;
        if (r < rows) Wh[((size_t)c * N + gx0 + r) * Mh + k] = r2c_post(lds, r * L, k, L, logM, tw);
    }
}

// ---- batched transpose of complex matrices, src (B, R, C) -> dst (B, C, R), 32 x 32 tiles through LDS
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int R, int C) {
    __shared__ cplx<T> tile[32][33];
    const size_t off = (size_t)blockIdx.z * R * C;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        if (r < R && c < C) tile[i][tx] = src[off + (size_t)r * C + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (r < R && c < C) dst[off + (size_t)c * R + r] = tile[tx][i];
    }
}


// kernel table value at column frequency kx (stored for kx <= M/2; `odd` flips the sign above M/2)
template <class T> __device__ inline T ktab(const T* row, int kx, int M, bool odd) {
    const int Mh = M / 2 + 1;
    if (kx < Mh) return row[kx];
    const T v = row[M - kx];
    return odd ? -v : v;
}

template <class T> __device__ inline void load_column(cplx<T>* lds, int base, const cplx<T>* __restrict__ src, int N, int M) {
    for (int i = threadIdx.x; i < M; i += FTPB) lds[swz<T>(base + i)] = i < N ? src[i] : cplx<T>{T(0), T(0)};
}
template <class T> __device__ inline void store_column(cplx<T>* __restrict__ dst, const cplx<T>* lds, int base, int N) {
    for (int g = threadIdx.x; g < N; g += FTPB) dst[g] = lds[swz<T>(base + g)];
}

// ---- columns. One workgroup per column ky: the three channel columns are transformed (rows >= N are
// zero), multiplied by the kernel spectra and transformed back; only the N live rows are stored. The
// channel-0 spectrum, used by every product, stays in registers (element k = tid + FTPB e of thread tid).
// NB = LDS buffers of M elements: 2 pairs the transforms (channels 1 and 2, and the two channel-0
// products), 1 runs them one at a time in half the LDS (two workgroups per CU at the largest M).
// CEN selects the product set: false = the reference's, true = the centred one (odd-kernel spectra
// i Tx, i Ty).
template <class T, int NB, bool CEN>
__global__ __launch_bounds__(FTPB, 1) void k_columns(const cplx<T>* __restrict__ Wh, int N, int logM,
                                                  const T* __restrict__ K1, const T* __restrict__ K2,
                                                  const T* __restrict__ TX, const T* __restrict__ TY,
                                                  cplx<T>* __restrict__ Ph, TwPtr<T> twp) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TwLDS<T> tw;
    constexpr int EMAX = MAX_M / FTPB;
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    fft_stage_twiddles(tw, twp.hi, twp.lo);
    const int M = 1 << logM, Mh = M / 2 + 1, ky = blockIdx.x;
    const int A = 0, B = NB == 2 ? M : 0;
    const int E = (M + FTPB - 1) / FTPB;
    const T* k1 = K1 + (size_t)ky * Mh;
    const T* k2 = K2 + (size_t)ky * Mh;
    const T* tx = CEN ? TX + (size_t)ky * Mh : nullptr;
    const T* ty = CEN ? TY + (size_t)ky * Mh : nullptr;
    auto col = [&](int ch) { return Wh + ((size_t)ch * Mh + ky) * N; };
    auto out = [&](int o) { return Ph + ((size_t)o * Mh + ky) * N; };
    auto at = [&](int buf, int k) -> cplx<T>& { return lds[swz<T>(buf + k)]; };
    // product of channel ch (1, 2) at frequency k with spectrum s: reference K2 s; centred i t s0 - K2 s
    auto prod = [&](int ch, int k, cplx<T> s0, cplx<T> s) {
        const T c2 = ktab(k2, k, M, false);
        if (!CEN) return cplx<T>{s.x * c2, s.y * c2};
        const T t = ch == 1 ? ktab(tx, k, M, true) : ktab(ty, k, M, false);
        return cplx<T>{-t * s0.y - c2 * s.x, t * s0.x - c2 * s.y};
    };
    cplx<T> s0[EMAX];
    // channel 0 forward, kept in registers
    load_column(lds, A, col(0), N, M);
    __syncthreads();
    lds_fft<T, -1>(lds, A, 1, logM, tw);
#pragma unroll
    for (int e = 0; e < EMAX; ++e) if (e < E) { const int k = threadIdx.x + e * FTPB; if (k < M) s0[e] = at(A, k); }
    __syncthreads();
    if (NB == 2) {
        // phi0 = K1 s0, phi1 = K2 s0, back together
#pragma unroll
        for (int e = 0; e < EMAX; ++e) {
            const int k = threadIdx.x + e * FTPB;
            if (e >= E || k >= M) continue;
            const T c1 = ktab(k1, k, M, false), c2 = ktab(k2, k, M, false);
            at(A, k) = {s0[e].x * c1, s0[e].y * c1};
            at(B, k) = {s0[e].x * c2, s0[e].y * c2};
        }
        __syncthreads();
        lds_fft_batch<T, +1>(lds, A, 2, logM, tw);
        store_column(out(0), lds, A, N);
        store_column(out(1), lds, B, N);
        __syncthreads();
        // channels 1 and 2 together
        load_column(lds, A, col(1), N, M);
        load_column(lds, B, col(2), N, M);
        __syncthreads();
        lds_fft_batch<T, -1>(lds, A, 2, logM, tw);
#pragma unroll
        for (int e = 0; e < EMAX; ++e) {
            const int k = threadIdx.x + e * FTPB;
            if (e >= E || k >= M) continue;
            at(A, k) = prod(1, k, s0[e], at(A, k));
            at(B, k) = prod(2, k, s0[e], at(B, k));
        }
        __syncthreads();
        lds_fft_batch<T, +1>(lds, A, 2, logM, tw);
        store_column(out(2), lds, A, N);
        store_column(out(3), lds, B, N);
    } else {
        for (int o = 0; o < 4; ++o) {
            if (o >= 2) {   // channel o - 1 forward
                load_column(lds, A, col(o - 1), N, M);
                __syncthreads();
                lds_fft<T, -1>(lds, A, 1, logM, tw);
            }
#pragma unroll
            for (int e = 0; e < EMAX; ++e) {
                const int k = threadIdx.x + e * FTPB;
                if (e >= E || k >= M) continue;
                if (o < 2) { const T c = ktab(o == 0 ? k1 : k2, k, M, false); at(A, k) = {s0[e].x * c, s0[e].y * c}; }
                else at(A, k) = prod(o - 1, k, s0[e], at(A, k));
            }
            __syncthreads();
            lds_fft<T, +1>(lds, A, 1, logM, tw);
            store_column(out(o), lds, A, N);
            __syncthreads();
        }
    }
}

// ---- rows inverse: CH channels per workgroup (4: all, interleaved store; 1: blockIdx.y = channel)
template <class T> struct vec4;
template <> struct vec4<float> { typedef float4 type; };
template <> struct vec4<double> { typedef double4 type; };

template <class T, int CH>
__global__ __launch_bounds__(FTPB, 1) void k_rows_inv(const cplx<T>* __restrict__ Ph, int N, int logM, int R,
                                                   T* __restrict__ phi, T scale, TwPtr<T> twp) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TwLDS<T> tw;
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    fft_stage_twiddles(tw, twp.hi, twp.lo);
    __syncthreads();
    const int L = 1 << (logM - 1), logL = logM - 1, Mh = L + 1;
    const int gx0 = blockIdx.x * R, o0 = CH == 4 ? 0 : blockIdx.y;
    const int rows = min(R, N - gx0);
    // Z_k = E + i O with E = X_k + conj X_{L-k}, O = (X_k - conj X_{L-k}) exp(+2 pi i k / M)
    for (int e = threadIdx.x; e < CH * R * L; e += FTPB) {
        const int oc = e / (R * L), rem = e - oc * (R * L);
        const int r = rem / L, k = rem - r * L;
        cplx<T> z = {T(0), T(0)};
        if (r < rows) {   // row-major input (channel, gx, ky)
            const size_t base = ((size_t)(o0 + oc) * N + gx0 + r) * Mh;
            const cplx<T> xk = Ph[base + k], xc = cconj(Ph[base + (L - k)]);
            const cplx<T> ev = cadd(xk, xc);
            const cplx<T> od = cmul(csub(xk, xc), twiddle<T, +1>(tw, (k << (LOG_MAX_M - logM)) & (MAX_M - 1)));
            z = {ev.x - od.y, ev.y + od.x};
        }
        lds[swz<T>((oc * R + r) * L + k)] = z;
    }
    __syncthreads();
    lds_fft_batch<T, +1>(lds, 0, CH * R, logL, tw);
    for (int e = threadIdx.x; e < R * N; e += FTPB) {
        const int r = e / N, gy = e - r * N;
        if (r >= rows) continue;
        const int j = gy >> 1;
        const bool odd = gy & 1;
        if (CH == 4) {
            typename vec4<T>::type v;
            const cplx<T> a = lds[swz<T>((0 * R + r) * L + j)], b = lds[swz<T>((1 * R + r) * L + j)];
            const cplx<T> c = lds[swz<T>((2 * R + r) * L + j)], d = lds[swz<T>((3 * R + r) * L + j)];
            v.x = (odd ? a.y : a.x) * scale; v.y = (odd ? b.y : b.x) * scale;
            v.z = (odd ? c.y : c.x) * scale; v.w = (odd ? d.y : d.x) * scale;
            reinterpret_cast<typename vec4<T>::type*>(phi)[(size_t)(gx0 + r) * N + gy] = v;
        } else {
            const cplx<T> a = lds[swz<T>(r * L + j)];
            phi[((size_t)(gx0 + r) * N + gy) * 4 + o0] = (odd ? a.y : a.x) * scale;
        }
    }
}

// ---- kernel spectra, computed in TC (double for the FP64 path, float for the FP32 path, whose
// reference counterpart rounds the kernel values to float32 and transforms them in single precision).
// Tables: 0 K1, 1 K2, 2 x K2, 3 y K2 (offsets in grid units * delta).
// Pass 1: per row a, the real transform along b (keep Re, or Im for the table odd in b: y K2); the row
// transform of x K2 is off_a times that of K2. Pass 2: per column kb, the real transform along a (keep
// Re, or Im for the table odd in a: x K2). Output (Mh, Mh) per table: [ky][kx], kx <= M/2.
__device__ inline void kernel_values(int a, int b, int M, double delta, double dof, const double* table, bool round_kernel,
                                     double& k1, double& k2, double& ky2) {
    const double ox = (a <= M / 2 ? a : a - M) * delta, oy = (b <= M / 2 ? b : b - M) * delta;
    if (table) {
        k1 = table[(size_t)a * M + b];
        k2 = table[(size_t)M * M + (size_t)a * M + b];
    } else if (dof == 1.0) {
        const double inv = 1.0 / (1.0 + ox * ox + oy * oy);
        k1 = inv; k2 = inv * inv;
    } else {
        const double base = 1.0 + (ox * ox + oy * oy) / dof;
        k1 = pow(base, -dof); k2 = pow(base, -dof - 1.0);
    }
    if (round_kernel) { k1 = (double)(float)k1; k2 = (double)(float)k2; }
    // the odd tables drop the Nyquist offset M/2 (no negative twin on the torus), which keeps them exactly
    // odd -- so their spectra are imaginary -- and is never used: the live offsets stay below N <= M/2
    ky2 = b == M / 2 ? 0.0 : oy * k2;
}

template <class TC>
__global__ __launch_bounds__(FTPB, 1) void k_spec_rows(int logM, int ntab, double delta, double dof,
                                                    const double* __restrict__ table, int round_kernel,
                                                    TC* __restrict__ tmp, TwPtr<TC> twp) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TwLDS<TC> tw;
    cplx<TC>* lds = reinterpret_cast<cplx<TC>*>(smem);
    fft_stage_twiddles(tw, twp.hi, twp.lo);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, a = blockIdx.x;
    const int nfft = ntab == 2 ? 2 : 3;   // K1, K2 (, y K2)
    for (int j = threadIdx.x; j < L; j += FTPB) {
        double e1, e2, ey, o1, o2, oy;
        kernel_values(a, 2 * j, M, delta, dof, table, round_kernel, e1, e2, ey);
        kernel_values(a, 2 * j + 1, M, delta, dof, table, round_kernel, o1, o2, oy);
        lds[swz<TC>(j)] = {(TC)e1, (TC)o1};
        lds[swz<TC>(L + j)] = {(TC)e2, (TC)o2};
        if (nfft == 3) lds[swz<TC>(2 * L + j)] = {(TC)ey, (TC)oy};
    }
    __syncthreads();
    lds_fft_batch<TC, -1>(lds, 0, nfft, logL, tw);
    const double ox = (a <= M / 2 ? a : a - M) * delta;
    const TC xa = a == M / 2 ? TC(0) : (TC)ox;
    for (int k = threadIdx.x; k < Mh; k += FTPB) {
        const TC r1 = r2c_post(lds, 0, k, L, logM, tw).x, r2 = r2c_post(lds, L, k, L, logM, tw).x;
        tmp[((size_t)0 * M + a) * Mh + k] = r1;
        tmp[((size_t)1 * M + a) * Mh + k] = r2;
        if (ntab == 4) {
            tmp[((size_t)2 * M + a) * Mh + k] = xa * r2;
            tmp[((size_t)3 * M + a) * Mh + k] = r2c_post(lds, 2 * L, k, L, logM, tw).y;
        }
    }
}

template <class TC, class T>
__global__ __launch_bounds__(FTPB, 1) void k_spec_cols(int logM, int ntab, const TC* __restrict__ tmp, T* __restrict__ K1,
                                                    T* __restrict__ K2, T* __restrict__ TX, T* __restrict__ TY,
                                                    TwPtr<TC> twp) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ TwLDS<TC> tw;
    cplx<TC>* lds = reinterpret_cast<cplx<TC>*>(smem);
    fft_stage_twiddles(tw, twp.hi, twp.lo);
    const int M = 1 << logM, L = M >> 1, logL = logM - 1, Mh = L + 1, kb = blockIdx.x;
    for (int w = 0; w < ntab; ++w) {
        const TC* src = tmp + (size_t)w * M * Mh + kb;
        for (int j = threadIdx.x; j < L; j += FTPB) lds[swz<TC>(w * L + j)] = {src[(size_t)(2 * j) * Mh], src[(size_t)(2 * j + 1) * Mh]};
    }
    __syncthreads();
    lds_fft_batch<TC, -1>(lds, 0, ntab, logL, tw);
    for (int w = 0; w < ntab; ++w) {
        T* out = (w == 0 ? K1 : w == 1 ? K2 : w == 2 ? TX : TY) + (size_t)kb * Mh;
        for (int k = threadIdx.x; k < Mh; k += FTPB) {
            const cplx<TC> X = r2c_post(lds, w * L, k, L, logM, tw);
            out[k] = (T)(w == 2 ? X.y : X.x);
        }
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
    if (g_tw[dev].hi64) return;
    const long double pi = 3.141592653589793238462643383279502884L;
    auto entry = [&](int m) {
        const long double ang = -2.0L * pi * (long double)m / (long double)MAX_M;
        return cplx<double>{(double)cosl(ang), (double)sinl(ang)};
    };
    std::vector<cplx<double>> hi64(TW_HI), lo64(TW_LO);
    std::vector<cplx<float>> hi32(TW_HI), lo32(TW_LO);
    for (int a = 0; a < TW_HI; ++a) { hi64[a] = entry(64 * a); hi32[a] = {(float)hi64[a].x, (float)hi64[a].y}; }
    for (int b = 0; b < TW_LO; ++b) { lo64[b] = entry(b); lo32[b] = {(float)lo64[b].x, (float)lo64[b].y}; }
    MG_HIP(hipMalloc(&g_tw[dev].hi64, TW_HI * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&g_tw[dev].lo64, TW_LO * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&g_tw[dev].hi32, TW_HI * sizeof(cplx<float>)));
    MG_HIP(hipMalloc(&g_tw[dev].lo32, TW_LO * sizeof(cplx<float>)));
    MG_HIP(hipMemcpy(g_tw[dev].hi64, hi64.data(), TW_HI * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g_tw[dev].lo64, lo64.data(), TW_LO * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g_tw[dev].hi32, hi32.data(), TW_HI * sizeof(cplx<float>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(g_tw[dev].lo32, lo32.data(), TW_LO * sizeof(cplx<float>), hipMemcpyHostToDevice));
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
    g_tw[dev].lds_max = lds - (int)sizeof(TwLDS<double>);   // room for the staged twiddle tables
}

static int g_lds_cap = 0;   // test knob: LDS bytes the FFT planner may assume (0 = the device's)
int fft_lds_max() {
    const int dev = g_tw[current_device()].lds_max;   // dynamic LDS a workgroup may use
    return g_lds_cap > 0 && g_lds_cap < dev ? g_lds_cap : dev;
}
void fft_set_lds_cap(int bytes) { g_lds_cap = bytes; }

// LDS buffers of M elements the column pass uses: 2 (paired transforms) when two fit in 64 KB, so that
// at least two workgroups share a CU, else 1 (0: the column does not fit at all)
int column_buffers(int M, int elem, bool centred) {
    (void)centred;
    const int lds = fft_lds_max(), one = M * elem;
    const int budget = lds < 65536 ? lds : 65536;
    if (2 * one <= budget) return 2;
    return one <= lds ? 1 : 0;
}

template <class T> void launch_rows_forward(const T* W, int N, int M, cplx<T>* Wh, hipStream_t st) {
    const int logM = log2i(M), L = M / 2;
    const int lds_max = fft_lds_max();
    const int per = L * (int)sizeof(cplx<T>);
    require(per <= lds_max, "FFT size too large for LDS", 4);
    const int budget = lds_max < 65536 ? lds_max : 65536;
    int R = budget / per; R = R < 1 ? 1 : (R > 4 ? 4 : R);
    const int bytes = R * per;
    set_lds(k_rows_fwd<T>, bytes);
    hipLaunchKernelGGL(k_rows_fwd<T>, dim3((N + R - 1) / R, 3), dim3(FTPB), bytes, st, W, N, logM, R, Wh, twiddles<T>(current_device()));
    MG_HIP(hipGetLastError());
}

template <class T> void launch_transpose(const cplx<T>* src, cplx<T>* dst, int batch, int R, int C, hipStream_t st) {
    hipLaunchKernelGGL(k_transpose<T>, dim3((C + 31) / 32, (R + 31) / 32, batch), dim3(256), 0, st, src, dst, R, C);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_columns(const cplx<T>* Wh, int N, int M, const T* const* K, bool centred, cplx<T>* Ph,
                                       cplx<T>* stash, hipStream_t st) {
    (void)stash;
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int one = M * (int)sizeof(cplx<T>);
    const int nb = column_buffers(M, (int)sizeof(cplx<T>), centred);
    require(nb > 0, "FFT size too large for LDS", 4);
    const TwPtr<T> tw = twiddles<T>(current_device());
#define MG_COL(NB, CEN) do { \
        set_lds(k_columns<T, NB, CEN>, NB * one); \
        hipLaunchKernelGGL((k_columns<T, NB, CEN>), dim3(Mh), dim3(FTPB), NB * one, st, Wh, N, logM, K[0], K[1], K[2], K[3], Ph, tw); \
    } while (0)
    if (nb == 2) { if (centred) MG_COL(2, true); else MG_COL(2, false); }
    else { if (centred) MG_COL(1, true); else MG_COL(1, false); }
#undef MG_COL
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_inverse(const cplx<T>* Ph, int N, int M, T* phi, hipStream_t st) {
    const int logM = log2i(M), L = M / 2;
    const int lds_max = fft_lds_max();
    const int per = L * (int)sizeof(cplx<T>);
    const T scale = (T)(1.0 / ((double)M * (double)M));
    const int budget = lds_max < 65536 ? lds_max : 65536;
    if (4 * per <= lds_max) {   // all four channels of a row: interleaved (gx, gy, 4) stores
        int R = budget / (4 * per); R = R < 1 ? 1 : (R > 4 ? 4 : R);
        const int bytes = 4 * R * per;
        set_lds(k_rows_inv<T, 4>, bytes);
        hipLaunchKernelGGL((k_rows_inv<T, 4>), dim3((N + R - 1) / R), dim3(FTPB), bytes, st, Ph, N, logM, R, phi, scale, twiddles<T>(current_device()));
    } else {
        require(per <= lds_max, "FFT size too large for LDS", 4);
        int R = budget / per; R = R < 1 ? 1 : (R > 4 ? 4 : R);
        const int bytes = R * per;
        set_lds(k_rows_inv<T, 1>, bytes);
        hipLaunchKernelGGL((k_rows_inv<T, 1>), dim3((N + R - 1) / R, 4), dim3(FTPB), bytes, st, Ph, N, logM, R, phi, scale, twiddles<T>(current_device()));
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_spectra(int M, double delta, double dof, const double* table, bool round_kernel, int ntab,
                                       double* tmp, T* const* K, hipStream_t st) {
    typedef T TC;   // the path's own precision
    const int logM = log2i(M), Mh = M / 2 + 1;
    const int nfft = ntab == 2 ? 2 : 3;
    const int bytes_rows = nfft * (M / 2) * (int)sizeof(cplx<TC>);
    const int bytes_cols = ntab * (M / 2) * (int)sizeof(cplx<TC>);
    require(bytes_rows <= fft_lds_max() && bytes_cols <= fft_lds_max(), "FFT size too large for LDS", 4);
    const TwPtr<TC> tw = twiddles<TC>(current_device());
    TC* t = reinterpret_cast<TC*>(tmp);
    set_lds(k_spec_rows<TC>, bytes_rows);
    hipLaunchKernelGGL(k_spec_rows<TC>, dim3(M), dim3(FTPB), bytes_rows, st, logM, ntab, delta, dof, table, (int)round_kernel, t, tw);
    MG_HIP(hipGetLastError());
    set_lds(k_spec_cols<TC, T>, bytes_cols);
    hipLaunchKernelGGL((k_spec_cols<TC, T>), dim3(Mh), dim3(FTPB), bytes_cols, st, logM, ntab, t, K[0], K[1], K[2], K[3], tw);
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template void launch_rows_forward<T>(const T*, int, int, cplx<T>*, hipStream_t); \
    template void launch_transpose<T>(const cplx<T>*, cplx<T>*, int, int, int, hipStream_t); \
    template void launch_columns<T>(const cplx<T>*, int, int, const T* const*, bool, cplx<T>*, cplx<T>*, hipStream_t); \
    template void launch_rows_inverse<T>(const cplx<T>*, int, int, T*, hipStream_t); \
    template void launch_spectra<T>(int, double, double, const double*, bool, int, double*, T* const*, hipStream_t);
MG_INST(float)
MG_INST(double)

}  // namespace mg
