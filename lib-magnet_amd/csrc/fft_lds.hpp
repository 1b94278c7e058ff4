// Workgroup FFT in LDS for gfx950: batched power-of-two complex transforms, Stockham auto-sort with
// register radix-16/8/4/2 passes. Each pass: every work item reads its R inputs (stride L/R) into
// registers, applies the twiddles exp(sign 2 pi i r k / (Ns R)), runs a radix-R DFT in registers and
// writes back in the expanded order; a barrier separates the reads from the writes, so one LDS buffer
// per transform suffices. Natural order in and out. A transform of 2^13 points takes 4 passes.
// All functions address one LDS region `lds`; a transform batch starts at element offset `base`.
//
// LDS layout: element i of a buffer lives at swz(i) = i ^ ((i >> S) & (2^S - 1)), with 2^S elements
// per 128 bytes (S = 4 for float complex, 3 for double complex). The XOR keeps every aligned block of
// 2^S elements in place as a whole and permutes inside it, so both the strided reads and the expanded
// writes of a pass hit distinct banks across 2^S consecutive lanes. Every LDS access of a transform
// buffer goes through swz.
//
// Twiddles: exp(-2 pi i m / MAX_M) = hi[m >> 6] * lo[m & 63], two small tables staged in LDS by
// fft_stage_twiddles (MAX_M / 64 + 64 entries); the product of two correctly rounded entries. A pass
// looks up w^1, w^2, w^4, w^8 and forms the other powers with at most three complex products.
#pragma once
#include "common.hpp"

namespace mg {

template <class T> struct Swz { static constexpr int S = sizeof(cplx<T>) == 8 ? 4 : 3; };
template <class T> __device__ __forceinline__ int swz(int i) {
    constexpr int S = Swz<T>::S;
    return i ^ ((i >> S) & ((1 << S) - 1));
}

constexpr int TW_HI = MAX_M / 64;   // 128 entries: exp(-2 pi i 64 a / MAX_M)
constexpr int TW_LO = 64;           // 64 entries:  exp(-2 pi i b / MAX_M)

// cos / sin of 2 pi k / 16, k = 0..15 (exact decimal expansions to 20 digits)
__device__ constexpr double kC16[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                        0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                        -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                                        0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__device__ constexpr double kS16[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                                        1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                        0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                        -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

// in-register DFT of size R: input in bit-reversed order, output natural, in place; log2(R) radix-2
// decimation-in-time stages with constant twiddles exp(SIGN 2 pi i k / len) from the 16-point table
template <class T, int R, int SIGN> __device__ __forceinline__ void reg_dft(cplx<T>* v) {
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; ++k) {
                const int q = k * (16 / len);
                const cplx<T> o = v[i + k + len / 2];
                cplx<T> t;
                if (q == 0) t = o;
                else if (q == 4) t = SIGN < 0 ? cplx<T>{o.y, -o.x} : cplx<T>{-o.y, o.x};
                else t = cmul(o, cplx<T>{(T)kC16[q], (T)(SIGN * kS16[q])});
                const cplx<T> e = v[i + k];
                v[i + k] = cadd(e, t);
                v[i + k + len / 2] = csub(e, t);
            }
        }
    }
}

template <int logR> __device__ constexpr int brev_c(int r) {
    int o = 0;
    for (int b = 0; b < logR; ++b) o |= ((r >> b) & 1) << (logR - 1 - b);
    return o;
}

// LDS twiddle tables of one workgroup
template <class T> struct TwLDS { cplx<T> hi[TW_HI]; cplx<T> lo[TW_LO]; };

template <class T> __device__ inline void fft_stage_twiddles(TwLDS<T>& s, const cplx<T>* __restrict__ hi,
                                                             const cplx<T>* __restrict__ lo) {
    for (int i = threadIdx.x; i < TW_HI; i += blockDim.x) s.hi[i] = hi[i];
    for (int i = threadIdx.x; i < TW_LO; i += blockDim.x) s.lo[i] = lo[i];
}
template <class T, int SIGN> __device__ __forceinline__ cplx<T> twiddle(const TwLDS<T>& s, int m) {
    cplx<T> w = cmul(s.hi[(m >> 6) & (TW_HI - 1)], s.lo[m & 63]);
    if (SIGN > 0) w.y = -w.y;
    return w;
}

// Elements each thread holds in one pass, at most: ntr * L <= FFT_MAXE * blockDim.x (16 complex values:
// 32 VGPRs in FP32, 64 in FP64, so a column kernel also holding a spectrum stays clear of spills).
constexpr int FFT_MAXE = 16;

// One Stockham pass of radix R = 2^logR over `ntr` transforms of length 2^logL stored back to back from
// element `base` of the swizzled region `lds`. Ns = 2^logNs is the product of the earlier radices.
template <class T, int logR, int SIGN>
__device__ inline void stockham_pass(cplx<T>* lds, int base, int ntr, int logL, int logNs, const TwLDS<T>& tw) {
    constexpr int R = 1 << logR;
    constexpr int MAXJ = FFT_MAXE / R;
    const int L = 1 << logL, stride = L >> logR;
    const int items = ntr * stride;
    cplx<T> v[MAXJ][R];   // v[q][brev(r)] holds input r: the in-register DFT takes bit-reversed input
    int out[MAXJ];
#pragma unroll
    for (int q = 0; q < MAXJ; ++q) {
        const int j = threadIdx.x + q * blockDim.x;
        out[q] = -1;
        if (j >= items) continue;
        const int tr = j >> (logL - logR), jj = j & (stride - 1);
        const int k = jj & ((1 << logNs) - 1);
        const int src = base + (tr << logL) + jj;
#pragma unroll
        for (int r = 0; r < R; ++r) v[q][brev_c<logR>(r)] = lds[swz<T>(src + r * stride)];
        if (logNs > 0) {
            // w^r for r < R from exact table entries at r = 1, 2, 4, 8 and at most three products
            const int step = k << (LOG_MAX_M - logNs - logR);
            cplx<T> p2[4];
#pragma unroll
            for (int b = 0; b < logR; ++b) p2[b] = twiddle<T, SIGN>(tw, (1 << b) * step);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cplx<T> w = {T(1), T(0)};
                bool first = true;
#pragma unroll
                for (int b = 0; b < logR; ++b)
                    if ((r >> b) & 1) { w = first ? p2[b] : cmul(w, p2[b]); first = false; }
                v[q][brev_c<logR>(r)] = cmul(v[q][brev_c<logR>(r)], w);
            }
        }
        reg_dft<T, R, SIGN>(v[q]);
        out[q] = base + (tr << logL) + ((jj >> logNs) << (logNs + logR)) + k;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXJ; ++q) {
        if (out[q] < 0) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[swz<T>(out[q] + (r << logNs))] = v[q][r];
    }
    __syncthreads();
}

// Full transform(s): radix-16 passes while at least 8 bits remain, then 8 / 4 / 2. Requires
// ntr * 2^logL <= FFT_MAXE * blockDim.x. Unnormalised; SIGN -1 forward, +1 inverse.
template <class T, int SIGN>
__device__ inline void lds_fft(cplx<T>* lds, int base, int ntr, int logL, const TwLDS<T>& tw) {
    int logNs = 0, rem = logL;
    while (rem > 0) {
        const int lr = rem >= 8 ? 4 : (rem == 7 ? 4 : (rem >= 5 ? 3 : rem));
        switch (lr) {
            case 4: stockham_pass<T, 4, SIGN>(lds, base, ntr, logL, logNs, tw); break;
            case 3: stockham_pass<T, 3, SIGN>(lds, base, ntr, logL, logNs, tw); break;
            case 2: stockham_pass<T, 2, SIGN>(lds, base, ntr, logL, logNs, tw); break;
            default: stockham_pass<T, 1, SIGN>(lds, base, ntr, logL, logNs, tw); break;
        }
        logNs += lr;
        rem -= lr;
    }
}

// Batched: any number of transforms (contiguous in the swizzled buffer), in groups that fit the register
// budget of one pass.
template <class T, int SIGN>
__device__ inline void lds_fft_batch(cplx<T>* lds, int base, int ntr, int logL, const TwLDS<T>& tw) {
    const int per = (FFT_MAXE * (int)blockDim.x) >> logL;
    const int g = per < 1 ? 1 : per;
    for (int t0 = 0; t0 < ntr; t0 += g) lds_fft<T, SIGN>(lds, base + (t0 << logL), ntr - t0 < g ? ntr - t0 : g, logL, tw);
}

}  // namespace mg
