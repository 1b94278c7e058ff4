// FFT convolution of the charge grids with the Cauchy kernels (src/fftconv.cpp:124-198), and the kernel
// spectra (magnet/tsne.py:48-71), on gfx950 with the compile-time plans of fft_core.hpp (no rocFFT).
// Included by one translation unit per precision (k_fft32.hip, k_fft64.hip).
//   rows forward : real rows of length M (N live values, zero padding) -> M/2 + 1 complex values, through
//                  one complex transform of length L = M/2 (even / odd samples as real / imaginary parts);
//                  stored by rows, (channel, gx, ky); a tiled transpose then lays them out (channel, ky, gx);
//   columns      : per ky, the three channel columns are transformed (rows >= N are zero), multiplied by
//                  the kernel spectra in registers and transformed back; only the N live rows are stored;
//   rows inverse : (after a transpose back to rows) Hermitian half rows -> real rows of length M through one
//                  complex transform of length M/2, scaled by 1/M^2; planar potentials (4, N, N).
// Two product sets. The reference's, used by the FP64 path: phi0 = K1*W0, phi1..3 = K2*W0..2 with the
// charges (1, x, y). The centred one, used by the FP32 path, where the charges are (1, x - x_node,
// y - y_node): phi0 = K1*W0, phi1 = K2*W0, phi2 = (x K2)*W0 - K2*W1', phi3 = (y K2)*W0 - K2*W2'. The odd
// kernels x K2, y K2 have imaginary spectra (i Tx, i Ty). Both sets give the same forces in exact arithmetic;
// the centred one avoids the cancellation of pot1 * x - pot2 at large |x|.
// Kernel spectra are stored (Mh, Mh), [ky][kx], kx <= M/2 (they are even in kx, or odd for Tx).
#pragma once
#include "common.hpp"
#include "fft_core.hpp"
#include "batch.hpp"
#include <algorithm>

namespace mg {
namespace {

template <class T> __device__ __forceinline__ cplx<T> czero() { return {T(0), T(0)}; }

// X_k (0 <= k <= L) of a real sequence of length 2L from Z = DFT_L(even + i odd) held in LDS:
// E = (Z_k + conj Z_{L-k}) / 2, O = -i (Z_k - conj Z_{L-k}) / 2, X_k = E + exp(-2 pi i k / 2L) O
template <class T, int L> __device__ inline cplx<T> r2c_post(const cplx<T>* lds, int base, int k, const cplx<T>* __restrict__ half) {
    const cplx<T> zk = lds[pad(base + (k == L ? 0 : k))];
    const cplx<T> zc = cconj(lds[pad(base + (k == 0 ? 0 : L - k))]);
    const cplx<T> e = {(zk.x + zc.x) * T(0.5), (zk.y + zc.y) * T(0.5)};
    const cplx<T> d = csub(zk, zc);
    const cplx<T> o = {d.y * T(0.5), -d.x * T(0.5)};
    return cadd(e, cmul(half[k], o));
}

// ---- rows forward: blockIdx.y = channel, fft_batch rows per workgroup
template <class T, int L>
__device__ __forceinline__ void rows_fwd_block(const T* __restrict__ W, int N, cplx<T>* __restrict__ out, const TwSet<T>& tws,
                                               int bx, int by) {
    using Pl = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), Mh = L + 1, R0 = Pl::R[0];
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, b = threadIdx.x / TH;
    const int c = by, gx = bx * B + b;
    const bool live = gx < N, even = (N & 1) == 0;
    const T* src = W + ((size_t)c * N + (live ? gx : 0)) * N;
    auto ld = [&](int e, int, int r) -> cplx<T> {
        cplx<T> z = czero<T>();
        if (2 * r >= R0) return z;   // e >= L/2: samples 2e >= L >= N are padding
        const int n0 = 2 * e;
        if (live && n0 < N) {
            if (even) { const V2 p = reinterpret_cast<const V2*>(src)[e]; z.x = p.x; z.y = p.y; }
            else { z.x = src[n0]; if (n0 + 1 < N) z.y = src[n0 + 1]; }
        }
        return z;
    };
    auto none = [](int, cplx<T>, int, int) {};
    fft_run<T, Pl, -1, TH, TO_LDS>(t, lds, b * L, tws.fwd, ld, none);
    __syncthreads();
    if (!live) return;
    cplx<T>* dst = out + ((size_t)c * N + gx) * Mh;
    for (int k = t; k < Mh; k += TH) dst[k] = r2c_post<T, L>(lds, b * L, k, tws.half);
}
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void k_rows_fwd(const T* __restrict__ W, int N, cplx<T>* __restrict__ out,
                                                             TwSet<T> tws) {
    rows_fwd_block<T, L>(W, N, out, tws, blockIdx.x, blockIdx.y);
}
// over a work list: virtual block v of instance k is the single launch's block (row group lb % nbx, channel
// lb / nbx) of that instance's live grid
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void kw_rows_fwd(const BInst<T>* __restrict__ I, const int* __restrict__ wl,
                                                              TwSet<T> tws) {
    constexpr int B = fft_batch<L>();
    const int total = wl[0];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int k = wl_find(wl, v);
        const BInst<T>& b = I[wl_inst(wl, k)];
        const int N = b.g.N, nbx = (N + B - 1) / B, lb = v - wl_start(wl, k);
        const T* W = b.W;
        cplx<T>* Tr = b.Tr;
        rows_fwd_block<T, L>(W, N, Tr, tws, lb % nbx, lb / nbx);
        __syncthreads();   // the LDS is taken by the next virtual block
    }
}

// ---- batched transpose of complex matrices, src (B, R, C) -> dst (B, C, R), 32 x 32 tiles through LDS
template <class T>
__device__ __forceinline__ void transpose_block(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int R, int C,
                                                int bx, int by, int bz) {
    __shared__ cplx<T> tile[32][33];
    const size_t off = (size_t)bz * R * C;
    const int c0 = bx * 32, r0 = by * 32;
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
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int R, int C) {
    transpose_block<T>(src, dst, R, C, blockIdx.x, blockIdx.y, blockIdx.z);
}
// the spectra's transpose (stage 1 -> 2, spec_tmp) of the clipped instances at size Ms whose tables are fresh
template <class T>
__global__ __launch_bounds__(256) void kb_transpose_spec(const BInst<T>* __restrict__ I, BMap m, int planes, int Ms) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active || !b.fresh || b.M != Ms) return;
    const int Mh = Ms / 2 + 1, nx = (Mh + 31) / 32, lb = blockIdx.x - m.start[k];
    (void)planes;
    transpose_block<T>(b.spec_tmp, b.spec_tmp + 2 * (size_t)Mh * Mh, Mh, Mh, lb % nx, (lb / nx) % nx, lb / (nx * nx));
}
// the grid passes' transposes over a work list (which 0: Tr -> Wh, 1: Ph -> Tr), tiles numbered over the live
// matrices; a transpose moves values only
template <class T>
__global__ __launch_bounds__(256) void kw_transpose(const BInst<T>* __restrict__ I, const int* __restrict__ wl, int which) {
    const int total = wl[0];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int k = wl_find(wl, v);
        const BInst<T>& b = I[wl_inst(wl, k)];
        const int Mh = b.M / 2 + 1, N = b.g.N;
        const cplx<T>* src = which == 0 ? b.Tr : b.Ph;
        cplx<T>* dst = which == 0 ? b.Wh : b.Tr;
        const int R = which == 0 ? N : Mh, C = which == 0 ? Mh : N;
        const int nx = (C + 31) / 32, ny = (R + 31) / 32, lb = v - wl_start(wl, k);
        transpose_block<T>(src, dst, R, C, lb % nx, (lb / nx) % ny, lb / (nx * ny));
        __syncthreads();   // the tile is taken by the next virtual block
    }
}

// kernel table value at column frequency kx (stored for kx <= M/2; `odd` flips the sign above M/2)
template <class T, int M> __device__ __forceinline__ T ktab(const T* __restrict__ row, int kx, bool odd) {
    if (kx <= M / 2) return row[kx];
    const T v = row[M - kx];
    return odd ? -v : v;
}

// ---- columns: one column ky per fft_th threads. Channel 0 is transformed once and its spectrum kept in
// registers (slot (q, r) of thread t holds frequency t + q TH + r M / 16); each output is a product formed
// in those registers and transformed back by the reversed plan, with no LDS round trip in between.
// STR: the input and output stay in the row passes' layout, (3, N, Mh) in and (4, N, Mh) out, read and written with
// stride Mh (for grids small enough to sit in L2); no transposes then, and the same values. Otherwise column ky of
// channel c starts at (c Mh + ky) LD (LD >= N: the folded row passes pad it to whole 32-byte segments)
template <class T, int L, bool CEN, bool STR = false>
__device__ __forceinline__ void columns_block(const cplx<T>* __restrict__ Wh, int N,
                                              const T* __restrict__ K1, const T* __restrict__ K2,
                                              const T* __restrict__ TX, const T* __restrict__ TY,
                                              cplx<T>* __restrict__ Ph, const TwSet<T>& tws, int bx, int LD) {
    using F = typename FwdPlan<T, L>::type;
    using I = Rev<F>;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), Mh = L / 2 + 1, R0 = F::R[0], RL = F::R[F::P - 1];
    constexpr int QL = L / RL / TH;
    static_assert((L / RL) % TH == 0, "the register link needs the last pass to cover every thread evenly");
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, b = threadIdx.x / TH;
    const int ky = bx * B + b;
    const bool live = ky < Mh;
    const int kyc = live ? ky : 0;
    cplx<T> s0[QL][RL], v[QL][RL];
    // FP64: each transform recomputes its thread's LDS addresses and reloads its twiddles (an index and an offset the
    // compiler cannot see through); kept across the seven transforms, they would hold the kernel above the registers
    // of three waves per SIMD (M = 6144: 518 against 585 us). FP32 keeps them (its kernel measured slower without)
    auto opaque = [&](int& tq, int& z) {
        if constexpr (sizeof(T) == 8) asm volatile("" : "+v"(tq), "+v"(z));
    };
    auto fwd = [&](int ch, cplx<T> (&dst)[QL][RL]) {
        int tq = t, z = 0;
        opaque(tq, z);
        const cplx<T>* twf = tws.fwd + z;
        const cplx<T>* col = STR ? Wh + (size_t)ch * N * Mh + kyc : Wh + ((size_t)ch * Mh + kyc) * LD;
        auto ld = [&](int e, int, int r) -> cplx<T> {
            if (2 * r >= R0 || e >= N) return czero<T>();   // rows >= N (all of r >= R0/2) are padding
            return STR ? col[(size_t)e * Mh] : col[e];
        };
        auto st = [&](int, cplx<T> x, int q, int r) { dst[q][r] = x; };
        fft_run<T, F, -1, TH, FN>(tq, lds, b * L, twf, ld, st);
    };
    auto inv = [&](int o, cplx<T> (&src)[QL][RL]) {
        int tq = t, z = 0;
        opaque(tq, z);
        const cplx<T>* twr = tws.rev + z;
        cplx<T>* dst = STR ? Ph + (size_t)o * N * Mh + kyc : Ph + ((size_t)o * Mh + kyc) * LD;
        auto ld = [&](int, int q, int r) { return src[q][r]; };
        auto st = [&](int e, cplx<T> x, int, int r) {
            if (2 * r < R0 && e < N && live) { if (STR) dst[(size_t)e * Mh] = x; else dst[e] = x; }   // only the N live rows
        };
        fft_run<T, I, +1, TH, FN>(tq, lds, b * L, twr, ld, st);
    };
    const T* k1 = K1 + (size_t)kyc * Mh;
    const T* k2 = K2 + (size_t)kyc * Mh;
    auto freq = [&](int q, int r) { return t + q * TH + r * (L / RL); };
    fwd(0, s0);
#pragma unroll
    for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) { const T c = ktab<T, L>(k1, freq(q, r), false); v[q][r] = {s0[q][r].x * c, s0[q][r].y * c}; }
    inv(0, v);
#pragma unroll
    for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) { const T c = ktab<T, L>(k2, freq(q, r), false); v[q][r] = {s0[q][r].x * c, s0[q][r].y * c}; }
    inv(1, v);
    for (int ch = 1; ch <= 2; ++ch) {
        fwd(ch, v);
        const T* tt = CEN ? (ch == 1 ? TX : TY) + (size_t)kyc * Mh : nullptr;
#pragma unroll
        for (int q = 0; q < QL; ++q)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int k = freq(q, r);
                const T c2 = ktab<T, L>(k2, k, false);
                const cplx<T> s = v[q][r];
                if (CEN) {   // i t s0 - K2 s, t odd in kx for x K2
                    const T tk = ktab<T, L>(tt, k, ch == 1);
                    v[q][r] = {-tk * s0[q][r].y - c2 * s.x, tk * s0[q][r].x - c2 * s.y};
                } else {
                    v[q][r] = {s.x * c2, s.y * c2};
                }
            }
        inv(ch + 1, v);
    }
}
template <class T, int L, bool CEN, bool STR>
__global__ __launch_bounds__(fft_block<L>()) void k_columns(const cplx<T>* __restrict__ Wh, int N,
                                                            const T* __restrict__ K1, const T* __restrict__ K2,
                                                            const T* __restrict__ TX, const T* __restrict__ TY,
                                                            cplx<T>* __restrict__ Ph, TwSet<T> tws, int LD) {
    columns_block<T, L, CEN, STR>(Wh, N, K1, K2, TX, TY, Ph, tws, blockIdx.x, LD);
}
template <class T, int L, bool CEN>
__global__ __launch_bounds__(fft_block<L>()) void kw_columns(const BInst<T>* __restrict__ I, const int* __restrict__ wl,
                                                             TwSet<T> tws) {
    const int total = wl[0];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int k = wl_find(wl, v);
        const BInst<T>& b = I[wl_inst(wl, k)];
        const cplx<T>* Wh = b.Wh;
        cplx<T>* Ph = b.Ph;
        const T *K0 = b.K[0], *K1 = b.K[1], *K2 = b.K[2], *K3 = b.K[3];
        columns_block<T, L, CEN>(Wh, b.g.N, K0, K1, K2, K3, Ph, tws, v - wl_start(wl, k), b.g.N);
        __syncthreads();
    }
}

// ---- rows inverse: blockIdx.y = channel, fft_batch rows per workgroup; planar output (4, N, N)
// Z_e = E + i O with E = X_e + conj X_{L-e}, O = (X_e - conj X_{L-e}) exp(+2 pi i e / 2L)
template <class T, int L>
__device__ __forceinline__ void rows_inv_block(const cplx<T>* __restrict__ X, int N, T* __restrict__ phi, T scale,
                                               const TwSet<T>& tws, int bx, int by) {
    using Pl = Rev<typename FwdPlan<T, L>::type>;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), Mh = L + 1, RL = Pl::R[Pl::P - 1];
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, b = threadIdx.x / TH;
    const int c = by, gx = bx * B + b;
    const bool live = gx < N, even = (N & 1) == 0;
    const cplx<T>* src = X + ((size_t)c * N + (live ? gx : 0)) * Mh;
    auto ld = [&](int e, int, int) -> cplx<T> {
        const cplx<T> xk = src[e], xc = cconj(src[L - e]);
        const cplx<T> ev = cadd(xk, xc);
        const cplx<T> od = cmul(csub(xk, xc), cconj(tws.half[e]));
        return {ev.x - od.y, ev.y + od.x};
    };
    T* dst = phi + ((size_t)c * N + (live ? gx : 0)) * N;
    auto st = [&](int e, cplx<T> v, int, int r) {
        if (2 * r >= RL) return;   // e >= L/2: samples 2e >= L >= N are padding
        const int n0 = 2 * e;
        if (!live || n0 >= N) return;
        if (even) { V2 p; p.x = v.x * scale; p.y = v.y * scale; reinterpret_cast<V2*>(dst)[e] = p; }
        else { dst[n0] = v.x * scale; if (n0 + 1 < N) dst[n0 + 1] = v.y * scale; }
    };
    fft_run<T, Pl, +1, TH, FN>(t, lds, b * L, tws.rev, ld, st);
}
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void k_rows_inv(const cplx<T>* __restrict__ X, int N, T* __restrict__ phi,
                                                             T scale, TwSet<T> tws) {
    rows_inv_block<T, L>(X, N, phi, scale, tws, blockIdx.x, blockIdx.y);
}
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void kw_rows_inv(const BInst<T>* __restrict__ I, const int* __restrict__ wl, T scale,
                                                              TwSet<T> tws) {
    constexpr int B = fft_batch<L>();
    const int total = wl[0];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int k = wl_find(wl, v);
        const BInst<T>& b = I[wl_inst(wl, k)];
        const int N = b.g.N, nbx = (N + B - 1) / B, lb = v - wl_start(wl, k);
        const cplx<T>* Tr = b.Tr;
        T* phi = b.phi;
        rows_inv_block<T, L>(Tr, N, phi, scale, tws, lb % nbx, lb / nbx);
        __syncthreads();
    }
}

// ---- the row passes with the transposes folded in (grids above the strided column pass). A block takes RB
// consecutive rows of one channel, its RB transforms in LDS at element offsets r S, and writes or reads the column
// layout (channel, ky, row) in segments of RB rows (32 bytes for RB x 8 and RB x 16 bytes). S = L + 16 m makes the
// padded starts of consecutive buffers differ by an odd multiple of 8 elements modulo 32 (the pad adds one element
// per 16, so buffer r starts at r (17 L / 16 + 17 m)): the 8 columns x 4 rows of FP32 elements (or 8 x 2 of FP64)
// that one half-wave stores or loads then fall on distinct banks. Blocks are numbered so that each of the 8 XCDs (the
// dispatcher hands consecutive blocks to them in turn) takes a contiguous run of row groups: the blocks writing or
// reading the other segments of a 128-byte line then run in the same L2 at about the same time. The column layout's
// stride LD rounds N up to 4 rows, so every segment is 32-byte aligned. The arithmetic is the row kernels'.
constexpr int kXcds = 8;
inline int xcd_grid(int total) { return (total + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_block(int total) {   // the block's index in row-group order; -1 past the end
    const int per = (int)gridDim.x / kXcds;              // gridDim.x = xcd_grid(total)
    const int v = ((int)blockIdx.x % kXcds) * per + (int)blockIdx.x / kXcds;
    return v < total ? v : -1;
}
template <class T, int L, int RB> constexpr int fold_block() { return RB * fft_th<L>(); }
// the smaller m >= 1 with 17 (L / 16 + m) = 8 or 24 (mod 32); 17 is its own inverse modulo 32. m >= 1 leaves the
// inverse pass the slot after each row's L elements for X_L (with m = 0 it would be the next row's X_0)
constexpr int fold_m(int L, int target) {
    const int m = (((target - (17 * L / 16) % 32) % 32 + 32) % 32) * 17 % 32;
    return m == 0 ? 32 : m;
}
template <int L> constexpr int fold_stride() { return L + 16 * (fold_m(L, 8) < fold_m(L, 24) ? fold_m(L, 8) : fold_m(L, 24)); }
template <int L> constexpr bool fold_stride_ok() {
    return (17 * fold_stride<L>() / 16) % 16 == 8 && fold_stride<L>() >= L + 16;
}
static_assert(fold_stride_ok<256>() && fold_stride_ok<384>() && fold_stride_ok<512>() && fold_stride_ok<768>() &&
              fold_stride_ok<1024>() && fold_stride_ok<1536>() && fold_stride_ok<2048>() && fold_stride_ok<3072>() &&
              fold_stride_ok<4096>(), "row buffers an odd multiple of 8 elements apart modulo 32, each with room for X_L");
template <class T, int L, int RB> constexpr int fold_lds_bytes() { return padded(RB * fold_stride<L>()) * (int)sizeof(cplx<T>); }

// W (3, N, N) -> Wh (3, Mh, LD): rows_fwd_block's transform per row, then the Mh outputs of the RB rows by segments
template <class T, int L, int RB>
__global__ __launch_bounds__((fold_block<T, L, RB>())) void k_rows_fwd_fold(const T* __restrict__ W, int N, cplx<T>* __restrict__ Wh,
                                                                         int LD, TwSet<T> tws) {
    using Pl = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), Mh = L + 1, R0 = Pl::R[0], S = fold_stride<L>();
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int G = (N + RB - 1) / RB;
    const int v = xcd_block(3 * G);
    if (v < 0) return;
    const int c = v / G, row0 = (v % G) * RB;
    const int t = threadIdx.x % TH, b = threadIdx.x / TH;
    const int gx = row0 + b;
    const bool live = gx < N, even = (N & 1) == 0;
    const T* src = W + ((size_t)c * N + (live ? gx : 0)) * N;
    auto ld = [&](int e, int, int r) -> cplx<T> {
        cplx<T> z = czero<T>();
        if (2 * r >= R0) return z;   // e >= L/2: samples 2e >= L >= N are padding
        const int n0 = 2 * e;
        if (live && n0 < N) {
            if (even) { const V2 p = reinterpret_cast<const V2*>(src)[e]; z.x = p.x; z.y = p.y; }
            else { z.x = src[n0]; if (n0 + 1 < N) z.y = src[n0 + 1]; }
        }
        return z;
    };
    auto none = [](int, cplx<T>, int, int) {};
    fft_run<T, Pl, -1, TH, TO_LDS>(t, lds, b * S, tws.fwd, ld, none);
    __syncthreads();
    cplx<T>* dst = Wh + (size_t)c * Mh * LD + row0;
    constexpr int NB = fold_block<T, L, RB>(), IT = (Mh * RB + NB - 1) / NB;
#pragma unroll
    for (int u = 0; u < IT; ++u) {   // unrolled: the twiddle loads of every step are in flight together
        const int i = threadIdx.x + u * NB, r = i % RB, k = i / RB;
        if (i < Mh * RB && row0 + r < N) dst[(size_t)k * LD + r] = r2c_post<T, L>(lds, r * S, k, tws.half);
    }
}

// Ph (4, Mh, LD) -> phi (4, N, N): the Mh inputs of the RB rows by segments into LDS, then rows_inv_block's transform
// per row with its inputs from there
template <class T, int L, int RB>
__global__ __launch_bounds__((fold_block<T, L, RB>())) void k_rows_inv_fold(const cplx<T>* __restrict__ Ph, int N, int LD,
                                                                         T* __restrict__ phi, T scale, TwSet<T> tws) {
    using Pl = Rev<typename FwdPlan<T, L>::type>;
    constexpr int TH = fft_th<L>(), Mh = L + 1, RL = Pl::R[Pl::P - 1], S = fold_stride<L>();
    static_assert(Pl::P > 1, "the first pass must read all its inputs before it writes the buffer");
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int G = (N + RB - 1) / RB;
    const int v = xcd_block(4 * G);
    if (v < 0) return;
    const int c = v / G, row0 = (v % G) * RB;
    const cplx<T>* src = Ph + (size_t)c * Mh * LD + row0;
    constexpr int NB = fold_block<T, L, RB>(), IT = (Mh * RB + NB - 1) / NB;
    {
        cplx<T> x[IT];   // every segment load of the thread issued before the first LDS store
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            const int i = threadIdx.x + u * NB, r = i % RB, k = i / RB;
            x[u] = i < Mh * RB && row0 + r < N ? src[(size_t)k * LD + r] : czero<T>();
        }
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            const int i = threadIdx.x + u * NB, r = i % RB, k = i / RB;
            if (i < Mh * RB) lds[pad(r * S + k)] = x[u];
        }
    }
    __syncthreads();
    const int t = threadIdx.x % TH, b = threadIdx.x / TH;
    const int gx = row0 + b, base = b * S;
    const bool live = gx < N, even = (N & 1) == 0;
    auto ld = [&](int e, int, int) -> cplx<T> {
        const cplx<T> xk = lds[pad(base + e)], xc = cconj(lds[pad(base + L - e)]);
        const cplx<T> ev = cadd(xk, xc);
        const cplx<T> od = cmul(csub(xk, xc), cconj(tws.half[e]));
        return {ev.x - od.y, ev.y + od.x};
    };
    T* dst = phi + ((size_t)c * N + (live ? gx : 0)) * N;
    auto st = [&](int e, cplx<T> v, int, int r) {
        if (2 * r >= RL) return;   // e >= L/2: samples 2e >= L >= N are padding
        const int n0 = 2 * e;
        if (!live || n0 >= N) return;
        if (even) { V2 p; p.x = v.x * scale; p.y = v.y * scale; reinterpret_cast<V2*>(dst)[e] = p; }
        else { dst[n0] = v.x * scale; if (n0 + 1 < N) dst[n0 + 1] = v.y * scale; }
    };
    fft_run<T, Pl, +1, TH, FN>(t, lds, base, tws.rev, ld, st);
}

// The inverse pass, persistent: one block per CU walks its XCD's run of row groups (the same order as above). Its
// twiddle tables sit in LDS, so the transforms issue no global load, and the next group's segment loads, held in
// registers, are in flight while the current group's transforms run (loads complete in order, so a global twiddle
// load inside the transforms would wait for them). For the sizes whose buffers and tables fit in LDS.
template <class T, int L, int RB> constexpr int fold_p_lds_bytes() {
    return (padded(RB * fold_stride<L>()) + plan_tw_size<Rev<typename FwdPlan<T, L>::type>>() + L + 1) * (int)sizeof(cplx<T>);
}
template <class T, int L, int RB> constexpr bool fold_p_fits() { return fold_p_lds_bytes<T, L, RB>() <= 160 * 1024; }
// The persistent forms pay where the per-group form holds one block per CU (its LDS above half of the CU's): M = 6144,
// 4 FP32 rows per block, 0.84 against 0.86 ms per convolution; where two blocks fit (M <= 4096) the per-group form
// hides more latency (M = 2048: 0.227 against 0.235 ms, M = 4096: 0.401 against 0.416 ms).
template <class T, int L, int RB> constexpr bool fold_persist_pays() { return 2 * fold_lds_bytes<T, L, RB>() > 160 * 1024; }

template <class T, int L, int RB>
__global__ __launch_bounds__((fold_block<T, L, RB>())) void k_rows_inv_fold_p(const cplx<T>* __restrict__ Ph, int N, int LD,
                                                                           T* __restrict__ phi, T scale, TwSet<T> tws) {
    using Pl = Rev<typename FwdPlan<T, L>::type>;
    constexpr int TH = fft_th<L>(), Mh = L + 1, RL = Pl::R[Pl::P - 1], S = fold_stride<L>();
    constexpr int NB = fold_block<T, L, RB>(), IT = (Mh * RB + NB - 1) / NB, NTW = plan_tw_size<Pl>();
    static_assert(Pl::P > 1, "the first pass must read all its inputs before it writes the buffer");
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* tw_l = lds + padded(RB * S);
    cplx<T>* half_l = tw_l + NTW;
    const int G = (N + RB - 1) / RB, total = 4 * G;
    const int per = (int)gridDim.x / kXcds, x = (int)blockIdx.x % kXcds, GX = (total + kXcds - 1) / kXcds;
    int jj = (int)blockIdx.x / kXcds;
    auto valid = [&](int j) { return j < GX && x * GX + j < total; };
    if (!valid(jj)) return;   // uniform over the block
    for (int i = threadIdx.x; i < NTW; i += NB) tw_l[i] = tws.rev[i];
    for (int i = threadIdx.x; i <= L; i += NB) half_l[i] = tws.half[i];
    cplx<T> xr[IT];
    auto load = [&](int v) {
        const int c = v / G, row0 = (v % G) * RB;
        const cplx<T>* src = Ph + (size_t)c * Mh * LD + row0;
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));   // per call: the slot addresses are not kept live across the loop
#pragma unroll
        for (int u = 0; u < IT; ++u) {
            const int i = tid + u * NB, r = i % RB, k = i / RB;
            xr[u] = i < Mh * RB && row0 + r < N ? src[(size_t)k * LD + r] : czero<T>();
        }
    };
    load(x * GX + jj);
    for (;;) {
        const int v = x * GX + jj;
        const int c = v / G, row0 = (v % G) * RB;
        {
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
#pragma unroll
            for (int u = 0; u < IT; ++u) {
                const int i = tid + u * NB, r = i % RB, k = i / RB;
                if (i < Mh * RB) lds[pad(r * S + k)] = xr[u];
            }
        }
        __syncthreads();   // also orders the twiddle copies before their first use
        const int jn = jj + per;
        const bool more = valid(jn);
        if (more) load(x * GX + jn);
        {
            // an offset the compiler cannot see through, renewed per group: the twiddle reads stay inside the loop
            // instead of being hoisted out of it and held in registers across every group
            int opaque = 0, tid = threadIdx.x;
            asm volatile("" : "+v"(opaque), "+v"(tid));
            const cplx<T>* twg = tw_l + opaque;
            const cplx<T>* halfg = half_l + opaque;
            const int t = tid % TH, b = tid / TH;
            const int gx = row0 + b, base = b * S;
            const bool live = gx < N, even = (N & 1) == 0;
            auto ld = [&](int e, int, int) -> cplx<T> {
                const cplx<T> xk = lds[pad(base + e)], xc = cconj(lds[pad(base + L - e)]);
                const cplx<T> ev = cadd(xk, xc);
                const cplx<T> od = cmul(csub(xk, xc), cconj(halfg[e]));
                return {ev.x - od.y, ev.y + od.x};
            };
            T* dst = phi + ((size_t)c * N + (live ? gx : 0)) * N;
            auto st = [&](int e, cplx<T> w, int, int r) {
                if (2 * r >= RL) return;   // e >= L/2: samples 2e >= L >= N are padding
                const int n0 = 2 * e;
                if (!live || n0 >= N) return;
                if (even) { V2 p; p.x = w.x * scale; p.y = w.y * scale; reinterpret_cast<V2*>(dst)[e] = p; }
                else { dst[n0] = w.x * scale; if (n0 + 1 < N) dst[n0 + 1] = w.y * scale; }
            };
            fft_run<T, Pl, +1, TH, FN>(t, lds, base, twg, ld, st);
        }
        __syncthreads();   // the buffers are staged again
        if (!more) break;
        jj = jn;
    }
}

// The forward pass, persistent in the same way: one block per CU walks its XCD's run of row groups with the twiddle
// tables in LDS; the next group's row samples (each thread's live inputs of the first radix pass, in registers) are
// loaded while the current group is transformed and its segments written.
template <class T, int L, int RB> constexpr int fold_pf_lds_bytes() {
    return (padded(RB * fold_stride<L>()) + plan_tw_size<typename FwdPlan<T, L>::type>() + L + 1) * (int)sizeof(cplx<T>);
}
template <class T, int L, int RB> constexpr bool fold_pf_fits() { return fold_pf_lds_bytes<T, L, RB>() <= 160 * 1024; }

template <class T, int L, int RB>
__global__ __launch_bounds__((fold_block<T, L, RB>())) void k_rows_fwd_fold_p(const T* __restrict__ W, int N, cplx<T>* __restrict__ Wh,
                                                                           int LD, TwSet<T> tws) {
    using Pl = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), Mh = L + 1, R0 = Pl::R[0], NI0 = L / R0, S = fold_stride<L>();
    constexpr int NB = fold_block<T, L, RB>(), IT = (Mh * RB + NB - 1) / NB, NTW = plan_tw_size<Pl>();
    static_assert(Pl::P > 1 && NI0 <= TH, "one first-pass item per thread, and the first pass has no twiddles");
    typedef typename vec2<T>::type V2;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    cplx<T>* tw_l = lds + padded(RB * S);
    cplx<T>* half_l = tw_l + NTW;
    const int G = (N + RB - 1) / RB, total = 3 * G;
    const int per = (int)gridDim.x / kXcds, x = (int)blockIdx.x % kXcds, GX = (total + kXcds - 1) / kXcds;
    int jj = (int)blockIdx.x / kXcds;
    auto valid = [&](int j) { return j < GX && x * GX + j < total; };
    if (!valid(jj)) return;   // uniform over the block
    for (int i = threadIdx.x; i < NTW; i += NB) tw_l[i] = tws.fwd[i];
    for (int i = threadIdx.x; i <= L; i += NB) half_l[i] = tws.half[i];
    const bool even = (N & 1) == 0;
    cplx<T> xr[R0 / 2];   // item t's inputs r < R0 / 2 (element t + r L / R0; the others are padding) of the next group
    auto load = [&](int v) {
        const int c = v / G, row0 = (v % G) * RB;
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));   // per call: the addresses are not kept live across the loop
        const int t = tid % TH, gx = row0 + tid / TH;
        const bool live = gx < N && t < NI0;
        const T* src = W + ((size_t)c * N + (gx < N ? gx : 0)) * N;
#pragma unroll
        for (int r = 0; r < R0 / 2; ++r) {
            const int e = t + r * NI0, n0 = 2 * e;
            cplx<T> z = czero<T>();
            if (live && n0 < N) {
                if (even) { const V2 p = reinterpret_cast<const V2*>(src)[e]; z.x = p.x; z.y = p.y; }
                else { z.x = src[n0]; if (n0 + 1 < N) z.y = src[n0 + 1]; }
            }
            xr[r] = z;
        }
    };
    load(x * GX + jj);
    for (;;) {
        const int v = x * GX + jj;
        const int c = v / G, row0 = (v % G) * RB;
        cplx<T> xc[R0 / 2];
#pragma unroll
        for (int r = 0; r < R0 / 2; ++r) xc[r] = xr[r];
        const int jn = jj + per;
        const bool more = valid(jn);
        if (more) load(x * GX + jn);
        {
            // an offset the compiler cannot see through, renewed per group: the twiddle reads stay inside the loop
            int opaque = 0, tid = threadIdx.x;
            asm volatile("" : "+v"(opaque), "+v"(tid));
            const int t = tid % TH, b = tid / TH;
            auto ld = [&](int, int, int r) -> cplx<T> { return 2 * r < R0 ? xc[r] : czero<T>(); };
            auto none = [](int, cplx<T>, int, int) {};
            fft_run<T, Pl, -1, TH, TO_LDS>(t, lds, b * S, tw_l + opaque, ld, none);
        }
        __syncthreads();   // the RB transforms are in LDS (the first group's barriers also order the table copies)
        {
            int opaque = 0, tid = threadIdx.x;
            asm volatile("" : "+v"(opaque), "+v"(tid));
            const cplx<T>* halfg = half_l + opaque;
            cplx<T>* dst = Wh + (size_t)c * Mh * LD + row0;
#pragma unroll
            for (int u = 0; u < IT; ++u) {
                const int i = tid + u * NB, r = i % RB, k = i / RB;
                if (i < Mh * RB && row0 + r < N) dst[(size_t)k * LD + r] = r2c_post<T, L>(lds, r * S, k, halfg);
            }
        }
        __syncthreads();   // the buffers are written again by the next group's transforms
        if (!more) break;
        jj = jn;
    }
}

// ---- kernel spectra. Offsets (a, b) on the M x M torus, a along x (gx), b along y (gy); every table is even
// in each offset except x K2 (odd in a) and y K2 (odd in b). Values in double, rounded to float first when
// round_kernel (the reference builds its tables in float32), transformed in T.
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

// Stage 1, per row a <= M/2, along b. FFT(K1_a + i K2_a): both rows are real and even, so the real and
// imaginary parts of the result are their (real) spectra T1, T2. FFT(yK2_a) = i S (odd row). Output
// P0[a][ky] = (T1, T2), P1[a][ky] = (S, x_a T2) -- the row transform of x K2 is x_a times that of K2.
template <class T, int L, int NT>
__device__ __forceinline__ void spec1_block(double delta, double dof, const double* __restrict__ table, int round_kernel,
                                            cplx<T>* __restrict__ P0, cplx<T>* __restrict__ P1, const TwSet<T>& tws, int bx) {
    using F = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), M = L, Mh = L / 2 + 1, RL = F::R[F::P - 1];
    constexpr int QL = L / RL / TH;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, bb = threadIdx.x / TH;
    const int a = bx * B + bb;
    const bool live = a < Mh;
    const int ac = live ? a : 0;
    cplx<T> za[QL][RL], zb[QL][RL];
    auto put = [](cplx<T> (&d)[QL][RL]) { return [&d](int, cplx<T> x, int q, int r) { d[q][r] = x; }; };
    // the first pass reads each input slot (q, r) once, at the same element in both transforms: the y-odd
    // kernel values of the second transform are kept from the first one's evaluation
    constexpr int R0 = F::R[0], Q0 = (L / R0 + TH - 1) / TH;
    T ky_slot[Q0][R0];
    {
        auto ld = [&](int e, int q, int r) -> cplx<T> {
            double k1, k2, ky;
            kernel_values(ac, e, M, delta, dof, table, round_kernel, k1, k2, ky);
            if constexpr (NT == 4) ky_slot[q][r] = (T)ky;
            return {(T)k1, (T)k2};
        };
        fft_run<T, F, -1, TH, FN>(t, lds, bb * L, tws.fwd, ld, put(za));
    }
    if constexpr (NT == 4) {
        auto ld = [&](int, int q, int r) -> cplx<T> { return {ky_slot[q][r], T(0)}; };
        fft_run<T, F, -1, TH, FN>(t, lds, bb * L, tws.fwd, ld, put(zb));
    }
    if (!live) return;
    const T xa = a == M / 2 ? T(0) : (T)(a * delta);
#pragma unroll
    for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const int k = t + q * TH + r * (L / RL);
            if (k >= Mh) continue;
            P0[(size_t)a * Mh + k] = za[q][r];
            if constexpr (NT == 4) P1[(size_t)a * Mh + k] = {zb[q][r].y, xa * za[q][r].y};
        }
}
template <class T, int L, int NT>
__global__ __launch_bounds__(fft_block<L>()) void k_spec1(double delta, double dof, const double* __restrict__ table,
                                                          int round_kernel, cplx<T>* __restrict__ P0, cplx<T>* __restrict__ P1,
                                                          TwSet<T> tws) {
    spec1_block<T, L, NT>(delta, dof, table, round_kernel, P0, P1, tws, blockIdx.x);
}
template <class T, int L, int NT>
__global__ __launch_bounds__(fft_block<L>()) void kb_spec1(const BInst<T>* __restrict__ I, BMap m, double dof,
                                                           const double* __restrict__ table, int round_kernel, TwSet<T> tws) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active || !b.fresh || b.M != L) return;
    const size_t plane = (size_t)(L / 2 + 1) * (L / 2 + 1);
    spec1_block<T, L, NT>(b.delta, dof, table, round_kernel, b.spec_tmp, b.spec_tmp + plane, tws, blockIdx.x - m.start[k]);
}

// Stage 2, per ky, along a (inputs transposed to [ky][a], a <= M/2, mirrored: even, and odd for x K2).
// FFT(T1 + i T2) gives K1 and K2 as real and imaginary parts; FFT(S + i X) = S^ - U with S^ even and U odd
// in kx (FFT(X) = i U), split through Z(kx) and Z(M - kx): Ty = S^, Tx = U.
template <class T, int L, int NT>
__device__ __forceinline__ void spec2_block(const cplx<T>* __restrict__ Q0, const cplx<T>* __restrict__ Q1,
                                            T* __restrict__ K1, T* __restrict__ K2, T* __restrict__ TX,
                                            T* __restrict__ TY, const TwSet<T>& tws, int bx) {
    using F = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), M = L, Mh = L / 2 + 1, RL = F::R[F::P - 1];
    constexpr int QL = L / RL / TH;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, bb = threadIdx.x / TH;
    const int ky = bx * B + bb;
    const bool live = ky < Mh;
    const size_t row = (size_t)(live ? ky : 0) * Mh;
    cplx<T> z[QL][RL];
    {
        auto ld = [&](int e, int, int) -> cplx<T> { return Q0[row + (e <= M / 2 ? e : M - e)]; };
        auto st = [&](int, cplx<T> x, int q, int r) { z[q][r] = x; };
        fft_run<T, F, -1, TH, FN>(t, lds, bb * L, tws.fwd, ld, st);
    }
    if (live) {
#pragma unroll
        for (int q = 0; q < QL; ++q)
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int k = t + q * TH + r * (L / RL);
                if (k >= Mh) continue;
                K1[row + k] = z[q][r].x;
                K2[row + k] = z[q][r].y;
            }
    }
    if constexpr (NT == 4) {
        auto ld = [&](int e, int, int) -> cplx<T> {
            if (e <= M / 2) return Q1[row + e];
            const cplx<T> p = Q1[row + (M - e)];
            return {p.x, -p.y};
        };
        auto none = [](int, cplx<T>, int, int) {};
        fft_run<T, F, -1, TH, TO_LDS>(t, lds, bb * L, tws.fwd, ld, none);
        __syncthreads();
        if (!live) return;
        for (int k = t; k < Mh; k += TH) {
            const T zk = lds[pad(bb * L + k)].x, zc = lds[pad(bb * L + (k == 0 ? 0 : M - k))].x;
            TY[row + k] = T(0.5) * (zk + zc);
            TX[row + k] = T(-0.5) * (zk - zc);
        }
    }
}
template <class T, int L, int NT>
__global__ __launch_bounds__(fft_block<L>()) void k_spec2(const cplx<T>* __restrict__ Q0, const cplx<T>* __restrict__ Q1,
                                                          T* __restrict__ K1, T* __restrict__ K2, T* __restrict__ TX,
                                                          T* __restrict__ TY, TwSet<T> tws) {
    spec2_block<T, L, NT>(Q0, Q1, K1, K2, TX, TY, tws, blockIdx.x);
}
template <class T, int L, int NT>
__global__ __launch_bounds__(fft_block<L>()) void kb_spec2(const BInst<T>* __restrict__ I, BMap m, TwSet<T> tws) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active || !b.fresh || b.M != L) return;
    const size_t plane = (size_t)(L / 2 + 1) * (L / 2 + 1);
    spec2_block<T, L, NT>(b.spec_tmp + 2 * plane, b.spec_tmp + 3 * plane, b.Kw[0], b.Kw[1], b.Kw[2], b.Kw[3], tws,
                          blockIdx.x - m.start[k]);
}

// Stage 2 of the centred set by pairs of columns, for kernels symmetric in x <-> y (the t-SNE kernel, radial tables):
// y K2 is then x K2 transposed, so only K1, K2 and Tx are transformed here and Ty is Tx transposed afterwards. Per pair
// (ky0 = 2p, ky1 = 2p + 1), FFT(T1 + i T2) of each column gives K1 and K2 as real and imaginary parts, as spec2_block
// does; one FFT(X0 + i X1) of the two columns' X = x_a T2 (odd in a, so FFT(X) = i U) gives Tx of both: U0 = Im Z,
// U1 = -Re Z. The two odd columns share a transform (like magnitudes); an even and an odd table never do.
template <class T, int L>
__device__ __forceinline__ void spec2p_block(const cplx<T>* __restrict__ Q0, double delta, T* __restrict__ K1,
                                             T* __restrict__ K2, T* __restrict__ TX, const TwSet<T>& tws, int bx) {
    using F = typename FwdPlan<T, L>::type;
    constexpr int TH = fft_th<L>(), B = fft_batch<L>(), M = L, Mh = L / 2 + 1, RL = F::R[F::P - 1];
    constexpr int QL = L / RL / TH;
    extern __shared__ __align__(16) unsigned char smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
    const int t = threadIdx.x % TH, bb = threadIdx.x / TH;
    const int pr = bx * B + bb, ky0 = 2 * pr, ky1 = 2 * pr + 1;
    const bool live0 = ky0 < Mh, live1 = ky1 < Mh;
    const size_t row0 = (size_t)(live0 ? ky0 : 0) * Mh, row1 = (size_t)(live1 ? ky1 : 0) * Mh;
    cplx<T> z[QL][RL];
    auto put = [&](int, cplx<T> x, int q, int r) { z[q][r] = x; };
    for (int c = 0; c < 2; ++c) {
        const size_t row = c == 0 ? row0 : row1;
        const bool live = c == 0 ? live0 : live1;
        auto ld = [&](int e, int, int) -> cplx<T> { return Q0[row + (e <= M / 2 ? e : M - e)]; };
        fft_run<T, F, -1, TH, FN>(t, lds, bb * L, tws.fwd, ld, put);
        if (live) {
#pragma unroll
            for (int q = 0; q < QL; ++q)
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int k = t + q * TH + r * (L / RL);
                    if (k >= Mh) continue;
                    K1[row + k] = z[q][r].x;
                    K2[row + k] = z[q][r].y;
                }
        }
    }
    auto ldx = [&](int e, int, int) -> cplx<T> {
        const int a = e <= M / 2 ? e : M - e;
        const T xa = a == M / 2 ? T(0) : (T)(a * delta);
        const T s = e <= M / 2 ? xa : -xa;
        return {s * Q0[row0 + a].y, s * Q0[row1 + a].y};
    };
    fft_run<T, F, -1, TH, FN>(t, lds, bb * L, tws.fwd, ldx, put);
#pragma unroll
    for (int q = 0; q < QL; ++q)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const int k = t + q * TH + r * (L / RL);
            if (k >= Mh) continue;
            if (live0) TX[row0 + k] = z[q][r].y;
            if (live1) TX[row1 + k] = -z[q][r].x;
        }
}
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void k_spec2p(const cplx<T>* __restrict__ Q0, double delta, T* __restrict__ K1,
                                                           T* __restrict__ K2, T* __restrict__ TX, TwSet<T> tws) {
    spec2p_block<T, L>(Q0, delta, K1, K2, TX, tws, blockIdx.x);
}
template <class T, int L>
__global__ __launch_bounds__(fft_block<L>()) void kb_spec2p(const BInst<T>* __restrict__ I, BMap m, TwSet<T> tws) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active || !b.fresh || b.M != L) return;
    const size_t plane = (size_t)(L / 2 + 1) * (L / 2 + 1);
    spec2p_block<T, L>(b.spec_tmp + 2 * plane, b.delta, b.Kw[0], b.Kw[1], b.Kw[2], tws, blockIdx.x - m.start[k]);
}

// real (R, C) -> (C, R), 32 x 32 tiles through LDS (Ty from Tx)
template <class T>
__device__ __forceinline__ void transpose_real_block(const T* __restrict__ src, T* __restrict__ dst, int R, int C, int bx, int by) {
    __shared__ T tile[32][33];
    const int c0 = bx * 32, r0 = by * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        if (r < R && c < C) tile[i][tx] = src[(size_t)r * C + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (r < R && c < C) dst[(size_t)c * R + r] = tile[tx][i];
    }
}
template <class T>
__global__ __launch_bounds__(256) void k_transpose_real(const T* __restrict__ src, T* __restrict__ dst, int R, int C) {
    transpose_real_block<T>(src, dst, R, C, blockIdx.x, blockIdx.y);
}
template <class T>
__global__ __launch_bounds__(256) void kb_transpose_real_spec(const BInst<T>* __restrict__ I, BMap m, int Ms) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active || !b.fresh || b.M != Ms) return;
    const int Mh = Ms / 2 + 1, nx = (Mh + 31) / 32, lb = blockIdx.x - m.start[k];
    transpose_real_block<T>(b.Kw[2], b.Kw[3], Mh, Mh, lb % nx, lb / nx);
}

// potentials planar (4, N, N) -> interleaved (N, N, 4)
template <class T>
__global__ void k_interleave4(const T* __restrict__ src, T* __restrict__ dst, i64 NN) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < NN; i += (i64)gridDim.x * blockDim.x)
        for (int c = 0; c < 4; ++c) dst[4 * i + c] = src[c * NN + i];
}

template <class F> void set_lds(F* f, int bytes) {
    if (bytes > 65536) kernel_lds_limit(reinterpret_cast<const void*>(f), bytes);
}

template <class T, int L> int lds_bytes() { return fft_batch<L>() * padded(L) * (int)sizeof(cplx<T>); }

}  // namespace

#define MG_ROW_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096)
#define MG_COL_SIZES(X) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096) X(6144) X(8192)

template <class T> void launch_rows_forward(const T* W, int N, int M, cplx<T>* out, hipStream_t st) {
    require(N <= M / 2, "FFT size below 2N");
#define MG_C(L_) case L_: { \
        const int by = lds_bytes<T, L_>(); set_lds(k_rows_fwd<T, L_>, by); \
        hipLaunchKernelGGL((k_rows_fwd<T, L_>), dim3((N + fft_batch<L_>() - 1) / fft_batch<L_>(), 3), dim3(fft_block<L_>()), by, st, W, N, out, fft_twiddle_set<T>(L_)); \
        break; }
    switch (M / 2) { MG_ROW_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
    MG_HIP(hipGetLastError());
}

template <class T> void launch_transpose(const cplx<T>* src, cplx<T>* dst, int batch, int R, int C, hipStream_t st) {
    hipLaunchKernelGGL(k_transpose<T>, dim3((C + 31) / 32, (R + 31) / 32, batch), dim3(256), 0, st, src, dst, R, C);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_columns(const cplx<T>* Wh, int N, int M, const T* const* K, bool centred, cplx<T>* Ph,
                                       hipStream_t st, bool strided, int LD) {
    require(N <= M / 2, "FFT size below 2N");
    if (LD <= 0) LD = N;
    require(LD >= N && LD <= M / 2, "column stride outside [N, M/2]");
    require(centred == (sizeof(T) == 4), "the FP32 path uses the centred product set, the FP64 path the reference's");
    constexpr bool CEN = sizeof(T) == 4;
    const int Mh = M / 2 + 1;
#define MG_L(L_, S_) do { \
        const int by = lds_bytes<T, L_>(); set_lds(k_columns<T, L_, CEN, S_>, by); \
        hipLaunchKernelGGL((k_columns<T, L_, CEN, S_>), dim3((Mh + fft_batch<L_>() - 1) / fft_batch<L_>()), dim3(fft_block<L_>()), by, st, \
                           Wh, N, K[0], K[1], K[2], K[3], Ph, fft_twiddle_set<T>(L_), LD); } while (0)
#define MG_C(L_) case L_: if (strided) MG_L(L_, true); else MG_L(L_, false); break;
    switch (M) { MG_COL_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
#undef MG_L
    MG_HIP(hipGetLastError());
}

#define MG_FOLD_SIZES(X) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096)
// rows per block of the folded row passes: 4 or 2 (FP32), 2 (FP64; its LDS holds no more)
template <class T> int fold_rows(int want) { return sizeof(T) == 8 ? 2 : (want >= 4 ? 4 : 2); }

template <class T> bool fold_supported(int M) {
    switch (M / 2) {
#define MG_C(L_) case L_: return true;
        MG_FOLD_SIZES(MG_C)
#undef MG_C
        default: return false;
    }
}

template <class T> void launch_rows_forward_fold(const T* W, int N, int M, cplx<T>* Wh, int rows, hipStream_t st) {
    require(N <= M / 2, "FFT size below 2N");
    const int LD = fold_stride_n(N);
    const int rb = fold_rows<T>(rows);
    if (tunables().fft_fold_persist & 2) {
        // one block per CU (the LDS holds one), at most as many as row groups per XCD, times the 8 XCDs
        const int cus = device_cus();
#define MG_P(L_, RB_) do { if constexpr (fold_pf_fits<T, L_, RB_>() && fold_persist_pays<T, L_, RB_>()) { \
            const int G = (N + RB_ - 1) / RB_, GX = (3 * G + kXcds - 1) / kXcds, per = std::min(GX, std::max(1, cus / kXcds)); \
            const int by = fold_pf_lds_bytes<T, L_, RB_>(); set_lds(k_rows_fwd_fold_p<T, L_, RB_>, by); \
            hipLaunchKernelGGL((k_rows_fwd_fold_p<T, L_, RB_>), dim3(per * kXcds), dim3(fold_block<T, L_, RB_>()), by, st, \
                               W, N, Wh, LD, fft_twiddle_set<T>(L_)); \
            MG_HIP(hipGetLastError()); return; } } while (0)
#define MG_C(L_) case L_: if constexpr (sizeof(T) == 4) { if (rb == 4) MG_P(L_, 4); else MG_P(L_, 2); } else MG_P(L_, 2); break;
        switch (M / 2) { MG_FOLD_SIZES(MG_C) default: break; }
#undef MG_C
#undef MG_P
    }
#define MG_R(L_, RB_) do { \
        const int by = fold_lds_bytes<T, L_, RB_>(); set_lds(k_rows_fwd_fold<T, L_, RB_>, by); \
        hipLaunchKernelGGL((k_rows_fwd_fold<T, L_, RB_>), dim3(xcd_grid(3 * ((N + RB_ - 1) / RB_))), dim3(fold_block<T, L_, RB_>()), by, st, \
                           W, N, Wh, LD, fft_twiddle_set<T>(L_)); } while (0)
#define MG_C(L_) case L_: if constexpr (sizeof(T) == 4) { if (rb == 4) MG_R(L_, 4); else MG_R(L_, 2); } else MG_R(L_, 2); break;
    switch (M / 2) { MG_FOLD_SIZES(MG_C) default: throw Error(ERR_UNSUP, "no folded row pass at this FFT size"); }
#undef MG_C
#undef MG_R
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_inverse_fold(const cplx<T>* Ph, int N, int M, T* phi, int rows, hipStream_t st) {
    require(N <= M / 2, "FFT size below 2N");
    const int LD = fold_stride_n(N);
    const int rb = fold_rows<T>(rows);
    const T scale = (T)(1.0 / ((double)M * (double)M));
    if (tunables().fft_fold_persist & 1) {
        // one block per CU (the LDS holds one), at most as many as row groups per XCD, times the 8 XCDs
        const int cus = device_cus();
#define MG_P(L_, RB_) do { if constexpr (fold_p_fits<T, L_, RB_>() && fold_persist_pays<T, L_, RB_>()) { \
            const int G = (N + RB_ - 1) / RB_, GX = (4 * G + kXcds - 1) / kXcds, per = std::min(GX, std::max(1, cus / kXcds)); \
            const int by = fold_p_lds_bytes<T, L_, RB_>(); set_lds(k_rows_inv_fold_p<T, L_, RB_>, by); \
            hipLaunchKernelGGL((k_rows_inv_fold_p<T, L_, RB_>), dim3(per * kXcds), dim3(fold_block<T, L_, RB_>()), by, st, \
                               Ph, N, LD, phi, scale, fft_twiddle_set<T>(L_)); \
            MG_HIP(hipGetLastError()); return; } } while (0)
#define MG_C(L_) case L_: if constexpr (sizeof(T) == 4) { if (rb == 4) MG_P(L_, 4); else MG_P(L_, 2); } else MG_P(L_, 2); break;
        switch (M / 2) { MG_FOLD_SIZES(MG_C) default: break; }
#undef MG_C
#undef MG_P
    }
#define MG_R(L_, RB_) do { \
        const int by = fold_lds_bytes<T, L_, RB_>(); set_lds(k_rows_inv_fold<T, L_, RB_>, by); \
        hipLaunchKernelGGL((k_rows_inv_fold<T, L_, RB_>), dim3(xcd_grid(4 * ((N + RB_ - 1) / RB_))), dim3(fold_block<T, L_, RB_>()), by, st, \
                           Ph, N, LD, phi, scale, fft_twiddle_set<T>(L_)); } while (0)
#define MG_C(L_) case L_: if constexpr (sizeof(T) == 4) { if (rb == 4) MG_R(L_, 4); else MG_R(L_, 2); } else MG_R(L_, 2); break;
    switch (M / 2) { MG_FOLD_SIZES(MG_C) default: throw Error(ERR_UNSUP, "no folded row pass at this FFT size"); }
#undef MG_C
#undef MG_R
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_inverse(const cplx<T>* X, int N, int M, T* phi, hipStream_t st) {
    require(N <= M / 2, "FFT size below 2N");
    const T scale = (T)(1.0 / ((double)M * (double)M));
#define MG_C(L_) case L_: { \
        const int by = lds_bytes<T, L_>(); set_lds(k_rows_inv<T, L_>, by); \
        hipLaunchKernelGGL((k_rows_inv<T, L_>), dim3((N + fft_batch<L_>() - 1) / fft_batch<L_>(), 4), dim3(fft_block<L_>()), by, st, X, N, phi, scale, fft_twiddle_set<T>(L_)); \
        break; }
    switch (M / 2) { MG_ROW_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
    MG_HIP(hipGetLastError());
}

template <class T> void launch_spectra(int M, double delta, double dof, const double* table, bool round_kernel, int ntab,
                                       cplx<T>* tmp, T* const* K, hipStream_t st, bool sym) {
    require(ntab == 2 || ntab == 4, "ntab must be 2 or 4");
    const int Mh = M / 2 + 1;
    const size_t plane = (size_t)Mh * Mh;
    cplx<T>* P0 = tmp;              // stage 1 output [a][ky]
    cplx<T>* P1 = tmp + plane;
    cplx<T>* Q0 = tmp + 2 * plane;  // transposed [ky][a]
    cplx<T>* Q1 = tmp + 3 * plane;
#define MG_S1(L_, NT) do { \
        const int by = lds_bytes<T, L_>(); set_lds(k_spec1<T, L_, NT>, by); \
        hipLaunchKernelGGL((k_spec1<T, L_, NT>), dim3((Mh + fft_batch<L_>() - 1) / fft_batch<L_>()), dim3(fft_block<L_>()), by, st, \
                           delta, dof, table, (int)round_kernel, P0, P1, fft_twiddle_set<T>(L_)); \
    } while (0)
#define MG_S2(L_, NT) do { \
        const int by = lds_bytes<T, L_>(); set_lds(k_spec2<T, L_, NT>, by); \
        hipLaunchKernelGGL((k_spec2<T, L_, NT>), dim3((Mh + fft_batch<L_>() - 1) / fft_batch<L_>()), dim3(fft_block<L_>()), by, st, \
                           Q0, Q1, K[0], K[1], K[2], K[3], fft_twiddle_set<T>(L_)); \
    } while (0)
#define MG_S2P(L_) do { \
        const int by = lds_bytes<T, L_>(); set_lds(k_spec2p<T, L_>, by); \
        const int np = (Mh + 1) / 2; \
        hipLaunchKernelGGL((k_spec2p<T, L_>), dim3((np + fft_batch<L_>() - 1) / fft_batch<L_>()), dim3(fft_block<L_>()), by, st, \
                           Q0, delta, K[0], K[1], K[2], fft_twiddle_set<T>(L_)); \
        MG_HIP(hipGetLastError()); \
        hipLaunchKernelGGL(k_transpose_real<T>, dim3((Mh + 31) / 32, (Mh + 31) / 32), dim3(256), 0, st, (const T*)K[2], K[3], Mh, Mh); \
    } while (0)
#define MG_C(L_) case L_: \
        if (ntab == 4 && sym) { MG_S1(L_, 2); launch_transpose<T>(P0, Q0, 1, Mh, Mh, st); MG_S2P(L_); } \
        else if (ntab == 4) { MG_S1(L_, 4); launch_transpose<T>(P0, Q0, 2, Mh, Mh, st); MG_S2(L_, 4); } \
        else { MG_S1(L_, 2); launch_transpose<T>(P0, Q0, 1, Mh, Mh, st); MG_S2(L_, 2); } \
        break;
    switch (M) { MG_COL_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
#undef MG_S1
#undef MG_S2
#undef MG_S2P
    MG_HIP(hipGetLastError());
}

template <class T> void launch_phi_interleave(const T* planar, T* out, i64 NN, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<i64>(4096, std::max<i64>(1, (NN + 255) / 256));
    hipLaunchKernelGGL(k_interleave4<T>, dim3(blocks), dim3(256), 0, st, planar, out, NN);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_forward_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st) {
    if (grid <= 0) return;
#define MG_C(L_) case L_: { \
        const int by = lds_bytes<T, L_>(); set_lds(kw_rows_fwd<T, L_>, by); \
        hipLaunchKernelGGL((kw_rows_fwd<T, L_>), dim3(grid), dim3(fft_block<L_>()), by, st, I, wl, fft_twiddle_set<T>(L_)); \
        break; }
    switch (M / 2) { MG_ROW_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
    MG_HIP(hipGetLastError());
}

template <class T> void launch_transpose_wl(const BInst<T>* I, const int* wl, int which, int grid, hipStream_t st) {
    require(which == 0 || which == 1, "grid transposes are 0 or 1");
    if (grid <= 0) return;
    hipLaunchKernelGGL(kw_transpose<T>, dim3(grid), dim3(256), 0, st, I, wl, which);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_columns_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st) {
    if (grid <= 0) return;
    constexpr bool CEN = sizeof(T) == 4;
#define MG_C(L_) case L_: { \
        const int by = lds_bytes<T, L_>(); set_lds(kw_columns<T, L_, CEN>, by); \
        hipLaunchKernelGGL((kw_columns<T, L_, CEN>), dim3(grid), dim3(fft_block<L_>()), by, st, I, wl, fft_twiddle_set<T>(L_)); \
        break; }
    switch (M) { MG_COL_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
    MG_HIP(hipGetLastError());
}

template <class T> void launch_rows_inverse_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st) {
    if (grid <= 0) return;
    const T scale = (T)(1.0 / ((double)M * (double)M));
#define MG_C(L_) case L_: { \
        const int by = lds_bytes<T, L_>(); set_lds(kw_rows_inv<T, L_>, by); \
        hipLaunchKernelGGL((kw_rows_inv<T, L_>), dim3(grid), dim3(fft_block<L_>()), by, st, I, wl, scale, fft_twiddle_set<T>(L_)); \
        break; }
    switch (M / 2) { MG_ROW_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
    MG_HIP(hipGetLastError());
}

// spectra of the instances in m1 (all of FFT size M): stage 1, transpose (blocks mt), stage 2 (the centred set by
// pairs of columns, blocks m2, then Ty from Tx, blocks mr; the t-SNE kernel is symmetric)
template <class T> void launch_spectra_batch(const BInst<T>* I, const BMap& m1, const BMap& mt, const BMap& m2, const BMap& mr,
                                             int M, double dof, bool round_kernel, int ntab, hipStream_t st) {
    if (m1.count == 0 || m1.start[m1.count] == 0) return;
    require(ntab == 2 || ntab == 4, "ntab must be 2 or 4");
    const dim3 grid((unsigned)m1.start[m1.count]);
#define MG_S(L_, NT) do { \
        const int by = lds_bytes<T, L_>(); set_lds(kb_spec1<T, L_, NT>, by); set_lds(kb_spec2<T, L_, NT>, by); \
        hipLaunchKernelGGL((kb_spec1<T, L_, NT>), grid, dim3(fft_block<L_>()), by, st, I, m1, dof, (const double*)nullptr, (int)round_kernel, fft_twiddle_set<T>(L_)); \
        MG_HIP(hipGetLastError()); \
        if (mt.count > 0 && mt.start[mt.count] > 0) \
            hipLaunchKernelGGL(kb_transpose_spec<T>, dim3((unsigned)mt.start[mt.count]), dim3(256), 0, st, I, mt, NT / 2, M); \
        MG_HIP(hipGetLastError()); \
        hipLaunchKernelGGL((kb_spec2<T, L_, NT>), grid, dim3(fft_block<L_>()), by, st, I, m1, fft_twiddle_set<T>(L_)); \
    } while (0)
#define MG_SP(L_) do { \
        const int by = lds_bytes<T, L_>(); set_lds(kb_spec1<T, L_, 2>, by); set_lds(kb_spec2p<T, L_>, by); \
        hipLaunchKernelGGL((kb_spec1<T, L_, 2>), grid, dim3(fft_block<L_>()), by, st, I, m1, dof, (const double*)nullptr, (int)round_kernel, fft_twiddle_set<T>(L_)); \
        MG_HIP(hipGetLastError()); \
        if (mt.count > 0 && mt.start[mt.count] > 0) \
            hipLaunchKernelGGL(kb_transpose_spec<T>, dim3((unsigned)mt.start[mt.count]), dim3(256), 0, st, I, mt, 1, M); \
        MG_HIP(hipGetLastError()); \
        if (m2.count > 0 && m2.start[m2.count] > 0) \
            hipLaunchKernelGGL((kb_spec2p<T, L_>), dim3((unsigned)m2.start[m2.count]), dim3(fft_block<L_>()), by, st, I, m2, fft_twiddle_set<T>(L_)); \
        MG_HIP(hipGetLastError()); \
        if (mr.count > 0 && mr.start[mr.count] > 0) \
            hipLaunchKernelGGL(kb_transpose_real_spec<T>, dim3((unsigned)mr.start[mr.count]), dim3(256), 0, st, I, mr, M); \
    } while (0)
#define MG_C(L_) case L_: if (ntab == 4) MG_SP(L_); else MG_S(L_, 2); break;
    switch (M) { MG_COL_SIZES(MG_C) default: throw Error(ERR_UNSUP, "FFT size not supported"); }
#undef MG_C
#undef MG_S
#undef MG_SP
    MG_HIP(hipGetLastError());
}

// Raises the LDS limit of every FFT kernel of every size, once, before a batch's host threads start: a kernel's first
// launch would otherwise set it (hipFuncSetAttribute) in whichever thread launches it first, while other threads
// launch and capture the same kernels.
template <class T> void prepare_fft_kernels() {
    constexpr bool CEN = sizeof(T) == 4;
#define MG_R(L_) { const int by = lds_bytes<T, L_>(); set_lds(k_rows_fwd<T, L_>, by); set_lds(kw_rows_fwd<T, L_>, by); \
                   set_lds(k_rows_inv<T, L_>, by); set_lds(kw_rows_inv<T, L_>, by); }
#define MG_K(L_) { const int by = lds_bytes<T, L_>(); set_lds(k_columns<T, L_, CEN, false>, by); \
                   set_lds(k_columns<T, L_, CEN, true>, by); set_lds(kw_columns<T, L_, CEN>, by); \
                   set_lds(k_spec1<T, L_, 2>, by); set_lds(k_spec1<T, L_, 4>, by); set_lds(k_spec2<T, L_, 2>, by); \
                   set_lds(k_spec2<T, L_, 4>, by); set_lds(k_spec2p<T, L_>, by); set_lds(kb_spec1<T, L_, 2>, by); \
                   set_lds(kb_spec2<T, L_, 2>, by); set_lds(kb_spec2p<T, L_>, by); }
#define MG_F(L_) { set_lds(k_rows_fwd_fold<T, L_, 2>, fold_lds_bytes<T, L_, 2>()); set_lds(k_rows_inv_fold<T, L_, 2>, fold_lds_bytes<T, L_, 2>()); \
                   if constexpr (fold_p_fits<T, L_, 2>()) set_lds(k_rows_inv_fold_p<T, L_, 2>, fold_p_lds_bytes<T, L_, 2>()); \
                   if constexpr (fold_pf_fits<T, L_, 2>()) set_lds(k_rows_fwd_fold_p<T, L_, 2>, fold_pf_lds_bytes<T, L_, 2>()); \
                   if constexpr (sizeof(T) == 4) { set_lds(k_rows_fwd_fold<T, L_, 4>, fold_lds_bytes<T, L_, 4>()); \
                                                   set_lds(k_rows_inv_fold<T, L_, 4>, fold_lds_bytes<T, L_, 4>()); \
                                                   if constexpr (fold_p_fits<T, L_, 4>()) \
                                                       set_lds(k_rows_inv_fold_p<T, L_, 4>, fold_p_lds_bytes<T, L_, 4>()); \
                                                   if constexpr (fold_pf_fits<T, L_, 4>()) \
                                                       set_lds(k_rows_fwd_fold_p<T, L_, 4>, fold_pf_lds_bytes<T, L_, 4>()); } }
    MG_ROW_SIZES(MG_R)
    MG_COL_SIZES(MG_K)
    MG_FOLD_SIZES(MG_F)
#undef MG_R
#undef MG_K
#undef MG_F
}

#define MG_FFT_INST(T) \
    template void prepare_fft_kernels<T>(); \
    template void launch_rows_forward<T>(const T*, int, int, cplx<T>*, hipStream_t); \
    template void launch_transpose<T>(const cplx<T>*, cplx<T>*, int, int, int, hipStream_t); \
    template void launch_columns<T>(const cplx<T>*, int, int, const T* const*, bool, cplx<T>*, hipStream_t, bool, int); \
    template void launch_rows_inverse<T>(const cplx<T>*, int, int, T*, hipStream_t); \
    template bool fold_supported<T>(int); \
    template void launch_rows_forward_fold<T>(const T*, int, int, cplx<T>*, int, hipStream_t); \
    template void launch_rows_inverse_fold<T>(const cplx<T>*, int, int, T*, int, hipStream_t); \
    template void launch_spectra<T>(int, double, double, const double*, bool, int, cplx<T>*, T* const*, hipStream_t, bool); \
    template void launch_phi_interleave<T>(const T*, T*, i64, hipStream_t); \
    template void launch_rows_forward_wl<T>(const BInst<T>*, const int*, int, int, hipStream_t); \
    template void launch_transpose_wl<T>(const BInst<T>*, const int*, int, int, hipStream_t); \
    template void launch_columns_wl<T>(const BInst<T>*, const int*, int, int, hipStream_t); \
    template void launch_rows_inverse_wl<T>(const BInst<T>*, const int*, int, int, hipStream_t); \
    template void launch_spectra_batch<T>(const BInst<T>*, const BMap&, const BMap&, const BMap&, const BMap&, int, double, bool, int, hipStream_t);

}  // namespace mg
