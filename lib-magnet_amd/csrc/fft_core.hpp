// Workgroup FFTs for gfx950 with plans fixed at compile time. Lengths L = 2^s or 3 * 2^s (64 <= L <= 8192),
// each transform held by TH = L / 16 threads. A plan is a list of radices (3 .. 32); every pass is a Stockham
// auto-sort step computed in registers: item j (< L / R) reads elements j + r L / R (r < R), multiplies element
// r by w^(r k) with w = exp(sign 2 pi i / (Ns R)), Ns the product of the earlier radices and k = j mod Ns, runs
// an R-point DFT and writes element r to (j - k) R + k + r Ns. Natural order in and out.
//
// Only the passes in between exchange data through LDS: the first pass takes its inputs from a caller's
// functor (global memory, registers) and the last pass hands its outputs to one. The last pass leaves element
// j + r L / R_last (j = t + q TH) in slot (q, r) of thread t, which is also where the first pass of a plan
// starting with R_last takes it from; so a forward transform, a pointwise product and an inverse transform
// whose plan is the reverse of the forward one chain through registers.
//
// LDS layout: element i of a buffer lives at pad(i) = i + i / 16 (one spare element after every 16). The
// strided reads of a pass (stride a multiple of 16 elements) and its expanded writes then fall on distinct
// banks across a wave's lanes for 8- and 16-byte elements alike, and pad(i + 16 k) = pad(i) + 17 k, so the
// R accesses of one item are one address plus compile-time offsets.
//
// Twiddles: per plan and pass, a table of exp(-2 pi i k 2^b / (Ns R)) for k < Ns and 2^b < R in global memory
// (consecutive k on consecutive lanes: coalesced, L1-resident), correctly rounded from long double on the
// host. A pass reads w^(2^b) and forms the other powers with products (at most three deep): in FP32 each from an
// earlier power, in FP64 each from the lookups of its set bits, so that only the lookups stay live.
#pragma once
#include "common.hpp"
#include <utility>

namespace mg {

__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
constexpr int padded(int n) { return n + n / 16; }   // LDS elements of a buffer of n (n a multiple of 16)

// twiddle tables of one length L: the forward plan's, the reversed plan's, and exp(-2 pi i k / 2L), k <= L
// (the real-transform post / pre step of the row passes)
template <class T> struct TwSet { const cplx<T>* fwd; const cplx<T>* rev; const cplx<T>* half; };
template <class T> TwSet<T> fft_twiddle_set(int L);   // this device's tables (after fft_init_twiddles)

// cos and sin of 2 pi m / 96 (every radix used divides 96), built from the first octant by symmetry
__device__ constexpr double kC96[96] = {
    1.0, 0.9978589232386035, 0.9914448613738104, 0.9807852804032304, 0.9659258262890683, 0.9469301294951057,
    0.9238795325112867, 0.8968727415326883, 0.8660254037844386, 0.8314696123025452, 0.7933533402912352, 0.7518398074789774,
    0.7071067811865476, 0.6593458151000688, 0.6087614290087207, 0.5555702330196022, 0.5, 0.4422886902190013,
    0.3826834323650898, 0.3214394653031616, 0.25881904510252074, 0.19509032201612828, 0.1305261922200516, 0.06540312923014306,
    0.0, -0.06540312923014306, -0.1305261922200516, -0.19509032201612828, -0.25881904510252074, -0.3214394653031616,
    -0.3826834323650898, -0.4422886902190013, -0.5, -0.5555702330196022, -0.6087614290087207, -0.6593458151000688,
    -0.7071067811865476, -0.7518398074789774, -0.7933533402912352, -0.8314696123025452, -0.8660254037844386, -0.8968727415326883,
    -0.9238795325112867, -0.9469301294951057, -0.9659258262890683, -0.9807852804032304, -0.9914448613738104, -0.9978589232386035,
    -1.0, -0.9978589232386035, -0.9914448613738104, -0.9807852804032304, -0.9659258262890683, -0.9469301294951057,
    -0.9238795325112867, -0.8968727415326883, -0.8660254037844386, -0.8314696123025452, -0.7933533402912352, -0.7518398074789774,
    -0.7071067811865476, -0.6593458151000688, -0.6087614290087207, -0.5555702330196022, -0.5, -0.4422886902190013,
    -0.3826834323650898, -0.3214394653031616, -0.25881904510252074, -0.19509032201612828, -0.1305261922200516, -0.06540312923014306,
    0.0, 0.06540312923014306, 0.1305261922200516, 0.19509032201612828, 0.25881904510252074, 0.3214394653031616,
    0.3826834323650898, 0.4422886902190013, 0.5, 0.5555702330196022, 0.6087614290087207, 0.6593458151000688,
    0.7071067811865476, 0.7518398074789774, 0.7933533402912352, 0.8314696123025452, 0.8660254037844386, 0.8968727415326883,
    0.9238795325112867, 0.9469301294951057, 0.9659258262890683, 0.9807852804032304, 0.9914448613738104, 0.9978589232386035};
__device__ constexpr double kS96[96] = {
    0.0, 0.06540312923014306, 0.1305261922200516, 0.19509032201612828, 0.25881904510252074, 0.3214394653031616,
    0.3826834323650898, 0.4422886902190013, 0.5, 0.5555702330196022, 0.6087614290087207, 0.6593458151000688,
    0.7071067811865476, 0.7518398074789774, 0.7933533402912352, 0.8314696123025452, 0.8660254037844386, 0.8968727415326883,
    0.9238795325112867, 0.9469301294951057, 0.9659258262890683, 0.9807852804032304, 0.9914448613738104, 0.9978589232386035,
    1.0, 0.9978589232386035, 0.9914448613738104, 0.9807852804032304, 0.9659258262890683, 0.9469301294951057,
    0.9238795325112867, 0.8968727415326883, 0.8660254037844386, 0.8314696123025452, 0.7933533402912352, 0.7518398074789774,
    0.7071067811865476, 0.6593458151000688, 0.6087614290087207, 0.5555702330196022, 0.5, 0.4422886902190013,
    0.3826834323650898, 0.3214394653031616, 0.25881904510252074, 0.19509032201612828, 0.1305261922200516, 0.06540312923014306,
    0.0, -0.06540312923014306, -0.1305261922200516, -0.19509032201612828, -0.25881904510252074, -0.3214394653031616,
    -0.3826834323650898, -0.4422886902190013, -0.5, -0.5555702330196022, -0.6087614290087207, -0.6593458151000688,
    -0.7071067811865476, -0.7518398074789774, -0.7933533402912352, -0.8314696123025452, -0.8660254037844386, -0.8968727415326883,
    -0.9238795325112867, -0.9469301294951057, -0.9659258262890683, -0.9807852804032304, -0.9914448613738104, -0.9978589232386035,
    -1.0, -0.9978589232386035, -0.9914448613738104, -0.9807852804032304, -0.9659258262890683, -0.9469301294951057,
    -0.9238795325112867, -0.8968727415326883, -0.8660254037844386, -0.8314696123025452, -0.7933533402912352, -0.7518398074789774,
    -0.7071067811865476, -0.6593458151000688, -0.6087614290087207, -0.5555702330196022, -0.5, -0.4422886902190013,
    -0.3826834323650898, -0.3214394653031616, -0.25881904510252074, -0.19509032201612828, -0.1305261922200516, -0.06540312923014306};

// x * exp(SIGN 2 pi i m / R); m and R are constants once the callers' loops are unrolled, quarter turns
// cost no multiplication
template <class T, int SIGN> __device__ __forceinline__ cplx<T> rot(cplx<T> x, int m, int R) {
    const int q = ((m % R) * (96 / R)) % 96;
    if (q == 0) return x;
    if (q == 48) return {-x.x, -x.y};
    if (q == 24) return SIGN > 0 ? cplx<T>{-x.y, x.x} : cplx<T>{x.y, -x.x};
    if (q == 72) return SIGN > 0 ? cplx<T>{x.y, -x.x} : cplx<T>{-x.y, x.x};
    return cmul(x, cplx<T>{(T)kC96[q], (T)(SIGN * kS96[q])});
}

// R-point DFT in registers, natural order in and out: X_k = sum_n x_n exp(SIGN 2 pi i n k / R).
// R = R1 R2 splits into R2 DFTs of size R1 over x[n2 + R2 n1], twiddles exp(SIGN 2 pi i n2 k1 / R), and R1
// DFTs of size R2 (Cooley-Tukey); R1 = 4 when 4 divides R, else 2 or 3.
template <class T, int R, int SIGN> __device__ __forceinline__ void dft(cplx<T>* x) {
    if constexpr (R == 2) {
        const cplx<T> a = x[0], b = x[1];
        x[0] = cadd(a, b); x[1] = csub(a, b);
    } else if constexpr (R == 3) {
        constexpr double h = SIGN * 0.86602540378443864676;   // SIGN sqrt(3) / 2
        const cplx<T> a = x[0], s = cadd(x[1], x[2]), d = csub(x[1], x[2]);
        const cplx<T> m = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
        const cplx<T> t = {(T)(-h) * d.y, (T)h * d.x};
        x[0] = cadd(a, s); x[1] = cadd(m, t); x[2] = csub(m, t);
    } else if constexpr (R == 4) {
        const cplx<T> t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]), t2 = cadd(x[1], x[3]);
        const cplx<T> d = csub(x[1], x[3]);
        const cplx<T> t3 = SIGN > 0 ? cplx<T>{-d.y, d.x} : cplx<T>{d.y, -d.x};
        x[0] = cadd(t0, t2); x[2] = csub(t0, t2); x[1] = cadd(t1, t3); x[3] = csub(t1, t3);
    } else {
        constexpr int R1 = R % 4 == 0 ? 4 : (R % 2 == 0 ? 2 : 3), R2 = R / R1;
        cplx<T> t[R];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            cplx<T> a[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) a[n1] = x[n2 + R2 * n1];
            dft<T, R1, SIGN>(a);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) t[n2 + R2 * k1] = rot<T, SIGN>(a[k1], n2 * k1, R);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            cplx<T> b[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) b[n2] = t[n2 + R2 * k1];
            dft<T, R2, SIGN>(b);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) x[k1 + R1 * k2] = b[k2];
        }
    }
}

constexpr int hibit(int r) { int h = 1; while (2 * h <= r) h *= 2; return h; }
constexpr int log2c(int r) { int l = 0; while ((1 << l) < r) ++l; return l; }

// a[r] *= w^r (r < R), w = exp(SIGN 2 pi i k / (Ns R)); tab[b Ns + k] = exp(-2 pi i k 2^b / (Ns R))
template <class T, int R, int Ns, int SIGN>
__device__ __forceinline__ void twiddle_apply(cplx<T>* a, const cplx<T>* __restrict__ tab, int k) {
    constexpr int NB = log2c(R);   // lookups at 2^b, b < NB (2^b <= R - 1)
    cplx<T> p[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        cplx<T> x = tab[b * Ns + k];
        if (SIGN > 0) x.y = -x.y;
        p[b] = x;
    }
    if constexpr (sizeof(T) == 8) {
        // w^r as the product of the lookups of r's set bits, lowest first: only the NB lookups stay live (FP64: the
        // registers this frees let the column pass hold more waves, which pays there)
#pragma unroll
        for (int r = 1; r < R; ++r) {
            cplx<T> w = {T(1), T(0)};
            bool first = true;
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (r & (1 << b)) { w = first ? p[b] : cmul(w, p[b]); first = false; }
            a[r] = cmul(a[r], w);
        }
    } else {
        // w^r = w^(r - h) w^h, h the highest power of two in r: one product per power
        cplx<T> w[R];
#pragma unroll
        for (int b = 0; b < NB; ++b) w[1 << b] = p[b];
#pragma unroll
        for (int r = 3; r < R; ++r)
            if (r != hibit(r)) w[r] = cmul(w[r - hibit(r)], w[hibit(r)]);
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
    }
}

// ---- plans
template <int... Rs> struct Plan {
    static constexpr int P = sizeof...(Rs);
    static constexpr int R[P] = {Rs...};
    static constexpr int L = (1 * ... * Rs);
};
template <class Pl> constexpr int plan_ns(int I) { int n = 1; for (int i = 0; i < I; ++i) n *= Pl::R[i]; return n; }
// offset of pass I's twiddle table in the plan's (passes 1 .. P-1: log2c(R) Ns entries each)
template <class Pl> constexpr int plan_tw_offset(int I) {
    int o = 0;
    for (int i = 1; i < I; ++i) o += log2c(Pl::R[i]) * plan_ns<Pl>(i);
    return o;
}
template <class Pl> constexpr int plan_tw_size() { return plan_tw_offset<Pl>(Pl::P); }
template <class Pl, class S> struct RevPlan;
template <class Pl, int... I> struct RevPlan<Pl, std::integer_sequence<int, I...>> { using type = Plan<Pl::R[Pl::P - 1 - I]...>; };
template <class Pl> using Rev = typename RevPlan<Pl, std::make_integer_sequence<int, Pl::P>>::type;

// forward plans: radix 16 at both ends (the first pass of a column transform then reads only the N <= L/2
// live rows, and the last one is the register link: 16 elements per thread, L / 16 threads per transform),
// the remaining factor in between; lengths below 512 take that factor first. The radix-16 ends also keep
// the column pass clear of register spills at the largest sizes.
template <class T, int L> struct FwdPlan;
#define MG_PLAN(L_, ...) \
    template <> struct FwdPlan<float, L_> { using type = Plan<__VA_ARGS__>; }; \
    template <> struct FwdPlan<double, L_> { using type = Plan<__VA_ARGS__>; };
MG_PLAN(64, 4, 16)
MG_PLAN(96, 6, 16)
MG_PLAN(128, 8, 16)
MG_PLAN(192, 12, 16)
MG_PLAN(256, 16, 16)
MG_PLAN(384, 24, 16)
MG_PLAN(512, 16, 2, 16)
MG_PLAN(768, 16, 3, 16)
MG_PLAN(1024, 16, 4, 16)
MG_PLAN(1536, 16, 6, 16)
MG_PLAN(2048, 16, 8, 16)
MG_PLAN(3072, 16, 12, 16)
MG_PLAN(4096, 16, 16, 16)
MG_PLAN(6144, 16, 24, 16)
MG_PLAN(8192, 16, 32, 16)
#undef MG_PLAN

template <int L> constexpr int fft_th() { return L / 16; }                                       // threads per transform
template <int L> constexpr int fft_batch() { return L / 16 >= 256 ? 1 : 256 / (L / 16); }         // transforms per workgroup
template <int L> constexpr int fft_block() { return fft_th<L>() * fft_batch<L>(); }

// One Stockham pass (see the top of the file). Inputs come from LDS (SRC = 1, transform at element offset
// `base`) or from ld(e, q, r) for slot (q, r); outputs go to LDS (DST = 1) or to st(e, v, q, r). BAR puts a
// barrier between the loads and the stores (both on the same LDS buffer). tab: this pass's twiddles.
template <class T, int L, int R, int Ns, int TH, int SIGN, bool SRC, bool DST, bool BAR, class Ld, class St>
__device__ __forceinline__ void stockham(int t, cplx<T>* lds, int base, const cplx<T>* __restrict__ tab, Ld&& ld, St&& st) {
    constexpr int NI = L / R, Q = (NI + TH - 1) / TH;
    constexpr bool exact = NI % TH == 0;
    cplx<T> v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * TH;
        if (!exact && j >= NI) continue;
        if constexpr (SRC) {
            if constexpr (NI % 16 == 0) {
                const int p0 = pad(base + j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = lds[p0 + r * padded(NI)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = lds[pad(base + j + r * NI)];
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = ld(j + r * NI, q, r);
        }
        if constexpr (Ns > 1) twiddle_apply<T, R, Ns, SIGN>(v[q], tab, j % Ns);
        dft<T, R, SIGN>(v[q]);
    }
    if constexpr (BAR) __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = t + q * TH;
        if (!exact && j >= NI) continue;
        const int k = Ns > 1 ? j % Ns : 0;
        const int o = (j - k) * R + k;
        if constexpr (DST) {
            if constexpr (Ns % 16 == 0) {
                const int p0 = pad(base + o);
#pragma unroll
                for (int r = 0; r < R; ++r) lds[p0 + r * padded(Ns)] = v[q][r];
            } else if constexpr (Ns == 1 && R % 16 == 0) {   // base + o is a multiple of 16
                const int p0 = pad(base + o);
#pragma unroll
                for (int r = 0; r < R; ++r) lds[p0 + r + (r >> 4)] = v[q][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) lds[pad(base + o + r * Ns)] = v[q][r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) st(o + r * Ns, v[q][r], q, r);
        }
    }
}

// Whole transform of plan Pl by the TH threads of one transform (t = index among them), LDS buffer at element
// offset `base` (a multiple of 16) of `lds`, twiddle tables of the plan at tw. Passes that write LDS are bracketed by barriers, so consecutive
// transforms on one buffer need no extra synchronisation. LAST: where the last pass puts its outputs --
// FN: st(e, v, q, r); FN_BAR: the same after a barrier (st writes the transform's own LDS); TO_LDS: the
// transform's LDS buffer in natural order, after a barrier. Unnormalised; SIGN -1 forward, +1 inverse.
enum FftLast { FN = 0, FN_BAR = 1, TO_LDS = 2 };
template <class T, class Pl, int SIGN, int TH, int LAST, int I = 0, class Ld, class St>
__device__ __forceinline__ void fft_run(int t, cplx<T>* lds, int base, const cplx<T>* __restrict__ tw, Ld&& ld, St&& st) {
    constexpr int P = Pl::P, L = Pl::L, R = Pl::R[I], Ns = plan_ns<Pl>(I);
    const cplx<T>* tab = tw + plan_tw_offset<Pl>(I);
    constexpr bool last_lds = LAST == TO_LDS, last_bar = LAST != FN;
    if constexpr (P == 1) {
        stockham<T, L, R, Ns, TH, SIGN, false, last_lds, last_bar>(t, lds, base, tab, ld, st);
    } else if constexpr (I == 0) {
        stockham<T, L, R, Ns, TH, SIGN, false, true, true>(t, lds, base, tab, ld, st);
        __syncthreads();
        fft_run<T, Pl, SIGN, TH, LAST, 1>(t, lds, base, tw, ld, st);
    } else if constexpr (I + 1 < P) {
        stockham<T, L, R, Ns, TH, SIGN, true, true, true>(t, lds, base, tab, ld, st);
        __syncthreads();
        fft_run<T, Pl, SIGN, TH, LAST, I + 1>(t, lds, base, tw, ld, st);
    } else {
        stockham<T, L, R, Ns, TH, SIGN, true, last_lds, last_bar>(t, lds, base, tab, ld, st);
    }
}

}  // namespace mg
