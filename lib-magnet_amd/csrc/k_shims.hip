// The reference's fallback kernels of the FFT repulsion (src/fftrep.cpp:24-125, 294-411), used by
// FFTRepulsion2D._scipy_call / _numpy_call (magnet/tsne.py:92-129) when FFTConv2D is not used, and by tests.
// Same box rule and weights as the reference: u = (x - xmin) / width, box = clamp(int(u)), product-form
// Lagrange weights at the clamped local coordinate; sums in the reference's order at one thread, with no
// contraction of products into fused multiply-adds, so the float32 results are those of the reference.
//   interp_scatter / scatter_padded_2d : points sorted by box (stable: index order within a box), then one
//                                        thread per grid node sums its box's points -- deterministic, no atomics;
//   interp_gather / gather_forces_2d   : one thread per point; Z as a fixed-order double reduction;
//   spectral_products                  : elementwise complex products.
#include "common.hpp"
#include <algorithm>

namespace mg {

namespace {

constexpr int SH_TPB = 256;
constexpr int SH_CMAX = 8;   // charge columns per pass of the node kernel

// product-form Lagrange weights (src/fftrep.cpp:24-35)
template <class T> __device__ inline void lagrange_prod(T u, int ni, T* w) {
#pragma clang fp contract(off)
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

template <class T> struct ShimBox { T xmin, ymin, width; int nb, ni, N; };

// Box::locate (src/fftrep.cpp:38-45)
template <class T> __device__ inline void locate(T x, T y, const ShimBox<T>& B, int& bx, int& by, T* wx, T* wy) {
#pragma clang fp contract(off)
    const T ux = (x - B.xmin) / B.width, uy = (y - B.ymin) / B.width;
    bx = min(B.nb - 1, max(0, (int)ux));
    by = min(B.nb - 1, max(0, (int)uy));
    lagrange_prod<T>(fmin(T(1), fmax(T(0), ux - (T)bx)), B.ni, wx);
    lagrange_prod<T>(fmin(T(1), fmax(T(0), uy - (T)by)), B.ni, wy);
}
template <class T> __device__ inline int box_of(T x, T y, const ShimBox<T>& B) {
    const T ux = (x - B.xmin) / B.width, uy = (y - B.ymin) / B.width;
    return min(B.nb - 1, max(0, (int)ux)) * B.nb + min(B.nb - 1, max(0, (int)uy));
}

template <class T> __global__ void k_shim_keys(const T* __restrict__ Y, i64 n, ShimBox<T> B, int* __restrict__ keys, int* __restrict__ vals) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        keys[i] = box_of<T>(Y[2 * i], Y[2 * i + 1], B);
        vals[i] = (int)i;
    }
}

// first / one-past-last sorted position of every box that holds points (others keep 0, 0)
__global__ void k_shim_ranges(const int* __restrict__ keys, i64 n, int* __restrict__ start, int* __restrict__ end) {
    for (i64 p = (i64)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (i64)gridDim.x * blockDim.x) {
        const int k = keys[p];
        if (p == 0 || keys[p - 1] != k) start[k] = (int)p;
        if (p == n - 1 || keys[p + 1] != k) end[k] = (int)(p + 1);
    }
}

__global__ void k_shim_zero_int(int* __restrict__ a, i64 count) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (i64)gridDim.x * blockDim.x) a[i] = 0;
}

// one thread per node of the live N x N grid: the sum over its box's points (in index order) of
// w * charge, for charge columns [c0, c0 + nc). Q == nullptr: the charges (1, x, y) of scatter_padded_2d.
// Output plane c at out + c * plane, row stride `ld` (N, or M for the padded grid).
template <class T>
__global__ __launch_bounds__(SH_TPB) void k_shim_scatter(const T* __restrict__ Y, const int* __restrict__ vals,
                                                         const int* __restrict__ start, const int* __restrict__ end,
                                                         ShimBox<T> B, const T* __restrict__ Q, int C, int c0, int nc,
                                                         T* __restrict__ out, int ld, i64 plane) {
#pragma clang fp contract(off)
    const i64 node = (i64)blockIdx.x * SH_TPB + threadIdx.x;
    if (node >= (i64)B.N * B.N) return;
    const int gx = (int)(node / B.N), gy = (int)(node % B.N);
    const int a = gx % B.ni, b = gy % B.ni;
    const int box = (gx / B.ni) * B.nb + gy / B.ni;
    T acc[SH_CMAX];
#pragma unroll
    for (int c = 0; c < SH_CMAX; ++c) acc[c] = T(0);
    T wx[8], wy[8];
    for (int p = start[box]; p < end[box]; ++p) {
        const int i = vals[p];
        const T x = Y[2 * (i64)i], y = Y[2 * (i64)i + 1];
        int bx, by;
        locate<T>(x, y, B, bx, by, wx, wy);
        const T w = wx[a] * wy[b];
        if (Q) {
#pragma unroll
            for (int c = 0; c < SH_CMAX; ++c) if (c < nc) acc[c] += w * Q[(i64)i * C + c0 + c];
        } else {
            acc[0] += w * T(1); acc[1] += w * x; acc[2] += w * y;
        }
    }
#pragma unroll
    for (int c = 0; c < SH_CMAX; ++c)
        if (c < nc) out[(i64)(c0 + c) * plane + (i64)gx * ld + gy] = acc[c];
}

// interp_gather (src/fftrep.cpp:94-125): out (n, C) from grid (C, N, N)
template <class T>
__global__ __launch_bounds__(SH_TPB) void k_shim_gather(const T* __restrict__ Y, i64 n, ShimBox<T> B, const T* __restrict__ G,
                                                        int C, T* __restrict__ out) {
#pragma clang fp contract(off)
    const i64 i = (i64)blockIdx.x * SH_TPB + threadIdx.x;
    if (i >= n) return;
    T wx[8], wy[8];
    int bx, by;
    locate<T>(Y[2 * i], Y[2 * i + 1], B, bx, by, wx, wy);
    const i64 NN = (i64)B.N * B.N;
    for (int c = 0; c < C; ++c) {
        T s = 0;
        for (int a = 0; a < B.ni; ++a) {
            const int gx = bx * B.ni + a;
            for (int b = 0; b < B.ni; ++b) {
                const int gy = by * B.ni + b;
                const T w = wx[a] * wy[b];
                s += w * G[c * NN + (i64)gx * B.N + gy];
            }
        }
        out[i * C + c] = s;
    }
}

// gather_forces_2d (src/fftrep.cpp:373-411): F (n, 2) from phi (4, M, M); per-block double partials of pot0
template <class T>
__global__ __launch_bounds__(SH_TPB) void k_shim_forces(const T* __restrict__ Y, i64 n, ShimBox<T> B, const T* __restrict__ phi,
                                                        int M, T* __restrict__ F, double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sh[SH_TPB];
    const i64 i = (i64)blockIdx.x * SH_TPB + threadIdx.x;
    double z = 0.0;
    if (i < n) {
        T wx[8], wy[8];
        int bx, by;
        const T x = Y[2 * i], y = Y[2 * i + 1];
        locate<T>(x, y, B, bx, by, wx, wy);
        const i64 MM = (i64)M * M;
        T pot[4] = {0, 0, 0, 0};
        for (int a = 0; a < B.ni; ++a) {
            const i64 gx = bx * B.ni + a;
            for (int b = 0; b < B.ni; ++b) {
                const i64 gy = by * B.ni + b;
                const T w = wx[a] * wy[b];
#pragma unroll
                for (int c = 0; c < 4; ++c) pot[c] += w * phi[c * MM + gx * M + gy];
            }
        }
        z = (double)pot[0];
        F[2 * i] = pot[1] * x - pot[2];
        F[2 * i + 1] = pot[1] * y - pot[3];
    }
    sh[threadIdx.x] = z;
    __syncthreads();
    for (int s = SH_TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// spectral_products (src/fftrep.cpp:346-371): complex (M, Mh) K1, K2 and (3, M, Mh) W -> (4, M, Mh)
template <class T>
__global__ __launch_bounds__(SH_TPB) void k_shim_products(const T* __restrict__ K1, const T* __restrict__ K2, const T* __restrict__ W,
                                                          i64 L, T* __restrict__ O) {
#pragma clang fp contract(off)
    for (i64 k = (i64)blockIdx.x * SH_TPB + threadIdx.x; k < L; k += (i64)gridDim.x * SH_TPB) {
        const T ar = K1[2 * k], ai = K1[2 * k + 1], br = W[2 * k], bi = W[2 * k + 1];
        O[2 * k] = ar * br - ai * bi; O[2 * k + 1] = ar * bi + ai * br;
        const T cr = K2[2 * k], ci = K2[2 * k + 1];
        for (int c = 0; c < 3; ++c) {
            const T wr = W[2 * (c * L + k)], wi = W[2 * (c * L + k) + 1];
            O[2 * ((c + 1) * L + k)] = cr * wr - ci * wi; O[2 * ((c + 1) * L + k) + 1] = cr * wi + ci * wr;
        }
    }
}

template <class T> __global__ void k_shim_fill(T* __restrict__ a, i64 count) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (i64)gridDim.x * blockDim.x) a[i] = T(0);
}

unsigned grid1(i64 count) { return (unsigned)std::min<i64>(8192, std::max<i64>(1, (count + SH_TPB - 1) / SH_TPB)); }

template <class T> ShimBox<T> make_box(double xmin, double ymin, double width, int nb, int ni) {
    require(nb >= 1 && ni >= 1 && ni <= 8 && width > 0, "bad grid parameters");
    require((i64)nb * ni <= 46340 && (i64)nb * nb < (i64)INT32_MAX, "grid too large");
    return {(T)xmin, (T)ymin, (T)width, nb, ni, nb * ni};
}

}  // namespace

// scratch for the box sort of the scatter: keys / vals ping-pong, digit histograms, box ranges
struct ShimScratch { int *keysA, *valsA, *keysB, *valsB, *hist, *start, *end; };
template <class T> int launch_key_sort(int* keysA, int* valsA, int* keysB, int* valsB, int* hist, i64 n, int bits, hipStream_t st);
int sort_hist_blocks(i64 n);

size_t shim_scratch_ints(i64 n, int nb) {
    return (size_t)4 * std::max<i64>(n, 1) + (size_t)sort_hist_ints(n) + (size_t)2 * nb * nb;
}

// the sorted point order and per-box ranges of the scatter, in `ws` (shim_scratch_ints ints)
template <class T> static const int* shim_sort(const T* Y, i64 n, const ShimBox<T>& B, int* ws, int*& start, int*& end, hipStream_t st) {
    const size_t nn = (size_t)std::max<i64>(n, 1);
    int* keysA = ws; int* valsA = ws + nn; int* keysB = ws + 2 * nn; int* valsB = ws + 3 * nn;
    int* hist = ws + 4 * nn;
    start = hist + (size_t)sort_hist_ints(n);
    end = start + (size_t)B.nb * B.nb;
    hipLaunchKernelGGL(k_shim_zero_int, dim3(grid1(2 * (i64)B.nb * B.nb)), dim3(SH_TPB), 0, st, start, 2 * (i64)B.nb * B.nb);
    if (n == 0) return valsA;
    hipLaunchKernelGGL(k_shim_keys<T>, dim3(grid1(n)), dim3(SH_TPB), 0, st, Y, n, B, keysA, valsA);
    int bits = 0;
    while ((1ll << bits) < (long long)B.nb * B.nb) ++bits;
    const int cur = launch_key_sort<T>(keysA, valsA, keysB, valsB, hist, n, bits, st);
    const int* keys = cur == 0 ? keysA : keysB;
    const int* vals = cur == 0 ? valsA : valsB;
    hipLaunchKernelGGL(k_shim_ranges, dim3(grid1(n)), dim3(SH_TPB), 0, st, keys, n, start, end);
    MG_HIP(hipGetLastError());
    return vals;
}

template <class T> void launch_interp_scatter(const T* Y, i64 n, const T* Q, int C, double xmin, double ymin, double width, int nb,
                                              int ni, T* out, int* ws, hipStream_t st) {
    const ShimBox<T> B = make_box<T>(xmin, ymin, width, nb, ni);
    require(C >= 1, "charges must be (n, T) with T >= 1");
    int *start, *end;
    const int* vals = shim_sort<T>(Y, n, B, ws, start, end, st);
    const i64 NN = (i64)B.N * B.N;
    for (int c0 = 0; c0 < C; c0 += SH_CMAX)
        hipLaunchKernelGGL(k_shim_scatter<T>, dim3((unsigned)((NN + SH_TPB - 1) / SH_TPB)), dim3(SH_TPB), 0, st, Y, vals, start, end, B, Q,
                           C, c0, std::min(SH_CMAX, C - c0), out, B.N, NN);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_scatter_padded(const T* Y, i64 n, double xmin, double ymin, double width, int nb, int ni, int M, T* out,
                                              int* ws, hipStream_t st) {
    const ShimBox<T> B = make_box<T>(xmin, ymin, width, nb, ni);
    require(M >= B.N, "bad grid parameters");
    const i64 MM = (i64)M * M;
    hipLaunchKernelGGL(k_shim_fill<T>, dim3(grid1(3 * MM)), dim3(SH_TPB), 0, st, out, 3 * MM);
    int *start, *end;
    const int* vals = shim_sort<T>(Y, n, B, ws, start, end, st);
    const i64 NN = (i64)B.N * B.N;
    hipLaunchKernelGGL(k_shim_scatter<T>, dim3((unsigned)((NN + SH_TPB - 1) / SH_TPB)), dim3(SH_TPB), 0, st, Y, vals, start, end, B,
                       (const T*)nullptr, 3, 0, 3, out, M, MM);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_interp_gather(const T* Y, i64 n, const T* G, int C, double xmin, double ymin, double width, int nb,
                                             int ni, T* out, hipStream_t st) {
    const ShimBox<T> B = make_box<T>(xmin, ymin, width, nb, ni);
    require(C >= 1, "grid must be (T, N, N) with T >= 1");
    if (n == 0) return;
    hipLaunchKernelGGL(k_shim_gather<T>, dim3((unsigned)((n + SH_TPB - 1) / SH_TPB)), dim3(SH_TPB), 0, st, Y, n, B, G, C, out);
    MG_HIP(hipGetLastError());
}

// returns the number of double partials written to `part` (the caller sums them in order)
template <class T> int launch_gather_forces(const T* Y, i64 n, const T* phi, int M, double xmin, double ymin, double width, int nb,
                                            int ni, T* F, double* part, hipStream_t st) {
    const ShimBox<T> B = make_box<T>(xmin, ymin, width, nb, ni);
    require(M >= B.N, "phi must be (4, M, M) with M >= N");
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)((n + SH_TPB - 1) / SH_TPB);
    hipLaunchKernelGGL(k_shim_forces<T>, dim3(blocks), dim3(SH_TPB), 0, st, Y, n, B, phi, M, F, part);
    MG_HIP(hipGetLastError());
    return (int)blocks;
}

template <class T> void launch_spectral_products(const T* K1, const T* K2, const T* W, i64 M, i64 Mh, T* out, hipStream_t st) {
    const i64 L = M * Mh;
    if (L == 0) return;
    hipLaunchKernelGGL(k_shim_products<T>, dim3(grid1(L)), dim3(SH_TPB), 0, st, K1, K2, W, L, out);
    MG_HIP(hipGetLastError());
}

#define MG_SHIM_INST(T) \
    template void launch_interp_scatter<T>(const T*, i64, const T*, int, double, double, double, int, int, T*, int*, hipStream_t); \
    template void launch_scatter_padded<T>(const T*, i64, double, double, double, int, int, int, T*, int*, hipStream_t); \
    template void launch_interp_gather<T>(const T*, i64, const T*, int, double, double, double, int, int, T*, hipStream_t); \
    template int launch_gather_forces<T>(const T*, i64, const T*, int, double, double, double, int, int, T*, double*, hipStream_t); \
    template void launch_spectral_products<T>(const T*, const T*, const T*, i64, i64, T*, hipStream_t);
MG_SHIM_INST(float)
MG_SHIM_INST(double)

}  // namespace mg
