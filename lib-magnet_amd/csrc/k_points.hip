// Per-point kernels of the t-SNE pass: bounding box, CSR attraction, gradient combination and the
// delta-bar-delta / momentum update with recentring. All reductions run in a fixed order (block tree,
// then the last block over the block partials in index order), so every result is deterministic.
#include "common.hpp"
#include "batch.hpp"
#include <algorithm>
#include <cmath>
#include <limits>

namespace mg {

namespace {

constexpr int TPB = 256;

// The last block of a grid of nblk blocks: every block publishes its partial, then bumps `counter`; the block
// that sees nblk - 1 reads all partials (in block order) and resets the counter for the next launch.
__device__ inline bool last_block(unsigned* counter, unsigned nblk) {
    __shared__ bool am_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) am_last = atomicAdd(counter, 1u) == nblk - 1;
    __syncthreads();
    if (am_last) __threadfence();
    return am_last;
}

template <class T> struct MinOp { __device__ T operator()(T a, T b) const { return a < b ? a : b; } };
template <class T> struct MaxOp { __device__ T operator()(T a, T b) const { return a > b ? a : b; } };
template <class T> struct AddOp { __device__ T operator()(T a, T b) const { return a + b; } };

// Tree reduction of one value per thread over a TPB-thread block; result valid in thread 0.
template <class T, class Op> __device__ T block_reduce(T v, T* sh, Op op) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = op(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const T r = sh[0];
    __syncthreads();
    return r;
}

template <class T> __device__ inline T pinf() { return std::numeric_limits<T>::infinity(); }

// compensated coordinate sums (for the mean) and bbox of a block: a butterfly within each wave (fixed order), then
// thread 0 merges the waves in index order; one barrier. The result is valid in thread 0.
template <class T> struct MeanBox { CSum sx, sy; T x0, y0, x1, y1; };
template <class T> __device__ inline MeanBox<T> block_meanbox(MeanBox<T> m) {
    __shared__ double sh[8][TPB / WAVE];
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        m.sx = csum_merge(m.sx, {__shfl_xor(m.sx.s, o), __shfl_xor(m.sx.c, o)});
        m.sy = csum_merge(m.sy, {__shfl_xor(m.sy.s, o), __shfl_xor(m.sy.c, o)});
        const T a = __shfl_xor(m.x0, o), b = __shfl_xor(m.y0, o), c = __shfl_xor(m.x1, o), d = __shfl_xor(m.y1, o);
        m.x0 = a < m.x0 ? a : m.x0; m.y0 = b < m.y0 ? b : m.y0;
        m.x1 = c > m.x1 ? c : m.x1; m.y1 = d > m.y1 ? d : m.y1;
    }
    const int w = threadIdx.x / WAVE;
    __syncthreads();   // the previous call's readers are done with sh
    if (threadIdx.x % WAVE == 0) {
        sh[0][w] = m.sx.s; sh[1][w] = m.sy.s; sh[2][w] = (double)m.x0; sh[3][w] = (double)m.y0; sh[4][w] = (double)m.x1;
        sh[5][w] = (double)m.y1; sh[6][w] = m.sx.c; sh[7][w] = m.sy.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < TPB / WAVE; ++k) {
            m.sx = csum_merge(m.sx, {sh[0][k], sh[6][k]}); m.sy = csum_merge(m.sy, {sh[1][k], sh[7][k]});
            const T a = (T)sh[2][k], b = (T)sh[3][k], c = (T)sh[4][k], d = (T)sh[5][k];
            m.x0 = a < m.x0 ? a : m.x0; m.y0 = b < m.y0 ? b : m.y0;
            m.x1 = c > m.x1 ? c : m.x1; m.y1 = d > m.y1 ? d : m.y1;
        }
    }
    return m;
}
// a block's partial: {sum x, sum y, xmin, ymin, xmax, ymax, error of sum x, error of sum y}
constexpr int MB_PART = 8;
template <class T> __device__ inline void write_meanbox(double* p, const MeanBox<T>& m) {
    p[0] = m.sx.s; p[1] = m.sy.s; p[2] = (double)m.x0; p[3] = (double)m.y0; p[4] = (double)m.x1; p[5] = (double)m.y1;
    p[6] = m.sx.c; p[7] = m.sy.c;
}

// ---- bbox_2d (src/fftrep.cpp:415-434)
template <class T>
__global__ __launch_bounds__(TPB) void k_bbox(const T* __restrict__ Y, i64 n, T* __restrict__ part,
                                              unsigned* counter, T* __restrict__ out) {
    __shared__ T sh[TPB];
    T x0 = pinf<T>(), y0 = pinf<T>(), x1 = -pinf<T>(), y1 = -pinf<T>();
    typedef typename vec2<T>::type V;
    const V* Y2 = reinterpret_cast<const V*>(Y);
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const V p = Y2[i];
        x0 = p.x < x0 ? p.x : x0; x1 = p.x > x1 ? p.x : x1;
        y0 = p.y < y0 ? p.y : y0; y1 = p.y > y1 ? p.y : y1;
    }
    x0 = block_reduce(x0, sh, MinOp<T>()); y0 = block_reduce(y0, sh, MinOp<T>());
    x1 = block_reduce(x1, sh, MaxOp<T>()); y1 = block_reduce(y1, sh, MaxOp<T>());
    if (threadIdx.x == 0) { T* p = part + 4 * blockIdx.x; p[0] = x0; p[1] = y0; p[2] = x1; p[3] = y1; }
    if (!last_block(counter, gridDim.x)) return;
    x0 = pinf<T>(); y0 = pinf<T>(); x1 = -pinf<T>(); y1 = -pinf<T>();
    for (unsigned b = threadIdx.x; b < gridDim.x; b += TPB) {
        const T* p = part + 4 * b;
        x0 = p[0] < x0 ? p[0] : x0; y0 = p[1] < y0 ? p[1] : y0;
        x1 = p[2] > x1 ? p[2] : x1; y1 = p[3] > y1 ? p[3] : y1;
    }
    x0 = block_reduce(x0, sh, MinOp<T>()); y0 = block_reduce(y0, sh, MinOp<T>());
    x1 = block_reduce(x1, sh, MaxOp<T>()); y1 = block_reduce(y1, sh, MaxOp<T>());
    if (threadIdx.x == 0) { out[0] = x0; out[1] = y0; out[2] = x1; out[3] = y1; *counter = 0u; }
}

// ---- tsne_attraction, 2-D path (src/fftrep.cpp:142-158): G lanes sweep one CSR row with stride G, then a
// butterfly over the G lanes (the same order on every run). Each lane takes U edges per round: their indices
// and weights are loaded together (streamed past the caches), then the U neighbour positions (L2-resident),
// so a round costs two memory latencies whatever the row length. Per lane the edges are summed in increasing
// order, as a one-edge-per-round sweep would.
template <class T, int G, int U>
__device__ __forceinline__ void attr2_block(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                            const T* __restrict__ P, const T* __restrict__ Y, i64 n, T inv_dof,
                                            T* __restrict__ out, int bid) {
    typedef typename vec2<T>::type V;
    const V* Y2 = reinterpret_cast<const V*>(Y);
    const i64 row = (i64)bid * (TPB / G) + threadIdx.x / G;
    const int sub = threadIdx.x % G;
    T ax = 0, ay = 0;
    if (row < n) {
        const V yi = Y2[row];
        const int e0 = ip[row], e1 = ip[row + 1];
        for (int base = e0 + sub; base < e1; base += G * U) {
            int j[U];
            T p[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = base + u * G;
                j[u] = e < e1 ? __builtin_nontemporal_load(ix + e) : -1;
                p[u] = e < e1 ? __builtin_nontemporal_load(P + e) : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j[u] < 0) break;
                const V yj = Y2[j[u]];
                const T dx = yi.x - yj.x, dy = yi.y - yj.y;
                const T w = p[u] / (T(1) + (dx * dx + dy * dy) * inv_dof);
                ax += w * dx; ay += w * dy;
            }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) { ax += __shfl_xor(ax, o, G); ay += __shfl_xor(ay, o, G); }
    if (row < n && sub == 0) { V r; r.x = ax; r.y = ay; reinterpret_cast<V*>(out)[row] = r; }
}
// Workgroups are dispatched to the 8 XCDs in turn; with a grid of a multiple of 8 blocks, XCD x takes the x-th
// eighth of the rows, so neighbouring rows (neighbouring points once the layout is relabelled in box order) share
// one XCD's L2.
template <class T, int G, int U>
__global__ __launch_bounds__(TPB) void k_attr2(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                               const T* __restrict__ P, const T* __restrict__ Y, i64 n, T inv_dof,
                                               T* __restrict__ out) {
    const int b = (gridDim.x % 8) == 0 ? (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8) : (int)blockIdx.x;
    attr2_block<T, G, U>(ip, ix, P, Y, n, inv_dof, out, b);
}
template <class T, int G, int U>
__global__ __launch_bounds__(TPB) void kb_attr2(const BInst<T>* __restrict__ I, BMap m, const T* __restrict__ Y, T inv_dof,
                                                T* __restrict__ out) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    attr2_block<T, G, U>(b.ip, b.ix, b.P, Y + 2 * b.off, b.n, inv_dof, out + 2 * b.off, blockIdx.x - m.start[k]);
}

// The same sweep with a window of the layout in LDS. Once the points are renumbered in the order of a
// space-filling curve over the grid boxes, the neighbours of a run of rows mostly sit within a few thousand
// positions of it. A block of TW threads takes RB consecutive rows (G lanes per row, TW / G rows at a time),
// first copies the WIN positions centred on them into LDS (coalesced), then serves every neighbour inside the
// window from LDS and only the others from L2: the random 8-byte gathers, whose request rate bounds the sweep,
// become LDS reads. Each lane sums its edges in the order of attr2_block, so the result is the same bit for bit.
template <class T, int G, int U, int WIN, int TW>
__global__ __launch_bounds__(TW) void k_attr2w(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                               const T* __restrict__ P, const T* __restrict__ Y, i64 n, T inv_dof,
                                               T* __restrict__ out, int RB) {
    typedef typename vec2<T>::type V;
    extern __shared__ __align__(16) unsigned char smem_w[];
    V* sY = reinterpret_cast<V*>(smem_w);
    const V* Y2 = reinterpret_cast<const V*>(Y);
    const int b = (gridDim.x % 8) == 0 ? (int)((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8) : (int)blockIdx.x;
    const i64 r0 = (i64)b * RB, r1 = r0 + RB < n ? r0 + RB : n;
    if (r0 >= n) return;
    i64 w0 = (r0 + r1) / 2 - WIN / 2;
    w0 = w0 > n - WIN ? n - WIN : w0;
    w0 = w0 < 0 ? 0 : w0;
    const int wn = WIN == 0 ? 0 : (int)(n - w0 < WIN ? n - w0 : WIN);   // WIN = 0: no window, rows only
    if (WIN > 0) {
        for (int k = threadIdx.x; k < wn; k += TW) sY[k] = Y2[w0 + k];
        __syncthreads();
    }
    const int sub = threadIdx.x % G;
    for (i64 row = r0 + threadIdx.x / G; row < r1; row += TW / G) {   // every lane of a row group takes the same rows
        T ax = 0, ay = 0;
        const i64 li = row - w0;
        const V yi = (li >= 0 && li < wn) ? sY[li] : Y2[row];
        const int e0 = ip[row], e1 = ip[row + 1];
        for (int base = e0 + sub; base < e1; base += G * U) {
            int j[U];
            T p[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = base + u * G;
                j[u] = e < e1 ? __builtin_nontemporal_load(ix + e) : -1;
                p[u] = e < e1 ? __builtin_nontemporal_load(P + e) : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j[u] < 0) break;
                const i64 lj = (i64)j[u] - w0;
                const V yj = (lj >= 0 && lj < wn) ? sY[lj] : Y2[j[u]];
                const T dx = yi.x - yj.x, dy = yi.y - yj.y;
                const T w = p[u] / (T(1) + (dx * dx + dy * dy) * inv_dof);
                ax += w * dx; ay += w * dy;
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) { ax += __shfl_xor(ax, o, G); ay += __shfl_xor(ay, o, G); }
        if (sub == 0) { V r; r.x = ax; r.y = ay; reinterpret_cast<V*>(out)[row] = r; }
    }
}

// any dim in [1, 64] (src/fftrep.cpp:160-176): one thread per row
template <class T>
__global__ __launch_bounds__(TPB) void k_attr_any(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                  const T* __restrict__ P, const T* __restrict__ Y, i64 n, int dim,
                                                  T inv_dof, T* __restrict__ out) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    T acc[64];
    for (int d = 0; d < dim; ++d) acc[d] = 0;
    const T* yi = Y + i * dim;
    for (int e = ip[i]; e < ip[i + 1]; ++e) {
        const T* yj = Y + (i64)ix[e] * dim;
        T d2 = 0;
        for (int d = 0; d < dim; ++d) { const T df = yi[d] - yj[d]; d2 += df * df; }
        const T w = P[e] / (T(1) + d2 * inv_dof);
        for (int d = 0; d < dim; ++d) acc[d] += w * (yi[d] - yj[d]);
    }
    for (int d = 0; d < dim; ++d) out[i * dim + d] = acc[d];
}

// delta-bar-delta gain, momentum and position of one coordinate (src/fftrep.cpp:245-255)
template <class T> __device__ inline T dbd(T g_, T& u, T& gain, T lr, T mom) {
    const bool same = (g_ > T(0)) == (u > T(0));
    T g = same ? gain * T(0.8) : gain + T(0.2);
    if (g < T(0.01)) g = T(0.01);
    gain = g;
    u = mom * u - lr * g * g_;
    return u;
}

// ---- fused step / update for 2-D layouts, then the mean (double block sums) and the bbox of the
// un-centred layout; the last block writes mb = {mean_x, mean_y, xmin, ymin, xmax, ymax} of the
// centred layout (rounding is monotone, so min(y_i - m) = min(y_i) - m exactly).
template <class T, int MODE>
__device__ __forceinline__ void update2_block(T* __restrict__ Y, T* __restrict__ U, T* __restrict__ Gn,
                                              const T* __restrict__ A, const T* __restrict__ F,
                                              const T* __restrict__ d_inv, T inv_h, T rho, T lr, T mom, i64 n,
                                              double* __restrict__ part, int bid, int nblk) {
    typedef typename vec2<T>::type V;
    const T inv = d_inv ? *d_inv : inv_h;
    CSum sx = {0.0, 0.0}, sy = {0.0, 0.0};
    T x0 = pinf<T>(), y0 = pinf<T>(), x1 = -pinf<T>(), y1 = -pinf<T>();
    V* Y2 = reinterpret_cast<V*>(Y); V* U2 = reinterpret_cast<V*>(U); V* G2 = reinterpret_cast<V*>(Gn);
    const V* A2 = reinterpret_cast<const V*>(A); const V* F2 = reinterpret_cast<const V*>(F);
    for (i64 i = (i64)bid * TPB + threadIdx.x; i < n; i += (i64)nblk * TPB) {
        const V a = A2[i];
        V gx;
        if (MODE == 0) { const V f = F2[i]; gx.x = rho * a.x - f.x * inv; gx.y = rho * a.y - f.y * inv; }
        else gx = a;
        V u = U2[i], g = G2[i], y = Y2[i];
        y.x = y.x + dbd(gx.x, u.x, g.x, lr, mom);
        y.y = y.y + dbd(gx.y, u.y, g.y, lr, mom);
        U2[i] = u; G2[i] = g; Y2[i] = y;
        csum_add(sx, (double)y.x); csum_add(sy, (double)y.y);
        x0 = y.x < x0 ? y.x : x0; x1 = y.x > x1 ? y.x : x1;
        y0 = y.y < y0 ? y.y : y0; y1 = y.y > y1 ? y.y : y1;
    }
    if (part == nullptr) return;   // partials taken in original point order by k_meanbox_perm
    MeanBox<T> m = block_meanbox<T>({sx, sy, x0, y0, x1, y1});
    if (threadIdx.x == 0) write_meanbox<T>(part + MB_PART * bid, m);
}
template <class T, int MODE>
__global__ __launch_bounds__(TPB) void k_update2(T* __restrict__ Y, T* __restrict__ U, T* __restrict__ Gn,
                                                 const T* __restrict__ A, const T* __restrict__ F,
                                                 const T* __restrict__ d_inv, T inv_h, T rho, T lr, T mom, i64 n,
                                                 double* __restrict__ part) {
    update2_block<T, MODE>(Y, U, Gn, A, F, d_inv, inv_h, rho, lr, mom, n, part, blockIdx.x, gridDim.x);
}

// The fused update of a relabelled layout (position inv[i] holds original point i) whose points all sit in L2: block b
// updates the points b * TPB + t (+ k * nblk * TPB) of the original numbering, where they now are, and takes the mean
// and bbox partials of the updated positions in that order -- the partials k_meanbox_perm takes after k_update2, in one
// launch. Each point's update is its own, so the order changes nothing in it.
template <class T, int MODE>
__global__ __launch_bounds__(TPB) void k_update2_perm(T* __restrict__ Y, T* __restrict__ U, T* __restrict__ Gn,
                                                      const T* __restrict__ A, const T* __restrict__ F,
                                                      const T* __restrict__ d_inv, T inv_h, T rho, T lr, T mom, i64 n,
                                                      const int* __restrict__ inv, double* __restrict__ part) {
    typedef typename vec2<T>::type V;
    const T invq = d_inv ? *d_inv : inv_h;
    const int bid = blockIdx.x, nblk = gridDim.x;
    CSum sx = {0.0, 0.0}, sy = {0.0, 0.0};
    T x0 = pinf<T>(), y0 = pinf<T>(), x1 = -pinf<T>(), y1 = -pinf<T>();
    V* Y2 = reinterpret_cast<V*>(Y); V* U2 = reinterpret_cast<V*>(U); V* G2 = reinterpret_cast<V*>(Gn);
    const V* A2 = reinterpret_cast<const V*>(A); const V* F2 = reinterpret_cast<const V*>(F);
    for (i64 i = (i64)bid * TPB + threadIdx.x; i < n; i += (i64)nblk * TPB) {
        const i64 q = inv[i];
        const V a = A2[q];
        V gx;
        if (MODE == 0) { const V f = F2[q]; gx.x = rho * a.x - f.x * invq; gx.y = rho * a.y - f.y * invq; }
        else gx = a;
        V u = U2[q], g = G2[q], y = Y2[q];
        y.x = y.x + dbd(gx.x, u.x, g.x, lr, mom);
        y.y = y.y + dbd(gx.y, u.y, g.y, lr, mom);
        U2[q] = u; G2[q] = g; Y2[q] = y;
        csum_add(sx, (double)y.x); csum_add(sy, (double)y.y);
        x0 = y.x < x0 ? y.x : x0; x1 = y.x > x1 ? y.x : x1;
        y0 = y.y < y0 ? y.y : y0; y1 = y.y > y1 ? y.y : y1;
    }
    const MeanBox<T> m = block_meanbox<T>({sx, sy, x0, y0, x1, y1});
    if (threadIdx.x == 0) write_meanbox<T>(part + MB_PART * bid, m);
}

// The update's block partials of a relabelled layout, in original point order: block b sums the points
// b * TPB + t (+ k * nblk * TPB) of the original numbering, as the update of the unrelabelled layout does.
template <class T>
__global__ __launch_bounds__(TPB) void k_meanbox_perm(const T* __restrict__ Y, const int* __restrict__ inv, i64 n,
                                                      double* __restrict__ part) {
    typedef typename vec2<T>::type V;
    const int bid = blockIdx.x, nblk = gridDim.x;
    CSum sx = {0.0, 0.0}, sy = {0.0, 0.0};
    T x0 = pinf<T>(), y0 = pinf<T>(), x1 = -pinf<T>(), y1 = -pinf<T>();
    for (i64 i = (i64)bid * TPB + threadIdx.x; i < n; i += (i64)nblk * TPB) {
        const V y = reinterpret_cast<const V*>(Y)[inv[i]];
        csum_add(sx, (double)y.x); csum_add(sy, (double)y.y);
        x0 = y.x < x0 ? y.x : x0; x1 = y.x > x1 ? y.x : x1;
        y0 = y.y < y0 ? y.y : y0; y1 = y.y > y1 ? y.y : y1;
    }
    const MeanBox<T> m = block_meanbox<T>({sx, sy, x0, y0, x1, y1});
    if (threadIdx.x == 0) write_meanbox<T>(part + MB_PART * bid, m);
}

// Recentring after the update: every block reduces the update's nblk block partials itself, in the same fixed
// order (so all blocks get the same mean), and block 0 writes mb = {mean_x, mean_y, xmin, ymin, xmax, ymax} of
// the centred layout (rounding is monotone, so min(y_i - m) = min(y_i) - m exactly). No block waits for another.
template <class T>
__device__ __forceinline__ void center2_block(T* __restrict__ Y, i64 n, const double* __restrict__ part, T* __restrict__ mb,
                                              int bid, int nblk, T* hmb = nullptr, int* htag = nullptr, int tag = 0) {
    __shared__ T s_m[2];
    typedef typename vec2<T>::type V;
    CSum sx = {0.0, 0.0}, sy = {0.0, 0.0};
    T x0 = pinf<T>(), y0 = pinf<T>(), x1 = -pinf<T>(), y1 = -pinf<T>();
    for (int b = threadIdx.x; b < nblk; b += TPB) {
        const double* p = part + MB_PART * b;
        sx = csum_merge(sx, {p[0], p[6]}); sy = csum_merge(sy, {p[1], p[7]});
        x0 = (T)p[2] < x0 ? (T)p[2] : x0; y0 = (T)p[3] < y0 ? (T)p[3] : y0;
        x1 = (T)p[4] > x1 ? (T)p[4] : x1; y1 = (T)p[5] > y1 ? (T)p[5] : y1;
    }
    const MeanBox<T> m = block_meanbox<T>({sx, sy, x0, y0, x1, y1});
    if (threadIdx.x == 0) {
        const T mx = (T)((m.sx.s + m.sx.c) / (double)n), my = (T)((m.sy.s + m.sy.c) / (double)n);
        s_m[0] = mx; s_m[1] = my;
        if (bid == 0) {
            mb[0] = mx; mb[1] = my; mb[2] = m.x0 - mx; mb[3] = m.y0 - my; mb[4] = m.x1 - mx; mb[5] = m.y1 - my;
            if (hmb) {   // the same into pinned host memory, then the tag the host waits for
                hmb[0] = mb[0]; hmb[1] = mb[1]; hmb[2] = mb[2]; hmb[3] = mb[3]; hmb[4] = mb[4]; hmb[5] = mb[5];
                __threadfence_system();
                __hip_atomic_store(htag, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
    __syncthreads();
    const T mx = s_m[0], my = s_m[1];
    V* Y2 = reinterpret_cast<V*>(Y);
    for (i64 i = (i64)bid * TPB + threadIdx.x; i < n; i += (i64)nblk * TPB) {
        V y = Y2[i]; y.x -= mx; y.y -= my; Y2[i] = y;
    }
}
template <class T>
__global__ __launch_bounds__(TPB) void k_center2(T* __restrict__ Y, i64 n, const double* __restrict__ part,
                                                 T* __restrict__ mb, T* __restrict__ hmb, int* __restrict__ htag, int tag) {
    center2_block<T>(Y, n, part, mb, blockIdx.x, gridDim.x, hmb, htag, tag);
}

// batched fused step of the t-SNE loop (gradient rho * attr - F / sumQ of each active instance, its iteration's
// exaggeration and momentum from the table) and its recentring
template <class T>
__global__ __launch_bounds__(TPB) void kb_update2(const BInst<T>* __restrict__ I, BMap m, T* __restrict__ Y, T* __restrict__ U,
                                                  T* __restrict__ Gn, const T* __restrict__ A, const T* __restrict__ F) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active) return;
    const i64 o = 2 * b.off, n = b.n;
    const T lr = b.lr, rho = b.rho, mom = b.mom;
    T* inv = b.inv_out;
    double* part = b.upart;
    update2_block<T, 0>(Y + o, U + o, Gn + o, A + o, F + o, inv, T(0), rho, lr, mom, n, part,
                        blockIdx.x - m.start[k], m.start[k + 1] - m.start[k]);
}
template <class T>
__global__ __launch_bounds__(TPB) void kb_center2(const BInst<T>* __restrict__ I, BMap m, T* __restrict__ Y) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active) return;
    const i64 off = b.off, n = b.n;
    const double* part = b.upart;
    T* mb = b.mb;
    center2_block<T>(Y + 2 * off, n, part, mb, blockIdx.x - m.start[k], m.start[k + 1] - m.start[k]);
}

// ---- any dim: elementwise update, then one block per dimension for the mean, then centring
template <class T, int MODE>
__global__ __launch_bounds__(TPB) void k_update_any(T* __restrict__ Y, T* __restrict__ U, T* __restrict__ Gn,
                                                    const T* __restrict__ A, const T* __restrict__ F, T inv, T rho,
                                                    T lr, T mom, i64 total) {
    for (i64 k = (i64)blockIdx.x * TPB + threadIdx.x; k < total; k += (i64)gridDim.x * TPB) {
        const T g_ = MODE == 0 ? rho * A[k] - F[k] * inv : A[k];
        T u = U[k], g = Gn[k];
        Y[k] = Y[k] + dbd(g_, u, g, lr, mom);
        U[k] = u; Gn[k] = g;
    }
}
template <class T>
__global__ __launch_bounds__(TPB) void k_colmean(const T* __restrict__ Y, i64 n, int dim, T* __restrict__ mean) {
    __shared__ double shd[TPB];
    const int d = blockIdx.x;
    double s = 0;
    for (i64 i = threadIdx.x; i < n; i += TPB) s += (double)Y[i * dim + d];
    s = block_reduce(s, shd, AddOp<double>());
    if (threadIdx.x == 0) mean[d] = (T)(s / (double)n);
}
template <class T>
__global__ __launch_bounds__(TPB) void k_center_any(T* __restrict__ Y, i64 total, int dim, const T* __restrict__ mean) {
    for (i64 k = (i64)blockIdx.x * TPB + threadIdx.x; k < total; k += (i64)gridDim.x * TPB) Y[k] -= mean[k % dim];
}

// ---- gradient of the non-fused path (magnet/tsne.py:185-198): rho * attr - F * inv, plus the
// mid-range spring w (Y_i - mean of the partners) of src/fftrep.cpp:278-285 when pairs != null.
template <class T>
__global__ __launch_bounds__(TPB) void k_grad(const T* __restrict__ A, const T* __restrict__ F, const T* __restrict__ d_inv,
                                              T inv_h, T rho, const T* __restrict__ Y, const int32_t* __restrict__ pairs,
                                              int m, T w, i64 n, int dim, int combine, T* __restrict__ grad) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const T inv = d_inv ? *d_inv : inv_h;
    for (int d = 0; d < dim; ++d) {
        const i64 k = i * dim + d;
        T g = combine ? rho * A[k] - F[k] * inv : grad[k];
        if (pairs) {
            T acc = 0;
            for (int q = 0; q < m; ++q) acc += Y[(i64)pairs[i * m + q] * dim + d];
            g += w * (Y[k] - acc / (T)m);
        }
        grad[k] = g;
    }
}

// ---- per-iteration diagnostics of the loop (no counterpart in the reference): the KL divergence of P and Q and the
// gradient norm, in FP64 with compensated sums, over the points and edges in original order (so a relabelled run
// gives the same values). KL = sum p log p - sum p log q with q_ij = (1 + d_ij^2 / dof)^-dof / Z, i.e.
// sum p log p + dof sum p log1p(d^2 / dof) + (sum p) log Z.
constexpr int TRACE_BLOCKS = 1024;
__device__ inline CSum block_csum_pts(CSum v, double* sh) {   // sh: 2 * TPB / WAVE doubles
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = csum_merge(v, {__shfl_xor(v.s, o), __shfl_xor(v.c, o)});
    __syncthreads();
    if (threadIdx.x % WAVE == 0) { sh[2 * (threadIdx.x / WAVE)] = v.s; sh[2 * (threadIdx.x / WAVE) + 1] = v.c; }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < TPB / WAVE; ++k) v = csum_merge(v, {sh[2 * k], sh[2 * k + 1]});
    return v;
}

// sum p log p and sum p over the CSR values: block partials {s, c} x 2
template <class T>
__global__ __launch_bounds__(TPB) void k_plogp(const T* __restrict__ P, i64 nnz, double* __restrict__ part) {
    __shared__ double sh[2 * TPB / WAVE];
    CSum a = {0, 0}, b = {0, 0};
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < nnz; e += (i64)gridDim.x * TPB) {
        const double p = (double)P[e];
        if (p > 0) csum_add(a, p * log(p));
        csum_add(b, p);
    }
    a = block_csum_pts(a, sh);
    b = block_csum_pts(b, sh);
    if (threadIdx.x == 0) { double* q = part + 4 * blockIdx.x; q[0] = a.s; q[1] = a.c; q[2] = b.s; q[3] = b.c; }
}

// sum_ij p_ij log1p(d_ij^2 / dof), one wave per row of the original CSR (Y read through inv: position of a point)
template <class T>
__global__ __launch_bounds__(TPB) void k_kl_edges(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                  const T* __restrict__ P, const T* __restrict__ Y, const int* __restrict__ inv,
                                                  i64 n, double dof, double* __restrict__ part) {
    __shared__ double sh[2 * TPB / WAVE];
    typedef typename vec2<T>::type V;
    const V* Y2 = reinterpret_cast<const V*>(Y);
    const int lane = threadIdx.x % WAVE;
    CSum acc = {0, 0};
    for (i64 row = (i64)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE; row < n; row += (i64)gridDim.x * (TPB / WAVE)) {
        const V yi = Y2[inv ? inv[row] : row];
        double s = 0;
        for (int e = ip[row] + lane; e < ip[row + 1]; e += WAVE) {
            const int j = ix[e];
            const V yj = Y2[inv ? inv[j] : j];
            const double dx = (double)yi.x - (double)yj.x, dy = (double)yi.y - (double)yj.y;
            s += (double)P[e] * log1p((dx * dx + dy * dy) / dof);
        }
        csum_add(acc, s);
    }
    acc = block_csum_pts(acc, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = acc.s; part[2 * blockIdx.x + 1] = acc.c; }
}

// sum_i |g_i|^2 of the fused gradient rho A - F / sumQ (F null: g = A, the gradient itself), points in original order
template <class T>
__global__ __launch_bounds__(TPB) void k_grad_sq(const T* __restrict__ A, const T* __restrict__ F, const T* __restrict__ d_inv,
                                                 T rho, const int* __restrict__ inv, i64 n, double* __restrict__ part) {
    __shared__ double sh[2 * TPB / WAVE];
    typedef typename vec2<T>::type V;
    const T iz = F ? *d_inv : T(0);
    CSum acc = {0, 0};
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const i64 k = inv ? inv[i] : i;
        const V a = reinterpret_cast<const V*>(A)[k];
        V g = a;
        if (F) { const V f = reinterpret_cast<const V*>(F)[k]; g.x = rho * a.x - f.x * iz; g.y = rho * a.y - f.y * iz; }
        csum_add(acc, (double)g.x * (double)g.x + (double)g.y * (double)g.y);
    }
    acc = block_csum_pts(acc, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = acc.s; part[2 * blockIdx.x + 1] = acc.c; }
}

// out = {KL, |g|}: the partials merged in block order; zout[0] is the iteration's sumQ = Z, pp the sum p log p and
// sum p partials
__global__ __launch_bounds__(TPB) void k_trace_final(const double* __restrict__ kl_part, const double* __restrict__ g_part,
                                                     const double* __restrict__ pp_part, const double* __restrict__ zout,
                                                     double dof, double* __restrict__ out) {
    __shared__ double sh[2 * TPB / WAVE];
    CSum k = {0, 0}, g = {0, 0}, pl = {0, 0}, ps = {0, 0};
    for (int b = threadIdx.x; b < TRACE_BLOCKS; b += TPB) {
        k = csum_merge(k, {kl_part[2 * b], kl_part[2 * b + 1]});
        g = csum_merge(g, {g_part[2 * b], g_part[2 * b + 1]});
        pl = csum_merge(pl, {pp_part[4 * b], pp_part[4 * b + 1]});
        ps = csum_merge(ps, {pp_part[4 * b + 2], pp_part[4 * b + 3]});
    }
    k = block_csum_pts(k, sh);
    g = block_csum_pts(g, sh);
    pl = block_csum_pts(pl, sh);
    ps = block_csum_pts(ps, sh);
    if (threadIdx.x == 0) {
        const double Z = zout[0] > 1e-300 ? zout[0] : 1e-300;
        CSum kl = csum_merge(pl, {dof * k.s, dof * k.c});
        kl = csum_merge(kl, {(ps.s + ps.c) * log(Z), 0.0});
        out[0] = kl.s + kl.c;
        out[1] = sqrt(g.s + g.c);
    }
}

template <class T>
__global__ __launch_bounds__(TPB) void k_fill(T* __restrict__ p, T v, i64 count) {
    for (i64 k = (i64)blockIdx.x * TPB + threadIdx.x; k < count; k += (i64)gridDim.x * TPB) p[k] = v;
}

inline unsigned grid_for(i64 work, unsigned cap = 4096) {
    i64 b = (work + TPB - 1) / TPB;
    if (b < 1) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace

int update_blocks(i64 n) { return (int)grid_for(n, 1024); }

template <class T> void launch_bbox(const T* Y, i64 n, T* part, unsigned* counter, T* out, hipStream_t st) {
    require(n >= 1, "Y must be (n, 2) with n >= 1");
    hipLaunchKernelGGL(k_bbox<T>, dim3(grid_for(n, 1024)), dim3(TPB), 0, st, Y, n, part, counter, out);
    MG_HIP(hipGetLastError());
}

// lanes per row: up to a whole wave, about four edges per lane; the long rows of perplexity affinities get a
// wave each (random neighbour reads from L2 bound the sweep: more rows per wave measured slower)
int attraction_lanes(i64 n, i64 nnz) {
    const double avg = (double)nnz / (double)(n > 0 ? n : 1);
    int G = 1;
    while (G < 64 && G * 4 < avg) G <<= 1;
    return G;
}

// window of the LDS sweep (points): 128 KiB of positions, the most one block may hold with room for nothing else
template <class T> constexpr int attr_window() { return (128 * 1024) / (2 * (int)sizeof(T)); }

template <class F> static void set_dyn_lds(F* f, int bytes) {
    if (bytes > 65536) kernel_lds_limit(reinterpret_cast<const void*>(f), bytes);
}

template <class T> void launch_attraction(const int32_t* ip, const int32_t* ix, const T* P, const T* Y, i64 n, int dim,
                                          T inv_dof, T* out, i64 nnz, hipStream_t st) {
    if (n == 0) return;
    const Tunables& tu = tunables();
    // measured: the window pays at 1M points (FP32 862 -> 682 us, FP64 1021 -> 924 us per sweep, in the loop's box
    // order), and costs at 70k, where too few blocks of its size fill the device
    const bool windowed = dim == 2 && (tu.attr_window > 0 || (tu.attr_window < 0 && n >= 262144));
    if (windowed) {
        const int G = attraction_lanes(n, nnz);
        const int big = tu.attr_win_kb >= 128, TW = tu.attr_tpb >= 1024 ? 1024 : tu.attr_tpb >= 512 ? 512 : 256;
        const int RB = std::max(TW / G, tu.attr_rows > 0 ? tu.attr_rows : 1024);
        const unsigned blocks = (unsigned)((n + RB - 1) / RB);
#define MG_ATTRW_LAUNCH(g, WIN, TWC) do { const int lds = (WIN) * 2 * (int)sizeof(T); set_dyn_lds(k_attr2w<T, g, 8, WIN, TWC>, lds); \
            hipLaunchKernelGGL((k_attr2w<T, g, 8, WIN, TWC>), dim3(blocks), dim3(TWC), lds, st, ip, ix, P, Y, n, inv_dof, out, RB); } while (0)
#define MG_ATTRW_CASE(g) case g: \
            if (tu.attr_win_kb == 0) { if (TW == 256) MG_ATTRW_LAUNCH(g, 0, 256); else if (TW == 512) MG_ATTRW_LAUNCH(g, 0, 512); else MG_ATTRW_LAUNCH(g, 0, 1024); } \
            else if (big) { if (TW == 1024) MG_ATTRW_LAUNCH(g, attr_window<T>(), 1024); else MG_ATTRW_LAUNCH(g, attr_window<T>(), 512); } \
            else { if (TW == 1024) MG_ATTRW_LAUNCH(g, attr_window<T>() / 2, 1024); else MG_ATTRW_LAUNCH(g, attr_window<T>() / 2, 512); } \
            break;
        switch (G) {
            MG_ATTRW_CASE(1) MG_ATTRW_CASE(2) MG_ATTRW_CASE(4) MG_ATTRW_CASE(8) MG_ATTRW_CASE(16) MG_ATTRW_CASE(32) MG_ATTRW_CASE(64)
        }
#undef MG_ATTRW_CASE
#undef MG_ATTRW_LAUNCH
    } else if (dim == 2) {
        const int G = attraction_lanes(n, nnz);
        const unsigned blocks = (unsigned)((attraction_blocks(n, G) + 7) / 8 * 8);   // rows past n: no work
        switch (G) {
#define MG_ATTR_CASE(g) case g: hipLaunchKernelGGL((k_attr2<T, g, 8>), dim3(blocks), dim3(TPB), 0, st, ip, ix, P, Y, n, inv_dof, out); break;
            MG_ATTR_CASE(1) MG_ATTR_CASE(2) MG_ATTR_CASE(4) MG_ATTR_CASE(8) MG_ATTR_CASE(16) MG_ATTR_CASE(32) MG_ATTR_CASE(64)
#undef MG_ATTR_CASE
        }
    } else {
        hipLaunchKernelGGL(k_attr_any<T>, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, ip, ix, P, Y, n, dim, inv_dof, out);
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_attraction_batch(const BInst<T>* I, const BMap& m, int G, const T* Y, T inv_dof, T* out,
                                                hipStream_t st) {
    if (m.count == 0 || m.start[m.count] == 0) return;
    const dim3 grid((unsigned)m.start[m.count]);
    switch (G) {
#define MG_ATTR_CASE(g) case g: hipLaunchKernelGGL((kb_attr2<T, g, 8>), grid, dim3(TPB), 0, st, I, m, Y, inv_dof, out); break;
        MG_ATTR_CASE(1) MG_ATTR_CASE(2) MG_ATTR_CASE(4) MG_ATTR_CASE(8) MG_ATTR_CASE(16) MG_ATTR_CASE(32) MG_ATTR_CASE(64)
#undef MG_ATTR_CASE
        default: throw Error(1, "lanes per row must be a power of two <= 64");
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_update_batch(const BInst<T>* I, const BMap& m, T* Y, T* U, T* G, const T* A, const T* F,
                                            hipStream_t st) {
    if (m.count == 0 || m.start[m.count] == 0) return;
    const dim3 grid((unsigned)m.start[m.count]);
    hipLaunchKernelGGL(kb_update2<T>, grid, dim3(TPB), 0, st, I, m, Y, U, G, A, F);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(kb_center2<T>, grid, dim3(TPB), 0, st, I, m, Y);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_update(int mode, T* Y, T* U, T* G, const T* A, const T* F, const T* d_inv, T inv_h, T rho,
                                      T lr, T mom, const int32_t* pairs, int m, T w_mid, i64 n, int dim, double* part,
                                      T* mean_bbox, const int* inv, hipStream_t st, T* host_mb, int* host_tag, int tag) {
    (void)pairs; (void)m; (void)w_mid;
    if (n == 0) return;
    if (dim == 2) {
        const unsigned blocks = grid_for(n, 1024);
        if (inv && n < 262144) {   // the layout sits in L2: the update in original order, partials in one launch
            if (mode == 0) hipLaunchKernelGGL((k_update2_perm<T, 0>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, d_inv, inv_h, rho, lr, mom, n, inv, part);
            else hipLaunchKernelGGL((k_update2_perm<T, 1>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, d_inv, inv_h, rho, lr, mom, n, inv, part);
            MG_HIP(hipGetLastError());
        } else {
            double* upart = inv ? nullptr : part;
            if (mode == 0) hipLaunchKernelGGL((k_update2<T, 0>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, d_inv, inv_h, rho, lr, mom, n, upart);
            else hipLaunchKernelGGL((k_update2<T, 1>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, d_inv, inv_h, rho, lr, mom, n, upart);
            MG_HIP(hipGetLastError());
            if (inv) {
                hipLaunchKernelGGL(k_meanbox_perm<T>, dim3(blocks), dim3(TPB), 0, st, (const T*)Y, inv, n, part);
                MG_HIP(hipGetLastError());
            }
        }
        hipLaunchKernelGGL(k_center2<T>, dim3(blocks), dim3(TPB), 0, st, Y, n, (const double*)part, mean_bbox, host_mb, host_tag, tag);
    } else {
        require(d_inv == nullptr && inv == nullptr, "device-side inv_sumQ and relabelling need a 2-D layout");
        const i64 total = n * dim;
        const unsigned blocks = grid_for(total, 4096);
        if (mode == 0) hipLaunchKernelGGL((k_update_any<T, 0>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, inv_h, rho, lr, mom, total);
        else hipLaunchKernelGGL((k_update_any<T, 1>), dim3(blocks), dim3(TPB), 0, st, Y, U, G, A, F, inv_h, rho, lr, mom, total);
        MG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_colmean<T>, dim3(dim), dim3(TPB), 0, st, Y, n, dim, mean_bbox);
        MG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_center_any<T>, dim3(blocks), dim3(TPB), 0, st, Y, total, dim, mean_bbox);
    }
    MG_HIP(hipGetLastError());
}

template <class T> void launch_grad(const T* A, const T* F, const T* d_inv, T inv_h, T rho, const T* Y, const int32_t* pairs,
                                    int m, T w, i64 n, int dim, int combine, T* grad, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_grad<T>, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, A, F, d_inv, inv_h, rho, Y, pairs, m, w, n, dim, combine, grad);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_mid_spring(const T* Y, const int32_t* pairs, T* grad, i64 n, int dim, int m, T w, hipStream_t st) {
    launch_grad<T>(nullptr, nullptr, nullptr, T(0), T(0), Y, pairs, m, w, n, dim, 0, grad, st);
}

int trace_scratch_doubles() { return 8 * TRACE_BLOCKS; }

template <class T> void launch_trace_plogp(const T* P, i64 nnz, double* scratch, hipStream_t st) {
    hipLaunchKernelGGL(k_plogp<T>, dim3(TRACE_BLOCKS), dim3(TPB), 0, st, P, nnz, scratch + 4 * TRACE_BLOCKS);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_trace(const int32_t* ip, const int32_t* ix, const T* P, const T* Y, const int* inv, i64 n,
                                     double dof, const T* A, const T* F, const T* d_inv, T rho, const double* zout,
                                     double* scratch, double* out, hipStream_t st) {
    double* kl = scratch;
    double* g = scratch + 2 * TRACE_BLOCKS;
    const double* pp = scratch + 4 * TRACE_BLOCKS;
    hipLaunchKernelGGL(k_kl_edges<T>, dim3(TRACE_BLOCKS), dim3(TPB), 0, st, ip, ix, P, Y, inv, n, dof, kl);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_grad_sq<T>, dim3(TRACE_BLOCKS), dim3(TPB), 0, st, A, F, d_inv, rho, inv, n, g);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_trace_final, dim3(1), dim3(TPB), 0, st, (const double*)kl, (const double*)g, pp, zout, dof, out);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_fill(T* p, T v, i64 count, hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(count, 4096)), dim3(TPB), 0, st, p, v, count);
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template void launch_bbox<T>(const T*, i64, T*, unsigned*, T*, hipStream_t); \
    template void launch_attraction<T>(const int32_t*, const int32_t*, const T*, const T*, i64, int, T, T*, i64, hipStream_t); \
    template void launch_update<T>(int, T*, T*, T*, const T*, const T*, const T*, T, T, T, T, const int32_t*, int, T, i64, int, double*, T*, const int*, hipStream_t, T*, int*, int); \
    template void launch_grad<T>(const T*, const T*, const T*, T, T, const T*, const int32_t*, int, T, i64, int, int, T*, hipStream_t); \
    template void launch_mid_spring<T>(const T*, const int32_t*, T*, i64, int, int, T, hipStream_t); \
    template void launch_fill<T>(T*, T, i64, hipStream_t); \
    template void launch_trace_plogp<T>(const T*, i64, double*, hipStream_t); \
    template void launch_trace<T>(const int32_t*, const int32_t*, const T*, const T*, const int*, i64, double, const T*, const T*, const T*, T, const double*, double*, double*, hipStream_t); \
    template void launch_attraction_batch<T>(const BInst<T>*, const BMap&, int, const T*, T, T*, hipStream_t); \
    template void launch_update_batch<T>(const BInst<T>*, const BMap&, T*, T*, T*, const T*, const T*, hipStream_t);
MG_INST(float)
MG_INST(double)
template void launch_fill<int>(int*, int, i64, hipStream_t);
template void launch_fill<unsigned>(unsigned*, unsigned, i64, hipStream_t);

}  // namespace mg
