// Point <-> grid kernels of the FFT repulsion (src/fftconv.cpp:83-123 and :200-229):
//   1. box sort: a stable LSD radix sort of (box key, point index), 8-bit digits, ranks by wave ballots
//      (no float atomics anywhere, so the order and every sum below are the same on every run);
//   2. spread: a segmented reduction of the Lagrange-weighted charges (1, x, y) over the sorted points.
//      A box's n_interp^2 nodes belong to that box only, so each box is written exactly once: by the
//      chunk that holds all its points, or by the span pass for boxes whose points cross chunk ends;
//   3. gather: potentials at the points in box order, forces, and Z by a fixed-order reduction.
#include "common.hpp"
#include "batch.hpp"

namespace mg {

namespace {

constexpr int TPB = 256;
// items per thread of a sort pass: 8, or 2 up to 262,144 keys, where the blocks of a pass, not its bytes, bound it; up to
// 2^20 keys the sort_items tunable overrides (the digit rows are sized for any setting: sort_hist_ints)
inline int rs_items(i64 n) {
    if (n > (1 << 20)) return 8;
    const int t = tunables().sort_items;
    return t > 0 ? t : (n <= 262144 ? 2 : 8);
}
constexpr int SP_PT = 4;                 // points per thread in the spread
constexpr int SP_CHUNK = TPB * SP_PT;    // points per spread chunk (one workgroup)

// local coordinates and weights of a point whose box is known from its key (src/fftconv.cpp:105-111).
// CEN (FP32 path): the local coordinates come from double arithmetic on the float inputs, and fxu / fyu
// return them unclamped, for the node-relative charges and forces.
template <class T, int NI, bool CEN> __device__ inline void weights(T x, T y, int key, const GridT<T>& g, int& bx, int& by,
                                                                   T* wx, T* wy, double& fxu, double& fyu) {
    bx = key / g.nb; by = key - bx * g.nb;
    T fx, fy;
    if (CEN) {
        fxu = ((double)x - g.xmin_d) / g.width_d - (double)bx;
        fyu = ((double)y - g.ymin_d) / g.width_d - (double)by;
        fx = (T)(fxu < 0.0 ? 0.0 : (fxu > 1.0 ? 1.0 : fxu));
        fy = (T)(fyu < 0.0 ? 0.0 : (fyu > 1.0 ? 1.0 : fyu));
    } else {
        const T ux = (x - g.xmin) / g.width, uy = (y - g.ymin) / g.width;
        fx = ux - (T)bx; fy = uy - (T)by;
        fx = fx < T(0) ? T(0) : (fx > T(1) ? T(1) : fx);
        fy = fy < T(0) ? T(0) : (fy > T(1) ? T(1) : fy);
        fxu = fyu = 0.0;
    }
    lagrange<T, NI>(fx, wx);
    lagrange<T, NI>(fy, wy);
}

// offset of a point from node k of its box, in coordinate units: width (f - (k + 1/2) / NI)
template <int NI> __device__ inline double node_offset(double f, double width, int k) {
    return width * (f - (k + 0.5) / NI);
}

// ---------------------------------------------------------------------------------------------- sort
// Stable LSD radix sort of the box keys (src/fftconv.cpp:83-97 sorts by x-box with a counting sort; the full
// key groups each box's points for the deterministic spread). Digits of DB bits (8 or 10: at most two passes
// for up to 1024 boxes per side). Per pass: per-block digit counts, hist[d][block]; an exclusive scan of each
// digit's row (one workgroup per digit); the stable scatter, whose blocks turn the digit totals into digit
// bases themselves.
// FIRST: keys from the points (the box of point src = inv ? inv[i] : i, value src), else keys_in
template <class T, bool FIRST, int DB>
__global__ __launch_bounds__(TPB) void k_rs_hist(const T* __restrict__ Y, GridT<T> g, const int* __restrict__ keys_in,
                                                 i64 n, int shift, int* __restrict__ hist, int nblocks, int items,
                                                 const int* __restrict__ inv) {
    constexpr int NBK = 1 << DB;
    __shared__ int h[NBK];
    for (int d = threadIdx.x; d < NBK; d += TPB) h[d] = 0;
    __syncthreads();
    typedef typename vec2<T>::type V;
    const i64 base = (i64)blockIdx.x * TPB * items;
    for (int r = 0; r < items; ++r) {
        const i64 i = base + r * TPB + threadIdx.x;
        if (i < n) {
            int key;
            if (FIRST) { const V p = reinterpret_cast<const V*>(Y)[inv ? (i64)inv[i] : i]; key = box_key<T>(p.x, p.y, g); }
            else key = keys_in[i];
            atomicAdd(&h[(key >> shift) & (NBK - 1)], 1);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NBK; d += TPB) hist[d * nblocks + blockIdx.x] = h[d];
}

// exclusive block-wide scan of one value per thread; returns the thread's prefix, *total the sum
__device__ inline int block_exscan(int v, int* sh, int* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < TPB; d <<= 1) {
        const int u = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int incl = sh[t];
    *total = sh[TPB - 1];
    __syncthreads();
    return incl - v;
}

// one workgroup per digit: exclusive scan of hist[d][0 .. nblocks) in place, digit total to tot[d]
__global__ __launch_bounds__(TPB) void k_rs_rowscan(int* __restrict__ hist, int nblocks, int* __restrict__ tot) {
    __shared__ int sh[TPB];
    int* row = hist + (size_t)blockIdx.x * nblocks;
    int carry = 0;
    for (int c0 = 0; c0 < nblocks; c0 += TPB) {
        const int k = c0 + threadIdx.x;
        const int v = k < nblocks ? row[k] : 0;
        int sum;
        const int ex = block_exscan(v, sh, &sum);
        if (k < nblocks) row[k] = carry + ex;
        carry += sum;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// stable scatter: an item's rank is (items of its digit in earlier blocks) + (in earlier rounds of this
// block) + (in earlier waves of this round) + (in earlier lanes of its wave, from DB ballots)
template <class T, bool FIRST, int DB>
__global__ __launch_bounds__(TPB) void k_rs_scatter(const T* __restrict__ Y, GridT<T> g, const int* __restrict__ keys_in,
                                                    const int* __restrict__ vals_in, i64 n, int shift,
                                                    const int* __restrict__ hist, const int* __restrict__ tot, int nblocks,
                                                    int* __restrict__ keys_out, int* __restrict__ vals_out, int kmask, int items,
                                                    const int* __restrict__ inv) {
    constexpr int NBK = 1 << DB, PER = NBK / TPB > 0 ? NBK / TPB : 1;
    __shared__ int off[NBK];
    __shared__ int wcnt[TPB / WAVE][NBK];
    __shared__ int sh[TPB];
    const int t = threadIdx.x, wave = t / WAVE, lane = t % WAVE;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    {   // digit bases: exclusive scan of the digit totals (thread t owns digits t*PER .. t*PER + PER - 1)
        int loc[PER], s = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) { const int d = t * PER + u; loc[u] = d < NBK ? tot[d] : 0; s += loc[u]; }
        int all;
        int ex = block_exscan(s, sh, &all);
#pragma unroll
        for (int u = 0; u < PER; ++u) { const int d = t * PER + u; if (d < NBK) off[d] = ex + hist[(size_t)d * nblocks + blockIdx.x]; ex += loc[u]; }
    }
    typedef typename vec2<T>::type V;
    const i64 base = (i64)blockIdx.x * TPB * items;
    for (int r = 0; r < items; ++r) {
        const i64 i = base + r * TPB + t;
        const bool valid = i < n;
        int key = 0, val = 0;
        if (valid) {
            if (FIRST) {
                const i64 src = inv ? (i64)inv[i] : i;
                const V p = reinterpret_cast<const V*>(Y)[src];
                key = box_key<T>(p.x, p.y, g); val = (int)src;
            }
            else { key = keys_in[i]; val = vals_in[i]; }
        }
        const int d = (key >> shift) & (NBK - 1);
        for (int k = t; k < (TPB / WAVE) * NBK; k += TPB) (&wcnt[0][0])[k] = 0;
        __syncthreads();
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = off[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
            keys_out[pos] = key & kmask; vals_out[pos] = val;
        }
        __syncthreads();
        for (int k = t; k < NBK; k += TPB) {
            int add = 0;
#pragma unroll
            for (int w = 0; w < TPB / WAVE; ++w) add += wcnt[w][k];
            off[k] += add;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------------------------------- spread
// combine of two segmented-scan elements, `a` before `b`
template <class T, int V> __device__ inline void seg_combine(bool fa, const T* va, bool& fb, T* vb) {
    if (!fb) {
#pragma unroll
        for (int v = 0; v < V; ++v) vb[v] = va[v] + vb[v];
    }
    fb = fa || fb;
}

template <class T, int NI, bool CEN>
__device__ __forceinline__ void spread_block(const int* __restrict__ keys, const int* __restrict__ vals,
                                             const T* __restrict__ Y, i64 n, const GridT<T>& g, T* __restrict__ W,
                                             int* __restrict__ item_key, int* __restrict__ item_valid,
                                             T* __restrict__ item_val, int bid) {
    constexpr int V = 3 * NI;
    __shared__ T sA[V][TPB];
    __shared__ int sF[TPB];
    typedef typename vec2<T>::type V2;
    const int t = threadIdx.x;
    const i64 c0 = (i64)bid * SP_CHUNK;
    const i64 cend = min(c0 + (i64)SP_CHUNK, n);
    const int key_first = keys[c0], key_last = keys[cend - 1];
    const bool starts_new = c0 == 0 || keys[c0 - 1] != key_first;
    const bool ends_closed = cend == n || keys[cend] != key_last;
    if (t == 0) {
        item_key[2 * bid] = key_first;
        item_valid[2 * bid] = !starts_new;
        item_key[2 * bid + 1] = key_last;
        item_valid[2 * bid + 1] = !ends_closed && (key_first != key_last || starts_new);
    }
    const i64 p0 = c0 + (i64)t * SP_PT;
    int key[SP_PT];
    bool valid[SP_PT], head[SP_PT], endseg[SP_PT];
    T x[SP_PT], y[SP_PT], wx[SP_PT][NI], wy[SP_PT][NI];
    double fxu[SP_PT], fyu[SP_PT];
    int first_head = SP_PT;
#pragma unroll
    for (int j = 0; j < SP_PT; ++j) {
        const i64 p = p0 + j;
        valid[j] = p < cend;
        key[j] = valid[j] ? keys[p] : -1;
        head[j] = false; endseg[j] = false;
        x[j] = y[j] = T(0);
        fxu[j] = fyu[j] = 0.0;
#pragma unroll
        for (int a = 0; a < NI; ++a) wx[j][a] = wy[j][a] = T(0);
        if (!valid[j]) continue;
        const V2 pt = reinterpret_cast<const V2*>(Y)[vals[p]];
        x[j] = pt.x; y[j] = pt.y;
        int bx, by;
        weights<T, NI, CEN>(x[j], y[j], key[j], g, bx, by, wx[j], wy[j], fxu[j], fyu[j]);
        head[j] = p == c0 || keys[p - 1] != key[j];
        endseg[j] = p == cend - 1 || keys[p + 1] != key[j];
        if (first_head == SP_PT && head[j]) first_head = j;
    }

    for (int a = 0; a < NI; ++a) {
        // local inclusive segmented scan of this thread's points (reset at its first point and at heads)
        T s[SP_PT][V];
#pragma unroll
        for (int j = 0; j < SP_PT; ++j) {
            const T qa = CEN ? (T)node_offset<NI>(fxu[j], g.width_d, a) : x[j];
#pragma unroll
            for (int b = 0; b < NI; ++b) {
                const T w = wx[j][a] * wy[j][b];   // zero weights for positions past the end
                const T qb = CEN ? (T)node_offset<NI>(fyu[j], g.width_d, b) : y[j];
                const T v0 = w, v1 = w * qa, v2 = w * qb;
                if (j == 0 || head[j]) {
                    s[j][3 * b + 0] = v0; s[j][3 * b + 1] = v1; s[j][3 * b + 2] = v2;
                } else {
                    s[j][3 * b + 0] = s[j - 1][3 * b + 0] + v0;
                    s[j][3 * b + 1] = s[j - 1][3 * b + 1] + v1;
                    s[j][3 * b + 2] = s[j - 1][3 * b + 2] + v2;
                }
            }
        }
        // block exclusive segmented scan of the per-thread aggregates (Hillis-Steele, fixed order)
        bool f = first_head < SP_PT;
        T agg[V];
#pragma unroll
        for (int v = 0; v < V; ++v) agg[v] = s[SP_PT - 1][v];
        sF[t] = f;
#pragma unroll
        for (int v = 0; v < V; ++v) sA[v][t] = agg[v];
        __syncthreads();
        for (int d = 1; d < TPB; d <<= 1) {
            bool fo = false;
            T vo[V];
            if (t >= d) {
                fo = sF[t - d];
#pragma unroll
                for (int v = 0; v < V; ++v) vo[v] = sA[v][t - d];
            }
            __syncthreads();
            if (t >= d) {
                seg_combine<T, V>(fo, vo, f, agg);
                sF[t] = f;
#pragma unroll
                for (int v = 0; v < V; ++v) sA[v][t] = agg[v];
            }
            __syncthreads();
        }
        T carry[V];
#pragma unroll
        for (int v = 0; v < V; ++v) carry[v] = t > 0 ? sA[v][t - 1] : T(0);
        __syncthreads();
        // segment ends: the whole segment is in this chunk -> grid; else -> the chunk's item slot
#pragma unroll
        for (int j = 0; j < SP_PT; ++j) {
            if (!endseg[j]) continue;
            T tot[V];
#pragma unroll
            for (int v = 0; v < V; ++v) tot[v] = j < first_head ? carry[v] + s[j][v] : s[j][v];
            const bool inc_start = key[j] == key_first, inc_end = p0 + j == cend - 1;
            const bool complete = (!inc_start || starts_new) && (!inc_end || ends_closed);
            if (complete) {
                const int bx = key[j] / g.nb, by = key[j] - bx * g.nb;
                const int gx = bx * NI + a;
#pragma unroll
                for (int b = 0; b < NI; ++b) {
                    const int gy = by * NI + b;
                    W[((size_t)0 * g.N + gx) * g.N + gy] = tot[3 * b + 0];
                    W[((size_t)1 * g.N + gx) * g.N + gy] = tot[3 * b + 1];
                    W[((size_t)2 * g.N + gx) * g.N + gy] = tot[3 * b + 2];
                }
            } else {
                const int slot = (inc_start && !starts_new) ? 2 * bid : 2 * bid + 1;
                T* dst = item_val + ((size_t)slot * NI + a) * V;
#pragma unroll
                for (int v = 0; v < V; ++v) dst[v] = tot[v];
            }
        }
    }
}

template <class T, int NI, bool CEN>
__global__ __launch_bounds__(TPB) void k_spread(const int* __restrict__ keys, const int* __restrict__ vals,
                                                const T* __restrict__ Y, i64 n, GridT<T> g, T* __restrict__ W,
                                                int* __restrict__ item_key, int* __restrict__ item_valid,
                                                T* __restrict__ item_val) {
    spread_block<T, NI, CEN>(keys, vals, Y, n, g, W, item_key, item_valid, item_val, blockIdx.x);
}

// boxes whose points cross chunk ends: the chunk where such a box starts sums its open last item with
// the first items of the following chunks holding the same key (found by binary search on the chunk
// first keys), in a fixed tree order
template <class T, int NI>
__device__ __forceinline__ void span_block(const int* __restrict__ keys, i64 n, int nchunks, const GridT<T>& g,
                                           T* __restrict__ W, const int* __restrict__ item_key,
                                           const int* __restrict__ item_valid, const T* __restrict__ item_val, int bid) {
    constexpr int V = 3 * NI;
    __shared__ T sA[V][TPB];
    __shared__ int s_c1;
    const int c = bid, t = threadIdx.x;
    if (!item_valid[2 * c + 1]) return;
    const int skey = item_key[2 * c + 1];
    if (t == 0) {
        int lo = c + 1, hi = nchunks - 1;   // keys[lo * CHUNK] == skey holds; find the last such chunk
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (keys[(i64)mid * SP_CHUNK] == skey) lo = mid; else hi = mid - 1;
        }
        s_c1 = lo;
    }
    __syncthreads();
    const int count = 1 + (s_c1 - c);   // item 0: slot 2c + 1; item k >= 1: slot 2(c + k)
    const int bx = skey / g.nb, by = skey - bx * g.nb;
    for (int a = 0; a < NI; ++a) {
        T acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = T(0);
        for (int k = t; k < count; k += TPB) {
            const int slot = k == 0 ? 2 * c + 1 : 2 * (c + k);
            const T* src = item_val + ((size_t)slot * NI + a) * V;
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += src[v];
        }
#pragma unroll
        for (int v = 0; v < V; ++v) sA[v][t] = acc[v];
        __syncthreads();
        for (int s = TPB / 2; s > 0; s >>= 1) {
            if (t < s) {
#pragma unroll
                for (int v = 0; v < V; ++v) sA[v][t] += sA[v][t + s];
            }
            __syncthreads();
        }
        if (t < V) {
            const int b = t / 3, ch = t % 3;
            W[((size_t)ch * g.N + bx * NI + a) * g.N + by * NI + b] = sA[t][0];
        }
        __syncthreads();
    }
}
template <class T, int NI>
__global__ __launch_bounds__(TPB) void k_spread_span(const int* __restrict__ keys, i64 n, int nchunks, GridT<T> g,
                                                     T* __restrict__ W, const int* __restrict__ item_key,
                                                     const int* __restrict__ item_valid, const T* __restrict__ item_val) {
    span_block<T, NI>(keys, n, nchunks, g, W, item_key, item_valid, item_val, blockIdx.x);
}

// batched spread and span passes: instance i's sorted points sit at [off, off + n) of keys / vals (vals index
// the concatenated Y), chunked from its own start
template <class T, int NI, bool CEN>
__global__ __launch_bounds__(TPB) void kb_spread(const BInst<T>* __restrict__ I, BMap m, const int* __restrict__ keys,
                                                 const int* __restrict__ vals, const T* __restrict__ Y) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active) return;
    const GridT<T> g = b.g;
    const i64 off = b.off, n = b.n;
    T* W = b.W;
    int* ik = b.item_key;
    int* iv = b.item_valid;
    T* ival = b.item_val;
    spread_block<T, NI, CEN>(keys + off, vals + off, Y, n, g, W, ik, iv, ival, blockIdx.x - m.start[k]);
}
template <class T, int NI>
__global__ __launch_bounds__(TPB) void kb_spread_span(const BInst<T>* __restrict__ I, BMap m, const int* __restrict__ keys) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active) return;
    const GridT<T> g = b.g;
    const i64 off = b.off, n = b.n;
    T* W = b.W;
    const int* ik = b.item_key;
    const int* iv = b.item_valid;
    const T* ival = b.item_val;
    span_block<T, NI>(keys + off, n, m.start[k + 1] - m.start[k], g, W, ik, iv, ival, blockIdx.x - m.start[k]);
}

// batched sort keys: (instance << box_bits) | box key, and the global point index; an inactive instance's points
// all take box 0 (they are sorted with the rest and not used)
template <class T>
__global__ __launch_bounds__(TPB) void kb_keys(const BInst<T>* __restrict__ I, BMap m, const T* __restrict__ Y, int box_bits,
                                               int* __restrict__ keys, int* __restrict__ vals) {
    const int k = bmap_find(m, blockIdx.x);
    const int inst = m.inst[k];
    const BInst<T>& b = I[inst];
    const i64 j = (i64)(blockIdx.x - m.start[k]) * TPB + threadIdx.x;
    if (j >= b.n) return;
    const i64 p = b.off + j;
    int box = 0;
    if (b.active) {
        const GridT<T> g = b.g;
        typedef typename vec2<T>::type V;
        const V y = reinterpret_cast<const V*>(Y)[p];
        box = box_key<T>(y.x, y.y, g);
    }
    keys[p] = (inst << box_bits) | box;
    vals[p] = (int)p;
}

// -------------------------------------------------------------------------------------------- gather
template <class T> struct vec4;
template <> struct vec4<float> { typedef float4 type; };
template <> struct vec4<double> { typedef double4 type; };


// compensated sum over a block: a butterfly of pairs within each wave (fixed order), then thread 0 merges the waves in
// index order; the result is valid in thread 0
__device__ inline CSum block_csum(CSum v, double* sh) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = csum_merge(v, {__shfl_xor(v.s, o), __shfl_xor(v.c, o)});
    __syncthreads();   // the previous call's readers are done with sh
    if (threadIdx.x % WAVE == 0) { sh[2 * (threadIdx.x / WAVE)] = v.s; sh[2 * (threadIdx.x / WAVE) + 1] = v.c; }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < TPB / WAVE; ++k) v = csum_merge(v, {sh[2 * k], sh[2 * k + 1]});
    return v;
}


template <class T, int NI, bool CEN>
__device__ __forceinline__ void gather_block(const int* __restrict__ keys, const int* __restrict__ vals,
                                             const T* __restrict__ Y, i64 n, const GridT<T>& g, const T* __restrict__ phi,
                                             T* __restrict__ F, double* __restrict__ zpart, int bid, int nblk,
                                             T* __restrict__ Wclr = nullptr) {
    __shared__ double shd[TPB];
    typedef typename vec2<T>::type V2;
    struct V4 { T x, y, z, w; };
    const size_t NN = (size_t)g.N * g.N;   // planar potentials (4, N, N)
    const T* __restrict__ ph0 = phi;
    const T* __restrict__ ph1 = phi + NN;
    const T* __restrict__ ph2 = phi + 2 * NN;
    const T* __restrict__ ph3 = phi + 3 * NN;
    CSum z = {0.0, 0.0};
    for (i64 p = (i64)bid * TPB + threadIdx.x; p < n; p += (i64)nblk * TPB) {
        const int key = keys[p], idx = vals[p];
        const V2 pt = reinterpret_cast<const V2*>(Y)[idx];
        int bx, by;
        T wx[NI], wy[NI];
        double fxu, fyu;
        weights<T, NI, CEN>(pt.x, pt.y, key, g, bx, by, wx, wy, fxu, fyu);
        if (Wclr != nullptr && (p == 0 || keys[p - 1] != key)) {
            // the box's first point: its charge nodes, the only ones the spread wrote, back to zero (the row pass has
            // read them), so that the next spread needs no fill of the whole grid
#pragma unroll
            for (int ch = 0; ch < 3; ++ch)
#pragma unroll
                for (int a = 0; a < NI; ++a)
#pragma unroll
                    for (int b = 0; b < NI; ++b) Wclr[((size_t)ch * g.N + bx * NI + a) * g.N + by * NI + b] = T(0);
        }
        V2 f;
        if (CEN) {
            // F = sum_g w(g) (phi1(g) (y_i - y_g) + psi(g)), offsets and sums in double
            double p0 = 0, fx = 0, fy = 0;
#pragma unroll
            for (int a = 0; a < NI; ++a) {
                const size_t row = (size_t)(bx * NI + a) * g.N + by * NI;
                const double ex = node_offset<NI>(fxu, g.width_d, a);
#pragma unroll
                for (int b = 0; b < NI; ++b) {
                    const double w = (double)(wx[a] * wy[b]);
                    const size_t o = row + b;
                    const V4 c = {ph0[o], ph1[o], ph2[o], ph3[o]};
                    const double ey = node_offset<NI>(fyu, g.width_d, b);
                    p0 += w * (double)c.x;
                    fx += w * ((double)c.y * ex + (double)c.z);
                    fy += w * ((double)c.y * ey + (double)c.w);
                }
            }
            csum_add(z, p0);
            f.x = (T)fx; f.y = (T)fy;
        } else {
            T p0 = 0, p1 = 0, p2 = 0, p3 = 0;
#pragma unroll
            for (int a = 0; a < NI; ++a) {
                const size_t row = (size_t)(bx * NI + a) * g.N + by * NI;
#pragma unroll
                for (int b = 0; b < NI; ++b) {
                    const T w = wx[a] * wy[b];
                    const size_t o = row + b;
                    const V4 c = {ph0[o], ph1[o], ph2[o], ph3[o]};
                    p0 += w * c.x; p1 += w * c.y; p2 += w * c.z; p3 += w * c.w;
                }
            }
            csum_add(z, (double)p0);
            f.x = p1 * pt.x - p2; f.y = p1 * pt.y - p3;
        }
        reinterpret_cast<V2*>(F)[idx] = f;
    }
    z = block_csum(z, shd);
    if (threadIdx.x == 0) { zpart[2 * bid] = z.s; zpart[2 * bid + 1] = z.c; }
}

// Z of the gather: the nblk compensated block partials merged in a fixed order; sumQ = Z - n (the subtraction
// taken into the compensated sum before it is rounded), 1 / max(sumQ, 1e-12) (one block)
__device__ inline void zfinal_block(const double* __restrict__ zpart, int nblk, i64 n, double* __restrict__ zout,
                                    void* inv_out, bool fp32, double* __restrict__ sumq_hist) {
    __shared__ double shd[2 * TPB];
    CSum s = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += TPB) s = csum_merge(s, {zpart[2 * b], zpart[2 * b + 1]});
    s = block_csum(s, shd);
    if (threadIdx.x == 0) {
        s = csum_merge(s, {-(double)n, 0.0});
        const double sumq = s.s + s.c;
        const double inv = 1.0 / (sumq > 1e-12 ? sumq : 1e-12);
        zout[0] = sumq; zout[1] = inv;
        if (fp32) *reinterpret_cast<float*>(inv_out) = (float)inv; else *reinterpret_cast<double*>(inv_out) = inv;
        if (sumq_hist) *sumq_hist = sumq;
    }
}
__global__ __launch_bounds__(TPB) void k_zfinal(const double* __restrict__ zpart, int nblk, i64 n, double* __restrict__ zout,
                                                void* inv_out, int fp32, double* __restrict__ sumq_hist) {
    zfinal_block(zpart, nblk, n, zout, inv_out, fp32 != 0, sumq_hist);
}

// (Z stays a kernel of its own: a last-block merge would need a device-scope fence in every gather block, which
// on this GPU writes each block's dirty L2 lines back across the XCDs -- measured 190 us more per 1M iteration)
template <class T, int NI, bool CEN>
__global__ __launch_bounds__(TPB) void k_gather(const int* __restrict__ keys, const int* __restrict__ vals,
                                                const T* __restrict__ Y, i64 n, GridT<T> g, const T* __restrict__ phi,
                                                T* __restrict__ F, double* __restrict__ zpart, T* __restrict__ Wclr) {
    gather_block<T, NI, CEN>(keys, vals, Y, n, g, phi, F, zpart, blockIdx.x, gridDim.x, Wclr);
}
template <class T, int NI, bool CEN>
__global__ __launch_bounds__(TPB) void kb_gather(const BInst<T>* __restrict__ I, BMap m, const int* __restrict__ keys,
                                                 const int* __restrict__ vals, const T* __restrict__ Y, T* __restrict__ F) {
    const int k = bmap_find(m, blockIdx.x);
    const BInst<T>& b = I[m.inst[k]];
    if (!b.active) return;
    const GridT<T> g = b.g;
    const i64 off = b.off, n = b.n;
    const T* phi = b.phi;
    double* zpart = b.zpart;
    gather_block<T, NI, CEN>(keys + off, vals + off, Y, n, g, phi, F, zpart, blockIdx.x - m.start[k], m.start[k + 1] - m.start[k],
                             b.W);   // the instance's charge grid stays zero between iterations
}
// one block per instance
template <class T>
__global__ __launch_bounds__(TPB) void kb_zfinal(const BInst<T>* __restrict__ I, BMap m) {
    const BInst<T>& b = I[m.inst[bmap_find(m, blockIdx.x)]];
    if (!b.active) return;
    const i64 n = b.n, nb = (n + TPB - 1) / TPB;   // the gather's blocks (gather_blocks)
    const int nblk = (int)(nb < 1024 ? (nb < 1 ? 1 : nb) : 1024);
    zfinal_block(b.zpart, nblk, n, b.zout, b.inv_out, sizeof(T) == 4, b.sumq);
}

// transforms per workgroup of the FFT passes of length L (fft_batch of fft_core.hpp)
__device__ inline int fft_batch_of(int L) { return L / 16 >= 256 ? 1 : 256 / (L / 16); }

// inclusive prefix sum over the TPB threads of a block (wave scans, then the wave totals through LDS)
__device__ inline int block_inclusive_scan(int v, int* sh) {
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == WAVE - 1) sh[wv] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wv; ++w) add += sh[w];
    __syncthreads();
    return v + add;
}

// The work lists of one iteration, one block per list: each candidate instance that runs the iteration (and, for the
// FFT passes, sits at the list's size) takes, in candidate order, the blocks its single launch would: rows
// ceil(N / batch) per channel, columns ceil(Mh / batch), transposes the 32 x 32 tiles of its live matrices.
template <class T>
__global__ __launch_bounds__(TPB) void kb_worklists(const BInst<T>* __restrict__ I, const WorkDesc* __restrict__ D,
                                                    const int* __restrict__ cand, int* __restrict__ wl) {
    __shared__ int sh[TPB / WAVE];
    const WorkDesc w = D[blockIdx.x];
    int* L = wl + w.off;
    int* inst = L + 3;
    int* start = L + 3 + w.cap;
    int n0 = 0, tot0 = 0;
    for (int base = 0; base < w.cap; base += TPB) {
        const int c = base + threadIdx.x;
        int i = -1, nb = 0;
        if (c < w.cap) {
            i = cand[w.cand + c];
            const BInst<T>& b = I[i];
            if (b.active) {
                const int N = b.g.N, M = b.M, Mh = M / 2 + 1;
                if (w.kind == WK_ROWS3 || w.kind == WK_ROWS4) {
                    if (M == w.M) { const int B = fft_batch_of(M / 2); nb = (N + B - 1) / B * (w.kind == WK_ROWS3 ? 3 : 4); }
                } else if (w.kind == WK_COLS) {
                    if (M == w.M) { const int B = fft_batch_of(M); nb = (Mh + B - 1) / B; }
                } else if (w.kind == WK_TR0) {
                    nb = ((Mh + 31) / 32) * ((N + 31) / 32) * 3;
                } else if (w.kind == WK_TR1) {
                    nb = ((N + 31) / 32) * ((Mh + 31) / 32) * 4;
                }
            }
        }
        const int f = nb > 0 ? 1 : 0;
        const int pos = block_inclusive_scan(f, sh) - f;
        const int off = block_inclusive_scan(nb, sh);
        if (f) {
            inst[n0 + pos] = i;
            start[n0 + pos + 1] = tot0 + off;
        }
        // the chunk's totals, from its last thread
        if (threadIdx.x == TPB - 1) { sh[0] = pos + f; sh[1] = off; }
        __syncthreads();
        n0 += sh[0];
        tot0 += sh[1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        start[0] = 0;
        L[0] = tot0;
        L[1] = n0;
        L[2] = w.cap;
    }
}

// The iteration's decision for every instance (one thread each): the grid of FFTRepulsion2D._grid
// (magnet/tsne.py:30-46) from the bbox of the last update (mb[2..5]), its FFT size against the plan's window, the
// spectra to recompute (the reference's cache key (M, width)), and the schedule's exaggeration and momentum
// (magnet/tsne.py:176-181). An instance whose size falls outside the window, or whose layout is not finite, does not
// run the iteration and keeps its state; the host sees the stall and plans again.
template <class T>
__global__ __launch_bounds__(TPB) void kb_decide(BInst<T>* __restrict__ I, BState* __restrict__ S, BSched q) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= q.count) return;
    BInst<T>& b = I[i];
    BState& s = S[i];
    b.active = 0;
    b.fresh = 0;
    if (s.next_it >= q.n_iter) return;
    const double bb[4] = {(double)b.mb[2], (double)b.mb[3], (double)b.mb[4], (double)b.mb[5]};
    if (!(isfinite(bb[0]) && isfinite(bb[1]) && isfinite(bb[2]) && isfinite(bb[3]))) { s.stall = 2; return; }
    const GridDecision gd = decide_grid(bb, q.box_width, q.min_boxes, q.max_boxes);
    const GridT<T> g = make_grid<T>(gd.xmin, gd.ymin, gd.width, gd.nb, q.ni);
    const int M = fft_size(g.N);
    s.M = M;
    if (M == 0 || M < b.Mlo || M > b.Mhi) { s.stall = 1; return; }
    s.stall = 0;
    const int it = s.next_it;
    s.next_it = it + 1;
    const bool early = it < q.early_iter;
    b.rho = (T)(early ? q.rho0 : q.rho1);
    b.mom = (T)(early ? q.mom0 : q.mom1);
    const double delta = gd.width / q.ni;
    if (delta == q.delta0 || b.Kw[0] == nullptr) {
        // an unclipped grid (width = box_width): the batch's shared spectra of size M
        int w = 0;
        for (int j = 1; j < 4; ++j) if (b.KM[j] == M) w = j;
        for (int t = 0; t < 4; ++t) b.K[t] = b.KS[w][t];
    } else {
        // a clipped grid at the largest size: the instance's own tables, recomputed when the width changed
        // (the reference's cache key (M, width))
        if (!(s.sM == M && s.sdelta == delta)) { b.fresh = 1; s.sM = M; s.sdelta = delta; }
        for (int t = 0; t < 4; ++t) b.K[t] = b.Kw[t];
    }
    b.g = g;
    b.M = M;
    b.delta = delta;
    b.sumq = b.sumq_base + it;
    b.active = 1;
}

}  // namespace

int sort_hist_blocks(i64 n) { const i64 tile = (i64)TPB * rs_items(n); return (int)((n + tile - 1) / tile); }
i64 sort_hist_ints(i64 n) {
    const i64 tile = (i64)TPB * (n > (1 << 20) ? 8 : 1);
    return (i64)1024 * std::max<i64>(1, (n + tile - 1) / tile) + 1024;
}
int spread_chunks(i64 n) { return (int)((n + SP_CHUNK - 1) / SP_CHUNK); }
int gather_blocks(i64 n) { i64 b = (n + TPB - 1) / TPB; return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024); }

// LSD passes of DB-bit digits. Box keys computed from Y in the first pass (keys0 == nullptr), or precomputed
// keys0 / vals0 (read by the first pass). Pass p writes to buffer (p + first) % 2 (0 = A, 1 = B), first = 0 when
// keys come from Y and 1 when they are precomputed in A; returns the buffer holding the result.
template <class T, int DB> static int sort_db(const T* Y, const GridT<T>* g, const int* keys0, const int* vals0, i64 n, int* keysA,
                                              int* valsA, int* keysB, int* valsB, int* hist, int passes, hipStream_t st,
                                              int kmask = -1, const int* inv = nullptr) {
    const int nblocks = sort_hist_blocks(n), items = rs_items(n);
    int* tot = hist + (size_t)(1 << DB) * nblocks;
    const int first = keys0 ? 1 : 0;
    const int* kin = keys0; const int* vin = vals0;
    int cur = 0;
    const GridT<T> gg = g ? *g : GridT<T>{};
    for (int p = 0; p < passes; ++p) {
        const int shift = DB * p;
        cur = (p + first) % 2;
        int* kout = cur == 0 ? keysA : keysB;
        int* vout = cur == 0 ? valsA : valsB;
        const bool fromY = p == 0 && !keys0;
        if (fromY) hipLaunchKernelGGL((k_rs_hist<T, true, DB>), dim3(nblocks), dim3(TPB), 0, st, Y, gg, kin, n, shift, hist, nblocks, items, inv);
        else hipLaunchKernelGGL((k_rs_hist<T, false, DB>), dim3(nblocks), dim3(TPB), 0, st, Y, gg, kin, n, shift, hist, nblocks, items, (const int*)nullptr);
        MG_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_rs_rowscan, dim3(1 << DB), dim3(TPB), 0, st, hist, nblocks, tot);
        MG_HIP(hipGetLastError());
        const int mk = p == passes - 1 ? kmask : -1;
        if (fromY) hipLaunchKernelGGL((k_rs_scatter<T, true, DB>), dim3(nblocks), dim3(TPB), 0, st, Y, gg, kin, vin, n, shift, hist, tot, nblocks, kout, vout, mk, items, inv);
        else hipLaunchKernelGGL((k_rs_scatter<T, false, DB>), dim3(nblocks), dim3(TPB), 0, st, Y, gg, kin, vin, n, shift, hist, tot, nblocks, kout, vout, mk, items, (const int*)nullptr);
        MG_HIP(hipGetLastError());
        kin = kout; vin = vout;
    }
    return cur;
}

template <class T> static int sort_keys_any(const T* Y, const GridT<T>* g, const int* keys0, const int* vals0, i64 n, int bits,
                                            int* kA, int* vA, int* kB, int* vB, int* hist, hipStream_t st, int kmask = -1,
                                            const int* inv = nullptr) {
    if (bits <= 8) return sort_db<T, 8>(Y, g, keys0, vals0, n, kA, vA, kB, vB, hist, 1, st, kmask, inv);
    // two passes either way up to 16 bits: 8-bit digits keep the digit table and its scan four times smaller
    if (bits <= 16 && tunables().sort_digit == 8) return sort_db<T, 8>(Y, g, keys0, vals0, n, kA, vA, kB, vB, hist, 2, st, kmask, inv);
    if (bits <= 30) return sort_db<T, 10>(Y, g, keys0, vals0, n, kA, vA, kB, vB, hist, (bits + 9) / 10, st, kmask, inv);
    return sort_db<T, 8>(Y, g, keys0, vals0, n, kA, vA, kB, vB, hist, (bits + 7) / 8, st, kmask, inv);
}

template <class T> int launch_box_sort(const T* Y, i64 n, const GridT<T>& g, RepBuffers<T>& b, hipStream_t st, const int* inv) {
    const long long nkeys = (long long)g.nb * g.nb;
    int bits = 0;
    while ((1ll << bits) < nkeys) ++bits;
    return sort_keys_any<T>(Y, &g, nullptr, nullptr, n, bits, b.keysA, b.valsA, b.keysB, b.valsB, b.hist, st, -1, inv);
}

// stable sort of precomputed keys (< 2^bits) with their values, both in A; hist as for the box sort
template <class T> int launch_key_sort(int* keysA, int* valsA, int* keysB, int* valsB, int* hist, i64 n, int bits, hipStream_t st) {
    return sort_keys_any<T>(nullptr, nullptr, keysA, valsA, n, bits, keysA, valsA, keysB, valsB, hist, st);
}

int launch_key_sort_masked(int* keysA, int* valsA, int* keysB, int* valsB, int* hist, i64 n, int bits, int keep_bits,
                           hipStream_t st) {
    require(bits >= 1 && bits <= 31 && keep_bits >= 1 && keep_bits <= bits, "bad key widths");
    const int kmask = keep_bits >= 31 ? -1 : (int)((1u << keep_bits) - 1u);
    return sort_keys_any<float>(nullptr, nullptr, keysA, valsA, n, bits, keysA, valsA, keysB, valsB, hist, st, kmask);
}

template <class T> void launch_worklists(const BInst<T>* I, const WorkDesc* d, int nd, const int* cand, int* wl, hipStream_t st) {
    if (nd <= 0) return;
    hipLaunchKernelGGL(kb_worklists<T>, dim3((unsigned)nd), dim3(TPB), 0, st, I, d, cand, wl);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_decide_batch(BInst<T>* I, BState* S, const BSched& q, hipStream_t st) {
    hipLaunchKernelGGL(kb_decide<T>, dim3((unsigned)((q.count + TPB - 1) / TPB)), dim3(TPB), 0, st, I, S, q);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_keys_batch(const BInst<T>* I, const BMap& m, const T* Y, int box_bits, int* keys, int* vals,
                                          hipStream_t st) {
    if (m.count == 0 || m.start[m.count] == 0) return;
    hipLaunchKernelGGL(kb_keys<T>, dim3((unsigned)m.start[m.count]), dim3(TPB), 0, st, I, m, Y, box_bits, keys, vals);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_spread_batch(const BInst<T>* I, const BMap& m, const int* keys, const int* vals, const T* Y,
                                            int ni, bool centred, hipStream_t st) {
    if (m.count == 0 || m.start[m.count] == 0) return;
    const dim3 grid((unsigned)m.start[m.count]);
#define MG_S(NI) do { \
        if (centred) hipLaunchKernelGGL((kb_spread<T, NI, true>), grid, dim3(TPB), 0, st, I, m, keys, vals, Y); \
        else hipLaunchKernelGGL((kb_spread<T, NI, false>), grid, dim3(TPB), 0, st, I, m, keys, vals, Y); \
        MG_HIP(hipGetLastError()); \
        hipLaunchKernelGGL((kb_spread_span<T, NI>), grid, dim3(TPB), 0, st, I, m, keys); \
    } while (0)
    switch (ni) {
        case 1: MG_S(1); break; case 2: MG_S(2); break; case 3: MG_S(3); break; case 4: MG_S(4); break;
        case 5: MG_S(5); break; case 6: MG_S(6); break; case 7: MG_S(7); break; case 8: MG_S(8); break;
        default: require(false, "n_interp must be in [1, 8]");
    }
#undef MG_S
    MG_HIP(hipGetLastError());
}

template <class T> void launch_gather_batch(const BInst<T>* I, const BMap& m, const int* keys, const int* vals, const T* Y,
                                            int ni, bool centred, T* F, hipStream_t st) {
    if (m.count == 0 || m.start[m.count] == 0) return;
    const dim3 grid((unsigned)m.start[m.count]);
#define MG_G(NI) do { \
        if (centred) hipLaunchKernelGGL((kb_gather<T, NI, true>), grid, dim3(TPB), 0, st, I, m, keys, vals, Y, F); \
        else hipLaunchKernelGGL((kb_gather<T, NI, false>), grid, dim3(TPB), 0, st, I, m, keys, vals, Y, F); \
    } while (0)
    switch (ni) {
        case 1: MG_G(1); break; case 2: MG_G(2); break; case 3: MG_G(3); break; case 4: MG_G(4); break;
        case 5: MG_G(5); break; case 6: MG_G(6); break; case 7: MG_G(7); break; case 8: MG_G(8); break;
        default: require(false, "n_interp must be in [1, 8]");
    }
#undef MG_G
    MG_HIP(hipGetLastError());
    BMap one = m;   // one block per instance
    for (int k = 0; k <= m.count; ++k) one.start[k] = k;
    hipLaunchKernelGGL(kb_zfinal<T>, dim3((unsigned)m.count), dim3(TPB), 0, st, I, one);
    MG_HIP(hipGetLastError());
}

template <class T, int NI, bool CEN> static void spread_ni(const int* keys, const int* vals, const T* Y, i64 n, const GridT<T>& g,
                                                 RepBuffers<T>& b, hipStream_t st) {
    const int nch = spread_chunks(n);
    hipLaunchKernelGGL((k_spread<T, NI, CEN>), dim3(nch), dim3(TPB), 0, st, keys, vals, Y, n, g, b.W, b.item_key, b.item_valid, b.item_val);
    MG_HIP(hipGetLastError());
    if (nch > 1) {
        hipLaunchKernelGGL((k_spread_span<T, NI>), dim3(nch), dim3(TPB), 0, st, keys, n, nch, g, b.W, b.item_key, b.item_valid, b.item_val);
        MG_HIP(hipGetLastError());
    }
}

template <class T> void launch_spread(const int* keys, const int* vals, const T* Y, i64 n, const GridT<T>& g, int ni,
                                      bool centred, RepBuffers<T>& b, hipStream_t st, bool fill) {
    if (fill) launch_fill<T>(b.W, T(0), (i64)3 * g.N * g.N, st);
    if (n == 0) return;
#define MG_S(NI) (centred ? spread_ni<T, NI, true>(keys, vals, Y, n, g, b, st) : spread_ni<T, NI, false>(keys, vals, Y, n, g, b, st))
    switch (ni) {
        case 1: MG_S(1); break; case 2: MG_S(2); break; case 3: MG_S(3); break; case 4: MG_S(4); break;
        case 5: MG_S(5); break; case 6: MG_S(6); break; case 7: MG_S(7); break; case 8: MG_S(8); break;
        default: require(false, "n_interp must be in [1, 8]");
    }
#undef MG_S
}

template <class T> void launch_gather(const int* keys, const int* vals, const T* Y, i64 n, const GridT<T>& g, int ni,
                                      bool centred, const T* phi, T* F, RepBuffers<T>& b, double* sumq_hist, hipStream_t st,
                                      T* Wclr) {
    const unsigned blocks = (unsigned)gather_blocks(n);
#define MG_G(NI) do { \
        if (centred) hipLaunchKernelGGL((k_gather<T, NI, true>), dim3(blocks), dim3(TPB), 0, st, keys, vals, Y, n, g, phi, F, b.zpart, Wclr); \
        else hipLaunchKernelGGL((k_gather<T, NI, false>), dim3(blocks), dim3(TPB), 0, st, keys, vals, Y, n, g, phi, F, b.zpart, Wclr); \
    } while (0)
    switch (ni) {
        case 1: MG_G(1); break; case 2: MG_G(2); break; case 3: MG_G(3); break; case 4: MG_G(4); break;
        case 5: MG_G(5); break; case 6: MG_G(6); break; case 7: MG_G(7); break; case 8: MG_G(8); break;
        default: require(false, "n_interp must be in [1, 8]");
    }
#undef MG_G
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_zfinal, dim3(1), dim3(TPB), 0, st, (const double*)b.zpart, (int)blocks, n, b.zout, (void*)b.inv_out,
                       (int)(sizeof(T) == 4), sumq_hist);
    MG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------- out-of-sample placement
// One thread per new point, all iterations in one launch (the points do not interact): the gradient
// rho * A_i - F_i / Z with A_i = sum_j (w_ij / n_train) q_ij (y_i - y_j) over the point's k training neighbours
// (w rows sum to one: the rows of the training P sum to about 1 / n_train) and F_i the repulsion of the fixed
// training layout, read from its potential field phi (4, N, N) exactly as the gather reads it (box clamped to the
// grid, local coordinate clamped to [0, 1]); Z the training layout's. Gains, momentum and step as in the training
// update (src/fftrep.cpp:245-255); no recentring (the training layout fixes the frame).
template <class T, int NI, bool CEN>
__global__ __launch_bounds__(TPB) void k_transform(const T* __restrict__ Ytr, T inv_ntr, const int32_t* __restrict__ nbr,
                                                   const T* __restrict__ w, int k, T* __restrict__ Yn, i64 m, GridT<T> g,
                                                   const T* __restrict__ phi, T inv_Z, int n_iter, T lr, T mom0, T mom1,
                                                   int mom_switch, T rho, T inv_dof, const double* __restrict__ centroid,
                                                   double n_train, double dof) {
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= m) return;
    typedef typename vec2<T>::type V2;
    const V2* Y2 = reinterpret_cast<const V2*>(Ytr);
    const size_t NN = (size_t)g.N * g.N;
    V2 y = reinterpret_cast<V2*>(Yn)[i];
    T ux = 0, uy = 0, gx_ = 1, gy_ = 1;
    for (int it = 0; it < n_iter; ++it) {
        T ax = 0, ay = 0;
        for (int j = 0; j < k; ++j) {
            const int t = nbr[i * k + j];
            const V2 yj = Y2[t];
            const T dx = y.x - yj.x, dy = y.y - yj.y;
            const T q = (w[i * k + j] * inv_ntr) / (T(1) + (dx * dx + dy * dy) * inv_dof);
            ax += q * dx; ay += q * dy;
        }
        int bx, by;
        T wx[NI], wy[NI];
        double fxu, fyu;
        T fx, fy;
        const double gxu = ((double)y.x - g.xmin_d) * g.inv_w_d, gyu = ((double)y.y - g.ymin_d) / g.width_d;
        if (!(gxu >= 0.0 && gxu < (double)g.nb && gyu >= 0.0 && gyu < (double)g.nb)) {
            // outside the training grid the field is not interpolated: the training layout acts as n_train points at
            // its centroid, sum_j K2(d_ij) (y - y_j) ~ n_train K2(|y - c|) (y - c), which decays like the exact sum
            const double dx = (double)y.x - centroid[0], dy = (double)y.y - centroid[1];
            const double k2 = pow(1.0 + (dx * dx + dy * dy) / dof, -dof - 1.0);
            fx = (T)(n_train * k2 * dx);
            fy = (T)(n_train * k2 * dy);
        } else {
        weights<T, NI, CEN>(y.x, y.y, box_key<T>(y.x, y.y, g), g, bx, by, wx, wy, fxu, fyu);
        if (CEN) {
            double sx = 0, sy = 0;
#pragma unroll
            for (int a = 0; a < NI; ++a) {
                const size_t row = (size_t)(bx * NI + a) * g.N + by * NI;
                const double ex = node_offset<NI>(fxu, g.width_d, a);
#pragma unroll
                for (int b = 0; b < NI; ++b) {
                    const double wv = (double)(wx[a] * wy[b]);
                    const size_t o = row + b;
                    const double ey = node_offset<NI>(fyu, g.width_d, b);
                    sx += wv * ((double)phi[NN + o] * ex + (double)phi[2 * NN + o]);
                    sy += wv * ((double)phi[NN + o] * ey + (double)phi[3 * NN + o]);
                }
            }
            fx = (T)sx; fy = (T)sy;
        } else {
            T p1 = 0, p2 = 0, p3 = 0;
#pragma unroll
            for (int a = 0; a < NI; ++a) {
                const size_t row = (size_t)(bx * NI + a) * g.N + by * NI;
#pragma unroll
                for (int b = 0; b < NI; ++b) {
                    const T wv = wx[a] * wy[b];
                    const size_t o = row + b;
                    p1 += wv * phi[NN + o]; p2 += wv * phi[2 * NN + o]; p3 += wv * phi[3 * NN + o];
                }
            }
            fx = p1 * y.x - p2; fy = p1 * y.y - p3;
        }
        }
        const T mom = it < mom_switch ? mom0 : mom1;
        const T g0 = rho * ax - fx * inv_Z, g1 = rho * ay - fy * inv_Z;
        // delta-bar-delta gains and momentum (src/fftrep.cpp:245-255)
        T gg = ((g0 > T(0)) == (ux > T(0))) ? gx_ * T(0.8) : gx_ + T(0.2);
        gx_ = gg < T(0.01) ? T(0.01) : gg;
        ux = mom * ux - lr * gx_ * g0;
        gg = ((g1 > T(0)) == (uy > T(0))) ? gy_ * T(0.8) : gy_ + T(0.2);
        gy_ = gg < T(0.01) ? T(0.01) : gg;
        uy = mom * uy - lr * gy_ * g1;
        y.x += ux; y.y += uy;
    }
    reinterpret_cast<V2*>(Yn)[i] = y;
}

// the mean of a layout (n, 2) in double, one block, fixed order: c[0], c[1]
template <class T>
__global__ __launch_bounds__(TPB) void k_centroid(const T* __restrict__ Y, i64 n, double* __restrict__ c) {
    __shared__ double sh[2][TPB];
    double sx = 0, sy = 0;
    for (i64 i = threadIdx.x; i < n; i += TPB) { sx += (double)Y[2 * i]; sy += (double)Y[2 * i + 1]; }
    sh[0][threadIdx.x] = sx;
    sh[1][threadIdx.x] = sy;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) { sh[0][threadIdx.x] += sh[0][threadIdx.x + o]; sh[1][threadIdx.x] += sh[1][threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { c[0] = sh[0][0] / (double)n; c[1] = sh[1][0] / (double)n; }
}

template <class T> void launch_centroid(const T* Y, i64 n, double* c, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_centroid<T>, dim3(1), dim3(TPB), 0, st, Y, n, c);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_transform(const T* Ytr, i64 n_train, const int32_t* nbr, const T* w, int k, T* Yn, i64 m,
                                         const GridT<T>& g, int ni, bool centred, const T* phi, double inv_Z, int n_iter,
                                         double lr, double mom0, double mom1, int mom_switch, double rho, double dof,
                                         const double* centroid, hipStream_t st) {
    if (m == 0) return;
    const unsigned blocks = (unsigned)((m + TPB - 1) / TPB);
    const T inv_ntr = (T)(1.0 / (double)n_train);
#define MG_T(NI) do { \
        if (centred) hipLaunchKernelGGL((k_transform<T, NI, true>), dim3(blocks), dim3(TPB), 0, st, Ytr, inv_ntr, nbr, w, k, Yn, m, g, phi, \
                                        (T)inv_Z, n_iter, (T)lr, (T)mom0, (T)mom1, mom_switch, (T)rho, (T)(1.0 / dof), centroid, \
                                        (double)n_train, dof); \
        else hipLaunchKernelGGL((k_transform<T, NI, false>), dim3(blocks), dim3(TPB), 0, st, Ytr, inv_ntr, nbr, w, k, Yn, m, g, phi, \
                                (T)inv_Z, n_iter, (T)lr, (T)mom0, (T)mom1, mom_switch, (T)rho, (T)(1.0 / dof), centroid, \
                                (double)n_train, dof); \
    } while (0)
    switch (ni) {
        case 1: MG_T(1); break; case 2: MG_T(2); break; case 3: MG_T(3); break; case 4: MG_T(4); break;
        case 5: MG_T(5); break; case 6: MG_T(6); break; case 7: MG_T(7); break; case 8: MG_T(8); break;
        default: require(false, "n_interp must be in [1, 8]");
    }
#undef MG_T
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template int launch_box_sort<T>(const T*, i64, const GridT<T>&, RepBuffers<T>&, hipStream_t, const int*); \
    template int launch_key_sort<T>(int*, int*, int*, int*, int*, i64, int, hipStream_t); \
    template void launch_spread<T>(const int*, const int*, const T*, i64, const GridT<T>&, int, bool, RepBuffers<T>&, hipStream_t, bool); \
    template void launch_gather<T>(const int*, const int*, const T*, i64, const GridT<T>&, int, bool, const T*, T*, RepBuffers<T>&, double*, hipStream_t, T*); \
    template void launch_transform<T>(const T*, i64, const int32_t*, const T*, int, T*, i64, const GridT<T>&, int, bool, const T*, double, int, double, double, double, int, double, double, const double*, hipStream_t); \
    template void launch_centroid<T>(const T*, i64, double*, hipStream_t); \
    template void launch_keys_batch<T>(const BInst<T>*, const BMap&, const T*, int, int*, int*, hipStream_t); \
    template void launch_worklists<T>(const BInst<T>*, const WorkDesc*, int, const int*, int*, hipStream_t); \
    template void launch_decide_batch<T>(BInst<T>*, BState*, const BSched&, hipStream_t); \
    template void launch_spread_batch<T>(const BInst<T>*, const BMap&, const int*, const int*, const T*, int, bool, hipStream_t); \
    template void launch_gather_batch<T>(const BInst<T>*, const BMap&, const int*, const int*, const T*, int, bool, T*, hipStream_t);
MG_INST(float)
MG_INST(double)

}  // namespace mg
