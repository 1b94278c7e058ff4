// Locality relabelling of the single-embedding loop: the points are renumbered from time to time in the Morton
// order of their grid boxes, so that the attraction's neighbour reads (random rows of Y) fall on nearby memory.
// The labelling never changes a result: the CSR of a relabelled point keeps its entries in their original order
// (so its attraction sums in the same order), the box sort breaks ties by original index, and the mean and bbox
// are summed over the points in original order with the original block decomposition.
#include "common.hpp"

namespace mg {

namespace {

constexpr int TPB = 256;

inline unsigned blocks_for(i64 work, unsigned cap = 8192) {
    i64 b = (work + TPB - 1) / TPB;
    return (unsigned)(b < 1 ? 1 : (b < cap ? b : cap));
}

// order[j] = perm[vals[j]]: the original point at sorted position j
__global__ __launch_bounds__(TPB) void k_order(const int* __restrict__ vals, const int* __restrict__ perm, i64 n,
                                               int* __restrict__ order, int* __restrict__ inv, int* __restrict__ len,
                                               const int32_t* __restrict__ ip) {
    for (i64 j = (i64)blockIdx.x * TPB + threadIdx.x; j < n; j += (i64)gridDim.x * TPB) {
        const int o = perm[vals[j]];
        order[j] = o;
        inv[o] = (int)j;
        len[j] = ip[o + 1] - ip[o];
    }
}

// exclusive scan in blocks of 1024 (4 per thread), then the block totals in one block, then the offsets
constexpr int SCAN_T = 1024;
__device__ inline int block_exscan_sum(int v, int* sh, int* total) {
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
__global__ __launch_bounds__(TPB) void k_scan_tiles(const int* __restrict__ in, i64 n, int* __restrict__ out, int* __restrict__ tot) {
    __shared__ int sh[TPB];
    const i64 base = (i64)blockIdx.x * SCAN_T + (i64)threadIdx.x * 4;
    int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[k] = base + k < n ? in[base + k] : 0; s += v[k]; }
    int all;
    int ex = block_exscan_sum(s, sh, &all);
#pragma unroll
    for (int k = 0; k < 4; ++k) { if (base + k < n) out[base + k] = ex; ex += v[k]; }
    if (threadIdx.x == 0) tot[blockIdx.x] = all;
}
__global__ __launch_bounds__(TPB) void k_scan_totals(int* __restrict__ tot, int ntiles, int* __restrict__ grand) {
    __shared__ int sh[TPB];
    int carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += TPB) {
        const int k = c0 + threadIdx.x;
        const int v = k < ntiles ? tot[k] : 0;
        int sum;
        const int ex = block_exscan_sum(v, sh, &sum);
        if (k < ntiles) tot[k] = carry + ex;
        carry += sum;
    }
    if (threadIdx.x == 0) *grand = carry;
}
__global__ __launch_bounds__(TPB) void k_scan_add(int* __restrict__ out, i64 n, const int* __restrict__ tot) {
    const i64 i = (i64)blockIdx.x * SCAN_T + threadIdx.x;
    const int add = tot[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) if (i + k * TPB < n) out[i + k * TPB] += add;
}

// relabelled CSR: new row j is original row order[j], entries in their original order, columns renumbered.
// One wave per row.
template <class T>
__global__ __launch_bounds__(TPB) void k_perm_csr(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix,
                                                  const T* __restrict__ P, const int* __restrict__ order,
                                                  const int* __restrict__ inv, const int32_t* __restrict__ ipn, i64 n,
                                                  int32_t* __restrict__ ixn, T* __restrict__ Pn) {
    const i64 row = (i64)blockIdx.x * (TPB / WAVE) + threadIdx.x / WAVE;
    if (row >= n) return;
    const int lane = threadIdx.x % WAVE;
    const int o = order[row];
    const int e0 = ip[o], e1 = ip[o + 1], d0 = ipn[row];
    for (int e = e0 + lane; e < e1; e += WAVE) {
        ixn[d0 + (e - e0)] = inv[ix[e]];
        Pn[d0 + (e - e0)] = P[e];
    }
}

// state in the new labelling: new position j takes original point order[j] from its old position inv_old[order[j]]
template <class T>
__global__ __launch_bounds__(TPB) void k_perm_state(const T* __restrict__ Y, const T* __restrict__ U, const T* __restrict__ G,
                                                    const int* __restrict__ inv_old, const int* __restrict__ order, i64 n,
                                                    T* __restrict__ Yn, T* __restrict__ Un, T* __restrict__ Gn) {
    typedef typename vec2<T>::type V;
    for (i64 j = (i64)blockIdx.x * TPB + threadIdx.x; j < n; j += (i64)gridDim.x * TPB) {
        const int s = inv_old[order[j]];
        reinterpret_cast<V*>(Yn)[j] = reinterpret_cast<const V*>(Y)[s];
        reinterpret_cast<V*>(Un)[j] = reinterpret_cast<const V*>(U)[s];
        reinterpret_cast<V*>(Gn)[j] = reinterpret_cast<const V*>(G)[s];
    }
}

// sort input in original order: key of original point i, value its current position
// Morton (Z-order) index of the box (bx, by) of a point on grid g, clamped to the grid: the bits of bx and by
// interleaved, so that boxes close in the plane are mostly close in the order
template <class T> __device__ inline int morton_key(T x, T y, const GridT<T>& g) {
    const unsigned bx = (unsigned)min(g.nb - 1, max(0, (int)(((double)x - g.xmin_d) * g.inv_w_d)));
    const unsigned by = (unsigned)min(g.nb - 1, max(0, (int)(((double)y - g.ymin_d) / g.width_d)));
    unsigned key = 0;
#pragma unroll
    for (int b = 0; b < 15; ++b) key |= (((bx >> b) & 1u) << (2 * b)) | (((by >> b) & 1u) << (2 * b + 1));
    return (int)key;
}

// relabelling sort input in original order: the Morton key of original point i, value its current position
template <class T>
__global__ __launch_bounds__(TPB) void k_morton_keys(const T* __restrict__ Y, const int* __restrict__ inv, i64 n, GridT<T> g,
                                                     int* __restrict__ keys, int* __restrict__ vals) {
    typedef typename vec2<T>::type V;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        const int p = inv[i];
        const V y = reinterpret_cast<const V*>(Y)[p];
        keys[i] = morton_key<T>(y.x, y.y, g);
        vals[i] = p;
    }
}

// original-order copy of one (n, 2) array: dst[i] = src[inv[i]]
template <class T>
__global__ __launch_bounds__(TPB) void k_unperm(const T* __restrict__ src, const int* __restrict__ inv, i64 n, T* __restrict__ dst) {
    typedef typename vec2<T>::type V;
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB)
        reinterpret_cast<V*>(dst)[i] = reinterpret_cast<const V*>(src)[inv[i]];
}

template <class T>
__global__ __launch_bounds__(TPB) void k_iota(int* __restrict__ a, int* __restrict__ b, i64 n) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) { a[i] = (int)i; b[i] = (int)i; }
}

}  // namespace

void launch_identity_labels(int* perm, int* inv, i64 n, hipStream_t st) {
    hipLaunchKernelGGL(k_iota<int>, dim3(blocks_for(n)), dim3(TPB), 0, st, perm, inv, n);
    MG_HIP(hipGetLastError());
}

// new labelling from the sorted positions `vals` (current labelling perm): order, inv_new, and the relabelled
// CSR (ipn n + 1, ixn / Pn nnz) from the original CSR; scratch: len (n), tiles (n / 1024 + 2)
template <class T> void launch_relabel(const int* vals, const int* perm, i64 n, const int32_t* ip, const int32_t* ix, const T* P,
                                       int* order, int* inv_new, int* len, int* tiles, int32_t* ipn, int32_t* ixn, T* Pn,
                                       hipStream_t st) {
    require(n >= 1 && n < (i64)INT32_MAX, "bad point count");
    hipLaunchKernelGGL(k_order, dim3(blocks_for(n)), dim3(TPB), 0, st, vals, perm, n, order, inv_new, len, ip);
    MG_HIP(hipGetLastError());
    const int ntiles = (int)((n + SCAN_T - 1) / SCAN_T);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(TPB), 0, st, (const int*)len, n, (int*)ipn, tiles);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(TPB), 0, st, tiles, ntiles, (int*)ipn + n);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(TPB), 0, st, (int*)ipn, n, (const int*)tiles);
    MG_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_perm_csr<T>, dim3((unsigned)((n + TPB / WAVE - 1) / (TPB / WAVE))), dim3(TPB), 0, st, ip, ix, P,
                       (const int*)order, (const int*)inv_new, (const int32_t*)ipn, n, ixn, Pn);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_perm_state(const T* Y, const T* U, const T* G, const int* inv_old, const int* order, i64 n,
                                          T* Yn, T* Un, T* Gn, hipStream_t st) {
    hipLaunchKernelGGL(k_perm_state<T>, dim3(blocks_for(n)), dim3(TPB), 0, st, Y, U, G, inv_old, order, n, Yn, Un, Gn);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_morton_keys(const T* Y, const int* inv, i64 n, const GridT<T>& g, int* keys, int* vals,
                                           hipStream_t st) {
    require(g.nb >= 1 && g.nb <= (1 << 15), "grid too large for a Morton key");
    hipLaunchKernelGGL(k_morton_keys<T>, dim3(blocks_for(n)), dim3(TPB), 0, st, Y, inv, n, g, keys, vals);
    MG_HIP(hipGetLastError());
}

template <class T> void launch_unperm(const T* src, const int* inv, i64 n, T* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_unperm<T>, dim3(blocks_for(n)), dim3(TPB), 0, st, src, inv, n, dst);
    MG_HIP(hipGetLastError());
}

#define MG_INST(T) \
    template void launch_relabel<T>(const int*, const int*, i64, const int32_t*, const int32_t*, const T*, int*, int*, int*, int*, int32_t*, int32_t*, T*, hipStream_t); \
    template void launch_perm_state<T>(const T*, const T*, const T*, const int*, const int*, i64, T*, T*, T*, hipStream_t); \
    template void launch_morton_keys<T>(const T*, const int*, i64, const GridT<T>&, int*, int*, hipStream_t); \
    template void launch_unperm<T>(const T*, const int*, i64, T*, hipStream_t);
MG_INST(float)
MG_INST(double)

}  // namespace mg
