// k smallest entries of each row of a distance tile (the selection step of the serving k-NN search): one workgroup
// per row; each thread keeps the KM smallest (distance, column) pairs of its strided share of the row, sorted, by a
// compare-exchange pass per candidate; then the TPB lists are merged pairwise through LDS. The order is total
// (distance, then column), so the result does not depend on the launch shape.
#include "common.hpp"
#include <cfloat>
#include <climits>
#include <cstdint>
#include <cmath>
#include <mutex>
#include <string>

namespace mg {

namespace {

__device__ __forceinline__ bool before(float a, int ia, float b, int ib) { return a < b || (a == b && ia < ib); }

template <int KM, int TPB>
__global__ __launch_bounds__(TPB) void k_topk_rows(const float* __restrict__ D, i64 cols, int k, int32_t* __restrict__ oi,
                                                   float* __restrict__ ov) {
    __shared__ float sd[TPB * KM];
    __shared__ int si[TPB * KM];
    const float* row = D + (i64)blockIdx.x * cols;
    float bd[KM];
    int bi[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) { bd[j] = FLT_MAX; bi[j] = INT_MAX; }
    auto offer = [&](float x, int ix) {
        if (!before(x, ix, bd[KM - 1], bi[KM - 1])) return;
#pragma unroll
        for (int j = 0; j < KM; ++j) {   // the candidate sinks to its place; the largest falls off the end
            const bool sw = before(x, ix, bd[j], bi[j]);
            const float td = bd[j]; const int ti = bi[j];
            bd[j] = sw ? x : td; bi[j] = sw ? ix : ti;
            x = sw ? td : x; ix = sw ? ti : ix;
        }
    };
    // the row in rounds of R float4 per thread (loads first, then the candidates in column order per thread)
    constexpr int R = 4;
    const bool vec = (cols % 4) == 0 && (reinterpret_cast<uintptr_t>(row) % 16) == 0;
    i64 c0 = 0;
    if (vec) {
        const float4* r4 = reinterpret_cast<const float4*>(row);
        const i64 n4 = cols / 4;
        for (i64 b = threadIdx.x; b < n4; b += (i64)TPB * R) {
            float4 v[R];
#pragma unroll
            for (int u = 0; u < R; ++u) { const i64 q = b + (i64)u * TPB; v[u] = q < n4 ? r4[q] : make_float4(FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX); }
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const i64 q = b + (i64)u * TPB;
                if (q >= n4) break;
                const int c = (int)(4 * q);
                offer(v[u].x, c); offer(v[u].y, c + 1); offer(v[u].z, c + 2); offer(v[u].w, c + 3);
            }
        }
        c0 = cols;
    }
    for (i64 b = c0 + threadIdx.x; b < cols; b += (i64)TPB * 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) { const i64 c = b + (i64)u * TPB; v[u] = c < cols ? row[c] : FLT_MAX; }
#pragma unroll
        for (int u = 0; u < 8; ++u) { const i64 c = b + (i64)u * TPB; if (c >= cols) break; offer(v[u], (int)c); }
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < KM; ++j) { sd[t * KM + j] = bd[j]; si[t * KM + j] = bi[j]; }
    __syncthreads();
    for (int s = TPB / 2; s >= 1; s >>= 1) {   // list t := first KM of merge(list t, list t + s)
        float md[KM];
        int mi[KM];
        if (t < s) {
            const float* ad = sd + t * KM; const int* ai = si + t * KM;
            const float* cd = sd + (t + s) * KM; const int* ci = si + (t + s) * KM;
            int p = 0, q = 0;
#pragma unroll
            for (int j = 0; j < KM; ++j) {
                const bool fa = before(ad[p], ai[p], cd[q], ci[q]);
                md[j] = fa ? ad[p] : cd[q]; mi[j] = fa ? ai[p] : ci[q];
                p += fa ? 1 : 0; q += fa ? 0 : 1;
            }
        }
        __syncthreads();
        if (t < s) {
#pragma unroll
            for (int j = 0; j < KM; ++j) { sd[t * KM + j] = md[j]; si[t * KM + j] = mi[j]; }
        }
        __syncthreads();
    }
    if (t < k) { oi[(i64)blockIdx.x * k + t] = si[t]; ov[(i64)blockIdx.x * k + t] = sd[t]; }
}


// ---- fused k-NN search for rows of up to 64 features: exact squared euclidean distances, summed over the
// features in order ((q - t)^2, zero padding adds nothing), so a pair's distance does not depend on the launch.
// Pass 1: block (query tile of 256, training chunk s): one query per thread, the chunk's training rows streamed
// through LDS in tiles of 64, each thread keeping the KM nearest of the chunk. Pass 2: per query, the S chunk
// lists merged through LDS.
constexpr int KNN_Q = 256, KNN_TR = 64, KNN_SMAX = 1024, KNN_MERGE = 256;

template <int DP, int KM>
__global__ __launch_bounds__(KNN_Q) void k_knn_chunks(const float* __restrict__ T, i64 nt, const float* __restrict__ Q, i64 m,
                                                      int d, i64 chunk, i64 self_offset, float* __restrict__ pd,
                                                      int* __restrict__ pi, int S) {
    // rows of DP floats (a multiple of 8), 16-byte aligned: every lane of a wave reads the same training
    // element (a broadcast), four features per LDS read
    static_assert(DP % 8 == 0, "DP must be a multiple of 8");
    __shared__ __align__(16) float sT[KNN_TR][DP];
    const i64 qi = (i64)blockIdx.x * KNN_Q + threadIdx.x;
    const int s = blockIdx.y;
    float q[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) q[j] = (qi < m && j < d) ? Q[qi * d + j] : 0.f;
    float bd[KM];
    int bi[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) { bd[j] = INFINITY; bi[j] = INT_MAX; }   // every finite distance goes before
    const i64 c0 = (i64)s * chunk, c1 = c0 + chunk < nt ? c0 + chunk : nt;
    const i64 self = (self_offset >= 0 && qi < m) ? qi + self_offset : -1;
    for (i64 t0 = c0; t0 < c1; t0 += KNN_TR) {
        __syncthreads();
        for (int e = threadIdx.x; e < KNN_TR * DP; e += KNN_Q) {   // tile rows, zero-padded
            const int r = e / DP, j = e % DP;
            sT[r][j] = (t0 + r < c1 && j < d) ? T[(t0 + r) * d + j] : 0.f;
        }
        __syncthreads();
        const int rows = (int)(c1 - t0 < KNN_TR ? c1 - t0 : KNN_TR);
        for (int r = 0; r < rows; ++r) {
            float acc = 0.f;
            const float4* tr = reinterpret_cast<const float4*>(&sT[r][0]);
#pragma unroll
            for (int j = 0; j < DP / 4; ++j) {   // features in order; the zero padding adds exact zeros
                const float4 t4 = tr[j];
                float df = q[4 * j] - t4.x; acc += df * df;
                df = q[4 * j + 1] - t4.y; acc += df * df;
                df = q[4 * j + 2] - t4.z; acc += df * df;
                df = q[4 * j + 3] - t4.w; acc += df * df;
            }
            const int ix = (int)(t0 + r);
            if (ix == self) continue;
            const float x = acc;
            // A thread visits its rows in increasing index order, so an entry of equal distance already in the
            // list has the smaller index and stays before the new one: the new row goes before entry j iff
            // x < bd[j]. Over the sorted list that test is false up to the insertion point and true after it,
            // so each slot takes the new row (first true), its left neighbour (later trues) or keeps its entry.
            // The entries that move keep their order, ties included.
            if (!(x < bd[KM - 1])) continue;
#pragma unroll
            for (int j = KM - 1; j >= 1; --j) {
                const bool cj = x < bd[j], cp = x < bd[j - 1];
                const float nd = cp ? bd[j - 1] : x;
                const int ni = cp ? bi[j - 1] : ix;
                bd[j] = cj ? nd : bd[j];
                bi[j] = cj ? ni : bi[j];
            }
            if (x < bd[0]) { bd[0] = x; bi[0] = ix; }
        }
    }
    if (qi >= m) return;
    float* od = pd + (qi * S + s) * KM;
    int* oi = pi + (qi * S + s) * KM;
#pragma unroll
    for (int j = 0; j < KM; ++j) { od[j] = bd[j]; oi[j] = bi[j]; }
}

// The same search for rows of more than 64 features: the features go through in chunks of 32 (the query's chunk
// in registers, the tile's in LDS), each thread keeping the running distances of the tile's 32 rows, so every
// distance is still the sum over the features in order; the tile's rows then enter the list in index order.
template <int KM>
__global__ __launch_bounds__(KNN_Q) void k_knn_chunks_wide(const float* __restrict__ T, i64 nt, const float* __restrict__ Q,
                                                           i64 m, int d, i64 chunk, i64 self_offset, float* __restrict__ pd,
                                                           int* __restrict__ pi, int S) {
    constexpr int FC = 32, TRW = 32;   // features per stage, training rows per tile
    __shared__ __align__(16) float sT[TRW][FC];
    const i64 qi = (i64)blockIdx.x * KNN_Q + threadIdx.x;
    const int s = blockIdx.y;
    const float* qrow = Q + (qi < m ? qi : 0) * d;
    float bd[KM];
    int bi[KM];
#pragma unroll
    for (int j = 0; j < KM; ++j) { bd[j] = INFINITY; bi[j] = INT_MAX; }
    const i64 c0 = (i64)s * chunk, c1 = c0 + chunk < nt ? c0 + chunk : nt;
    const i64 self = (self_offset >= 0 && qi < m) ? qi + self_offset : -1;
    for (i64 t0 = c0; t0 < c1; t0 += TRW) {
        float acc[TRW];
#pragma unroll
        for (int r = 0; r < TRW; ++r) acc[r] = 0.f;
        for (int f0 = 0; f0 < d; f0 += FC) {
            float q[FC];
#pragma unroll
            for (int j = 0; j < FC; ++j) q[j] = (qi < m && f0 + j < d) ? qrow[f0 + j] : 0.f;
            __syncthreads();
            for (int e = threadIdx.x; e < TRW * FC; e += KNN_Q) {
                const int r = e / FC, j = e % FC;
                sT[r][j] = (t0 + r < c1 && f0 + j < d) ? T[(t0 + r) * d + f0 + j] : 0.f;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < TRW; ++r) {
                const float4* tr = reinterpret_cast<const float4*>(&sT[r][0]);
                float a = acc[r];
#pragma unroll
                for (int j = 0; j < FC / 4; ++j) {   // features in order; the zero padding adds exact zeros
                    const float4 t4 = tr[j];
                    float df = q[4 * j] - t4.x; a += df * df;
                    df = q[4 * j + 1] - t4.y; a += df * df;
                    df = q[4 * j + 2] - t4.z; a += df * df;
                    df = q[4 * j + 3] - t4.w; a += df * df;
                }
                acc[r] = a;
            }
        }
        const int rows = (int)(c1 - t0 < TRW ? c1 - t0 : TRW);
#pragma unroll
        for (int r = 0; r < TRW; ++r) {
            const int ix = (int)(t0 + r);
            if (r >= rows || ix == self) continue;
            const float x = acc[r];
            if (!(x < bd[KM - 1])) continue;   // insertion as in k_knn_chunks: ties keep index order
#pragma unroll
            for (int j = KM - 1; j >= 1; --j) {
                const bool cj = x < bd[j], cp = x < bd[j - 1];
                const float nd = cp ? bd[j - 1] : x;
                const int ni = cp ? bi[j - 1] : ix;
                bd[j] = cj ? nd : bd[j];
                bi[j] = cj ? ni : bi[j];
            }
            if (x < bd[0]) { bd[0] = x; bi[0] = ix; }
        }
    }
    if (qi >= m) return;
    float* od = pd + (qi * S + s) * KM;
    int* oi = pi + (qi * S + s) * KM;
#pragma unroll
    for (int j = 0; j < KM; ++j) { od[j] = bd[j]; oi[j] = bi[j]; }
}

// one block of P = min(S, 256) threads per query (S and P powers of two): thread t first merges the chunk lists
// t, t + P, t + 2P, ... in turn, then the P lists are merged as a tree through LDS
template <int KM>
__device__ __forceinline__ void knn_merge2(const float* ad, const int* ai, const float* cd, const int* ci, float (&md)[KM],
                                           int (&mi)[KM]) {
    int p = 0, r = 0;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
        const bool fa = before(ad[p], ai[p], cd[r], ci[r]);
        md[j] = fa ? ad[p] : cd[r]; mi[j] = fa ? ai[p] : ci[r];
        p += fa ? 1 : 0; r += fa ? 0 : 1;
    }
}
template <int KM>
__global__ void k_knn_merge(const float* __restrict__ pd, const int* __restrict__ pi, int S, int k, int32_t* __restrict__ oi,
                            float* __restrict__ ov) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int P = (int)blockDim.x;
    float* sd = reinterpret_cast<float*>(smem);
    int* si = reinterpret_cast<int*>(sd + P * KM);
    const i64 qi = blockIdx.x;
    const int t = threadIdx.x;
    for (int j = 0; j < KM; ++j) { sd[t * KM + j] = pd[(qi * S + t) * KM + j]; si[t * KM + j] = pi[(qi * S + t) * KM + j]; }
    for (int l = t + P; l < S; l += P) {   // only this thread's LDS slot: no barrier needed
        float md[KM];
        int mi[KM];
        knn_merge2<KM>(sd + t * KM, si + t * KM, pd + (qi * S + l) * KM, pi + (qi * S + l) * KM, md, mi);
#pragma unroll
        for (int j = 0; j < KM; ++j) { sd[t * KM + j] = md[j]; si[t * KM + j] = mi[j]; }
    }
    __syncthreads();
    for (int s = P / 2; s >= 1; s >>= 1) {
        float md[KM];
        int mi[KM];
        if (t < s) knn_merge2<KM>(sd + t * KM, si + t * KM, sd + (t + s) * KM, si + (t + s) * KM, md, mi);
        __syncthreads();
        if (t < s) {
#pragma unroll
            for (int j = 0; j < KM; ++j) { sd[t * KM + j] = md[j]; si[t * KM + j] = mi[j]; }
        }
        __syncthreads();
    }
    for (int j = t; j < k; j += P) { oi[qi * k + j] = si[j]; ov[qi * k + j] = sd[j]; }   // P may be below k
}
}  // namespace

void launch_topk_rows(const float* D, i64 rows, i64 cols, int k, int32_t* oi, float* ov, hipStream_t st) {
    require(rows >= 0 && cols >= 1 && cols < (i64)INT32_MAX, "bad tile shape");
    require(k >= 1 && k <= 32 && k <= cols, "k must be in [1, min(32, cols)]");
    if (rows == 0) return;
    if (k <= 16) hipLaunchKernelGGL((k_topk_rows<16, 256>), dim3((unsigned)rows), dim3(256), 0, st, D, cols, k, oi, ov);
    else hipLaunchKernelGGL((k_topk_rows<32, 128>), dim3((unsigned)rows), dim3(128), 0, st, D, cols, k, oi, ov);
    MG_HIP(hipGetLastError());
}

}  // namespace mg

namespace mg {
// k nearest training rows of each query row (any number of features): indices and squared distances (m, k), sorted by
// (distance, index); self_offset >= 0 leaves out training row q + self_offset for query q. Scratch: part_d / part_i,
// m * S * 32 each, S = knn_chunks(m, nt).
int knn_chunks(i64 m, i64 nt) {
    const i64 tiles = (m + KNN_Q - 1) / KNN_Q;
    int S = 1;
    while (S < KNN_SMAX && tiles * S < 1024 && (i64)S * 2 * KNN_TR <= nt) S *= 2;   // about 1024 blocks, >= 64 rows per chunk
    return S;
}
// the search with S training chunks per query (a power of two); the result does not depend on S: each chunk list is
// its chunk's exact nearest KM by (distance, index), and the merge keeps that order
static void knn_with_chunks(const float* T, i64 nt, const float* Q, i64 m, int d, int k, i64 self_offset, int S,
                            float* part_d, int* part_i, int32_t* oi, float* ov, hipStream_t st) {
    const i64 chunk = (nt + S - 1) / S;
    const dim3 g1((unsigned)((m + KNN_Q - 1) / KNN_Q), (unsigned)S);
#define MG_K(DP, KM) hipLaunchKernelGGL((k_knn_chunks<DP, KM>), g1, dim3(KNN_Q), 0, st, T, nt, Q, m, d, chunk, self_offset, part_d, part_i, S)
    // the query row in registers, padded to the next multiple of 8 features
#define MG_KD(KM) switch ((d + 7) / 8) { \
        case 1: MG_K(8, KM); break; case 2: MG_K(16, KM); break; case 3: MG_K(24, KM); break; \
        case 4: MG_K(32, KM); break; case 5: MG_K(40, KM); break; case 6: MG_K(48, KM); break; \
        case 7: MG_K(56, KM); break; default: MG_K(64, KM); break; }
    if (d > 64) {
        if (k <= 16) hipLaunchKernelGGL((k_knn_chunks_wide<16>), g1, dim3(KNN_Q), 0, st, T, nt, Q, m, d, chunk, self_offset, part_d, part_i, S);
        else hipLaunchKernelGGL((k_knn_chunks_wide<32>), g1, dim3(KNN_Q), 0, st, T, nt, Q, m, d, chunk, self_offset, part_d, part_i, S);
    } else if (k <= 16) { MG_KD(16) } else { MG_KD(32) }
#undef MG_KD
#undef MG_K
    MG_HIP(hipGetLastError());
    const int KM = k <= 16 ? 16 : 32;
    const int P = S < KNN_MERGE ? S : KNN_MERGE;
    const size_t lds = (size_t)P * KM * (sizeof(float) + sizeof(int));   // at most 64 KiB
    if (KM == 16) hipLaunchKernelGGL(k_knn_merge<16>, dim3((unsigned)m), dim3(P), lds, st, part_d, part_i, S, k, oi, ov);
    else hipLaunchKernelGGL(k_knn_merge<32>, dim3((unsigned)m), dim3(P), lds, st, part_d, part_i, S, k, oi, ov);
    MG_HIP(hipGetLastError());
}

static void knn_check(i64 nt, int d, int k, i64 self_offset) {
    require(d >= 1, "rows must have at least one feature");
    require(k >= 1 && k <= 32 && k <= nt - (self_offset >= 0 ? 1 : 0), "k must be in [1, min(32, rows)]");
}

void launch_knn(const float* T, i64 nt, const float* Q, i64 m, int d, int k, i64 self_offset, float* part_d, int* part_i,
                int32_t* oi, float* ov, hipStream_t st) {
    knn_check(nt, d, k, self_offset);
    if (m == 0) return;
    knn_with_chunks(T, nt, Q, m, d, k, self_offset, knn_chunks(m, nt), part_d, part_i, oi, ov, st);
}

// ---- the allocation-free search (mg_knn_noalloc): its chunk lists live in `scratch`, a fixed array of this file,
// KNN_SCRATCH_ENTRIES (distance, index) pairs (8 MiB: the distances, then the indices). Each device holds its own
// copy until the process ends. Every list slot is written by the chunk pass before the merge reads it, so nothing
// needs clearing between calls; calls that overlap on one device share the array and race.
constexpr i64 KNN_SCRATCH_ENTRIES = (i64)1 << 20;
static __device__ __align__(256) unsigned char scratch[KNN_SCRATCH_ENTRIES * (sizeof(float) + sizeof(int))];

constexpr int KNN_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[KNN_MAX_DEVICES] = {};

// The device address of this device's copy of `scratch`, looked up once per device with that device current. The
// host-side name `scratch` is only a stand-in: kernels get this address.
static unsigned char* knn_scratch_on_current_device() {
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < KNN_MAX_DEVICES, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "no device address for the k-NN scratch");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

i64 knn_noalloc_max_queries(int k) { return KNN_SCRATCH_ENTRIES / (k <= 16 ? 16 : 32); }

void launch_knn_noalloc(const float* T, i64 nt, const float* Q, i64 m, int d, int k, i64 self_offset, int32_t* oi, float* ov,
                        hipStream_t st) {
    knn_check(nt, d, k, self_offset);
    const i64 KM = k <= 16 ? 16 : 32, cap = knn_noalloc_max_queries(k);
    require(m <= cap, "mg_knn_noalloc takes at most " + std::to_string(cap) + " query rows for k = " + std::to_string(k) +
                          " (its fixed scratch holds " + std::to_string(KNN_SCRATCH_ENTRIES) + " list entries); got " +
                          std::to_string(m));
    if (m == 0) return;
    unsigned char* base = knn_scratch_on_current_device();
    int S = knn_chunks(m, nt);
    while (S > 1 && m * S * KM > KNN_SCRATCH_ENTRIES) S /= 2;   // fewer, longer chunks: the same result
    float* part_d = reinterpret_cast<float*>(base);
    int* part_i = reinterpret_cast<int*>(base + KNN_SCRATCH_ENTRIES * sizeof(float));
    knn_with_chunks(T, nt, Q, m, d, k, self_offset, S, part_d, part_i, oi, ov, st);
}
}  // namespace mg

namespace mg {
namespace {
// UMAP's smooth k-NN memberships of one row of sorted neighbour distances (the numpy restatement of the
// reference, magnet/knn.py:165-198, in float32): rho = the nearest non-zero distance, sigma by bisection so that
// sum exp(-(d - rho) / sigma) = log2(n_neighbors) * bandwidth, floored at 1e-3 of the row mean (of `mean_all`
// when the row has no non-zero distance).
template <int KM>
__global__ __launch_bounds__(256) void k_smooth_knn(const float* __restrict__ dist, i64 m, int k, float target, int n_iter,
                                                    const float* __restrict__ mean_all, float* __restrict__ vals) {
    const i64 row = (i64)blockIdx.x * 256 + threadIdx.x;
    if (row >= m) return;
    float d[KM];
    float rho = 0.f, sum = 0.f;
    bool has = false;
#pragma unroll
    for (int j = 0; j < KM; ++j) {
        d[j] = j < k ? dist[row * k + j] : 0.f;
        if (j < k) { sum += d[j]; if (!has && d[j] > 0.f) { rho = d[j]; has = true; } }
    }
#pragma unroll
    for (int j = 0; j < KM; ++j) d[j] = j < k ? fmaxf(d[j] - rho, 0.f) : 0.f;
    float lo = 0.f, hi = INFINITY, mid = 1.f, sigma = 1.f;
    bool done = false;
    for (int it = 0; it < n_iter && !done; ++it) {
        float psum = 0.f;
#pragma unroll
        for (int j = 0; j < KM; ++j) if (j < k) psum += expf(-d[j] / mid);
        if (fabsf(psum - target) < 1e-5f) { sigma = mid; done = true; break; }
        const bool gt = psum > target;
        hi = gt ? mid : hi;
        lo = gt ? lo : mid;
        mid = gt ? (lo + hi) / 2.f : (isinf(hi) ? mid * 2.f : (lo + hi) / 2.f);
    }
    if (!done) sigma = mid;
    const float mean_row = sum / (float)k;
    sigma = has ? fmaxf(sigma, 1e-3f * mean_row) : fmaxf(sigma, 1e-3f * mean_all[row]);
    for (int j = 0; j < k; ++j) vals[row * k + j] = d[j] <= 0.f ? 1.f : expf(-d[j] / sigma);
}
}  // namespace

void launch_smooth_knn(const float* dist, i64 m, int k, int n_neighbors, const float* mean_all, float* vals, hipStream_t st) {
    require(k >= 1 && k <= 32, "k must be in [1, 32]");
    if (m == 0) return;
    const float target = (float)std::log2((double)n_neighbors);
    const dim3 grid((unsigned)((m + 255) / 256));
    if (k <= 16) hipLaunchKernelGGL(k_smooth_knn<16>, grid, dim3(256), 0, st, dist, m, k, target, 64, mean_all, vals);
    else hipLaunchKernelGGL(k_smooth_knn<32>, grid, dim3(256), 0, st, dist, m, k, target, 64, mean_all, vals);
    MG_HIP(hipGetLastError());
}
}  // namespace mg
