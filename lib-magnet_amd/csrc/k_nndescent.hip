// NN-descent construction of the k-NN graph on the device: the reference's nn_descent (src/nndescent.cpp:390-583,
// 636-645): random-projection trees whose leaves are joined as cliques, random neighbours for the slots still open,
// then rounds of sampling new and old candidates (forward and reverse), joining every node's candidates pairwise and
// keeping the k nearest, until a round changes at most delta * n * k entries. Squared euclidean distances.
//
// The device form:
// - A node's list is k keys (squared distance bits << 32 | neighbour), sorted, empty slots last; a flag per slot marks
//   entries not yet joined ("new").
// - Trees are built level by level: every point of a segment larger than the leaf size goes to the side of the nearer
//   of two pivots drawn from it (the perpendicular bisector of the reference's plane, src/nndescent.cpp:238-330), and
//   one stable radix sort by (segment, side) splits every segment in place.
// - A join takes a group of nodes (a leaf, or one node's new and old candidates), stages their rows in LDS by
//   chunks of features, and forms the distances of its new-new and new-old pairs. A first pass counts the pairs
//   that beat an end's current worst entry, a scan turns the counts into positions, and a second pass writes
//   those proposals there, so the buffer is sized exactly and filled in a fixed order.
// - Proposals are sorted by target with the stable radix sort of the repulsion pass, and one wave per node merges
//   its proposals into its list 64 at a time (a wave sort, then a merge by rank that keeps k).
// - Candidates: the reverse lists come from a sort of all entries by neighbour; per node and kind (new / old) a wave
//   keeps the max_candidates records of smallest hashed priority, forward and reverse together (the reference's
//   priority heaps, src/nndescent.cpp:350-385, 491-513); the forward new entries picked are then unflagged.
// Every distance is the same sequence of fused multiply-adds over the features in order, wherever it is formed, and
// counts use integer atomics only, so a run is deterministic. Random numbers come from a hash of (seed, round,
// indices) rather than the reference's per-thread generators: the graph agrees with the reference's in law (recall
// against exact search), not in value.
#include "common.hpp"
#include <algorithm>
#include <chrono>
#include <vector>

namespace mg {

namespace {

constexpr int TPB = 256;
constexpr int SMAX = 128;                                        // members of one join pass
constexpr int DC = 32;                                           // features per LDS stage
constexpr int PPT_MAX = (SMAX * (SMAX - 1) / 2 + TPB - 1) / TPB;  // pairs per thread at most
constexpr int KMAX = 128;                                        // list length at most
constexpr uint64_t LIST_EMPTY = 0x7F800000FFFFFFFFULL;           // (inf, -1)
constexpr uint64_t CAND_EMPTY = ~0ULL;

__device__ __host__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return mix64(seed ^ mix64(a ^ mix64(b ^ mix64(c))));
}
__device__ inline uint64_t pack(float d, int idx) { return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)idx; }
__device__ inline int key_node(uint64_t k) { return (int)(uint32_t)k; }

// squared distance of two rows: one fused multiply-add per feature, in feature order (the join's sequence)
__device__ inline float row_dist(const float* __restrict__ X, int d, int a, int b) {
    const float* xa = X + (i64)a * d;
    const float* xb = X + (i64)b * d;
    float acc = 0.f;
    for (int f = 0; f < d; ++f) {
        const float t = xa[f] - xb[f];
        acc = fmaf(t, t, acc);
    }
    return acc;
}

__global__ __launch_bounds__(TPB) void k_fill_u64(uint64_t* __restrict__ a, i64 n, uint64_t v) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) a[i] = v;
}
__global__ __launch_bounds__(TPB) void k_fill_i32(int* __restrict__ a, i64 n, int v) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) a[i] = v;
}
__global__ __launch_bounds__(TPB) void k_fill_u8(uint8_t* __restrict__ a, i64 n, uint8_t v) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) a[i] = v;
}
__global__ __launch_bounds__(TPB) void k_iota(int* __restrict__ a, i64 n) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) a[i] = (int)i;
}

// ------------------------------------------------------------------------------------------------ trees
// tb trees of n points: position q holds point q % n, in segment q / n (tree q / n's root)
__global__ __launch_bounds__(TPB) void k_rp_init(int* __restrict__ perm, int* __restrict__ seg_of, int* __restrict__ seg_start,
                                                 int n, int tb) {
    const int N = n * tb;
    for (int q = blockIdx.x * TPB + threadIdx.x; q < N; q += gridDim.x * TPB) {
        perm[q] = q % n;
        seg_of[q] = q / n;
        if (q % n == 0) seg_start[q / n] = q;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) seg_start[tb] = N;
}
// key of position p (N positions, trees of n) = 2 * segment + side; segments of at most `leaf` points keep side 0
__global__ __launch_bounds__(TPB) void k_rp_keys(const float* __restrict__ X, int d, const int* __restrict__ perm,
                                                 const int* __restrict__ seg_of, const int* __restrict__ seg_start, int N, int n,
                                                 int leaf, uint64_t seed, int t0, int depth, int* __restrict__ keys) {
    for (int p = blockIdx.x * TPB + threadIdx.x; p < N; p += gridDim.x * TPB) {
        const int s = seg_of[p], b0 = seg_start[s], sz = seg_start[s + 1] - b0;
        const int tree = t0 + p / n;
        int side = 0;
        if (sz > leaf) {
            const uint64_t h = hash4(seed, 0x7265u + t0, depth, s);
            const int ia = b0 + (int)((h & 0xffffffffu) % (uint32_t)sz);
            int ib = b0 + (int)((h >> 32) % (uint32_t)sz);
            if (ib == ia) ib = b0 + (ia - b0 + 1) % sz;
            const int x = perm[p];
            const float da = row_dist(X, d, x, perm[ia]), db = row_dist(X, d, x, perm[ib]);
            side = da < db ? 0 : da > db ? 1 : (int)(hash4(seed ^ 0x5bd1e995u, tree, depth, x) & 1u);
        }
        keys[p] = 2 * s + side;
    }
}
// first position of every run of equal keys
__global__ __launch_bounds__(TPB) void k_rp_marks(const int* __restrict__ keys, int n, int* __restrict__ flag) {
    for (int p = blockIdx.x * TPB + threadIdx.x; p < n; p += gridDim.x * TPB) flag[p] = p == 0 || keys[p] != keys[p - 1];
}
// segment of every position and the segments' starts; ex = exclusive scan of the marks (ex[n] = segment count)
__global__ __launch_bounds__(TPB) void k_rp_segments(const int* __restrict__ flag, const int* __restrict__ ex, int n,
                                                     int* __restrict__ seg_of, int* __restrict__ seg_start) {
    for (int p = blockIdx.x * TPB + threadIdx.x; p < n; p += gridDim.x * TPB) {
        const int s = ex[p] + flag[p] - 1;
        seg_of[p] = s;
        if (flag[p]) seg_start[s] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) seg_start[ex[n]] = n;
}
__global__ __launch_bounds__(TPB) void k_rp_maxseg(const int* __restrict__ seg_start, const int* __restrict__ nseg_p,
                                                   int* __restrict__ mx) {
    const int nseg = *nseg_p;
    int m = 0;
    for (int s = blockIdx.x * TPB + threadIdx.x; s < nseg; s += gridDim.x * TPB) m = max(m, seg_start[s + 1] - seg_start[s]);
    for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if (lane_id() == 0) atomicMax(mx, m);
}
// leaf groups: every segment, all members new
__global__ __launch_bounds__(TPB) void k_leaf_groups(const int* __restrict__ seg_start, int nseg, int* __restrict__ gbeg,
                                                     int* __restrict__ gsize, int* __restrict__ gnew) {
    for (int s = blockIdx.x * TPB + threadIdx.x; s < nseg; s += gridDim.x * TPB) {
        gbeg[s] = seg_start[s];
        gsize[s] = gnew[s] = seg_start[s + 1] - seg_start[s];
    }
}

// ------------------------------------------------------------------------------------------------ join
__device__ inline void pair_of(int p, int nn, int s, int& a, int& b) {
    const int T = nn * (nn - 1) / 2;
    if (p < T) {
        int bb = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
        while (bb * (bb - 1) / 2 > p) --bb;
        while ((bb + 1) * bb / 2 <= p) ++bb;
        a = p - bb * (bb - 1) / 2;
        b = bb;
    } else {
        const int q = p - T, no = s - nn;
        a = q / no;
        b = nn + q % no;
    }
}

// exclusive prefix over the block's threads (a thread-count exchange)
__device__ inline int block_exscan_i(int v, int* sh, int* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < TPB; o <<= 1) {
        const int u = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int r = sh[t] - v;
    *total = sh[TPB - 1];
    __syncthreads();
    return r;
}

// One group per block: members[gbeg[g] .. + gsize[g]), the first gnew[g] of them new. Pairs new-new and new-old; a
// group larger than SMAX is joined as cliques of SMAX consecutive members. WRITE = false: count[g] = proposals that
// beat an end's worst entry; WRITE = true: those proposals at offs[g] on, in a fixed order, and tcount[target] += 1.
template <bool WRITE, int PPT>
__global__ __launch_bounds__(TPB) void k_join(const float* __restrict__ X, int d, const int* __restrict__ gbeg, int gstride,
                                              const int* __restrict__ gsize, const int* __restrict__ gnew,
                                              const int* __restrict__ members, int g0, int ng, const uint64_t* __restrict__ L,
                                              int K, int* __restrict__ count, const int* __restrict__ offs,
                                              int* __restrict__ Ptgt, int* __restrict__ Psrc, float* __restrict__ Pdist,
                                              int* __restrict__ tcount) {
    __shared__ float xs[SMAX][DC + 1];
    __shared__ int mem[SMAX];
    __shared__ uint64_t wk[SMAX];
    __shared__ int sh[TPB];
    const int g = g0 + blockIdx.x;
    if (blockIdx.x >= ng) return;
    const i64 beg = gbeg ? (i64)gbeg[g] : (i64)g * gstride;
    const int size = gsize[g], nnew = gnew[g];
    if (nnew == 0) {   // no new member: no pair to join
        if (!WRITE && threadIdx.x == 0) count[blockIdx.x] = 0;
        return;
    }
    int base = WRITE ? offs[blockIdx.x] : 0;
    int total_count = 0;
    for (int c0 = 0; c0 < size; c0 += SMAX) {
        const int s = min(SMAX, size - c0), nn = min(max(nnew - c0, 0), s);
        const int P = nn * (nn - 1) / 2 + nn * (s - nn);
        for (int j = threadIdx.x; j < s; j += TPB) {
            const int v = members[beg + c0 + j];
            mem[j] = v;
            wk[j] = L[(i64)v * K + K - 1];
        }
        int pab[PPT];
        float acc[PPT];
#pragma unroll
        for (int r = 0; r < PPT; ++r) {
            const int p = threadIdx.x + r * TPB;
            int a = 0, b = 0;
            if (p < P) pair_of(p, nn, s, a, b);
            pab[r] = a | (b << 8);
            acc[r] = 0.f;
        }
        for (int f0 = 0; f0 < d; f0 += DC) {
            const int fc = min(DC, d - f0);
            __syncthreads();
            for (int e = threadIdx.x; e < s * DC; e += TPB) {
                const int j = e / DC, f = e % DC;
                if (f < fc) xs[j][f] = X[(i64)mem[j] * d + f0 + f];
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < PPT; ++r) {
                if (threadIdx.x + r * TPB < P) {
                    const int a = pab[r] & 0xff, b = pab[r] >> 8;
                    float v = acc[r];
                    for (int f = 0; f < fc; ++f) {
                        const float t = xs[a][f] - xs[b][f];
                        v = fmaf(t, t, v);
                    }
                    acc[r] = v;
                }
            }
        }
        // proposals of this chunk: a <- b when (dist, b) beats a's worst, and b <- a likewise
        int mine = 0;
#pragma unroll
        for (int r = 0; r < PPT; ++r) {
            if (threadIdx.x + r * TPB < P) {
                const int a = pab[r] & 0xff, b = pab[r] >> 8;
                const int ia = mem[a], ib = mem[b];
                if (ia != ib) mine += (pack(acc[r], ib) < wk[a]) + (pack(acc[r], ia) < wk[b]);
            }
        }
        int chunk_total;
        const int ex = block_exscan_i(mine, sh, &chunk_total);
        if (WRITE) {
            int pos = base + ex;
#pragma unroll
            for (int r = 0; r < PPT; ++r) {
                if (threadIdx.x + r * TPB < P) {
                    const int a = pab[r] & 0xff, b = pab[r] >> 8;
                    const int ia = mem[a], ib = mem[b];
                    if (ia == ib) continue;
                    if (pack(acc[r], ib) < wk[a]) { Ptgt[pos] = ia; Psrc[pos] = ib; Pdist[pos] = acc[r]; atomicAdd(&tcount[ia], 1); ++pos; }
                    if (pack(acc[r], ia) < wk[b]) { Ptgt[pos] = ib; Psrc[pos] = ia; Pdist[pos] = acc[r]; atomicAdd(&tcount[ib], 1); ++pos; }
                }
            }
        }
        base += chunk_total;
        total_count += chunk_total;
        __syncthreads();
    }
    if (!WRITE && threadIdx.x == 0) count[blockIdx.x] = total_count;
}

// ------------------------------------------------------------------------------------------------ wave merge
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int o) {
    const int lo = __shfl_xor((int)(uint32_t)v, o), hi = __shfl_xor((int)(uint32_t)(v >> 32), o);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
__device__ inline uint64_t shfl_up_u64(uint64_t v, int o) {
    const int lo = __shfl_up((int)(uint32_t)v, o), hi = __shfl_up((int)(uint32_t)(v >> 32), o);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
// ascending bitonic sort of one key per lane across the wave
__device__ inline uint64_t wave_sort(uint64_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int size = 2; size <= WAVE; size <<= 1) {
#pragma unroll
        for (int stride = size / 2; stride > 0; stride >>= 1) {
            const uint64_t o = shfl_xor_u64(v, stride);
            const bool up = (lane & size) == 0, low = (lane & stride) == 0;
            v = (low == up) ? (v < o ? v : o) : (v < o ? o : v);
        }
    }
    return v;
}
__device__ inline int lower_bound(const uint64_t* a, int n, uint64_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Merges one key per lane (or `empty`) into the sorted buffer A (cap slots, empties last) keeping the cap smallest;
// keys whose node (low 32 bits) is already in A, and repeats within the chunk, are dropped. New entries get flag 1.
// One wave per block: the block barriers order the LDS traffic. Returns the entries inserted (wave-uniform).
__device__ inline int merge_chunk(uint64_t key, uint64_t empty, uint64_t*& A, uint64_t*& B, uint8_t*& fA, uint8_t*& fB,
                                  int cap, uint64_t* surv) {
    const int lane = lane_id();
    bool valid = key < A[cap - 1];
    if (valid)
        for (int c = 0; c < cap; ++c)
            if ((uint32_t)A[c] == (uint32_t)key) { valid = false; break; }
    key = wave_sort(valid ? key : CAND_EMPTY);
    const uint64_t prev = shfl_up_u64(key, 1);
    const bool keep = key != CAND_EMPTY && !(lane > 0 && (uint32_t)prev == (uint32_t)key);
    const uint64_t bal = __ballot(keep);
    const int c = __popcll(bal);
    if (c == 0) return 0;
    const int pos = __popcll(bal & ((1ull << lane) - 1ull));
    if (keep) surv[pos] = key;
    __syncthreads();
    int ins = 0;
    if (lane < c) {
        const uint64_t kj = surv[lane];
        const int p = lower_bound(A, cap, kj) + lane;
        if (p < cap) { B[p] = kj; if (fB) fB[p] = 1; ins = 1; }
    }
    for (int i = lane; i < cap; i += WAVE) {
        const uint64_t ka = A[i];
        const int p = i + lower_bound(surv, c, ka);
        if (p < cap) { B[p] = ka; if (fB) fB[p] = fA[i]; }
    }
    __syncthreads();
    uint64_t* t = A; A = B; B = t;
    uint8_t* u = fA; fA = fB; fB = u;
    (void)empty;
    for (int o = WAVE / 2; o > 0; o >>= 1) ins += __shfl_xor(ins, o);
    return ins;
}

// one node per 64-thread block: its proposals (sorted positions [tptr[i], tptr[i+1]) of sv) into its list
__global__ __launch_bounds__(WAVE) void k_merge(const int* __restrict__ sv, const int* __restrict__ tptr,
                                                const int* __restrict__ Psrc, const float* __restrict__ Pdist, int n, int K,
                                                uint64_t* __restrict__ L, uint8_t* __restrict__ F, int* __restrict__ updates) {
    __shared__ uint64_t bufA[KMAX], bufB[KMAX], surv[WAVE];
    __shared__ uint8_t flA[KMAX], flB[KMAX];
    const int i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const int q0 = tptr[i], q1 = tptr[i + 1];
    if (q1 == q0) return;
    uint64_t *A = bufA, *B = bufB;
    uint8_t *fA = flA, *fB = flB;
    for (int c = lane; c < K; c += WAVE) { A[c] = L[(i64)i * K + c]; fA[c] = F[(i64)i * K + c]; }
    __syncthreads();
    int ins = 0;
    for (int q = q0; q < q1; q += WAVE) {
        uint64_t key = CAND_EMPTY;
        if (q + lane < q1) { const int pi = sv[q + lane]; key = pack(Pdist[pi], Psrc[pi]); }
        ins += merge_chunk(key, LIST_EMPTY, A, B, fA, fB, K, surv);
    }
    for (int c = lane; c < K; c += WAVE) { L[(i64)i * K + c] = A[c]; F[(i64)i * K + c] = fA[c]; }
    if (lane == 0 && ins) atomicAdd(updates, ins);
}

// ------------------------------------------------------------------------------------------------ candidates
// reverse records: key = neighbour of entry e (n for an empty slot), value e; rcount[neighbour] += 1
__global__ __launch_bounds__(TPB) void k_rev_keys(const uint64_t* __restrict__ L, i64 nK, int n, int* __restrict__ keys,
                                                  int* __restrict__ vals, int* __restrict__ rcount) {
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < nK; e += (i64)gridDim.x * TPB) {
        const uint64_t k = L[e];
        const int j = k == LIST_EMPTY ? n : key_node(k);
        keys[e] = j;
        vals[e] = (int)e;
        if (j < n) atomicAdd(&rcount[j], 1);
    }
}
// per node and kind, the m records of smallest priority among its forward entries and its reverse records (i in
// the list of j); the group of node i: members[i * 2m ..] = new candidates, then old ones
__global__ __launch_bounds__(WAVE) void k_candidates(const uint64_t* __restrict__ L, const uint8_t* __restrict__ F, int K,
                                                     const int* __restrict__ rptr, const int* __restrict__ rsv, int n, int m,
                                                     uint64_t seed, int it, int* __restrict__ members, int* __restrict__ gnew,
                                                     int* __restrict__ gsize) {
    __shared__ uint64_t bufA[WAVE], bufB[WAVE], surv[WAVE];
    const int i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    int got[2];
    for (int kind = 1; kind >= 0; --kind) {
        uint64_t *A = bufA, *B = bufB;
        uint8_t *fA = nullptr, *fB = nullptr;
        A[lane] = CAND_EMPTY;
        __syncthreads();
        for (int c0 = 0; c0 < K; c0 += WAVE) {   // forward entries
            const int c = c0 + lane;
            uint64_t key = CAND_EMPTY;
            if (c < K) {
                const i64 e = (i64)i * K + c;
                const uint64_t k = L[e];
                if (k != LIST_EMPTY && F[e] == kind)
                    key = ((hash4(seed, 0x63616e64u + it, (uint64_t)e, 0) >> 32) << 32) | (uint32_t)key_node(k);
            }
            merge_chunk(key, CAND_EMPTY, A, B, fA, fB, m, surv);
        }
        for (int r0 = rptr[i]; r0 < rptr[i + 1]; r0 += WAVE) {   // reverse records
            const int r = r0 + lane;
            uint64_t key = CAND_EMPTY;
            if (r < rptr[i + 1]) {
                const int e = rsv[r];
                if (F[e] == kind) key = ((hash4(seed, 0x63616e64u + it, (uint64_t)e, 0) >> 32) << 32) | (uint32_t)(e / K);
            }
            merge_chunk(key, CAND_EMPTY, A, B, fA, fB, m, surv);
        }
        const int cnt = __popcll(__ballot(lane < m && A[lane] != CAND_EMPTY));
        const int off = kind == 1 ? 0 : got[1];
        if (lane < cnt) members[(i64)i * 2 * m + off + lane] = key_node(A[lane]);
        got[kind] = cnt;
        __syncthreads();
    }
    if (lane == 0) { gnew[i] = got[1]; gsize[i] = got[1] + got[0]; }
}
// forward new entries picked as new candidates are no longer new (src/nndescent.cpp:517-522)
__global__ __launch_bounds__(WAVE) void k_unflag(const uint64_t* __restrict__ L, uint8_t* __restrict__ F, int K,
                                                 const int* __restrict__ members, const int* __restrict__ gnew, int n, int m) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const int nn = gnew[i];
    for (int c = threadIdx.x; c < K; c += WAVE) {
        const i64 e = (i64)i * K + c;
        if (!F[e]) continue;
        const int j = key_node(L[e]);
        for (int t = 0; t < nn; ++t)
            if (members[(i64)i * 2 * m + t] == j) { F[e] = 0; break; }
    }
}

// ------------------------------------------------------------------------------------------------ random fill
// slots (i, s < 2K): for a node with `need` open slots, the first 2 need slots propose random other nodes
__global__ __launch_bounds__(TPB) void k_fill_random(const float* __restrict__ X, int d, const uint64_t* __restrict__ L, int n,
                                                     int K, uint64_t seed, int round, int* __restrict__ Ptgt,
                                                     int* __restrict__ Psrc, float* __restrict__ Pdist, int* __restrict__ tcount) {
    const i64 total = (i64)n * 2 * K;
    for (i64 q = (i64)blockIdx.x * TPB + threadIdx.x; q < total; q += (i64)gridDim.x * TPB) {
        const int i = (int)(q / (2 * K)), s = (int)(q % (2 * K));
        const uint64_t* row = L + (i64)i * K;
        const int filled = lower_bound(row, K, LIST_EMPTY);
        int tgt = n;
        if (s < 2 * (K - filled)) {
            int j = (int)(hash4(seed, 0x66696c6cu + round, (uint64_t)i, (uint64_t)s) % (uint64_t)n);
            if (j == i) j = (j + 1) % n;
            tgt = i;
            Psrc[q] = j;
            Pdist[q] = row_dist(X, d, i, j);
            atomicAdd(&tcount[i], 1);
        }
        Ptgt[q] = tgt;
    }
}
// every other node proposed to each node whose list is not full (small n only)
__global__ __launch_bounds__(TPB) void k_fill_all(const float* __restrict__ X, int d, const uint64_t* __restrict__ L, int n,
                                                  int K, int* __restrict__ Ptgt, int* __restrict__ Psrc, float* __restrict__ Pdist,
                                                  int* __restrict__ tcount) {
    const i64 total = (i64)n * n;
    for (i64 q = (i64)blockIdx.x * TPB + threadIdx.x; q < total; q += (i64)gridDim.x * TPB) {
        const int i = (int)(q / n), j = (int)(q % n);
        int tgt = n;
        if (j != i && L[(i64)i * K + K - 1] == LIST_EMPTY) {
            tgt = i;
            Psrc[q] = j;
            Pdist[q] = row_dist(X, d, i, j);
            atomicAdd(&tcount[i], 1);
        }
        Ptgt[q] = tgt;
    }
}
__global__ __launch_bounds__(TPB) void k_count_open(const uint64_t* __restrict__ L, int n, int K, int* __restrict__ open) {
    int c = 0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) c += L[(i64)i * K + K - 1] == LIST_EMPTY;
    for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0 && c) atomicAdd(open, c);
}
__global__ __launch_bounds__(TPB) void k_unpack(const uint64_t* __restrict__ L, i64 nK, int32_t* __restrict__ idx,
                                                float* __restrict__ dist) {
    for (i64 e = (i64)blockIdx.x * TPB + threadIdx.x; e < nK; e += (i64)gridDim.x * TPB) {
        const uint64_t k = L[e];
        idx[e] = k == LIST_EMPTY ? -1 : key_node(k);
        dist[e] = __uint_as_float((uint32_t)(k >> 32));
    }
}

// PaCMAP-style mid-near partners (src/nndescent.cpp:730-759): for pair p of row i, the second nearest, by (squared
// distance, index), of n_cand rows drawn by hash(seed, i, p, c) (a draw equal to i moves to i + 1)
__global__ __launch_bounds__(TPB) void k_mid_near(const float* __restrict__ X, int d, int n, int n_pairs, int n_cand,
                                                  uint64_t seed, int32_t* __restrict__ out) {
    const i64 total = (i64)n * n_pairs;
    for (i64 q = (i64)blockIdx.x * TPB + threadIdx.x; q < total; q += (i64)gridDim.x * TPB) {
        const int i = (int)(q / n_pairs), p = (int)(q % n_pairs);
        uint64_t best = ~0ULL, second = ~0ULL;
        for (int c = 0; c < n_cand; ++c) {
            int j = (int)(hash4(seed, (uint64_t)i, (uint64_t)p, (uint64_t)c) % (uint64_t)n);
            if (j == i) j = (j + 1) % n;
            const uint64_t k = pack(row_dist(X, d, i, j), j);
            if (k < best) { second = best; best = k; }
            else if (k < second) second = k;
        }
        out[q] = key_node(second);
    }
}

// ------------------------------------------------------------------------------------------------ exact graph
// Exact k nearest other rows of every row, bounded by an approximate graph: tau[q] = the largest distance of q's
// approximate list, so q's true k nearest all lie within tau[q]. A scan over all (query, row) pairs, 64 x 64 per
// tile (4 x 4 per thread, features through LDS 32 at a time, each distance one fused multiply-add per feature in
// order, as NN-descent forms it), appends every row within tau[q] to q's buffer; the buffer is then sorted by
// (distance, index) and its first k kept.
constexpr int XQ = 64, XT = 64, XF = 32, XSPAN = 4096;
__global__ __launch_bounds__(TPB) void k_knn_threshold(const float* __restrict__ X, int n, int d, const float* __restrict__ tau,
                                                       int cap, int* __restrict__ cnt, uint64_t* __restrict__ buf) {
    __shared__ float sQ[XQ][XF + 1], sT[XT][XF + 1];
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    const int qb = blockIdx.x * XQ, tb = blockIdx.y * XSPAN, te = min(n, tb + XSPAN);
    int qi[4];
    float tq[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        qi[i] = qb + ty * 4 + i;
        tq[i] = qi[i] < n ? tau[qi[i]] : -1.0f;
    }
    for (int t0 = tb; t0 < te; t0 += XT) {
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
        for (int f0 = 0; f0 < d; f0 += XF) {
            const int fc = min(XF, d - f0);
            __syncthreads();
            for (int e = threadIdx.x; e < XQ * XF; e += TPB) {
                const int r = e / XF, f = e % XF;
                sQ[r][f] = (qb + r < n && f < fc) ? X[(i64)(qb + r) * d + f0 + f] : 0.f;
                sT[r][f] = (t0 + r < te && f < fc) ? X[(i64)(t0 + r) * d + f0 + f] : 0.f;
            }
            __syncthreads();
            for (int f = 0; f < fc; ++f) {
                float a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) { a[i] = sQ[ty * 4 + i][f]; b[i] = sT[tx * 4 + i][f]; }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float t = a[i] - b[j];
                        acc[i][j] = fmaf(t, t, acc[i][j]);
                    }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int t = t0 + tx * 4 + j;
                if (qi[i] < n && t < te && t != qi[i] && acc[i][j] <= tq[i]) {
                    const int pos = atomicAdd(&cnt[qi[i]], 1);
                    if (pos < cap) buf[(i64)qi[i] * cap + pos] = pack(acc[i][j], t);
                }
            }
    }
}
__global__ __launch_bounds__(TPB) void k_tau(const float* __restrict__ dist, int n, int K, float* __restrict__ tau) {
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) tau[i] = dist[(i64)i * K + K - 1];
}
__global__ __launch_bounds__(TPB) void k_max_i32(const int* __restrict__ a, int n, int* __restrict__ mx) {
    int m = 0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) m = max(m, a[i]);
    for (int o = WAVE / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if (lane_id() == 0) atomicMax(mx, m);
}
// one row per block: its min(cnt, cap) candidates sorted by (distance, index) in LDS (bitonic, P2 slots), first K out
template <int P2>
__global__ __launch_bounds__(TPB) void k_select_sorted(const uint64_t* __restrict__ buf, const int* __restrict__ cnt, int cap,
                                                       int K, int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
    __shared__ uint64_t sk[P2];
    const int q = blockIdx.x;
    const int c = min(cnt[q], cap);
    for (int e = threadIdx.x; e < P2; e += TPB) sk[e] = e < c ? buf[(i64)q * cap + e] : ~0ULL;
    __syncthreads();
    for (int size = 2; size <= P2; size <<= 1)
        for (int stride = size / 2; stride > 0; stride >>= 1) {
            for (int e = threadIdx.x; e < P2 / 2; e += TPB) {
                const int lo = 2 * e - (e & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = sk[lo], b = sk[hi];
                if ((a > b) == up) { sk[lo] = b; sk[hi] = a; }
            }
            __syncthreads();
        }
    for (int j = threadIdx.x; j < K; j += TPB) {
        const uint64_t k = sk[j];
        out_idx[(i64)q * K + j] = k == ~0ULL ? -1 : key_node(k);
        out_dist[(i64)q * K + j] = __uint_as_float((uint32_t)(k >> 32));
    }
}

unsigned grid_for(i64 n) {
    const i64 b = (n + TPB - 1) / TPB;
    return (unsigned)(b < 1 ? 1 : b > 65536 ? 65536 : b);
}
int bits_for(i64 v) {   // bits to hold keys 0 .. v
    int b = 1;
    while ((1ll << b) <= v) ++b;
    return b;
}

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
    template <class U> U* get(size_t count) {
        const size_t b = std::max<size_t>(256, count * sizeof(U));
        if (b > bytes) {
            if (p) { MG_HIP(hipFree(p)); p = nullptr; bytes = 0; }
            MG_HIP(hipMalloc(&p, b));
            bytes = b;
        }
        return reinterpret_cast<U*>(p);
    }
};

struct NND {
    const float* X;
    int n, d, K, m;
    hipStream_t st;
    uint64_t* L;
    uint8_t* F;
    Buf bL, bF, bPt, bPs, bPd, bKa, bVa, bKb, bVb, bHist, bCnt, bTp, bTmp, bMisc, bMem, bGn, bGs, bGb;
    int* misc;   // [0] updates, [1] open rows, [2] max segment

    int host_int(const int* dptr) {
        int v = 0;
        MG_HIP(hipMemcpyAsync(&v, dptr, sizeof(int), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        return v;
    }

    // sorts proposals [0, P) by target (targets == n sorted last) and merges them; returns the entries inserted
    int apply(i64 P, int* Ptgt, int* Psrc, float* Pdist, int* tcount) {
        if (P == 0) return 0;
        int* ka = bKa.get<int>(P);
        int* va = bVa.get<int>(P);
        int* kb = bKb.get<int>(P);
        int* vb = bVb.get<int>(P);
        int* hist = bHist.get<int>((size_t)sort_hist_ints(P));
        MG_HIP(hipMemcpyAsync(ka, Ptgt, (size_t)P * sizeof(int), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_iota, dim3(grid_for(P)), dim3(TPB), 0, st, va, P);
        const int cur = launch_key_sort<float>(ka, va, kb, vb, hist, P, bits_for(n), st);
        const int* sv = cur == 0 ? va : vb;
        int* tptr = bTp.get<int>((size_t)n + 1);
        launch_exscan(tcount, n, tptr, bTmp.get<int>(scan_tmp_ints(n)), st);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(TPB), 0, st, misc, (i64)1, 0);
        hipLaunchKernelGGL(k_merge, dim3((unsigned)n), dim3(WAVE), 0, st, sv, tptr, Psrc, Pdist, n, K, L, F, misc);
        MG_HIP(hipGetLastError());
        return host_int(misc);
    }

    // joins groups [0, ng): counts, then ranges of groups whose proposals fit `cap`, each written, sorted and merged
    // the join kernel for groups of at most smax members (its pairs per thread)
    template <bool WRITE>
    void launch_join(int smax, int g0, int ng, const int* gbeg, int gstride, const int* gsize, const int* gnew,
                     const int* members, int* count, const int* offs, int* Pt, int* Ps, float* Pd, int* tcount) {
        const int s = std::min(smax, SMAX), pmax = s * (s - 1) / 2;
#define MG_JOIN(P_) hipLaunchKernelGGL((k_join<WRITE, P_>), dim3((unsigned)ng), dim3(TPB), 0, st, X, d, gbeg, gstride, gsize, gnew, \
                                       members, g0, ng, L, K, count, offs, Pt, Ps, Pd, tcount)
        if (pmax <= 4 * TPB) MG_JOIN(4);
        else if (pmax <= 8 * TPB) MG_JOIN(8);
        else if (pmax <= 20 * TPB) MG_JOIN(20);
        else MG_JOIN(PPT_MAX);
#undef MG_JOIN
        MG_HIP(hipGetLastError());
    }

    i64 join(int ng, int smax, const int* gbeg, int gstride, const int* gsize, const int* gnew, const int* members) {
        if (ng == 0) return 0;
        int* cnt = bCnt.get<int>((size_t)ng + 1);
        launch_join<false>(smax, 0, ng, gbeg, gstride, gsize, gnew, members, cnt, nullptr, nullptr, nullptr, nullptr, nullptr);
        std::vector<int> hc(ng);
        MG_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)ng * sizeof(int), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        const i64 cap = std::max<i64>((i64)64 << 20, (i64)16 * n);
        i64 updates = 0;
        int g0 = 0;
        while (g0 < ng) {
            i64 tot = 0;
            int g1 = g0;
            while (g1 < ng && (g1 == g0 || tot + hc[g1] <= cap)) tot += hc[g1++];
            // counts of this range against the lists as they are now: those of the first pass for the first range,
            // taken again after earlier ranges have merged (no more than before)
            int* c2 = bCnt.get<int>((size_t)ng + 1);
            if (g0 > 0) launch_join<false>(smax, g0, g1 - g0, gbeg, gstride, gsize, gnew, members, c2, nullptr, nullptr, nullptr, nullptr, nullptr);
            int* offs = bGb.get<int>((size_t)(g1 - g0) + 1);
            launch_exscan(c2, g1 - g0, offs, bTmp.get<int>(scan_tmp_ints(std::max(g1 - g0, n))), st);
            const int P = host_int(offs + (g1 - g0));
            int* Pt = bPt.get<int>(P);
            int* Ps = bPs.get<int>(P);
            float* Pd = bPd.get<float>(P);
            int* tcount = bMisc.get<int>((size_t)n + 8);
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, st, tcount, (i64)n, 0);
            launch_join<true>(smax, g0, g1 - g0, gbeg, gstride, gsize, gnew, members, nullptr, offs, Pt, Ps, Pd, tcount);
            updates += apply(P, Pt, Ps, Pd, tcount);
            // the range's counts were re-taken: hc of later groups stays an upper bound
            for (int g = g0; g < g1; ++g) hc[g] = 0;
            g0 = g1;
            (void)tot;
        }
        return updates;
    }
};

}  // namespace

void mid_near_pairs(const float* X, i64 n, int d, int n_pairs, int n_cand, uint64_t seed, int32_t* out, hipStream_t st) {
    require(n_pairs >= 1 && n_cand >= 2 && n >= n_cand + 1, "need n_pairs >= 1, n_cand >= 2 and n > n_cand");
    require(n < (1ll << 31) && d >= 1, "at most 2^31 - 1 rows of at least one column");
    hipLaunchKernelGGL(k_mid_near, dim3(grid_for(n * n_pairs)), dim3(TPB), 0, st, X, d, (int)n, n_pairs, n_cand, seed, out);
    MG_HIP(hipGetLastError());
}

void knn_self_exact(const float* X, i64 n_, int d, int K, int32_t* out_idx, float* out_dist, hipStream_t st) {
    require(n_ >= 2 && n_ < (1ll << 30) && d >= 1, "the exact graph takes 2 to 2^30 - 1 rows of at least one column");
    require(K >= 1 && K < n_ && K <= KMAX, "k must be between 1 and min(n - 1, 128)");
    require(n_ <= (i64)65535 * XSPAN, "the exact graph takes at most 65535 * 4096 rows");
    const int n = (int)n_;
    const int n_trees = (int)std::max(2.0, std::min(8.0, std::round(std::sqrt((double)n) / 40.0)));
    nn_descent(X, n, d, K, 50, 10, 0.001f, n_trees, 30, 0, out_idx, out_dist, 0, st);   // the bound
    Buf btau, bcnt, bbuf, bmx;
    float* tau = btau.get<float>(n);
    int* cnt = bcnt.get<int>(n);
    int* mx = bmx.get<int>(1);
    hipLaunchKernelGGL(k_tau, dim3(grid_for(n)), dim3(TPB), 0, st, out_dist, n, K, tau);
    int cap = std::max(2 * K, 64);
    for (int attempt = 0; attempt < 2; ++attempt) {
        uint64_t* buf = bbuf.get<uint64_t>((size_t)n * cap);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, st, cnt, (i64)n, 0);
        hipLaunchKernelGGL(k_knn_threshold, dim3((unsigned)((n + XQ - 1) / XQ), (unsigned)((n + XSPAN - 1) / XSPAN)), dim3(TPB), 0, st,
                           X, n, d, tau, cap, cnt, buf);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(TPB), 0, st, mx, (i64)1, 0);
        hipLaunchKernelGGL(k_max_i32, dim3(grid_for(n)), dim3(TPB), 0, st, cnt, n, mx);
        MG_HIP(hipGetLastError());
        int h_mx = 0;
        MG_HIP(hipMemcpyAsync(&h_mx, mx, sizeof(int), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        require(h_mx >= K, "the exact search found fewer than k rows within the bound");
        if (h_mx <= cap) {
#define MG_SEL(P) hipLaunchKernelGGL((k_select_sorted<P>), dim3((unsigned)n), dim3(TPB), 0, st, buf, cnt, cap, K, out_idx, out_dist)
            if (cap <= 256) MG_SEL(256);
            else if (cap <= 1024) MG_SEL(1024);
            else MG_SEL(4096);
#undef MG_SEL
            MG_HIP(hipGetLastError());
            MG_HIP(hipStreamSynchronize(st));
            return;
        }
        require(h_mx <= 4096, "more than 4096 rows tie within a row's k-th distance");
        cap = h_mx;   // a second scan with room for every row's candidates (ties of duplicate rows)
    }
    require(false, "the exact search did not settle");
}

void nn_descent(const float* X, i64 n_, int d, int K, int max_candidates, int n_iters, float delta, int n_trees,
                int leaf_size, uint64_t seed, int32_t* out_idx, float* out_dist, int verbose, hipStream_t st) {
    require(n_ >= 2 && n_ < (1ll << 30), "NN-descent takes 2 to 2^30 - 1 points");
    require(d >= 1, "X must have at least one column");
    require(K >= 1 && K < n_, "k must be between 1 and n - 1");
    require(K <= KMAX, "k is at most 128 on this path");
    require(max_candidates >= 1 && n_iters >= 0 && leaf_size >= 2 && n_trees >= 0, "invalid NN-descent parameters");
    const int n = (int)n_;
    auto clock = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    double t_mark = clock();
    auto lap = [&]() { MG_HIP(hipStreamSynchronize(st)); const double t = clock(), dt = t - t_mark; t_mark = t; return dt; };
    NND s;
    s.X = X; s.n = n; s.d = d; s.K = K; s.m = std::min(std::min(max_candidates, K), SMAX / 2); s.st = st;
    const i64 nK = (i64)n * K;
    s.L = s.bL.get<uint64_t>(nK);
    s.F = s.bF.get<uint8_t>(nK);
    s.misc = s.bMisc.get<int>((size_t)n + 8);   // reused as tcount by join; misc counters live past n
    s.misc = s.misc + n;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nK)), dim3(TPB), 0, st, s.L, nK, LIST_EMPTY);
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(nK)), dim3(TPB), 0, st, s.F, nK, (uint8_t)0);

    // ---- trees (src/nndescent.cpp:238-330, 414-461), several at once: tree t of a batch owns positions [t n, (t + 1) n)
    // of one order, its segments never cross into another's, so one sort per level splits every tree's segments
    if (n_trees > 0) {
        const int max_depth = (int)std::ceil(std::log2(std::max(2.0, (double)n / leaf_size))) + 8;
        const int TB = (int)std::max<i64>(1, std::min<i64>(n_trees, 0x3fffffffLL / n));   // trees per batch (keys < 2^31)
        const i64 NT = (i64)TB * n;
        Buf bperm, bsegof, bsegst, bflag, bex, bk2, bv2, bhist, btmp;
        int* perm = bperm.get<int>(NT);
        int* seg_of = bsegof.get<int>(NT);
        int* seg_start = bsegst.get<int>((size_t)NT + 1);
        int* flag = bflag.get<int>(NT);
        int* ex = bex.get<int>((size_t)NT + 1);
        int* keys = bk2.get<int>(2 * (size_t)NT);
        int* vals2 = bv2.get<int>(NT);
        int* hist = bhist.get<int>((size_t)sort_hist_ints(NT));
        int* tmp = btmp.get<int>(scan_tmp_ints(NT));
        int* mx = s.misc + 2;
        for (int t0 = 0; t0 < n_trees; t0 += TB) {
            const int tb = std::min(TB, n_trees - t0);
            const int N = tb * n;
            hipLaunchKernelGGL(k_rp_init, dim3(grid_for(N)), dim3(TPB), 0, st, perm, seg_of, seg_start, n, tb);
            int nseg = tb, maxseg = n;
            for (int depth = 0; depth < max_depth && maxseg > leaf_size; ++depth) {
                hipLaunchKernelGGL(k_rp_keys, dim3(grid_for(N)), dim3(TPB), 0, st, X, d, perm, seg_of, seg_start, N, n, leaf_size,
                                   seed, t0, depth, keys);
                int* kb = keys + N;
                const int cur = launch_key_sort<float>(keys, perm, kb, vals2, hist, N, bits_for(2 * (i64)nseg), st);
                int* sk = cur == 0 ? keys : kb;
                if (cur != 0) MG_HIP(hipMemcpyAsync(perm, vals2, (size_t)N * sizeof(int), hipMemcpyDeviceToDevice, st));
                hipLaunchKernelGGL(k_rp_marks, dim3(grid_for(N)), dim3(TPB), 0, st, sk, N, flag);
                launch_exscan(flag, N, ex, tmp, st);
                hipLaunchKernelGGL(k_rp_segments, dim3(grid_for(N)), dim3(TPB), 0, st, flag, ex, N, seg_of, seg_start);
                hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(TPB), 0, st, mx, (i64)1, 0);
                hipLaunchKernelGGL(k_rp_maxseg, dim3(grid_for(N)), dim3(TPB), 0, st, seg_start, ex + N, mx);
                MG_HIP(hipGetLastError());
                nseg = s.host_int(ex + N);
                maxseg = s.host_int(mx);
            }
            // leaf cliques, tree after tree (oversized leaves: cliques of SMAX consecutive members)
            int* gbeg = s.bGs.get<int>((size_t)3 * nseg);
            int* gsize = gbeg + nseg;
            int* gnew = gsize + nseg;
            hipLaunchKernelGGL(k_leaf_groups, dim3(grid_for(nseg)), dim3(TPB), 0, st, seg_start, nseg, gbeg, gsize, gnew);
            const double t_split = verbose ? lap() : 0.0;
            const i64 u = s.join(nseg, maxseg, gbeg, 0, gsize, gnew, perm);
            if (verbose) fprintf(stderr, "nn-descent trees %d-%d: %d leaves (%.3f s), %lld insertions (%.3f s)\n", t0, t0 + tb - 1, nseg,
                                 t_split, (long long)u, lap());
        }
    }
    // ---- random neighbours for open slots, then every node for rows still open (small n)
    for (int round = 0; round < 8; ++round) {
        int* open = s.misc + 1;
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(TPB), 0, st, open, (i64)1, 0);
        hipLaunchKernelGGL(k_count_open, dim3(grid_for(n)), dim3(TPB), 0, st, s.L, n, K, open);
        if (s.host_int(open) == 0) break;
        const i64 P = (i64)n * 2 * K;
        int* Pt = s.bPt.get<int>(P);
        int* Ps = s.bPs.get<int>(P);
        float* Pd = s.bPd.get<float>(P);
        int* tcount = s.bCnt.get<int>((size_t)n + 1);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, st, tcount, (i64)n, 0);
        hipLaunchKernelGGL(k_fill_random, dim3(grid_for(P)), dim3(TPB), 0, st, X, d, s.L, n, K, seed, round, Pt, Ps, Pd, tcount);
        s.apply(P, Pt, Ps, Pd, tcount);
    }
    {
        int* open = s.misc + 1;
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(TPB), 0, st, open, (i64)1, 0);
        hipLaunchKernelGGL(k_count_open, dim3(grid_for(n)), dim3(TPB), 0, st, s.L, n, K, open);
        if (s.host_int(open) > 0) {
            require((i64)n * n <= (i64)1 << 28, "NN-descent could not fill every neighbour list");
            const i64 P = (i64)n * n;
            int* Pt = s.bPt.get<int>(P);
            int* Ps = s.bPs.get<int>(P);
            float* Pd = s.bPd.get<float>(P);
            int* tcount = s.bCnt.get<int>((size_t)n + 1);
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, st, tcount, (i64)n, 0);
            hipLaunchKernelGGL(k_fill_all, dim3(grid_for(P)), dim3(TPB), 0, st, X, d, s.L, n, K, Pt, Ps, Pd, tcount);
            s.apply(P, Pt, Ps, Pd, tcount);
        }
    }
    if (verbose) fprintf(stderr, "nn-descent random fill: %.3f s\n", lap());
    // ---- rounds (src/nndescent.cpp:485-560)
    {
        const int m = s.m;
        Buf bk, bv, bk2, bv2, bhist, brc, brp, btmp;
        int* rk = bk.get<int>(nK);
        int* rv = bv.get<int>(nK);
        int* rk2 = bk2.get<int>(nK);
        int* rv2 = bv2.get<int>(nK);
        int* hist = bhist.get<int>((size_t)sort_hist_ints(nK));
        int* rcount = brc.get<int>((size_t)n + 1);
        int* rptr = brp.get<int>((size_t)n + 1);
        int* tmp = btmp.get<int>(scan_tmp_ints(n));
        int* members = s.bMem.get<int>((size_t)n * 2 * m);
        int* gnew = s.bGn.get<int>(n);
        int* gsize = s.bGs.get<int>(n);
        for (int it = 0; it < n_iters; ++it) {
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, st, rcount, (i64)n, 0);
            hipLaunchKernelGGL(k_rev_keys, dim3(grid_for(nK)), dim3(TPB), 0, st, s.L, nK, n, rk, rv, rcount);
            const int cur = launch_key_sort<float>(rk, rv, rk2, rv2, hist, nK, bits_for(n), st);
            const int* rsv = cur == 0 ? rv : rv2;
            launch_exscan(rcount, n, rptr, tmp, st);
            hipLaunchKernelGGL(k_candidates, dim3((unsigned)n), dim3(WAVE), 0, st, s.L, s.F, K, rptr, rsv, n, m, seed, it, members,
                               gnew, gsize);
            hipLaunchKernelGGL(k_unflag, dim3((unsigned)n), dim3(WAVE), 0, st, s.L, s.F, K, members, gnew, n, m);
            MG_HIP(hipGetLastError());
            const double t_cand = verbose ? lap() : 0.0;
            const i64 u = s.join(n, 2 * m, nullptr, 2 * m, gsize, gnew, members);
            if (verbose) fprintf(stderr, "nn-descent iteration %d: candidates %.3f s, join %.3f s, %lld updates\n", it + 1, t_cand, lap(),
                                 (long long)u);
            if (u <= (i64)(delta * (double)n * K)) break;
        }
    }
    hipLaunchKernelGGL(k_unpack, dim3(grid_for(nK)), dim3(TPB), 0, st, s.L, nK, out_idx, out_dist);
    MG_HIP(hipGetLastError());
    MG_HIP(hipStreamSynchronize(st));
}

}  // namespace mg
