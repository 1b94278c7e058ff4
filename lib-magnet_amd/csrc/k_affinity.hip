// Affinities upstream of the t-SNE pass, on the device:
//
// - perplexity calibration: for each row of k squared neighbour distances, the Gaussian precision beta whose
//   conditional distribution p_j|i = exp(-beta d_ij) / sum has entropy log(perplexity), by bisection (the textbook
//   rule of t-SNE implementations; the reference computes these affinities only through openTSNE in its benchmark
//   scripts, benchmarks/opentsne_audit.py:18-23);
// - symmetrisation of a directed k-NN graph into a CSR matrix: every unordered pair {i, j} of the list gets
//   a = (sum of its i -> j values) and b = (sum of its j -> i values), combined either as (a + b) * scale (the
//   perplexity affinities (P + P^T) / 2n) or as the fuzzy union mixed with the intersection,
//   r (a + b - a b) + (1 - r) a b (magnet/knn.py:215-230). Entries on the diagonal and entries equal to zero are
//   dropped; rows are sorted by column.
//
// The symmetrisation sorts the 2nk directed entries by (row, column) with the stable radix sort of the repulsion
// pass (column first, then row), so equal pairs are adjacent and keep the order forward-then-reverse; a mark pass
// keeps the first entry of each pair, a block scan turns the marks into output positions, and the row pointers come
// from per-row counts. Nothing depends on the launch: the output is the same bit for bit on every run.
#include "common.hpp"
#include <algorithm>

// a + b - ab and its mix rounded operation by operation, as numpy evaluates them
#pragma clang fp contract(off)

namespace mg {

namespace {

constexpr int TPB = 256;
constexpr int ITEMS = 8;                 // items per thread of the mark / count / write passes
constexpr int TILE = TPB * ITEMS;

template <int G> __device__ inline double seg_sum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}
template <int G> __device__ inline double seg_min(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, G));
    return v;
}

// G lanes per row, KPL entries per lane (entry c of a row sits in lane c % G, slot c / G). The lanes of a row form
// their sums by an xor butterfly, which leaves the same value in every lane, so the row's bisection steps are
// uniform across its lanes.
template <int G, int KPL>
__global__ __launch_bounds__(TPB) void k_perplexity(const float* __restrict__ d2, i64 n, int k, double target, double tol,
                                                    int n_iter, double* __restrict__ P, double* __restrict__ beta_out) {
    const i64 row = ((i64)blockIdx.x * TPB + threadIdx.x) / G;
    if (row >= n) return;   // whole segments leave together (G divides the block)
    const int sub = threadIdx.x % G;
    const float* dr = d2 + row * k;
    double d[KPL];
    double mn = INFINITY;
#pragma unroll
    for (int u = 0; u < KPL; ++u) {
        const int c = sub + u * G;
        d[u] = c < k ? (double)dr[c] : 0.0;
        if (c < k) mn = fmin(mn, d[u]);
    }
    mn = seg_min<G>(mn);
#pragma unroll
    for (int u = 0; u < KPL; ++u) d[u] -= mn;   // shift by the nearest distance: cancels in the normalisation
    double beta = 1.0, lo = -INFINITY, hi = INFINITY;
    for (int it = 0; it < n_iter; ++it) {
        double s = 0, sd = 0;
#pragma unroll
        for (int u = 0; u < KPL; ++u) {
            if (sub + u * G < k) {
                const double w = exp(-d[u] * beta);
                s += w;
                sd += d[u] * w;
            }
        }
        s = seg_sum<G>(s);
        sd = seg_sum<G>(sd);
        const double diff = log(s) + beta * sd / s - target;
        if (fabs(diff) < tol) break;
        if (diff > 0) {   // entropy too high: a narrower kernel
            lo = beta;
            beta = isinf(hi) ? beta * 2.0 : (beta + hi) / 2.0;
        } else {
            hi = beta;
            beta = isinf(lo) ? beta / 2.0 : (beta + lo) / 2.0;
        }
    }
    double w[KPL], s = 0;
#pragma unroll
    for (int u = 0; u < KPL; ++u) {
        w[u] = sub + u * G < k ? exp(-d[u] * beta) : 0.0;
        s += w[u];
    }
    s = seg_sum<G>(s);
#pragma unroll
    for (int u = 0; u < KPL; ++u) {
        const int c = sub + u * G;
        if (c < k) P[row * k + c] = w[u] / s;
    }
    if (beta_out && sub == 0) beta_out[row] = beta;
}

// the pair of entry e of the doubled list: e < E is (e / k, idx[e]); e >= E the reverse of entry e - E
__device__ inline void sym_pair(const int32_t* __restrict__ idx, int E, int k, int e, int& r, int& c) {
    if (e < E) { r = e / k; c = idx[e]; }
    else { const int f = e - E; r = idx[f]; c = f / k; }
}

__global__ __launch_bounds__(TPB) void k_sym_check(const int32_t* __restrict__ idx, int E, int n, int* __restrict__ bad) {
    for (int e = blockIdx.x * TPB + threadIdx.x; e < E; e += gridDim.x * TPB)
        if (idx[e] < 0 || idx[e] >= n) atomicOr(bad, 1);
}

// first sort's input: key = column of entry e, value e
__global__ __launch_bounds__(TPB) void k_sym_keys_col(const int32_t* __restrict__ idx, int E, int k, int* __restrict__ keys,
                                                      int* __restrict__ vals) {
    for (int e = blockIdx.x * TPB + threadIdx.x; e < 2 * E; e += gridDim.x * TPB) {
        int r, c;
        sym_pair(idx, E, k, e, r, c);
        keys[e] = c;
        vals[e] = e;
    }
}

// second sort's input: key = row of the entry at each position of the column order
__global__ __launch_bounds__(TPB) void k_sym_keys_row(const int32_t* __restrict__ idx, int E, int k, const int* __restrict__ vin,
                                                      int* __restrict__ keys, int* __restrict__ vals) {
    for (int p = blockIdx.x * TPB + threadIdx.x; p < 2 * E; p += gridDim.x * TPB) {
        const int e = vin[p];
        int r, c;
        sym_pair(idx, E, k, e, r, c);
        keys[p] = r;
        vals[p] = e;
    }
}

// value of the pair whose first entry (in (row, column) order) sits at position p
template <int MODE>
__device__ inline double sym_value(const int32_t* __restrict__ idx, const double* __restrict__ val, const int* __restrict__ S,
                                   int E, int k, int p, int r, int c, double param) {
    double a = 0, b = 0;
    for (int q = p; q < 2 * E; ++q) {
        const int e = S[q];
        int rq, cq;
        sym_pair(idx, E, k, e, rq, cq);
        if (rq != r || cq != c) break;
        if (e < E) a += val[e]; else b += val[e - E];
    }
    if (MODE == 0) return (a + b) * param;
    const double both = a * b;
    return param * (a + b - both) + (1.0 - param) * both;
}

// keep[p] = 1 for the first entry of a pair off the diagonal whose value is not zero; per-block counts; per-row counts
template <int MODE>
__global__ __launch_bounds__(TPB) void k_sym_mark(const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                  const int* __restrict__ S, int E, int k, double param, int* __restrict__ keep,
                                                  int* __restrict__ block_count, int* __restrict__ row_count) {
    __shared__ int sh[TPB / WAVE];
    int cnt = 0;
    const int base = blockIdx.x * TILE;
    for (int it = 0; it < ITEMS; ++it) {
        const int p = base + it * TPB + threadIdx.x;
        if (p >= 2 * E) break;
        int r, c;
        sym_pair(idx, E, k, S[p], r, c);
        bool head = true;
        if (p > 0) {
            int r0, c0;
            sym_pair(idx, E, k, S[p - 1], r0, c0);
            head = r0 != r || c0 != c;
        }
        int kp = 0;
        if (head && r != c) {
            const double v = sym_value<MODE>(idx, val, S, E, k, p, r, c, param);
            kp = MODE == 0 ? v != 0.0 : (float)v != 0.0f;   // the sum: zero in double; the fuzzy graph: zero in float
        }
        keep[p] = kp;
        if (kp) atomicAdd(&row_count[r], 1);
        cnt += kp;
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (lane_id() == 0) sh[threadIdx.x / WAVE] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < TPB / WAVE; ++w) t += sh[w];
        block_count[blockIdx.x] = t;
    }
}

// exclusive scan of a[0 .. m) in place by one block, total to a[m]
__global__ __launch_bounds__(1024) void k_scan_one_block(int* __restrict__ a, int m) {
    __shared__ int sh[1024];
    int carry = 0;
    for (int c0 = 0; c0 < m; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int v = i < m ? a[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int u = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
            __syncthreads();
            sh[threadIdx.x] += u;
            __syncthreads();
        }
        if (i < m) a[i] = carry + sh[threadIdx.x] - v;
        carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) a[m] = carry;
}

// block-level pass of the scan over n values: per-block sums
__global__ __launch_bounds__(TPB) void k_block_sums(const int* __restrict__ a, i64 n, int* __restrict__ sums) {
    __shared__ int sh[TPB / WAVE];
    int s = 0;
    const i64 base = (i64)blockIdx.x * TILE;
    for (int it = 0; it < ITEMS; ++it) {
        const i64 i = base + it * TPB + threadIdx.x;
        if (i < n) s += a[i];
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane_id() == 0) sh[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < TPB / WAVE; ++w) t += sh[w];
        sums[blockIdx.x] = t;
    }
}

// exclusive prefix within a block's tile, in item order (item it of thread t is position base + it * TPB + t)
__device__ inline int tile_exscan(int v, int* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < TPB; d <<= 1) {
        const int u = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int r = sh[t] - v;
    __syncthreads();
    return r;
}

// out[i] = offs[block] + exclusive prefix of a within the block (row pointers from row counts); out[n] = total
__global__ __launch_bounds__(TPB) void k_scan_apply(const int* __restrict__ a, i64 n, const int* __restrict__ offs,
                                                    int nblocks, int* __restrict__ out) {
    __shared__ int sh[TPB];
    __shared__ int tot;
    const i64 base = (i64)blockIdx.x * TILE;
    int carry = offs[blockIdx.x];
    for (int it = 0; it < ITEMS; ++it) {
        const i64 i = base + it * TPB + threadIdx.x;
        const int v = i < n ? a[i] : 0;
        const int ex = tile_exscan(v, sh);
        if (i < n) out[i] = carry + ex;
        if (threadIdx.x == TPB - 1) tot = ex + v;
        __syncthreads();
        carry += tot;
        __syncthreads();
    }
    if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) out[n] = offs[nblocks];
}

// writes the kept pairs: position from the block offset and the prefix of the marks within the block
template <int MODE>
__global__ __launch_bounds__(TPB) void k_sym_write(const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                   const int* __restrict__ S, int E, int k, double param,
                                                   const int* __restrict__ keep, const int* __restrict__ offs,
                                                   int32_t* __restrict__ out_idx, float* __restrict__ out_val) {
    __shared__ int sh[TPB];
    __shared__ int tot;
    const int base = blockIdx.x * TILE;
    int carry = offs[blockIdx.x];
    for (int it = 0; it < ITEMS; ++it) {
        const int p = base + it * TPB + threadIdx.x;
        const int kp = p < 2 * E ? keep[p] : 0;
        const int ex = tile_exscan(kp, sh);
        if (kp) {
            int r, c;
            sym_pair(idx, E, k, S[p], r, c);
            const int q = carry + ex;
            out_idx[q] = c;
            out_val[q] = (float)sym_value<MODE>(idx, val, S, E, k, p, r, c, param);
        }
        if (threadIdx.x == TPB - 1) tot = ex + kp;
        __syncthreads();
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(TPB) void k_fill_int(int* __restrict__ a, i64 n, int v) {
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) a[i] = v;
}

unsigned grid_for(i64 n) {
    const i64 b = (n + TPB - 1) / TPB;
    return (unsigned)(b < 1 ? 1 : b > 65536 ? 65536 : b);
}

// exclusive scan of a[0 .. n) into out[0 .. n], out[n] the total; tmp holds scan_tmp_ints(n) ints
void exscan(const int* a, i64 n, int* out, int* tmp, hipStream_t st) {
    const int nb = (int)std::max<i64>(1, (n + TILE - 1) / TILE);
    hipLaunchKernelGGL(k_block_sums, dim3(nb), dim3(TPB), 0, st, a, n, tmp);
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, st, tmp, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(TPB), 0, st, a, n, tmp, nb, out);
    MG_HIP(hipGetLastError());
}

}  // namespace

i64 scan_tmp_ints(i64 n) { return std::max<i64>(1, (n + TILE - 1) / TILE) + 1; }

void launch_exscan(const int* a, i64 n, int* out, int* tmp, hipStream_t st) { exscan(a, n, out, tmp, st); }

void launch_perplexity(const float* d2, i64 n, int k, double perplexity, double tol, int n_iter, double* P, double* beta,
                       hipStream_t st) {
    require(k >= 1 && k <= 512, "the perplexity calibration takes 1 to 512 neighbours per row");
    require(perplexity > 0 && tol > 0 && n_iter >= 0, "perplexity and tol must be positive, n_iter >= 0");
    if (n == 0) return;
    const double target = std::log(perplexity);
#define MG_PERP(G, KPL) do { const i64 rows_per_block = TPB / (G); \
        hipLaunchKernelGGL((k_perplexity<G, KPL>), dim3((unsigned)((n + rows_per_block - 1) / rows_per_block)), dim3(TPB), 0, st, \
                           d2, n, k, target, tol, n_iter, P, beta); } while (0)
    if (k <= 16) MG_PERP(8, 2);
    else if (k <= 64) MG_PERP(16, 4);
    else if (k <= 128) MG_PERP(16, 8);
    else if (k <= 256) MG_PERP(32, 8);
    else MG_PERP(64, 8);
#undef MG_PERP
    MG_HIP(hipGetLastError());
}

i64 symmetrize_scratch_ints(i64 n, int k) {
    const i64 E2 = 2 * n * (i64)k;
    return 5 * E2 + (sort_hist_ints(E2)) + scan_tmp_ints(E2) + scan_tmp_ints(n) + n + 8;
}

i64 symmetrize_knn(int mode, const int32_t* idx, const double* val, i64 n, int k, double param, int32_t* indptr,
                   int32_t* out_idx, float* out_val, i64 cap, int* scratch, hipStream_t st) {
    require(mode == 0 || mode == 1, "mode must be 0 (sum) or 1 (fuzzy union)");
    require(n >= 1 && k >= 1, "n and k must be positive");
    require(2 * n * (i64)k <= 0x7fffffffLL, "at most 2^31 - 1 directed entries (2 n k)");
    const int E = (int)(n * k), E2 = 2 * E;
    int bits = 1;
    while ((1ll << bits) < n) ++bits;
    int* keysA = scratch;
    int* valsA = keysA + E2;
    int* keysB = valsA + E2;
    int* valsB = keysB + E2;
    int* keep = valsB + E2;
    int* hist = keep + E2;
    int* offs = hist + (sort_hist_ints(E2));
    int* rtmp = offs + scan_tmp_ints(E2);
    int* rcount = rtmp + scan_tmp_ints(n);
    int* bad = rcount + n;

    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(TPB), 0, st, bad, (i64)1, 0);
    hipLaunchKernelGGL(k_sym_check, dim3(grid_for(E)), dim3(TPB), 0, st, idx, E, (int)n, bad);
    int h_bad = 0;
    MG_HIP(hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    require(h_bad == 0, "neighbour indices must lie in [0, n)");

    // (row, column) order: a stable sort by column, then a stable sort by row
    hipLaunchKernelGGL(k_sym_keys_col, dim3(grid_for(E2)), dim3(TPB), 0, st, idx, E, k, keysA, valsA);
    MG_HIP(hipGetLastError());
    int cur = launch_key_sort<float>(keysA, valsA, keysB, valsB, hist, E2, bits, st);
    const int* v1 = cur == 0 ? valsA : valsB;
    int* k2 = cur == 0 ? keysB : keysA;
    int* v2 = cur == 0 ? valsB : valsA;
    hipLaunchKernelGGL(k_sym_keys_row, dim3(grid_for(E2)), dim3(TPB), 0, st, idx, E, k, v1, k2, v2);
    MG_HIP(hipGetLastError());
    // the second sort starts from the other buffer pair: run it with that pair as "A"
    int* kA2 = k2; int* vA2 = v2;
    int* kB2 = cur == 0 ? keysA : keysB;
    int* vB2 = cur == 0 ? valsA : valsB;
    cur = launch_key_sort<float>(kA2, vA2, kB2, vB2, hist, E2, bits, st);
    const int* S = cur == 0 ? vA2 : vB2;

    const int nb = (E2 + TILE - 1) / TILE;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(n)), dim3(TPB), 0, st, rcount, n, 0);
    if (mode == 0) hipLaunchKernelGGL(k_sym_mark<0>, dim3(nb), dim3(TPB), 0, st, idx, val, S, E, k, param, keep, offs, rcount);
    else hipLaunchKernelGGL(k_sym_mark<1>, dim3(nb), dim3(TPB), 0, st, idx, val, S, E, k, param, keep, offs, rcount);
    MG_HIP(hipGetLastError());
    // block offsets of the marks, row pointers from the row counts
    hipLaunchKernelGGL(k_scan_one_block, dim3(1), dim3(1024), 0, st, offs, nb);
    exscan(rcount, n, indptr, rtmp, st);
    int h_nnz = 0;
    MG_HIP(hipMemcpyAsync(&h_nnz, offs + nb, sizeof(int), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    if (h_nnz > cap) return -(i64)h_nnz;
    if (mode == 0) hipLaunchKernelGGL(k_sym_write<0>, dim3(nb), dim3(TPB), 0, st, idx, val, S, E, k, param, keep, offs, out_idx, out_val);
    else hipLaunchKernelGGL(k_sym_write<1>, dim3(nb), dim3(TPB), 0, st, idx, val, S, E, k, param, keep, offs, out_idx, out_val);
    MG_HIP(hipGetLastError());
    return h_nnz;
}

}  // namespace mg
