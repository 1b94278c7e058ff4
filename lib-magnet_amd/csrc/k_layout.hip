// Sampled layout optimisation on the device: the reference's optimize_layout (src/layout.cpp:118-276, 343-472) for the
// loss of MAGNET's sampled pass ("umap": attraction along graph edges sampled by weight, repulsion from random nodes)
// with the power kernel q = 1 / (1 + a d^2b), plain SGD steps clipped at 4 alpha, optional mid-range pairs, and fixed
// leading nodes (n_fixed, the out-of-sample placement). It is the reference's "owner" parallel mode
// (src/layout.cpp:118-128, 343-360): within an epoch every node is moved by one thread only, which reads every other
// node from the layout as it stood at the start of the epoch and writes its own position to the next layout; the
// attraction and the mid-range spring move the head of an edge only, the attraction with the reference's owner
// factor of 1.5 (or a caller's factor, 1 for edges that have no reverse), and the partners' share of the springs is delivered through the reverse pairs. No two threads write
// one position, so the result is deterministic and does not depend on the launch. The random streams are per node and
// epoch (the generator of src/rng.hpp, seeded from the node, the epoch and the caller's seed), not the reference's
// per-chunk streams: runs agree with the reference in law, not in value.
#include "common.hpp"
#include "../../include/magnet_hip.h"
#include <algorithm>
#include <cmath>

// Every product and sum rounded on its own, as the reference's build does: the repulsion of nearly coincident points
// is clipped in a direction that a last-bit difference can turn, so a layout run amplifies any change of rounding,
// and with these rules the float arithmetic here is the one the reference performs (divisions and square roots are
// correctly rounded on both), the power function aside.
#pragma clang fp contract(off)

namespace mg {

namespace {

constexpr int TPB = 256;

// xorshift64* seeded through splitmix64, as src/rng.hpp:6-30
struct Rng {
    uint64_t s;
    __device__ explicit Rng(uint64_t seed) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        s = (z ^ (z >> 31)) | 1ULL;
    }
    __device__ uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
    __device__ uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    __device__ float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
};

struct LayoutArgs {
    const float* Ys;          // layout at the start of the epoch (n, 2)
    float* Yn;                // layout at its end
    i64 n, n_fixed;
    const int32_t* eptr;      // (n + 1): edges of head u are [eptr[u], eptr[u + 1])
    const int32_t* tail;
    const float* eps;
    float* eons;
    float* eonns;
    const int32_t* pairs;     // (n, n_mid) or null
    int n_mid;
    const int32_t* rev_ptr;   // who chose each node as a partner (CSR), or null
    const int32_t* rev_idx;
    float a, b, rep_a, rep_b, two_ab, two_b, nsr, exag, mid_rate, att;
    int b_is_one, rep_b_is_one, clip_mode;
    uint64_t seed;
};

// one SGD step: alpha g, clipped at 4 alpha per coordinate (clip_mode 0) or in norm (1) (src/layout.cpp:80-116)
__device__ inline void step(float& x, float& y, float gx, float gy, float alpha, int clip_mode) {
    float sx = alpha * gx, sy = alpha * gy;
    const float cap = 4.0f * alpha;
    if (clip_mode == 1) {
        const float n2 = sx * sx + sy * sy;
        if (n2 > cap * cap) { const float f = cap / sqrtf(n2); sx *= f; sy *= f; }
    } else {
        sx = sx > cap ? cap : (sx < -cap ? -cap : sx);
        sy = sy > cap ? cap : (sy < -cap ? -cap : sy);
    }
    x += sx; y += sy;
}

// one epoch for node u (src/layout.cpp:129-260, owner mode, loss 0, power kernel)
__global__ __launch_bounds__(TPB) void k_layout_epoch(LayoutArgs A, int epoch, float alpha, float mid_w) {
    const i64 u = (i64)blockIdx.x * TPB + threadIdx.x;
    if (u >= A.n) return;
    float x = A.Ys[2 * u], y = A.Ys[2 * u + 1];
    if (u >= A.n_fixed) {
        Rng rng(A.seed * 0x9E3779B1ULL + (uint64_t)epoch * 0x85EBCA6BULL + (uint64_t)u * 0xC2B2AE35ULL);
        const float fepoch = (float)epoch;
        for (int e = A.eptr[u]; e < A.eptr[u + 1]; ++e) {
            if (A.eons[e] > fepoch) continue;
            const int32_t v = A.tail[e];
            float dx = x - A.Ys[2 * (i64)v], dy = y - A.Ys[2 * (i64)v + 1];
            const float d2 = dx * dx + dy * dy;
            float gx = 0.f, gy = 0.f;
            if (d2 > 0.f) {   // attraction, gradient of log q
                const float pw = A.b_is_one ? d2 : powf(d2, A.b);
                const float c = -A.att * A.exag * A.two_ab * pw / (d2 * (A.a * pw + 1.f));
                gx += c * dx; gy += c * dy;
            }
            A.eons[e] += A.eps[e];
            const float epn = A.eps[e] / A.nsr;   // repulsion from random nodes, gradient of log(1 - q)
            const int n_neg = (int)((fepoch - A.eonns[e]) / epn);
            for (int p = 0; p < n_neg; ++p) {
                const i64 k = rng.below((uint32_t)A.n);
                if (k == u) continue;
                const float kx = x - A.Ys[2 * k], ky = y - A.Ys[2 * k + 1];
                const float dd = kx * kx + ky * ky;
                if (dd > 0.f) {
                    const float pw = A.rep_b_is_one ? dd : powf(dd, A.rep_b);
                    const float c = A.two_b / ((0.001f + dd) * (A.rep_a * pw + 1.f));
                    gx += c * kx; gy += c * ky;
                } else {
                    const float s = 1e-3f / fmaxf(alpha, 1e-6f);
                    gx += (rng.uniform() - 0.5f) * s;
                    gy += (rng.uniform() - 0.5f) * s;
                }
            }
            A.eonns[e] += n_neg * epn;
            if (mid_w > 0.f && A.pairs && rng.uniform() < A.mid_rate) {   // mid-range pull towards a partner
                const int32_t w = A.pairs[u * A.n_mid + rng.below((uint32_t)A.n_mid)];
                if (w != u) { gx += -mid_w * (x - A.Ys[2 * (i64)w]); gy += -mid_w * (y - A.Ys[2 * (i64)w + 1]); }
            }
            step(x, y, gx, gy, alpha, A.clip_mode);
        }
        // the partners' pulls: u as the partner of the nodes that chose it, each as a step of its own
        if (mid_w > 0.f && A.rev_ptr) {
            const int rs = A.rev_ptr[u], re = A.rev_ptr[u + 1];
            float p = A.mid_rate * (float)(re - rs) / (float)A.n_mid;
            while (p > 0.f && re > rs) {
                if (rng.uniform() < p) {
                    const int32_t c = A.rev_idx[rs + (int)rng.below((uint32_t)(re - rs))];
                    if (c != u) step(x, y, -mid_w * (x - A.Ys[2 * (i64)c]), -mid_w * (y - A.Ys[2 * (i64)c + 1]), alpha, A.clip_mode);
                }
                p -= 1.f;
            }
        }
    }
    A.Yn[2 * u] = x; A.Yn[2 * u + 1] = y;
}

// eptr[u] = first edge whose head is >= u, for heads sorted non-decreasingly; flag bit 1: heads out of order, or a
// head or tail out of range
__global__ __launch_bounds__(TPB) void k_edge_ptr(const int32_t* __restrict__ head, const int32_t* __restrict__ tail, i64 m,
                                                  i64 n, int32_t* __restrict__ eptr, int* __restrict__ flag) {
    const i64 u = (i64)blockIdx.x * TPB + threadIdx.x;
    if (u < m && (head[u] < 0 || head[u] >= n || (u > 0 && head[u - 1] > head[u]) || tail[u] < 0 || tail[u] >= n))
        atomicOr(flag, 1);
    if (u > n) return;
    i64 lo = 0, hi = m;
    while (lo < hi) { const i64 mid = (lo + hi) / 2; if (head[mid] < u) lo = mid + 1; else hi = mid; }
    eptr[u] = (int32_t)lo;
}

// flag bit 2: an index of a (count) outside [0, n); bit 4: a non-positive edge period
__global__ __launch_bounds__(TPB) void k_check_idx(const int32_t* __restrict__ a, i64 count, i64 n, int* __restrict__ flag) {
    for (i64 k = (i64)blockIdx.x * TPB + threadIdx.x; k < count; k += (i64)gridDim.x * TPB)
        if (a[k] < 0 || a[k] >= n) atomicOr(flag, 2);
}
__global__ __launch_bounds__(TPB) void k_check_eps(const float* __restrict__ eps, i64 m, int* __restrict__ flag) {
    for (i64 k = (i64)blockIdx.x * TPB + threadIdx.x; k < m; k += (i64)gridDim.x * TPB)
        if (!(eps[k] > 0.f)) atomicOr(flag, 4);
}

}  // namespace

size_t layout_scratch_bytes(i64 n) { return (size_t)n * 2 * sizeof(float) + (size_t)(n + 1) * sizeof(int32_t) + 256; }

void launch_layout(const mg_layout_params& p, float* Y, i64 n, const int32_t* head, const int32_t* tail, const float* eps,
                   float* eons, float* eonns, i64 m, const int32_t* pairs, int n_mid, const int32_t* rev_ptr,
                   const int32_t* rev_idx, void* scratch, hipStream_t st) {
    float* Yb = reinterpret_cast<float*>(scratch);
    int32_t* eptr = reinterpret_cast<int32_t*>(Yb + 2 * n);
    int* flag = reinterpret_cast<int*>(eptr + n + 1);
    MG_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_edge_ptr, dim3((unsigned)((std::max(n + 1, m) + TPB - 1) / TPB)), dim3(TPB), 0, st, head, tail, m, n,
                       eptr, flag);
    MG_HIP(hipGetLastError());
    const dim3 cg(1024);
    hipLaunchKernelGGL(k_check_eps, cg, dim3(TPB), 0, st, eps, m, flag);
    if (n_mid > 0) hipLaunchKernelGGL(k_check_idx, cg, dim3(TPB), 0, st, pairs, n * n_mid, n, flag);
    int rev_nnz = 0;
    if (rev_ptr) {
        int32_t ends[2] = {0, 0};
        MG_HIP(hipMemcpyAsync(&ends[0], rev_ptr, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MG_HIP(hipMemcpyAsync(&ends[1], rev_ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        require(ends[0] == 0 && ends[1] >= 0, "reverse pairs must be a CSR over the n nodes");
        rev_nnz = ends[1];
        validate_csr(rev_ptr, rev_idx, n, rev_nnz, st);
    }
    MG_HIP(hipGetLastError());
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    require(!(h & 1), "edges must be sorted by head, with heads and tails in [0, n)");
    require(!(h & 2), "mid-range pair index out of range");
    require(!(h & 4), "epochs_per_sample must be positive");
    LayoutArgs A;
    A.n = n; A.n_fixed = p.n_fixed; A.eptr = eptr; A.tail = tail; A.eps = eps; A.eons = eons; A.eonns = eonns;
    A.pairs = n_mid > 0 ? pairs : nullptr; A.n_mid = n_mid; A.rev_ptr = rev_ptr; A.rev_idx = rev_idx;
    const float rep_a = p.rep_b > 0 ? p.rep_a : p.a, rep_b = p.rep_b > 0 ? p.rep_b : p.b;
    A.a = p.a; A.b = p.b; A.rep_a = rep_a; A.rep_b = rep_b; A.two_ab = 2.f * p.a * p.b;
    A.two_b = 2.f * rep_b * p.repulsion_strength; A.nsr = p.negative_sample_rate; A.exag = p.exaggeration;
    A.mid_rate = p.mid_rate; A.b_is_one = std::fabs(p.b - 1.f) < 1e-6f; A.rep_b_is_one = std::fabs(rep_b - 1.f) < 1e-6f;
    A.clip_mode = p.clip_mode; A.seed = p.seed;
    A.att = p.attraction_scale > 0.f ? p.attraction_scale : 1.5f;   // the owner mode's factor (src/layout.cpp:31-38)
    const int span = std::max(1, p.epoch_end - p.epoch_start);
    float* cur = Y;
    float* nxt = Yb;
    const dim3 grid((unsigned)((n + TPB - 1) / TPB));
    for (int e = p.epoch_start; e < p.epoch_end; ++e) {   // rates linear over [epoch_start, epoch_end) (src/layout.cpp:278-285)
        const float frac = (float)(e - p.epoch_start) / (float)span;
        const float alpha = p.alpha0 + (p.alpha1 - p.alpha0) * frac;
        const float mid_w = p.mid_w0 + (p.mid_w1 - p.mid_w0) * frac;
        A.Ys = cur; A.Yn = nxt;
        hipLaunchKernelGGL(k_layout_epoch, grid, dim3(TPB), 0, st, A, e, alpha, mid_w);
        MG_HIP(hipGetLastError());
        std::swap(cur, nxt);
    }
    if (cur != Y) MG_HIP(hipMemcpyAsync(Y, cur, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToDevice, st));
}

}  // namespace mg
