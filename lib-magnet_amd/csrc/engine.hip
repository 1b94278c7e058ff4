// Host side of libmagnet_hip: the FFT repulsion object (FFTConv2D, src/fftconv.cpp:43-237), the
// whole-loop optimiser (optimize_tsne, magnet/tsne.py:132-214) and the C-ABI of include/magnet_hip.h.
// Every kernel is issued on one stream; the only host round trip per iteration is the 48-byte
// read-back of the layout's bounding box, which the grid decision of the next iteration needs.
#include "common.hpp"
#include "batch.hpp"
#include "../../include/magnet_hip.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <chrono>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace mg {

int update_blocks(i64 n);
void launch_layout(const mg_layout_params& p, float* Y, i64 n, const int32_t* head, const int32_t* tail, const float* eps,
                   float* eons, float* eonns, i64 m, const int32_t* pairs, int n_mid, const int32_t* rev_ptr,
                   const int32_t* rev_idx, void* scratch, hipStream_t st);
size_t shim_scratch_ints(i64 n, int nb);
template <class T> void launch_interp_scatter(const T*, i64, const T*, int, double, double, double, int, int, T*, int*, hipStream_t);
template <class T> void launch_scatter_padded(const T*, i64, double, double, double, int, int, int, T*, int*, hipStream_t);
template <class T> void launch_interp_gather(const T*, i64, const T*, int, double, double, double, int, int, T*, hipStream_t);
template <class T> int launch_gather_forces(const T*, i64, const T*, int, double, double, double, int, int, T*, double*, hipStream_t);
template <class T> void launch_spectral_products(const T*, const T*, const T*, i64, i64, T*, hipStream_t);

static thread_local std::string g_err;
static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
static Tunables g_tunables;
Tunables& tunables() { return g_tunables; }
// The stream the entry points issue on: the calling thread's own (mg_set_thread_stream) if it set one, else
// the process-wide one (mg_set_stream, default the null stream).
static hipStream_t g_stream_all = nullptr;
static thread_local hipStream_t t_stream = nullptr;
static thread_local bool t_stream_set = false;
static inline hipStream_t lib_stream() { return t_stream_set ? t_stream : g_stream_all; }

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    const int code = (e == hipErrorOutOfMemory) ? MG_ERR_NOMEM : MG_ERR_HIP;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

// grow-only device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void* ensure(size_t b) {
        if (b <= bytes && p) return p;
        if (p) { MG_HIP(hipFree(p)); p = nullptr; bytes = 0; }
        const size_t want = b < 256 ? 256 : b;
        MG_HIP(hipMalloc(&p, want));
        bytes = want;
        return p;
    }
    // as ensure, but the old allocation goes to `dead` instead of hipFree, which would wait for the device;
    // the owner frees `dead` once the stream is known to be idle
    void* ensure_keep(size_t b, std::vector<void*>& dead) {
        if (b <= bytes && p) return p;
        if (p) { dead.push_back(p); p = nullptr; bytes = 0; }
        return ensure(b);
    }
};
template <class U> inline U* dptr(const DevBuf& b) { return reinterpret_cast<U*>(b.p); }

struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    void* ensure(size_t b) {
        if (b <= bytes && p) return p;
        if (p) { MG_HIP(hipHostFree(p)); p = nullptr; }
        MG_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
        return p;
    }
};


// ---- CSR check on the device (validate_csr, src/layout.hpp:21-32): bit 1 indptr[0] != 0,
// 2 decreasing indptr, 4 indptr[n] != nnz, 8 column index out of range
__global__ void k_validate_csr(const int32_t* __restrict__ ip, const int32_t* __restrict__ ix, i64 n, i64 nnz, int* flag) {
    const i64 stride = (i64)gridDim.x * blockDim.x;
    for (i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x; k < (n > nnz ? n : nnz); k += stride) {
        if (k < n && ip[k + 1] < ip[k]) atomicOr(flag, 2);
        if (k < nnz && (ix[k] < 0 || ix[k] >= n)) atomicOr(flag, 8);
        if (k == 0) { if (ip[0] != 0) atomicOr(flag, 1); if ((i64)ip[n] != nnz) atomicOr(flag, 4); }
    }
}

void validate_csr(const int32_t* ip, const int32_t* ix, i64 n, i64 nnz, hipStream_t st) {
    require(ip != nullptr && (nnz == 0 || ix != nullptr), "indptr / indices must be device arrays");
    require(nnz >= 0 && nnz < (i64)INT32_MAX, "nnz out of range");
    DevBuf flag;
    int* f = reinterpret_cast<int*>(flag.ensure(sizeof(int)));
    MG_HIP(hipMemsetAsync(f, 0, sizeof(int), st));
    const i64 work = n > nnz ? n : nnz;
    const unsigned blocks = (unsigned)std::min<i64>(4096, std::max<i64>(1, (work + 255) / 256));
    hipLaunchKernelGGL(k_validate_csr, dim3(blocks), dim3(256), 0, st, ip, ix, n, nnz, f);
    MG_HIP(hipGetLastError());
    int h = 0;
    MG_HIP(hipMemcpyAsync(&h, f, sizeof(int), hipMemcpyDeviceToHost, st));
    MG_HIP(hipStreamSynchronize(st));
    require(!(h & 1), "indptr[0] must be 0");
    require(!(h & 2), "indptr must be non-decreasing");
    require(!(h & 4), "indices must have indptr[n] entries");
    require(!(h & 8), "column index out of range");
}

// ---- per-stage timing with HIP events on the working stream (the reference's MAGNET_DEBUG laps,
// src/fftconv.cpp:69-70): enabled by mg_profile_enable or the MAGNET_DEBUG environment variable
enum Stage { ST_ATTR = 0, ST_SORT, ST_SPREAD, ST_ROWS_FWD, ST_TR_FWD, ST_COLUMNS, ST_TR_INV, ST_ROWS_INV, ST_GATHER,
             ST_UPDATE, ST_SPECTRA, NSTAGE };
static const char* kStageNames[NSTAGE] = {"attraction", "box_sort", "spread", "rows_forward", "transpose_forward", "columns",
                                          "transpose_inverse", "rows_inverse", "gather", "update", "spectra"};
struct Prof {
    bool on = false, debug = false;
    hipEvent_t ev[2 * NSTAGE] = {};
    bool pending[NSTAGE] = {};
    double total[NSTAGE] = {};
    double bytes[NSTAGE] = {};     // algorithmic HBM bytes of the launches timed (inputs read once, outputs written once)
    double pend_bytes[NSTAGE] = {};
    long long count[NSTAGE] = {};
    double n_sum = 0; int n_max = 0, m_max = 0; long long iters = 0;   // live grid side N, padded M
    // the same per stage and padded FFT size (size 0 for stages whose kernels do not depend on it)
    int pend_M[NSTAGE] = {};
    struct Sized { int stage, M; double total, bytes; long long count; };
    std::vector<Sized> sized;
    void add_sized(int s, int M, double ms, double by) {
        for (auto& z : sized)
            if (z.stage == s && z.M == M) { z.total += ms; z.bytes += by; ++z.count; return; }
        sized.push_back({s, M, ms, by, 1});
    }
    void grid(int N, int M) { if (!on) return; n_sum += N; n_max = N > n_max ? N : n_max; m_max = M > m_max ? M : m_max; ++iters; }
    void begin(int s, hipStream_t st) {
        if (!on) return;
        if (!ev[2 * s]) { MG_HIP(hipEventCreate(&ev[2 * s])); MG_HIP(hipEventCreate(&ev[2 * s + 1])); }
        MG_HIP(hipEventRecord(ev[2 * s], st));
    }
    void end(int s, hipStream_t st, double algo_bytes = 0, int M = 0) {
        if (!on) return;
        MG_HIP(hipEventRecord(ev[2 * s + 1], st));
        pending[s] = true;
        pend_bytes[s] = algo_bytes;
        pend_M[s] = M;
    }
    // after the stream has passed every recorded event
    void collect(int it) {
        if (!on) return;
        for (int s = 0; s < NSTAGE; ++s) {
            if (!pending[s]) continue;
            float ms = 0.f;
            MG_HIP(hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]));
            total[s] += ms; ++count[s]; pending[s] = false; bytes[s] += pend_bytes[s];
            add_sized(s, pend_M[s], ms, pend_bytes[s]);
            if (debug) fprintf(stderr, "  [it %d] %s %.3f ms\n", it, kStageNames[s], ms);
        }
    }
    void reset() {
        for (int s = 0; s < NSTAGE; ++s) { total[s] = 0; bytes[s] = 0; count[s] = 0; pending[s] = false; }
        n_sum = 0; n_max = 0; m_max = 0; iters = 0;
        sized.clear();
    }
};
static Prof g_prof;

// ------------------------------------------------------------------------------ repulsion engine
template <class T> struct Repulsion {
    int ni;
    int M_user;
    // kernel definition: t-SNE (delta from width, dof) or a value table on an M_in torus
    bool have_kernel = false, tsne_kernel = true;
    double kwidth = 0.7, kdof = 1.0;
    int M_in = 0;
    std::vector<double> kvals;   // (2, M_in, M_in)
    bool kvals_sym = false;      // the value tables are symmetric in x <-> y
    // node spacing of a value table's grid: the centred product set's odd tables x K2, y K2 need the offsets in
    // layout units, which a value table does not carry (set from the grid of each compute; -1 when unused)
    double vdelta = -1.0;
    struct Spec { int M; double delta, dof; DevBuf K[4]; unsigned long use; };
    // FP32: the centred product set (node-relative charges); FP64: the reference's
    static constexpr bool centred = sizeof(T) == 4;
    static constexpr int ntab = centred ? 4 : 2;
    std::list<Spec> cache;
    unsigned long tick = 0;
    DevBuf keysA, valsA, keysB, valsB, hist, item_key, item_valid, item_val, W, Wh, Ph, Tr, phi;
    DevBuf zpart, zout, inv_out, spec_tmp, table, centroid;
    // allocations replaced while the loop runs (grid buffers, spectra, as the layout grows): freed by
    // release_dead() after a synchronisation, so that no growth waits for the device
    std::vector<void*> dead;
    // overlap (set by the loop): spectra that must be recomputed run on a second stream (default priority), beside
    // the box sort, the spread and the row pass, which leave most of the device idle at the loop's sizes. The side
    // stream starts after everything issued before the call (ev_go: no table in use is overwritten) and the column
    // pass waits for it (ev_spec).
    bool overlap = false;
    // clear_w (set by the loop): the gather sets the boxes the spread wrote back to zero, so W stays zero between
    // calls and the spread needs no fill of the grid; w_zero is W.p while that holds (a fill of the whole buffer
    // establishes it after an allocation, and after a call that issued the spread but not the gather)
    bool clear_w = false;
    const void* w_zero = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t ev_go = nullptr, ev_spec = nullptr;
    int last_N = 0, last_M = 0, sorted_buf = 0;
    GridT<T> last_g{};
    i64 last_n = 0;

    Repulsion(int M, int ni_) : ni(ni_), M_user(M) {
        require(ni >= 1 && ni <= 8, "n_interp must be in [1, 8]");
        fft_init_twiddles();
        zout.ensure(4 * sizeof(double));
        inv_out.ensure(4 * sizeof(T));
    }
    Repulsion(const Repulsion&) = delete;
    Repulsion& operator=(const Repulsion&) = delete;
    ~Repulsion() {
        release_dead();
        if (side) (void)hipStreamSynchronize(side);
        if (ev_go) (void)hipEventDestroy(ev_go);
        if (ev_spec) (void)hipEventDestroy(ev_spec);
        if (side) (void)hipStreamDestroy(side);
    }
    // only once the work that used them is complete (the caller has synchronised the stream)
    void release_dead() {
        for (void* q : dead) (void)hipFree(q);
        dead.clear();
    }

    bool have_spectra(int M) const {
        const double delta = tsne_kernel ? kwidth / ni : vdelta;
        for (const auto& s : cache)
            if (s.M == M && s.delta == delta && s.dof == kdof) return true;
        return false;
    }

    // spectra for the device size M (cached; the reference keeps up to 9, magnet/tsne.py:52-54)
    Spec& spectra(int M, hipStream_t st, bool* fresh = nullptr) {
        require(have_kernel, "kernels not set");
        if (fresh) *fresh = false;
        const double delta = tsne_kernel ? kwidth / ni : vdelta;
        for (auto& s : cache)
            if (s.M == M && s.delta == delta && s.dof == kdof) { s.use = ++tick; return s; }
        const int Mh = M / 2 + 1;
        Spec* slot = nullptr;
        if (cache.size() >= 9) {
            // the least recently used entry whose buffers hold the new tables takes them (while the grid clips,
            // every iteration brings a new width at one M: one entry is then reused without reallocation);
            // if none is large enough, the least recently used entry, its buffers grown
            auto fits = [&](const Spec& c) { return c.K[0].bytes >= (size_t)Mh * Mh * sizeof(T); };
            for (auto& c : cache)
                if (fits(c) && (!slot || c.use < slot->use)) slot = &c;
            if (!slot) {
                slot = &cache.front();
                for (auto& c : cache) if (c.use < slot->use) slot = &c;
            }
        } else {
            cache.emplace_back();
            slot = &cache.back();
        }
        Spec& s = *slot;
        s.M = M; s.delta = delta; s.dof = kdof; s.use = ++tick;
        if (fresh) *fresh = true;
        T* Kp[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int t = 0; t < ntab; ++t) Kp[t] = reinterpret_cast<T*>(s.K[t].ensure_keep((size_t)Mh * Mh * sizeof(T), dead));
        cplx<T>* tmp = reinterpret_cast<cplx<T>*>(spec_tmp.ensure_keep((size_t)4 * Mh * Mh * sizeof(cplx<T>), dead));
        const double* tab = nullptr;
        if (!tsne_kernel) {
            // resample the caller's table (offsets up to M_in / 2) onto the M x M torus
            std::vector<double> h((size_t)2 * M * M, 0.0);
            const int lim = std::min(M_in / 2, M / 2);
            for (int w = 0; w < 2; ++w)
                for (int a = 0; a < M; ++a) {
                    const int oa = a <= M / 2 ? a : a - M;
                    if (std::abs(oa) > lim) continue;
                    const int ia = (oa + M_in) % M_in;
                    for (int b = 0; b < M; ++b) {
                        const int ob = b <= M / 2 ? b : b - M;
                        if (std::abs(ob) > lim) continue;
                        h[((size_t)w * M + a) * M + b] = kvals[((size_t)w * M_in + ia) * M_in + (ob + M_in) % M_in];
                    }
                }
            double* d = reinterpret_cast<double*>(table.ensure(h.size() * sizeof(double)));
            MG_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, st));
            MG_HIP(hipStreamSynchronize(st));
            tab = d;
        }
        launch_spectra<T>(M, delta, kdof, tab, tsne_kernel && sizeof(T) == 4, ntab, tmp, Kp, st, tsne_kernel || kvals_sym);
        return s;
    }

    void ensure(i64 n, int N, int M) {
        const size_t nk = (size_t)std::max<i64>(n, 1);
        keysA.ensure(nk * sizeof(int)); valsA.ensure(nk * sizeof(int));
        keysB.ensure(nk * sizeof(int)); valsB.ensure(nk * sizeof(int));
        hist.ensure(((size_t)sort_hist_ints(n)) * sizeof(int));   // digit rows + totals
        const int nch = std::max(1, spread_chunks(n));
        item_key.ensure((size_t)2 * nch * sizeof(int));
        item_valid.ensure((size_t)2 * nch * sizeof(int));
        item_val.ensure((size_t)2 * nch * ni * 3 * ni * sizeof(T));
        // the grid buffers hold the largest grid of this FFT size (N <= M / 2): a layout that grows by a few
        // boxes then reuses them, instead of a hipFree (which waits for the device) and a hipMalloc per growth
        const int Mh = M / 2 + 1;
        const size_t Nc = (size_t)std::max(N, M / 2);
        W.ensure_keep(3 * Nc * Nc * sizeof(T), dead);
        Wh.ensure_keep(3 * Mh * Nc * sizeof(cplx<T>), dead);
        Ph.ensure_keep(4 * Mh * Nc * sizeof(cplx<T>), dead);
        Tr.ensure_keep(4 * Mh * Nc * sizeof(cplx<T>), dead);

        phi.ensure_keep(4 * Nc * Nc * sizeof(T), dead);
        zpart.ensure((size_t)2 * gather_blocks(n) * sizeof(double));   // compensated pairs
    }

    RepBuffers<T> bufs() {
        RepBuffers<T> b;
        b.keysA = dptr<int>(keysA); b.valsA = dptr<int>(valsA); b.keysB = dptr<int>(keysB); b.valsB = dptr<int>(valsB);
        b.hist = dptr<int>(hist); b.item_key = dptr<int>(item_key); b.item_valid = dptr<int>(item_valid);
        b.item_val = dptr<T>(item_val); b.W = dptr<T>(W); b.Wh = dptr<cplx<T>>(Wh); b.Ph = dptr<cplx<T>>(Ph);
        b.Tr = dptr<cplx<T>>(Tr); b.phi = dptr<T>(phi); b.zpart = dptr<double>(zpart);
        b.zout = dptr<double>(zout); b.inv_out = dptr<T>(inv_out);
        return b;
    }

    int device_M(int N) const { return fft_size(N); }

    // enqueue the whole repulsion; sumQ and 1/sumQ stay on the device (zout, inv_out)
    // inv != null: Y is relabelled (position inv[i] holds original point i); the box sort then takes the points in
    // original order, so its ties, and every sum after it, are those of the unrelabelled layout
    void compute(const T* Y, i64 n, const GridT<T>& g, T* F, double* sumq_hist, hipStream_t st, const int* inv = nullptr) {
        require(n >= 1, "Y must be (n, 2) with n >= 1");
        require(g.N >= 1 && g.width > T(0), "grid larger than the buffer");
        require(n < (i64)INT32_MAX, "too many points");
        if (!tsne_kernel) require(M_in >= 2 * g.N - 1, "padded size below 2N - 1 (aliased convolution) is not supported", MG_ERR_UNSUP);
        const int M = device_M(g.N);
        require(M > 0, "grid too large for the device FFT (n_interp * n_boxes > 4096)", MG_ERR_UNSUP);
        if (!tsne_kernel) vdelta = centred ? g.width_d / ni : -1.0;
        ensure(n, g.N, M);
        Prof& P = g_prof;
        const double ts = sizeof(T), NN = (double)g.N * g.N, cs = 2 * ts, Mh = M / 2 + 1, dn = (double)n;
        const bool side_run = overlap && tsne_kernel && !have_spectra(M);
        hipStream_t sst = st;
        if (side_run) {
            if (!side) {
                MG_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
                MG_HIP(hipEventCreateWithFlags(&ev_go, hipEventDisableTiming));
                MG_HIP(hipEventCreateWithFlags(&ev_spec, hipEventDisableTiming));
            }
            MG_HIP(hipEventRecord(ev_go, st));   // before this call's first launch
            MG_HIP(hipStreamWaitEvent(side, ev_go, 0));
            sst = side;
        }
        // the spectra have slack until the column pass: they are issued first, after the box sort (default), or after
        // the row pass (tunable spec_issue 0, 1, 2), so that the launches of the chain before them reach the device first
        const int spec_at = side_run ? tunables().spec_issue : 0;
        Spec* sp = nullptr;
        auto issue_spectra = [&]() {
            P.begin(ST_SPECTRA, sst);
            bool fresh = false;
            sp = &spectra(M, sst, &fresh);
            P.end(ST_SPECTRA, sst, fresh ? ntab * Mh * Mh * ts : 0.0, M);   // tables written when the grid changed
            if (side_run) MG_HIP(hipEventRecord(ev_spec, side));
        };
        if (spec_at == 0) issue_spectra();
        RepBuffers<T> b = bufs();
        P.begin(ST_SORT, st);
        sorted_buf = launch_box_sort<T>(Y, n, g, b, st, inv);
        P.end(ST_SORT, st, 2 * dn * ts + 8 * dn);
        if (spec_at == 1) issue_spectra();
        const int* keys = sorted_buf == 0 ? b.keysA : b.keysB;
        const int* vals = sorted_buf == 0 ? b.valsA : b.valsB;
        P.begin(ST_SPREAD, st);
        if (clear_w && w_zero != W.p) launch_fill<T>(b.W, T(0), (i64)(W.bytes / sizeof(T)), st);
        w_zero = nullptr;
        launch_spread<T>(keys, vals, Y, n, g, ni, centred, b, st, !clear_w);
        P.end(ST_SPREAD, st, 8 * dn + 2 * dn * ts + 3 * NN * ts);
        // small grids: the column pass reads and writes the row layout in place of the two transposes (L2-resident);
        // larger ones: the row passes write and read the column layout themselves (folded), or through transposes
        const bool strided = M <= tunables().fft_strided_max;
        const bool folded = !strided && tunables().fft_fold > 0 && fold_supported<T>(M);
        P.begin(ST_ROWS_FWD, st);
        if (folded) launch_rows_forward_fold<T>(b.W, g.N, M, b.Wh, tunables().fft_fold, st);
        else launch_rows_forward<T>(b.W, g.N, M, b.Tr, st);
        P.end(ST_ROWS_FWD, st, 3 * NN * ts + 3 * Mh * g.N * cs, M);
        if (!strided && !folded) {
            P.begin(ST_TR_FWD, st);   // layout change only: no algorithmic traffic
            launch_transpose<T>(b.Tr, b.Wh, 3, g.N, M / 2 + 1, st);
            P.end(ST_TR_FWD, st, 0, M);
        }
        if (spec_at == 2) issue_spectra();
        if (side_run) MG_HIP(hipStreamWaitEvent(st, ev_spec, 0));
        const Spec& s = *sp;
        P.begin(ST_COLUMNS, st);
        const T* Kc[4] = {dptr<T>(s.K[0]), dptr<T>(s.K[1]), dptr<T>(s.K[2]), dptr<T>(s.K[3])};
        launch_columns<T>(strided ? b.Tr : b.Wh, g.N, M, Kc, centred, b.Ph, st, strided, folded ? fold_stride_n(g.N) : 0);
        P.end(ST_COLUMNS, st, 3 * Mh * g.N * cs + ntab * Mh * Mh * ts + 4 * Mh * g.N * cs, M);
        if (!strided && !folded) {
            P.begin(ST_TR_INV, st);
            launch_transpose<T>(b.Ph, b.Tr, 4, M / 2 + 1, g.N, st);
            P.end(ST_TR_INV, st, 0, M);
        }
        P.begin(ST_ROWS_INV, st);
        if (folded) launch_rows_inverse_fold<T>(b.Ph, g.N, M, b.phi, tunables().fft_fold, st);
        else launch_rows_inverse<T>(strided ? b.Ph : b.Tr, g.N, M, b.phi, st);
        P.end(ST_ROWS_INV, st, 4 * Mh * g.N * cs + 4 * NN * ts, M);
        P.begin(ST_GATHER, st);
        launch_gather<T>(keys, vals, Y, n, g, ni, centred, b.phi, F, b, sumq_hist, st, clear_w ? b.W : nullptr);
        if (clear_w) w_zero = W.p;
        P.end(ST_GATHER, st, 8 * dn + 2 * dn * ts + 4 * NN * ts + 2 * dn * ts);
        last_N = g.N; last_M = M; last_n = n; last_g = g;
        P.grid(g.N, M);
    }
};

// ------------------------------------------------------------------------------ whole loop
// The schedule's grid limits, checked before any work: the largest grid the loop may reach (n_interp * max_boxes
// nodes per side) must fit the largest device FFT, so that a run cannot fail hundreds of iterations in, when the
// layout first spans that many boxes.
void check_schedule(const mg_tsne_params* p) {
    require(p != nullptr, "params must not be null");
    require(p->n_interp >= 1 && p->n_interp <= 8, "n_interp must be in [1, 8]");
    require(p->box_width > 0 && p->min_boxes >= 1 && p->max_boxes >= p->min_boxes, "bad grid parameters");
    require(fft_size(p->n_interp * p->max_boxes) > 0,
            "n_interp * max_boxes = " + std::to_string(p->n_interp * p->max_boxes) +
                " exceeds 4096, the largest grid side of the device FFT (8192 points): lower max_boxes to " +
                std::to_string(4096 / p->n_interp) + " or below",
            MG_ERR_UNSUP);
}

template <class T> struct Optimizer {
    Repulsion<T> rep;
    DevBuf attr, F, grad, part, mb, bb_part, bb_counter, bb_out, sumq;
    PinnedBuf host_mb;
    // per-iteration diagnostics (KL divergence, gradient norm), when the caller asks for them
    double *h_kl = nullptr, *h_gnorm = nullptr;
    DevBuf trace_scratch, trace_out;
    bool tracing() const { return h_kl || h_gnorm; }
    Optimizer(int ni) : rep(0, ni) {}

    // one run: [it0, it1) of the schedule p on the state (Y, U, G), issued on stream st
    const mg_tsne_params* p = nullptr;
    const int32_t *ip = nullptr, *ix = nullptr, *pairs = nullptr;
    const T* data = nullptr;
    T *Y = nullptr, *U = nullptr, *G = nullptr;
    i64 nnz = 0, n = 0;
    int m = 0, it0 = 0, it1 = 0;
    double* h_sumQ = nullptr;
    mg_tsne_callback cb = nullptr;
    void* user = nullptr;
    hipStream_t st = nullptr;
    double lr = 0, attr_bytes = 0;
    bool fused = true;
    T inv_dof = T(1);
    double* dsum = nullptr;
    T* hmb = nullptr;
    int* htag = nullptr;
    PinnedBuf host_tag;
    // locality relabelling: the state lives in Yw / Uw / Gw in the order of a recent box sort (position inv[i]
    // holds original point i, perm is its inverse); the CSR is renumbered to match (cip, cix, cP)
    bool relabel = false;
    int relabel_every = 25, since = 0;
    T *uY = nullptr, *uU = nullptr, *uG = nullptr;
    const int32_t *cip = nullptr, *cix = nullptr;
    const T* cP = nullptr;
    DevBuf Yw, Uw, Gw, Yt, Ut, Gt, perm, inv, order, inv_new, lenb, tiles, ipn, ixn, Pn, Ycb;

    // the next iteration's attraction, on its own stream (default priority): it needs the updated layout but no
    // grid, so it runs beside that iteration's repulsion; ev_a0 orders it after everything issued so far
    // (the update, a relabelling), and the update that reads its output waits for ev_a1
    void attraction() {
        MG_HIP(hipEventRecord(ev_a0, st));
        MG_HIP(hipStreamWaitEvent(aside, ev_a0, 0));
        g_prof.begin(ST_ATTR, aside);
        launch_attraction<T>(cip, cix, cP, Y, n, 2, inv_dof, dptr<T>(attr), nnz, aside);
        g_prof.end(ST_ATTR, aside, attr_bytes);
        MG_HIP(hipEventRecord(ev_a1, aside));
        a_pending = true;
    }

    void start_relabelling() {
        const size_t b = (size_t)2 * n * sizeof(T);
        T* y = reinterpret_cast<T*>(Yw.ensure(b)); T* u = reinterpret_cast<T*>(Uw.ensure(b)); T* g = reinterpret_cast<T*>(Gw.ensure(b));
        Yt.ensure(b); Ut.ensure(b); Gt.ensure(b);
        MG_HIP(hipMemcpyAsync(y, uY, b, hipMemcpyDeviceToDevice, st));
        MG_HIP(hipMemcpyAsync(u, uU, b, hipMemcpyDeviceToDevice, st));
        MG_HIP(hipMemcpyAsync(g, uG, b, hipMemcpyDeviceToDevice, st));
        Y = y; U = u; G = g;
        const size_t ni = (size_t)n * sizeof(int);
        perm.ensure(ni); inv.ensure(ni); order.ensure(ni); inv_new.ensure(ni); lenb.ensure(ni);
        tiles.ensure((size_t)(n / 1024 + 2) * sizeof(int));
        ipn.ensure((size_t)(n + 1) * sizeof(int32_t)); ixn.ensure((size_t)std::max<i64>(nnz, 1) * sizeof(int32_t));
        Pn.ensure((size_t)std::max<i64>(nnz, 1) * sizeof(T));
        launch_identity_labels(dptr<int>(perm), dptr<int>(inv), n, st);
    }

    // renumber the points in the Morton order of their boxes on the last grid (ties in original point order), and the
    // CSR (from the caller's) to match; the sort uses the repulsion's key buffers, free until its next box sort
    void relabel_now() {
        RepBuffers<T> b = rep.bufs();
        int bits = 1;
        while ((1 << bits) < rep.last_g.nb) ++bits;
        launch_morton_keys<T>(Y, dptr<int>(inv), n, rep.last_g, b.keysA, b.valsA, st);
        const int buf = launch_key_sort<T>(b.keysA, b.valsA, b.keysB, b.valsB, b.hist, n, 2 * bits, st);
        const int* vals = buf == 0 ? b.valsA : b.valsB;
        launch_relabel<T>(vals, dptr<int>(perm), n, ip, ix, data, dptr<int>(order), dptr<int>(inv_new), dptr<int>(lenb),
                          dptr<int>(tiles), dptr<int32_t>(ipn), dptr<int32_t>(ixn), dptr<T>(Pn), st);
        launch_perm_state<T>(Y, U, G, dptr<int>(inv), dptr<int>(order), n, dptr<T>(Yt), dptr<T>(Ut), dptr<T>(Gt), st);
        std::swap(Yw.p, Yt.p); std::swap(Uw.p, Ut.p); std::swap(Gw.p, Gt.p);
        std::swap(perm.p, order.p); std::swap(inv.p, inv_new.p);
        Y = dptr<T>(Yw); U = dptr<T>(Uw); G = dptr<T>(Gw);
        cip = dptr<int32_t>(ipn); cix = dptr<int32_t>(ixn); cP = dptr<T>(Pn);
    }

    // checks, buffers, the starting layout's bbox (read back), and the first iteration's attraction
    void begin(const mg_tsne_params* p_, const int32_t* ip_, const int32_t* ix_, const T* data_, i64 nnz_, i64 n_, T* Y_, T* U_,
               T* G_, const int32_t* pairs_, int m_, int it0_, int it1_, double* h_sumQ_, mg_tsne_callback cb_, void* user_,
               hipStream_t st_, bool check_csr = true) {
        p = p_; ip = ip_; ix = ix_; data = data_; nnz = nnz_; n = n_; Y = Y_; U = U_; G = G_; pairs = pairs_; m = m_;
        uY = Y_; uU = U_; uG = G_; cip = ip_; cix = ix_; cP = data_; since = 0;
        it0 = it0_; it1 = it1_; h_sumQ = h_sumQ_; cb = cb_; user = user_; st = st_;
        require(p != nullptr, "params must not be null");
        require(n >= 1 && Y && U && G, "Y, U, G must be (n, 2) device arrays");
        require(p->n_iter >= 0 && it0 >= 0 && it1 >= it0 && it1 <= p->n_iter, "iteration range outside the schedule");
        require(p->dof > 0, "dof must be positive");
        require(p->box_width > 0 && p->min_boxes >= 1 && p->max_boxes >= p->min_boxes, "bad grid parameters");
        require(!pairs || m >= 1, "pairs must be (n, m) with m >= 1");
        if (check_csr) validate_csr(ip, ix, n, nnz, st);
        rep.have_kernel = true; rep.tsne_kernel = true; rep.kdof = p->dof; rep.kwidth = p->box_width;
        rep.overlap = true;
        rep.clear_w = true;
        lr = p->learning_rate > 0 ? p->learning_rate : std::max(200.0, (double)n / p->early_exaggeration);
        fused = !(pairs && m > 0 && p->mid_weight > 0);
        attr.ensure((size_t)2 * n * sizeof(T));
        F.ensure((size_t)2 * n * sizeof(T));
        if (!fused) grad.ensure((size_t)2 * n * sizeof(T));
        part.ensure((size_t)8 * 1024 * sizeof(double));   // compensated block partials
        unsigned* bc = reinterpret_cast<unsigned*>(bb_counter.ensure(64));
        if (!counters_zeroed) { MG_HIP(hipMemsetAsync(bc, 0, 64, st)); counters_zeroed = true; }
        mb.ensure(8 * sizeof(T));
        bb_part.ensure((size_t)4 * 1024 * sizeof(T));
        bb_out.ensure(4 * sizeof(T));
        dsum = reinterpret_cast<double*>(sumq.ensure((size_t)std::max(1, it1 - it0) * sizeof(double)));
        hmb = reinterpret_cast<T*>(host_mb.ensure(8 * sizeof(T)));
        htag = reinterpret_cast<int*>(host_tag.ensure(sizeof(int)));
        __atomic_store_n(htag, -1, __ATOMIC_RELEASE);
        launch_bbox<T>(Y, n, dptr<T>(bb_part), bc, dptr<T>(bb_out), st);
        MG_HIP(hipMemcpyAsync(hmb + 2, dptr<T>(bb_out), 4 * sizeof(T), hipMemcpyDeviceToHost, st));
        MG_HIP(hipStreamSynchronize(st));
        inv_dof = T(1) / (T)p->dof;
        attr_bytes = 4.0 * (n + 1) + nnz * (4.0 + sizeof(T)) + 4.0 * n * sizeof(T);
        if (!mb_ready) MG_HIP(hipEventCreateWithFlags(&mb_ready, hipEventDisableTiming));
        if (!aside) {
            MG_HIP(hipStreamCreateWithFlags(&aside, hipStreamNonBlocking));
            MG_HIP(hipEventCreateWithFlags(&ev_a0, hipEventDisableTiming));
            MG_HIP(hipEventCreateWithFlags(&ev_a1, hipEventDisableTiming));
        }
        a_pending = false;
        if (tracing() && it1 > it0) {
            trace_scratch.ensure((size_t)trace_scratch_doubles() * sizeof(double));
            trace_out.ensure((size_t)2 * (it1 - it0) * sizeof(double));
            launch_trace_plogp<T>(data, nnz, dptr<double>(trace_scratch), st);
        }
        relabel = fused && p->relabel_every >= 0 && n >= 2 && it1 - it0 > 1;
        relabel_every = p->relabel_every > 0 ? p->relabel_every : 25;
        if (relabel) start_relabelling();
        // a run longer than the interval relabels after its first iteration (the first grid), then at the interval: a
        // layout in random point order would otherwise run its first iterations without locality (1M points in random
        // order: 1.63 against 1.07 ms per iteration)
        since = relabel && it1 - it0 > relabel_every ? relabel_every - 1 : 0;
        if (it0 < it1) attraction();
    }

    // iteration it: grid from the last bbox, repulsion, update, bbox read-back; then the next iteration's
    // attraction, which needs no grid, so the device runs it while the host waits for the bbox
    void issue(int it) {
        const double bb[4] = {(double)hmb[2], (double)hmb[3], (double)hmb[4], (double)hmb[5]};
        require(std::isfinite(bb[0]) && std::isfinite(bb[1]) && std::isfinite(bb[2]) && std::isfinite(bb[3]),
                "layout is not finite");
        const GridDecision gd = decide_grid(bb, p->box_width, p->min_boxes, p->max_boxes);
        const GridT<T> g = make_grid<T>(gd.xmin, gd.ymin, gd.width, gd.nb, p->n_interp);
        rep.kwidth = gd.width;
        const bool early = it < p->early_iter;
        const T rho = (T)(early ? p->early_exaggeration : p->exaggeration);
        const T mom = (T)(early ? p->momentum0 : p->momentum1);
        const int* labels = relabel ? dptr<int>(inv) : nullptr;
        rep.compute(Y, n, g, dptr<T>(F), dsum + (it - it0), st, labels);
        if (a_pending) { MG_HIP(hipStreamWaitEvent(st, ev_a1, 0)); a_pending = false; }   // attr of this iteration
        const T* dinv = dptr<T>(rep.inv_out);
        double* dtrace = tracing() ? dptr<double>(trace_out) + 2 * (it - it0) : nullptr;
        if (dtrace && fused)
            launch_trace<T>(ip, ix, data, Y, labels, n, p->dof, dptr<T>(attr), dptr<T>(F), dinv, rho, dptr<double>(rep.zout),
                            dptr<double>(trace_scratch), dtrace, st);
        g_prof.begin(ST_UPDATE, st);
        if (fused) {
            launch_update<T>(0, Y, U, G, dptr<T>(attr), dptr<T>(F), dinv, T(0), rho, (T)lr, mom, nullptr, 0, T(0), n, 2,
                             dptr<double>(part), dptr<T>(mb), labels, st, hmb, htag, it);
        } else {
            const double w_mid = p->mid_weight * std::max(0.0, 1.0 - it / std::max(1.0, p->mid_phase * p->n_iter)) / (double)n;
            launch_grad<T>(dptr<T>(attr), dptr<T>(F), dinv, T(0), rho, Y, w_mid > 0 ? pairs : nullptr, m, (T)w_mid, n, 2, 1,
                           dptr<T>(grad), st);
            if (dtrace)
                launch_trace<T>(ip, ix, data, Y, nullptr, n, p->dof, dptr<T>(grad), nullptr, nullptr, T(0),
                                dptr<double>(rep.zout), dptr<double>(trace_scratch), dtrace, st);
            launch_update<T>(1, Y, U, G, dptr<T>(grad), nullptr, nullptr, T(0), T(0), (T)lr, mom, nullptr, 0, T(0), n, 2,
                             dptr<double>(part), dptr<T>(mb), nullptr, st, hmb, htag, it);
        }
        g_prof.end(ST_UPDATE, st, (fused ? 20.0 : 24.0) * n * sizeof(T));
        MG_HIP(hipEventRecord(mb_ready, st));   // the recentring kernel wrote the bbox to hmb and `it` to *htag
        if (relabel && it + 1 < it1 && ++since % relabel_every == 0) relabel_now();
        if (it + 1 < it1) attraction();
    }

    // waits for iteration it; returns true when the callback asks to stop
    // the bbox of iteration it is in hmb once *htag reads it (written after it by the recentring kernel); Y, U, G are
    // final once mb_ready has passed
    void wait_bbox(int it) {
        for (unsigned spin = 1;; ++spin) {
            if (__atomic_load_n(htag, __ATOMIC_ACQUIRE) == it) return;
            if ((spin & 255) == 0) {
                const hipError_t q = hipEventQuery(mb_ready);
                if (q == hipSuccess) {
                    if (__atomic_load_n(htag, __ATOMIC_ACQUIRE) != it)   // not seen yet: read the device copy
                        MG_HIP(hipMemcpy(hmb, dptr<T>(mb), 6 * sizeof(T), hipMemcpyDeviceToHost));
                    return;
                }
                if (q != hipErrorNotReady) MG_HIP(q);
            }
        }
    }
    bool complete(int it) {
        const bool cb_due = cb && (it % 50 == 0 || it == p->n_iter - 1);
        if (cb_due || g_prof.on) MG_HIP(hipEventSynchronize(mb_ready));   // Y, U, G and the bbox of iteration it are final
        else wait_bbox(it);
        if (g_prof.on) { MG_HIP(hipStreamSynchronize(st)); MG_HIP(hipStreamSynchronize(aside)); g_prof.collect(it); }
        if (cb_due) {
            if (!relabel) return cb(it, Y, n, user) != 0;
            // the layout in original point order; Y itself is final for iteration it (the event), and the
            // relabelling that may follow it in the stream is not waited for
            T* y = reinterpret_cast<T*>(Ycb.ensure((size_t)2 * n * sizeof(T)));
            MG_HIP(hipStreamSynchronize(st));
            launch_unperm<T>(Y, dptr<int>(inv), n, y, st);
            MG_HIP(hipStreamSynchronize(st));
            return cb(it, y, n, user) != 0;
        }
        return false;
    }

    void finish(int it_end) {
        if (h_sumQ && it_end > it0)
            MG_HIP(hipMemcpyAsync(h_sumQ, dsum, (size_t)(it_end - it0) * sizeof(double), hipMemcpyDeviceToHost, st));
        std::vector<double> tr;
        if (tracing() && it_end > it0) {
            tr.resize((size_t)2 * (it_end - it0));
            MG_HIP(hipMemcpyAsync(tr.data(), trace_out.p, tr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        if (relabel) {   // the state back to the caller's buffers, in original point order
            launch_unperm<T>(Y, dptr<int>(inv), n, uY, st);
            launch_unperm<T>(U, dptr<int>(inv), n, uU, st);
            launch_unperm<T>(G, dptr<int>(inv), n, uG, st);
        }
        MG_HIP(hipStreamSynchronize(st));
        MG_HIP(hipStreamSynchronize(aside));   // an attraction queued before a callback stopped the loop
        a_pending = false;
        rep.release_dead();
        for (size_t k = 0; 2 * k < tr.size(); ++k) {
            if (h_kl) h_kl[k] = tr[2 * k];
            if (h_gnorm) h_gnorm[k] = tr[2 * k + 1];
        }
    }

    int run(const mg_tsne_params* p_, const int32_t* ip_, const int32_t* ix_, const T* data_, i64 nnz_, i64 n_, T* Y_, T* U_,
            T* G_, const int32_t* pairs_, int m_, int it0_, int it1_, double* h_sumQ_, mg_tsne_callback cb_, void* user_,
            hipStream_t st_, bool check_csr = true) {
        if (getenv("MAGNET_DEBUG")) { g_prof.on = true; g_prof.debug = true; }
        begin(p_, ip_, ix_, data_, nnz_, n_, Y_, U_, G_, pairs_, m_, it0_, it1_, h_sumQ_, cb_, user_, st_, check_csr);
        int stopped = 0, it_end = it1;
        for (int it = it0; it < it1; ++it) {
            issue(it);
            if (complete(it)) { stopped = 1; it_end = it + 1; break; }
        }
        finish(it_end);
        return stopped;
    }
    bool counters_zeroed = false;
    hipEvent_t mb_ready = nullptr;
    hipStream_t aside = nullptr;   // the attraction's stream
    hipEvent_t ev_a0 = nullptr, ev_a1 = nullptr;
    bool a_pending = false;
    ~Optimizer() {
        if (aside) (void)hipStreamSynchronize(aside);
        if (mb_ready) (void)hipEventDestroy(mb_ready);
        if (ev_a0) (void)hipEventDestroy(ev_a0);
        if (ev_a1) (void)hipEventDestroy(ev_a1);
        if (aside) (void)hipStreamDestroy(aside);
    }
};

// ------------------------------------------------------------------------------ batched engine
// Many embeddings on one schedule, each iteration's stages launched once for all of them (batched launches,
// batch.hpp). Point arrays are concatenated (instance i owns rows [off_i, off_i + n_i)); the box sort is one
// stable sort of (instance, box) keys over all points, which leaves every instance's points in the order its
// own box sort gives; the spread, FFT, gather and update stages run each instance's blocks with its own
// grid, buffers and block counts, so its result is the single-instance loop's result bit for bit. FFT stages
// run one launch per FFT size in use.
static int bits_for(long long v) { int b = 0; while ((1ll << b) < v) ++b; return b; }

// Groups plan and finish from several host threads. A stream capture must not overlap a call that may synchronise the
// device (hipFree, graph destruction) or allocate in another thread, so plan() -- its allocations, staging and
// capture -- and the releases of finish() hold this lock.
static std::mutex g_graph_mu;

// arena growth with headroom; the old allocation goes to `dead` (freed after the run's last synchronisation)
static void grow(DevBuf& b, size_t need, std::vector<void*>& dead) { if (need > b.bytes) b.ensure_keep(need + need / 4, dead); }

template <class F> std::vector<BMap> make_maps(const std::vector<int>& which, F blocks) {
    std::vector<BMap> out;
    BMap m;
    m.count = 0; m.start[0] = 0;
    for (int i : which) {
        const int b = blocks(i);
        if (b <= 0) continue;
        if (m.count == BMAX) { out.push_back(m); m.count = 0; m.start[0] = 0; }
        m.inst[m.count] = i;
        m.start[m.count + 1] = m.start[m.count] + b;
        ++m.count;
    }
    if (m.count) out.push_back(m);
    return out;
}

template <class T> struct BatchEngine {
    static constexpr bool centred = sizeof(T) == 4;
    static constexpr int ntab = centred ? 4 : 2;
    const mg_tsne_params* p = nullptr;
    std::vector<const mg_tsne_instance*> user;
    int count = 0, ni = 3, Mmin = 128, Mmax = 0;
    hipStream_t st = nullptr;
    std::vector<i64> n, off;
    i64 ntot = 0;
    T inv_dof = T(1);
    BSched sched{};
    DevBuf Yc, Uc, Gc, Ac, Fc, keysA, valsA, keysB, valsB, hist, atable, mbd, sumq, counters, zout, inv, upart, state;
    DevBuf aW, aWh, aPh, aTr, aphi, aik, aiv, aival, azp, spec_stage;
    const void* w_zero = nullptr;   // aW.p once its charge grids are zero (then kept zero by the gathers)
    DevBuf table[2];          // the plans' tables, in turn (a new plan's table is written while the last one's
    PinnedBuf htab[2], hrep[2];   // iterations still run); the states read back after each stretch, in turn
    std::vector<void*> dead;  // replaced allocations, freed once the stream is idle (at the end of the run)
    // kernel spectra of unclipped grids (cell width box_width / ni), per FFT size, shared by all instances
    struct Spec { int M; DevBuf K[4]; };
    std::list<Spec> shared;
    // per instance whose window holds the largest size: its own spectra of a clipped grid, and their staging
    std::vector<std::unique_ptr<DevBuf[]>> Kc;
    std::vector<BMap> attr_maps;
    std::vector<int> attr_G;
    hipStream_t aside = nullptr, cst = nullptr;   // the attraction's stream, the capture stream
    hipEvent_t ev_a0 = nullptr, ev_a1 = nullptr, ev_chunk[2] = {nullptr, nullptr};
    // ---- the plan: each instance's window of FFT sizes, the block ranges of every launch, the captured iteration
    std::vector<int> Mcur, Mlo, Mhi;
    std::vector<int> sizes;   // the FFT sizes of the windows, ascending
    int box_bits = 1, sorted = 0, tcur = 0;
    struct Launches {
        std::vector<BMap> spec1, spect, spec2, specr;   // clipped grids at the largest size
        std::vector<BMap> keys, spread, gather, update;
        // work lists: offsets of rows / columns / inverse rows per size, and of the transposes, in the list buffer,
        // with the fixed grid of each launch
        std::vector<int> rows_f, cols, rows_i, g_rows_f, g_cols, g_rows_i;
        int tr0 = 0, tr1 = 0, g_tr0 = 0, g_tr1 = 0;
        int ndesc = 0;
    } L;
    DevBuf wdesc[2], wcand[2], wlist[2];   // per plan, in turn with the tables
    PinnedBuf hdesc[2], hcand[2];          // their host staging (the copies run after the iterations queued before)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipGraph_t, hipGraphExec_t>> old_graphs;   // destroyed at the end of the run
    bool use_graph = true;
    int chunk = 8;
    bool done = false;
    // stretches queued and not yet read back (oldest first): their report slot and iteration count
    std::vector<std::pair<int, int>> pending;
    int slot = 0;
    std::vector<int> last_next;   // next_it of every instance at the last read-back
    // counters of the run (mg_tsne_batch_stats)
    long long n_plans = 0, n_replays = 0, n_launches = 0, n_checks = 0, n_launches_per_iter = 0, n_stalls = 0;
    long long n_iters_issued = 0;
    double plan_s = 0;

    ~BatchEngine() {
        if (st) (void)hipStreamSynchronize(st);
        if (aside) (void)hipStreamSynchronize(aside);
        drop_graphs();
        release_dead();
        for (hipEvent_t e : {ev_a0, ev_a1, ev_chunk[0], ev_chunk[1]}) if (e) (void)hipEventDestroy(e);
        if (aside) (void)hipStreamDestroy(aside);
        if (cst) (void)hipStreamDestroy(cst);
    }
    void drop_graphs() {
        if (exec || graph) old_graphs.push_back({graph, exec});
        for (auto& g : old_graphs) {
            if (g.second) (void)hipGraphExecDestroy(g.second);
            if (g.first) (void)hipGraphDestroy(g.first);
        }
        old_graphs.clear();
        graph = nullptr; exec = nullptr;
    }
    BInst<T>* host_table() { return reinterpret_cast<BInst<T>*>(htab[tcur].p); }
    BInst<T>* dev_table() const { return reinterpret_cast<BInst<T>*>(table[tcur].p); }
    const int* sorted_keys() const { return dptr<int>(sorted == 0 ? keysA : keysB); }
    const int* sorted_vals() const { return dptr<int>(sorted == 0 ? valsA : valsB); }

    // the window of sizes around M that the schedule can reach: the size before it and the two after it (layouts
    // mostly grow; a grid that leaves the window stalls its instance until the host's next plan, which comes a
    // stretch or two of iterations later)
    int window_lo(int M) const {
        if (!tunables().batch_window_down) return M;
        const int lo = fft_size_prev(M);
        return lo >= Mmin ? lo : M;
    }
    int window_hi(int M) const {
        int hi = M;
        for (int k = 0; k < tunables().batch_window_up; ++k) { const int nx = fft_size_next(hi); if (nx > 0 && nx <= Mmax) hi = nx; }
        return hi;
    }

    void begin(const mg_tsne_params* p_, const std::vector<const mg_tsne_instance*>& inst, hipStream_t st_) {
        p = p_; user = inst; st = st_;
        count = (int)inst.size();
        ni = p->n_interp;
        require(count >= 1 && count <= 1024, "a batch holds 1 to 1024 instances");
        require(ni >= 1 && ni <= 8, "n_interp must be in [1, 8]");
        require(p->dof > 0, "dof must be positive");
        require(p->n_iter >= 0, "n_iter must be >= 0");
        require(p->box_width > 0 && p->min_boxes >= 1 && p->max_boxes >= p->min_boxes, "bad grid parameters");
        require(p->max_boxes <= 1024, "max_boxes above 1024 is not supported in a batch", MG_ERR_UNSUP);
        fft_init_twiddles();
        const Tunables& tu = tunables();
        use_graph = tu.batch_graph != 0;
        chunk = tu.batch_chunk;
        Mmin = fft_size(ni * std::min(p->min_boxes, p->max_boxes));
        Mmax = fft_size(ni * p->max_boxes);
        n.resize(count); off.resize(count);
        ntot = 0;
        for (int i = 0; i < count; ++i) {
            const mg_tsne_instance& I = *inst[i];
            require(I.n >= 1 && I.Y && I.U && I.G, "Y, U, G must be (n, 2) device arrays");
            require(I.nnz == 0 || I.data != nullptr, "data must be a device array");
            validate_csr(I.indptr, I.indices, I.n, I.nnz, st);
            n[i] = I.n; off[i] = ntot; ntot += I.n;
        }
        require(ntot < (i64)INT32_MAX, "too many points in the batch");
        const size_t pb = (size_t)2 * ntot * sizeof(T);
        T* Y = reinterpret_cast<T*>(Yc.ensure(pb));
        T* U = reinterpret_cast<T*>(Uc.ensure(pb));
        T* G = reinterpret_cast<T*>(Gc.ensure(pb));
        Ac.ensure(pb); Fc.ensure(pb);
        for (int i = 0; i < count; ++i) {
            const size_t b = (size_t)2 * n[i] * sizeof(T);
            MG_HIP(hipMemcpyAsync(Y + 2 * off[i], inst[i]->Y, b, hipMemcpyDeviceToDevice, st));
            MG_HIP(hipMemcpyAsync(U + 2 * off[i], inst[i]->U, b, hipMemcpyDeviceToDevice, st));
            MG_HIP(hipMemcpyAsync(G + 2 * off[i], inst[i]->G, b, hipMemcpyDeviceToDevice, st));
        }
        const size_t nk = (size_t)ntot * sizeof(int);
        keysA.ensure(nk); valsA.ensure(nk); keysB.ensure(nk); valsB.ensure(nk);
        hist.ensure(((size_t)sort_hist_ints(ntot)) * sizeof(int));
        unsigned* cnt = reinterpret_cast<unsigned*>(counters.ensure(sizeof(unsigned)));   // the bbox's last-block counter
        MG_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), st));
        zout.ensure((size_t)2 * count * sizeof(double));
        inv.ensure((size_t)4 * count * sizeof(T));
        upart.ensure((size_t)count * 8 * 1024 * sizeof(double));
        T* mb = reinterpret_cast<T*>(mbd.ensure((size_t)8 * count * sizeof(T)));
        sumq.ensure((size_t)count * std::max(1, p->n_iter) * sizeof(double));
        for (int k = 0; k < 2; ++k) {
            table[k].ensure((size_t)count * sizeof(BInst<T>));
            htab[k].ensure((size_t)count * sizeof(BInst<T>));
            hrep[k].ensure((size_t)count * sizeof(BState));
            // work-list staging for any plan: 3 lists per FFT size (13 sizes) and 3 more, each instance a candidate
            // of at most 4 sizes' lists and of the 3 others
            hdesc[k].ensure((size_t)(3 * 13 + 3) * sizeof(WorkDesc));
            hcand[k].ensure((size_t)count * (3 * 4 + 3) * sizeof(int));
        }
        BState* S = reinterpret_cast<BState*>(state.ensure((size_t)count * sizeof(BState)));
        BState* hs = reinterpret_cast<BState*>(hrep[0].p);
        PinnedBuf hmb;
        T* hm = reinterpret_cast<T*>(hmb.ensure((size_t)8 * count * sizeof(T)));
        Kc.clear();
        for (int i = 0; i < count; ++i) Kc.emplace_back(new DevBuf[5]);
        // starting layouts' bboxes (the update of each iteration writes the next ones)
        for (int i = 0; i < count; ++i)
            launch_bbox<T>(Y + 2 * off[i], n[i], reinterpret_cast<T*>(upart.p), cnt, mb + 8 * i + 2, st);
        MG_HIP(hipMemcpyAsync(hm, mb, (size_t)8 * count * sizeof(T), hipMemcpyDeviceToHost, st));
        inv_dof = T(1) / (T)p->dof;
        sched = {p->n_iter, p->early_iter, ni, p->min_boxes, p->max_boxes, count, p->box_width,
                 p->early_exaggeration, p->exaggeration, p->momentum0, p->momentum1, p->box_width / ni};
        // the fixed part of the table, and the attraction launches (rows grouped by lanes per row)
        BInst<T>* h = reinterpret_cast<BInst<T>*>(htab[0].p);
        std::memset((void*)h, 0, (size_t)count * sizeof(BInst<T>));
        std::vector<std::vector<int>> byG(7);
        for (int i = 0; i < count; ++i) {
            const mg_tsne_instance& I = *inst[i];
            h[i].ip = I.indptr; h[i].ix = I.indices; h[i].P = reinterpret_cast<const T*>(I.data);
            h[i].n = n[i]; h[i].off = off[i];
            h[i].lr = (T)(p->learning_rate > 0 ? p->learning_rate : std::max(200.0, (double)n[i] / p->early_exaggeration));
            h[i].zout = dptr<double>(zout) + 2 * i;
            h[i].inv_out = dptr<T>(inv) + 4 * i;
            h[i].upart = dptr<double>(upart) + (size_t)i * 8 * 1024;
            h[i].mb = dptr<T>(mbd) + 8 * i;
            h[i].sumq_base = dptr<double>(sumq) + (size_t)i * std::max(1, p->n_iter);
            const int Gl = attraction_lanes(n[i], I.nnz);
            byG[bits_for(Gl)].push_back(i);
        }
        std::memcpy(htab[1].p, htab[0].p, (size_t)count * sizeof(BInst<T>));
        attr_maps.clear(); attr_G.clear();
        for (int b = 0; b < 7; ++b) {
            const int Gl = 1 << b;
            for (const BMap& m : make_maps(byG[b], [&](int i) { return attraction_blocks(n[i], Gl); })) {
                attr_maps.push_back(m); attr_G.push_back(Gl);
            }
        }
        MG_HIP(hipMemcpyAsync(atable.ensure((size_t)count * sizeof(BInst<T>)), h, (size_t)count * sizeof(BInst<T>),
                              hipMemcpyHostToDevice, st));
        for (int i = 0; i < count; ++i) hs[i] = {0, 0, 0, 0, -1.0};
        MG_HIP(hipMemcpyAsync(S, hs, (size_t)count * sizeof(BState), hipMemcpyHostToDevice, st));
        MG_HIP(hipStreamSynchronize(st));
        last_next.assign(count, 0);
        // the first plan from the starting grids
        Mcur.assign(count, 0);
        for (int i = 0; i < count; ++i) {
            const double bb[4] = {(double)hm[8 * i + 2], (double)hm[8 * i + 3], (double)hm[8 * i + 4], (double)hm[8 * i + 5]};
            require(std::isfinite(bb[0]) && std::isfinite(bb[1]) && std::isfinite(bb[2]) && std::isfinite(bb[3]),
                    "layout is not finite");
            const GridDecision gd = decide_grid(bb, p->box_width, p->min_boxes, p->max_boxes);
            Mcur[i] = fft_size(gd.nb * ni);
            require(Mcur[i] > 0, "grid too large for the device FFT (n_interp * n_boxes > 4096)", MG_ERR_UNSUP);
        }
        if (!aside) {
            MG_HIP(hipStreamCreateWithFlags(&aside, hipStreamNonBlocking));
            MG_HIP(hipStreamCreateWithFlags(&cst, hipStreamNonBlocking));
            MG_HIP(hipEventCreateWithFlags(&ev_a0, hipEventDisableTiming));
            MG_HIP(hipEventCreateWithFlags(&ev_a1, hipEventDisableTiming));
            MG_HIP(hipEventCreateWithFlags(&ev_chunk[0], hipEventDisableTiming));
            MG_HIP(hipEventCreateWithFlags(&ev_chunk[1], hipEventDisableTiming));
        }
        done = p->n_iter == 0;
        tcur = 1;   // plan() takes the other table
        if (!done) plan();
    }

    // the shared spectra of size M (computed on the stream the first time a window holds M)
    const Spec& shared_spectra(int M) {
        for (const Spec& s : shared) if (s.M == M) return s;
        shared.emplace_back();
        Spec& s = shared.back();
        s.M = M;
        const size_t Mh = M / 2 + 1;
        T* Kp[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int t = 0; t < ntab; ++t) Kp[t] = reinterpret_cast<T*>(s.K[t].ensure(Mh * Mh * sizeof(T)));
        cplx<T>* tmp = reinterpret_cast<cplx<T>*>(spec_stage.ensure_keep(4 * Mh * Mh * sizeof(cplx<T>), dead));
        launch_spectra<T>(M, sched.delta0, p->dof, nullptr, centred, ntab, tmp, Kp, st, true);
        return s;
    }

    // windows around Mcur, buffers for their largest grids, spectra of new sizes, block ranges and the table (the one
    // the last plan did not use, so the iterations still queued keep theirs); then the captured iteration. Nothing
    // here waits for the device: replaced buffers stay allocated until the run ends.
    void plan() {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        const double t0 = now_s();
        tcur ^= 1;
        Mlo.assign(count, 0); Mhi.assign(count, 0);
        sizes.clear();
        int nbmax = 1;
        size_t oW = 0, oWh = 0, oPh = 0, oTr = 0, ophi = 0, oit = 0, oiv = 0, ozp = 0;
        auto al = [](size_t x) { return (x + 63) / 64 * 64; };
        std::vector<size_t> pW(count), pWh(count), pPh(count), pTr(count), pphi(count), pit(count), piv(count), pzp(count);
        for (int i = 0; i < count; ++i) {
            Mlo[i] = window_lo(Mcur[i]); Mhi[i] = window_hi(Mcur[i]);
            for (int S = Mlo[i]; S > 0 && S <= Mhi[i]; S = fft_size_next(S))
                if (std::find(sizes.begin(), sizes.end(), S) == sizes.end()) sizes.push_back(S);
            const size_t Nc = (size_t)Mhi[i] / 2, Mh = Nc + 1, nch = spread_chunks(n[i]);
            nbmax = std::max(nbmax, std::min(p->max_boxes, (int)(Nc / ni)));
            pW[i] = oW; oW += al(3 * Nc * Nc);
            pWh[i] = oWh; oWh += al(3 * Mh * Nc);
            pPh[i] = oPh; oPh += al(4 * Mh * Nc);
            pTr[i] = oTr; oTr += al(4 * Mh * Nc);
            pphi[i] = ophi; ophi += al(4 * Nc * Nc);
            pit[i] = oit; oit += al(2 * nch);
            piv[i] = oiv; oiv += al(2 * nch * ni * 3 * ni);
            pzp[i] = ozp; ozp += al(2 * gather_blocks(n[i]));   // compensated pairs
            // a window that holds the largest size may clip: the instance's own tables, allocated once at that size
            if (Mhi[i] == Mmax && !Kc[i][0].p) {
                const size_t Mhx = Mmax / 2 + 1;
                for (int t = 0; t < ntab; ++t) Kc[i][t].ensure(Mhx * Mhx * sizeof(T));
                Kc[i][4].ensure(4 * Mhx * Mhx * sizeof(cplx<T>));
            }
        }
        std::sort(sizes.begin(), sizes.end());
        box_bits = bits_for((long long)nbmax * nbmax);
        grow(aW, oW * sizeof(T), dead); grow(aWh, oWh * sizeof(cplx<T>), dead); grow(aPh, oPh * sizeof(cplx<T>), dead);
        grow(aTr, oTr * sizeof(cplx<T>), dead); grow(aphi, ophi * sizeof(T), dead); grow(aik, oit * sizeof(int), dead);
        grow(aiv, oit * sizeof(int), dead); grow(aival, oiv * sizeof(T), dead); grow(azp, ozp * sizeof(double), dead);
        // the charge grids stay zero between iterations (each gather zeroes the boxes its spread wrote): a new buffer
        // is cleared once, after the iterations queued before, which use the old one
        if (w_zero != aW.p) { launch_fill<T>(dptr<T>(aW), T(0), (i64)(aW.bytes / sizeof(T)), st); w_zero = aW.p; }
        BInst<T>* h = host_table();
        for (int i = 0; i < count; ++i) {
            BInst<T>& b = h[i];
            b.Mlo = Mlo[i]; b.Mhi = Mhi[i];
            b.item_key = dptr<int>(aik) + pit[i]; b.item_valid = dptr<int>(aiv) + pit[i]; b.item_val = dptr<T>(aival) + piv[i];
            b.W = dptr<T>(aW) + pW[i]; b.Wh = dptr<cplx<T>>(aWh) + pWh[i]; b.Ph = dptr<cplx<T>>(aPh) + pPh[i];
            b.Tr = dptr<cplx<T>>(aTr) + pTr[i]; b.phi = dptr<T>(aphi) + pphi[i];
            int w = 0;
            for (int S = Mlo[i]; S > 0 && S <= Mhi[i] && w < 4; S = fft_size_next(S), ++w) {
                const Spec& sp = shared_spectra(S);
                b.KM[w] = S;
                for (int t = 0; t < 4; ++t) b.KS[w][t] = t < ntab ? dptr<T>(sp.K[t]) : nullptr;
            }
            for (; w < 4; ++w) { b.KM[w] = b.KM[w - 1]; for (int t = 0; t < 4; ++t) b.KS[w][t] = b.KS[w - 1][t]; }
            for (int t = 0; t < 4; ++t) b.Kw[t] = (t < ntab && Kc[i][t].p) ? dptr<T>(Kc[i][t]) : nullptr;
            b.spec_tmp = Kc[i][4].p ? dptr<cplx<T>>(Kc[i][4]) : nullptr;
            b.zpart = dptr<double>(azp) + pzp[i];
        }
        MG_HIP(hipMemcpyAsync(table[tcur].p, h, (size_t)count * sizeof(BInst<T>), hipMemcpyHostToDevice, st));
        // block ranges
        std::vector<int> all(count), clip;
        for (int i = 0; i < count; ++i) { all[i] = i; if (Mhi[i] == Mmax && Kc[i][0].p) clip.push_back(i); }
        L = Launches();
        const int Mhx = Mmax / 2 + 1;
        L.spec1 = make_maps(clip, [&](int) { return fft_col_blocks(Mmax); });
        L.spect = make_maps(clip, [&](int) { return transpose_blocks(Mhx, Mhx, 1); });
        L.spec2 = make_maps(clip, [&](int) { return spec_pair_blocks(Mmax); });
        L.specr = make_maps(clip, [&](int) { return transpose_blocks(Mhx, Mhx, 1); });
        // work lists: candidates and descriptors; a launch's grid is its most blocks (the largest grids of the
        // windows), at most wl_grid
        std::vector<WorkDesc> desc;
        std::vector<int> cand;
        int woff = 0;
        const int gmax = tunables().batch_wl_grid;
        auto add = [&](int kind, int M, const std::vector<int>& c, long long cap_blocks) {
            desc.push_back({kind, M, (int)c.size(), woff, (int)cand.size()});
            cand.insert(cand.end(), c.begin(), c.end());
            woff += wl_ints((int)c.size());
            return std::make_pair(desc.back().off, (int)std::min<long long>(cap_blocks, gmax));
        };
        for (int S : sizes) {
            std::vector<int> w;
            long long br = 0, bc = 0, bi = 0;
            for (int i = 0; i < count; ++i)
                if (Mlo[i] <= S && S <= Mhi[i]) {
                    w.push_back(i);
                    br += fft_row_blocks(S / 2, S, 3); bc += fft_col_blocks(S); bi += fft_row_blocks(S / 2, S, 4);
                }
            auto a = add(WK_ROWS3, S, w, br); L.rows_f.push_back(a.first); L.g_rows_f.push_back(a.second);
            a = add(WK_COLS, S, w, bc); L.cols.push_back(a.first); L.g_cols.push_back(a.second);
            a = add(WK_ROWS4, S, w, bi); L.rows_i.push_back(a.first); L.g_rows_i.push_back(a.second);
        }
        long long b0 = 0, b1 = 0;
        for (int i = 0; i < count; ++i) { b0 += transpose_cap_blocks(Mhi[i], 0); b1 += transpose_cap_blocks(Mhi[i], 1); }
        auto a = add(WK_TR0, 0, all, b0); L.tr0 = a.first; L.g_tr0 = a.second;
        a = add(WK_TR1, 0, all, b1); L.tr1 = a.first; L.g_tr1 = a.second;
        L.ndesc = (int)desc.size();
        require(desc.size() * sizeof(WorkDesc) <= hdesc[tcur].bytes && cand.size() * sizeof(int) <= hcand[tcur].bytes,
                "work-list staging too small");
        void* hd = hdesc[tcur].p;
        void* hc = hcand[tcur].p;
        std::memcpy(hd, desc.data(), desc.size() * sizeof(WorkDesc));
        std::memcpy(hc, cand.data(), cand.size() * sizeof(int));
        MG_HIP(hipMemcpyAsync(wdesc[tcur].ensure_keep(desc.size() * sizeof(WorkDesc), dead), hd, desc.size() * sizeof(WorkDesc),
                              hipMemcpyHostToDevice, st));
        MG_HIP(hipMemcpyAsync(wcand[tcur].ensure_keep(cand.size() * sizeof(int), dead), hc, cand.size() * sizeof(int),
                              hipMemcpyHostToDevice, st));
        wlist[tcur].ensure_keep((size_t)woff * sizeof(int), dead);
        L.keys = make_maps(all, [&](int i) { return key_blocks(n[i]); });
        L.spread = make_maps(all, [&](int i) { return spread_chunks(n[i]); });
        L.gather = make_maps(all, [&](int i) { return gather_blocks(n[i]); });
        L.update = make_maps(all, [&](int i) { return update_blocks(n[i]); });
        if (exec || graph) old_graphs.push_back({graph, exec});
        graph = nullptr; exec = nullptr;
        if (use_graph) {   // one capture at a time in the process (g_graph_mu)
            MG_HIP(hipStreamBeginCapture(cst, hipStreamCaptureModeRelaxed));
            try {
                iteration(cst);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cst, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            MG_HIP(hipStreamEndCapture(cst, &graph));
            MG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        } else {
            n_launches_per_iter = count_launches();
        }
        ++n_plans;
        plan_s += now_s() - t0;
    }
    long long count_launches() const {
        return 2 + (long long)attr_maps.size() + (ntab == 4 ? 4 : 3) * (long long)L.spec1.size() + (long long)L.keys.size()
               + 3 * ((box_bits + bits_for(count) + 9) / 10) + 2 + 2 * L.spread.size() + 2 * L.gather.size()
               + 2 * L.update.size() + 3 * (long long)sizes.size();
    }

    // one iteration of every instance on stream s: the decision, the attraction beside the repulsion, the update
    void iteration(hipStream_t s) {
        BInst<T>* I = dev_table();
        launch_decide_batch<T>(I, dptr<BState>(state), sched, s);
        launch_worklists<T>(I, dptr<WorkDesc>(wdesc[tcur]), L.ndesc, dptr<int>(wcand[tcur]), dptr<int>(wlist[tcur]), s);
        // the attraction needs the last update only: it runs on its own stream beside the repulsion
        MG_HIP(hipEventRecord(ev_a0, s));
        MG_HIP(hipStreamWaitEvent(aside, ev_a0, 0));
        const BInst<T>* IA = reinterpret_cast<const BInst<T>*>(atable.p);
        for (size_t a = 0; a < attr_maps.size(); ++a)
            launch_attraction_batch<T>(IA, attr_maps[a], attr_G[a], dptr<T>(Yc), inv_dof, dptr<T>(Ac), aside);
        MG_HIP(hipEventRecord(ev_a1, aside));
        // kernel spectra of the clipped grids whose width changed
        for (size_t q = 0; q < L.spec1.size(); ++q)
            launch_spectra_batch<T>(I, L.spec1[q], L.spect[q], L.spec2[q], L.specr[q], Mmax, p->dof, centred, ntab, s);
        // box sort of all points by (instance, box)
        for (const BMap& m : L.keys) launch_keys_batch<T>(I, m, dptr<T>(Yc), box_bits, dptr<int>(keysA), dptr<int>(valsA), s);
        sorted = launch_key_sort_masked(dptr<int>(keysA), dptr<int>(valsA), dptr<int>(keysB), dptr<int>(valsB), dptr<int>(hist),
                                        ntot, box_bits + bits_for(count), box_bits, s);
        // spread
        const int* wl = dptr<int>(wlist[tcur]);
        for (const BMap& m : L.spread) launch_spread_batch<T>(I, m, sorted_keys(), sorted_vals(), dptr<T>(Yc), ni, centred, s);
        // FFT convolution
        for (size_t z = 0; z < sizes.size(); ++z) launch_rows_forward_wl<T>(I, wl + L.rows_f[z], sizes[z], L.g_rows_f[z], s);
        launch_transpose_wl<T>(I, wl + L.tr0, 0, L.g_tr0, s);
        for (size_t z = 0; z < sizes.size(); ++z) launch_columns_wl<T>(I, wl + L.cols[z], sizes[z], L.g_cols[z], s);
        launch_transpose_wl<T>(I, wl + L.tr1, 1, L.g_tr1, s);
        for (size_t z = 0; z < sizes.size(); ++z) launch_rows_inverse_wl<T>(I, wl + L.rows_i[z], sizes[z], L.g_rows_i[z], s);
        // gather, Z, and the update
        for (const BMap& m : L.gather) launch_gather_batch<T>(I, m, sorted_keys(), sorted_vals(), dptr<T>(Yc), ni, centred, dptr<T>(Fc), s);
        MG_HIP(hipStreamWaitEvent(s, ev_a1, 0));
        for (const BMap& m : L.update) launch_update_batch<T>(I, m, dptr<T>(Yc), dptr<T>(Uc), dptr<T>(Gc), dptr<T>(Ac), dptr<T>(Fc), s);
        n_launches_per_iter = count_launches();
    }

    // queue the next stretch of iterations and the read-back of the instances' states (at most two stretches
    // outstanding: the device runs one while the host reads the other and plans)
    void launch() {
        if (done || pending.size() >= 2) return;
        int left = 0;
        for (int i = 0; i < count; ++i) left = std::max(left, p->n_iter - last_next[i]);
        for (const auto& q : pending) left -= q.second;
        if (left <= 0 && !pending.empty()) return;
        const int k = std::max(1, std::min(chunk, left));
        for (int r = 0; r < k; ++r) {
            ++n_iters_issued;
            if (use_graph) { MG_HIP(hipGraphLaunch(exec, st)); ++n_replays; }
            else iteration(st);
            n_launches += n_launches_per_iter;
        }
        MG_HIP(hipMemcpyAsync(hrep[slot].p, state.p, (size_t)count * sizeof(BState), hipMemcpyDeviceToHost, st));
        MG_HIP(hipEventRecord(ev_chunk[slot], st));
        pending.push_back({slot, k});
        slot ^= 1;
    }

    // wait for the oldest stretch; plan again when an instance stalled or reached the edge of its window
    void check() {
        if (done || pending.empty()) return;
        const int sl = pending.front().first;
        pending.erase(pending.begin());
        MG_HIP(hipEventSynchronize(ev_chunk[sl]));
        ++n_checks;
        const BState* hs = reinterpret_cast<const BState*>(hrep[sl].p);
        bool all_done = true, replan = false;
        for (int i = 0; i < count; ++i) {
            require(hs[i].stall != 2, "layout is not finite");
            last_next[i] = hs[i].next_it;
            if (hs[i].next_it >= p->n_iter) continue;
            all_done = false;
            const int M = hs[i].M;
            require(M > 0 && M <= Mmax, "grid too large for the device FFT (n_interp * n_boxes > 4096)", MG_ERR_UNSUP);
            if (hs[i].stall == 1) ++n_stalls;
            if (hs[i].stall == 1 || M != Mcur[i]) replan = true;
        }
        if (all_done) { done = true; return; }
        if (replan) {
            for (int i = 0; i < count; ++i)
                if (hs[i].next_it < p->n_iter) Mcur[i] = hs[i].M;
            plan();
        }
    }

    void finish() {
        const T* Y = dptr<T>(Yc); const T* U = dptr<T>(Uc); const T* G = dptr<T>(Gc);
        for (int i = 0; i < count; ++i) {
            const size_t b = (size_t)2 * n[i] * sizeof(T);
            MG_HIP(hipMemcpyAsync(user[i]->Y, Y + 2 * off[i], b, hipMemcpyDeviceToDevice, st));
            MG_HIP(hipMemcpyAsync(user[i]->U, U + 2 * off[i], b, hipMemcpyDeviceToDevice, st));
            MG_HIP(hipMemcpyAsync(user[i]->G, G + 2 * off[i], b, hipMemcpyDeviceToDevice, st));
            if (user[i]->h_sumQ && p->n_iter > 0)
                MG_HIP(hipMemcpyAsync(user[i]->h_sumQ, dptr<double>(sumq) + (size_t)i * p->n_iter,
                                      (size_t)p->n_iter * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        MG_HIP(hipStreamSynchronize(st));
        MG_HIP(hipStreamSynchronize(aside));
        pending.clear();
        std::lock_guard<std::mutex> lk(g_graph_mu);
        drop_graphs();
        release_dead();
    }
    void release_dead() {
        for (void* q : dead) (void)hipFree(q);
        dead.clear();
    }
};

// counters of the last batched run (summed over its groups)
static std::atomic<long long> g_bs_plans{0}, g_bs_replays{0}, g_bs_launches{0}, g_bs_checks{0}, g_bs_iters{0}, g_bs_stalls{0};
static std::atomic<long long> g_bs_plan_us{0};

// The batch in n_groups groups of about equal point counts (largest instances dealt first), each group one
// batched engine on its own HIP stream; group s is driven by host thread s mod n_threads. A thread queues a stretch
// of iterations for each of its groups, then checks each in turn, so the device runs one group while the host
// plans another.
template <class T> void optimize_batch(const mg_tsne_params* p, int count, const mg_tsne_instance* inst, int n_groups,
                                      int n_threads) {
    check_schedule(p);
    require(count >= 1 && inst != nullptr, "no instances");
    require(n_groups >= 1 && n_groups <= 64, "n_streams must be in [1, 64]");
    require(n_threads >= 1 && n_threads <= n_groups, "n_threads must be in [1, n_streams]");
    const int S = std::min(n_groups, count), NT = std::min(n_threads, S);
    std::vector<int> order(count);
    for (int i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].n > inst[b].n; });
    std::vector<std::vector<const mg_tsne_instance*>> groups(S);
    std::vector<i64> load(S, 0);
    for (int i : order) {
        const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groups[s].push_back(&inst[i]);
        load[s] += inst[i].n;
    }
    fft_init_twiddles();
    prepare_fft_kernels<T>();   // no kernel attribute is set while the groups' threads launch and capture
    std::vector<hipStream_t> streams(S, nullptr);
    struct Guard { std::vector<hipStream_t>& s; ~Guard() { for (auto x : s) if (x) (void)hipStreamDestroy(x); } } guard{streams};
    for (auto& s : streams) MG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<std::unique_ptr<BatchEngine<T>>> eng;
    for (int s = 0; s < S; ++s) {
        eng.emplace_back(new BatchEngine<T>());
        eng.back()->begin(p, groups[s], streams[s]);
    }
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    std::atomic<bool> failed{false};
    std::vector<std::string> errs(NT);
    std::vector<int> codes(NT, 0);
    auto work = [&](int t) {
        try {
            MG_HIP(hipSetDevice(dev));
            std::vector<BatchEngine<T>*> mine;
            for (int s = t; s < S; s += NT) mine.push_back(eng[s].get());
            for (auto* e : mine) e->launch();
            for (;;) {
                bool busy = false;
                for (auto* e : mine) if (!e->done) { e->launch(); busy = true; }
                if (!busy || failed.load(std::memory_order_relaxed)) break;
                for (auto* e : mine) e->check();
            }
            for (auto* e : mine) e->finish();
        } catch (const Error& e) {
            codes[t] = e.code; errs[t] = e.what(); failed = true;
        } catch (const std::exception& e) {
            codes[t] = 1; errs[t] = e.what(); failed = true;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < NT; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (auto x : streams) (void)hipStreamSynchronize(x);
    long long pl = 0, rp = 0, la = 0, ch = 0, us = 0, sl = 0, it = 0;
    for (auto& e : eng) {
        pl += e->n_plans; rp += e->n_replays; la += e->n_launches; ch += e->n_checks; us += (long long)(e->plan_s * 1e6);
        sl += e->n_stalls; it += e->n_iters_issued;
    }
    g_bs_plans = pl; g_bs_replays = rp; g_bs_launches = la; g_bs_checks = ch; g_bs_plan_us = us; g_bs_iters = it; g_bs_stalls = sl;
    for (int t = 0; t < NT; ++t) if (codes[t]) throw Error(codes[t], errs[t]);
}

}  // namespace mg

// ================================================================================== C-ABI
using namespace mg;

// A reusable whole loop: the CSR is validated once, and the optimiser's buffers, streams and spectra cache persist
// from one run to the next.
struct mg_tsne_session {
    int prec = MG_F32;
    mg_tsne_params p{};
    const int32_t *ip = nullptr, *ix = nullptr, *pairs = nullptr;
    const void* data = nullptr;
    int64_t nnz = 0, n = 0;
    int m = 0;
    std::unique_ptr<Optimizer<float>> f32;
    std::unique_ptr<Optimizer<double>> f64;
};

struct mg_fftconv {
    int prec;
    std::unique_ptr<Repulsion<float>> f32;
    std::unique_ptr<Repulsion<double>> f64;
};

#define MG_API_BEGIN try {
#define MG_API_END                                                           \
    }                                                                        \
    catch (const mg::Error& e) { g_err = e.what(); return e.code; }          \
    catch (const std::bad_alloc&) { g_err = "host allocation failed"; return MG_ERR_NOMEM; } \
    catch (const std::exception& e) { g_err = e.what(); return MG_ERR_HIP; } \
    return MG_OK;

static void check_prec(int prec) { require(prec == MG_F32 || prec == MG_F64, "prec must be MG_F32 or MG_F64"); }

template <class T> static void step_impl(int mode, void* Y, void* U, void* G, const void* A, const void* F, int64_t n, int dim,
                                         double rho, double inv, double lr, double mom) {
    require(n >= 1 && dim >= 1 && dim <= 64, "arrays must match Y");
    require(Y && U && G && A && (mode == 1 || F), "null array");
    DevBuf part, mb;
    launch_update<T>(mode, (T*)Y, (T*)U, (T*)G, (const T*)A, (const T*)F, nullptr, (T)inv, (T)rho, (T)lr, (T)mom, nullptr, 0,
                     T(0), n, dim, (double*)part.ensure(8 * 1024 * 8), (T*)mb.ensure(std::max(8, dim) * sizeof(T)), nullptr, lib_stream());
    MG_HIP(hipStreamSynchronize(lib_stream()));
}

// The device spectra use the symmetry of the kernel in each offset separately (K(a, b) = K(-a, b) = K(a, -b)),
// which every radial kernel has; a table without it is refused rather than convolved wrongly.
template <class T> static void set_values(Repulsion<T>& r, int M_in, const double* K1, const double* K2) {
    require(M_in >= 1 && K1 && K2, "K1 / K2 must be (M, M) tables");
    for (const double* K : {K1, K2}) {
        double scale = 0, dev = 0;
        for (int a = 0; a < M_in; ++a)
            for (int b = 0; b < M_in; ++b) {
                const double v = K[(size_t)a * M_in + b];
                scale = std::max(scale, std::fabs(v));
                dev = std::max(dev, std::fabs(v - K[(size_t)((M_in - a) % M_in) * M_in + b]));
                dev = std::max(dev, std::fabs(v - K[(size_t)a * M_in + (M_in - b) % M_in]));
            }
        require(dev <= 1e-5 * scale, "kernel tables must be even in each offset (radial kernels)", MG_ERR_UNSUP);
    }
    r.kvals.assign((size_t)2 * M_in * M_in, 0.0);
    std::memcpy(r.kvals.data(), K1, (size_t)M_in * M_in * sizeof(double));
    std::memcpy(r.kvals.data() + (size_t)M_in * M_in, K2, (size_t)M_in * M_in * sizeof(double));
    bool sym = true;
    for (int a = 0; a < M_in && sym; ++a)
        for (int b = 0; b < a && sym; ++b)
            sym = K1[(size_t)a * M_in + b] == K1[(size_t)b * M_in + a] && K2[(size_t)a * M_in + b] == K2[(size_t)b * M_in + a];
    r.kvals_sym = sym;
    r.M_in = M_in; r.tsne_kernel = false; r.have_kernel = true; r.cache.clear();
}

extern "C" {

const char* mg_last_error(void) { return g_err.c_str(); }
int mg_version(void) { return 100; }

int mg_set_stream(void* s) { MG_API_BEGIN g_stream_all = reinterpret_cast<hipStream_t>(s); MG_API_END }
int mg_set_thread_stream(void* s) {
    MG_API_BEGIN t_stream = reinterpret_cast<hipStream_t>(s); t_stream_set = true; MG_API_END
}
int mg_synchronize(void) { MG_API_BEGIN MG_HIP(hipStreamSynchronize(lib_stream())); MG_API_END }
int mg_malloc(void** p, size_t bytes) {
    MG_API_BEGIN
    require(p != nullptr, "null output pointer");
    *p = nullptr;
    MG_HIP(hipMalloc(p, bytes ? bytes : 1));
    MG_API_END
}
int mg_free(void* p) { MG_API_BEGIN if (p) MG_HIP(hipFree(p)); MG_API_END }
int mg_memcpy_htod(void* d, const void* h, size_t b) { MG_API_BEGIN if (b) { MG_HIP(hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, lib_stream())); MG_HIP(hipStreamSynchronize(lib_stream())); } MG_API_END }
int mg_memcpy_dtoh(void* h, const void* d, size_t b) { MG_API_BEGIN if (b) { MG_HIP(hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, lib_stream())); MG_HIP(hipStreamSynchronize(lib_stream())); } MG_API_END }
int mg_memcpy_dtod(void* d, const void* s, size_t b) { MG_API_BEGIN if (b) MG_HIP(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, lib_stream())); MG_API_END }
int mg_mem_info(size_t* f, size_t* t) { MG_API_BEGIN require(f && t, "null output pointer"); MG_HIP(hipMemGetInfo(f, t)); MG_API_END }

int mg_tune_set(const char* name, int value) {
    MG_API_BEGIN
    require(name != nullptr, "null name");
    const std::string k(name);
    if (k == "attr_window") { require(value >= -1 && value <= 1, "attr_window must be -1, 0 or 1"); g_tunables.attr_window = value; }
    else if (k == "attr_rows") { require(value >= 64 && value <= 65536, "attr_rows must be in [64, 65536]"); g_tunables.attr_rows = value; }
    else if (k == "attr_win_kb") { require(value == 0 || value == 64 || value == 128, "attr_win_kb must be 0, 64 or 128"); g_tunables.attr_win_kb = value; }
    else if (k == "attr_tpb") { require(value == 256 || value == 512 || value == 1024, "attr_tpb must be 256, 512 or 1024"); g_tunables.attr_tpb = value; }
    else if (k == "sort_items") { require(value == 0 || value == 1 || value == 2 || value == 4 || value == 8, "sort_items must be 0, 1, 2, 4 or 8"); g_tunables.sort_items = value; }
    else if (k == "sort_digit") { require(value == 8 || value == 10, "sort_digit must be 8 or 10"); g_tunables.sort_digit = value; }
    else if (k == "fft_strided_max") { require(value >= 0 && value <= 8192, "fft_strided_max must be in [0, 8192]"); g_tunables.fft_strided_max = value; }
    else if (k == "fft_fold") { require(value == 0 || value == 2 || value == 4, "fft_fold must be 0, 2 or 4"); g_tunables.fft_fold = value; }
    else if (k == "spec_issue") { require(value >= 0 && value <= 2, "spec_issue must be 0, 1 or 2"); g_tunables.spec_issue = value; }
    else if (k == "fft_fold_persist") { require(value >= 0 && value <= 3, "fft_fold_persist must be in [0, 3]"); g_tunables.fft_fold_persist = value; }
    else if (k == "batch_graph") { require(value == 0 || value == 1, "batch_graph must be 0 or 1"); g_tunables.batch_graph = value; }
    else if (k == "batch_chunk") { require(value >= 1 && value <= 64, "batch_chunk must be in [1, 64]"); g_tunables.batch_chunk = value; }
    else if (k == "batch_window_down") { require(value == 0 || value == 1, "batch_window_down must be 0 or 1"); g_tunables.batch_window_down = value; }
    else if (k == "batch_window_up") { require(value >= 0 && value <= 2, "batch_window_up must be 0, 1 or 2"); g_tunables.batch_window_up = value; }
    else if (k == "batch_wl_grid") { require(value >= 64 && value <= 65536, "batch_wl_grid must be in [64, 65536]"); g_tunables.batch_wl_grid = value; }
    else require(false, "unknown tunable: " + k);
    MG_API_END
}

int mg_tune_get(const char* name, int* value) {
    MG_API_BEGIN
    require(name != nullptr && value != nullptr, "null argument");
    const std::string k(name);
    if (k == "attr_window") *value = g_tunables.attr_window;
    else if (k == "attr_rows") *value = g_tunables.attr_rows;
    else if (k == "attr_win_kb") *value = g_tunables.attr_win_kb;
    else if (k == "attr_tpb") *value = g_tunables.attr_tpb;
    else if (k == "sort_items") *value = g_tunables.sort_items;
    else if (k == "sort_digit") *value = g_tunables.sort_digit;
    else if (k == "fft_strided_max") *value = g_tunables.fft_strided_max;
    else if (k == "fft_fold") *value = g_tunables.fft_fold;
    else if (k == "spec_issue") *value = g_tunables.spec_issue;
    else if (k == "fft_fold_persist") *value = g_tunables.fft_fold_persist;
    else if (k == "batch_graph") *value = g_tunables.batch_graph;
    else if (k == "batch_chunk") *value = g_tunables.batch_chunk;
    else if (k == "batch_window_down") *value = g_tunables.batch_window_down;
    else if (k == "batch_window_up") *value = g_tunables.batch_window_up;
    else if (k == "batch_wl_grid") *value = g_tunables.batch_wl_grid;
    else require(false, "unknown tunable: " + k);
    MG_API_END
}

int mg_profile_enable(int on) {
    MG_API_BEGIN
    g_prof.on = on != 0;
    MG_API_END
}
int mg_profile_reset(void) { MG_API_BEGIN g_prof.reset(); MG_API_END }
int mg_profile_num_stages(void) { return NSTAGE; }
int mg_fft_size(int N) { return N >= 1 ? fft_size(N) : 0; }
int mg_profile_grid(double* mean_N, int* max_N, int* max_M) {
    MG_API_BEGIN
    if (mean_N) *mean_N = g_prof.iters ? g_prof.n_sum / (double)g_prof.iters : 0.0;
    if (max_N) *max_N = g_prof.n_max;
    if (max_M) *max_M = g_prof.m_max;
    MG_API_END
}
int mg_profile_get(int stage, const char** name, double* total_ms, int64_t* count, double* algo_bytes) {
    MG_API_BEGIN
    require(stage >= 0 && stage < NSTAGE, "stage out of range");
    if (name) *name = kStageNames[stage];
    if (total_ms) *total_ms = g_prof.total[stage];
    if (algo_bytes) *algo_bytes = g_prof.bytes[stage];
    if (count) *count = g_prof.count[stage];
    MG_API_END
}

int mg_profile_sized(int stage, int max_sizes, int* n_sizes, int* M, double* total_ms, int64_t* count, double* algo_bytes) {
    MG_API_BEGIN
    require(stage >= 0 && stage < NSTAGE, "stage out of range");
    require(n_sizes != nullptr && (max_sizes <= 0 || (M && total_ms && count && algo_bytes)), "null output pointer");
    int k = 0;
    for (const auto& z : g_prof.sized) {
        if (z.stage != stage) continue;
        if (k < max_sizes) { M[k] = z.M; total_ms[k] = z.total; count[k] = z.count; algo_bytes[k] = z.bytes; }
        ++k;
    }
    *n_sizes = k;
    MG_API_END
}

int mg_bbox_2d(int prec, const void* Y, int64_t n, double* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(Y && out && n >= 1, "Y must be (n, 2) with n >= 1");
    DevBuf part, cnt, res;
    unsigned* c = reinterpret_cast<unsigned*>(cnt.ensure(64));
    MG_HIP(hipMemsetAsync(c, 0, 64, lib_stream()));
    if (prec == MG_F64) {
        launch_bbox<double>((const double*)Y, n, (double*)part.ensure(4 * 1024 * 8), c, (double*)res.ensure(32), lib_stream());
        MG_HIP(hipMemcpyAsync(out, res.p, 32, hipMemcpyDeviceToHost, lib_stream()));
        MG_HIP(hipStreamSynchronize(lib_stream()));
    } else {
        float h[4];
        launch_bbox<float>((const float*)Y, n, (float*)part.ensure(4 * 1024 * 4), c, (float*)res.ensure(16), lib_stream());
        MG_HIP(hipMemcpyAsync(h, res.p, 16, hipMemcpyDeviceToHost, lib_stream()));
        MG_HIP(hipStreamSynchronize(lib_stream()));
        for (int k = 0; k < 4; ++k) out[k] = h[k];
    }
    MG_API_END
}

int mg_tsne_attraction(int prec, const int32_t* ip, const int32_t* ix, const void* data, int64_t nnz, const void* Y,
                       int64_t n, int dim, double dof, void* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(dof > 0, "dof must be positive");
    require(n >= 0 && dim >= 1 && dim <= 64, "dim out of range");
    require(n == 0 || (Y && out && (nnz == 0 || data)), "null array");
    validate_csr(ip, ix, n, nnz, lib_stream());
    if (prec == MG_F64) launch_attraction<double>(ip, ix, (const double*)data, (const double*)Y, n, dim, 1.0 / dof, (double*)out, nnz, lib_stream());
    else launch_attraction<float>(ip, ix, (const float*)data, (const float*)Y, n, dim, 1.f / (float)dof, (float*)out, nnz, lib_stream());
    MG_API_END
}

int mg_tsne_step(int prec, void* Y, void* U, void* G, const void* A, const void* F, int64_t n, int dim, double rho,
                 double inv, double lr, double mom) {
    MG_API_BEGIN
    check_prec(prec);
    if (prec == MG_F64) step_impl<double>(0, Y, U, G, A, F, n, dim, rho, inv, lr, mom);
    else step_impl<float>(0, Y, U, G, A, F, n, dim, rho, inv, lr, mom);
    MG_API_END
}

int mg_tsne_update(int prec, void* Y, void* U, void* G, const void* grad, int64_t n, int dim, double lr, double mom) {
    MG_API_BEGIN
    check_prec(prec);
    if (prec == MG_F64) step_impl<double>(1, Y, U, G, grad, nullptr, n, dim, 0, 0, lr, mom);
    else step_impl<float>(1, Y, U, G, grad, nullptr, n, dim, 0, 0, lr, mom);
    MG_API_END
}

int mg_mid_spring(int prec, const void* Y, const int32_t* pairs, void* grad, int64_t n, int dim, int m, double w) {
    MG_API_BEGIN
    check_prec(prec);
    require(n >= 0 && dim >= 1 && dim <= 64 && m >= 1, "bad shapes");
    require(n == 0 || (Y && pairs && grad), "null array");
    if (prec == MG_F64) launch_mid_spring<double>((const double*)Y, pairs, (double*)grad, n, dim, m, w, lib_stream());
    else launch_mid_spring<float>((const float*)Y, pairs, (float*)grad, n, dim, m, (float)w, lib_stream());
    MG_API_END
}

int mg_fftconv_create(mg_fftconv** out, int prec, int M, int n_interp) {
    MG_API_BEGIN
    check_prec(prec);
    require(out != nullptr, "null output pointer");
    require(M >= 1 && n_interp >= 1 && n_interp <= 8, "bad grid parameters");
    std::unique_ptr<mg_fftconv> c(new mg_fftconv());
    c->prec = prec;
    if (prec == MG_F64) c->f64.reset(new Repulsion<double>(M, n_interp));
    else c->f32.reset(new Repulsion<float>(M, n_interp));
    *out = c.release();
    MG_API_END
}

int mg_fftconv_destroy(mg_fftconv* c) { MG_API_BEGIN delete c; MG_API_END }

int mg_fftconv_set_kernel_values(mg_fftconv* c, int M_in, const double* K1, const double* K2) {
    MG_API_BEGIN
    require(c != nullptr, "null object");
    if (c->prec == MG_F64) set_values(*c->f64, M_in, K1, K2); else set_values(*c->f32, M_in, K1, K2);
    MG_API_END
}

int mg_fftconv_set_kernels_tsne(mg_fftconv* c, double width, double dof) {
    MG_API_BEGIN
    require(c != nullptr, "null object");
    require(width > 0 && dof > 0, "width and dof must be positive");
    auto set = [&](auto& r) { r.tsne_kernel = true; r.have_kernel = true; r.kwidth = width; r.kdof = dof; };
    if (c->prec == MG_F64) set(*c->f64); else set(*c->f32);
    MG_API_END
}

int mg_fftconv_compute(mg_fftconv* c, const void* Y, int64_t n, double xmin, double ymin, double width, int n_boxes,
                       void* F, double* h_sumQ) {
    MG_API_BEGIN
    require(c != nullptr, "null object");
    require(Y && F, "null array");
    require(n_boxes >= 1 && width > 0, "grid larger than the buffer");
    auto run = [&](auto& r, auto tag) {
        typedef decltype(tag) T;
        require(n_boxes * r.ni <= r.M_user, "grid larger than the buffer");
        const GridT<T> g = make_grid<T>(xmin, ymin, width, n_boxes, r.ni);
        r.compute((const T*)Y, n, g, (T*)F, nullptr, lib_stream());
        launch_centroid<T>((const T*)Y, n, reinterpret_cast<double*>(r.centroid.ensure(2 * sizeof(double))), lib_stream());
        double z[2];
        MG_HIP(hipMemcpyAsync(z, r.zout.p, sizeof(z), hipMemcpyDeviceToHost, lib_stream()));
        MG_HIP(hipStreamSynchronize(lib_stream()));
        r.release_dead();
        if (h_sumQ) *h_sumQ = z[0];
    };
    if (c->prec == MG_F64) run(*c->f64, double(0)); else run(*c->f32, float(0));
    MG_API_END
}

int mg_fftconv_grids(mg_fftconv* c, void* d_W, void* d_phi) {
    MG_API_BEGIN
    require(c != nullptr, "null object");
    auto copy = [&](auto& r, auto zero) {
        typedef decltype(zero) T;
        require(r.last_N > 0, "no compute yet");
        const size_t NN = (size_t)r.last_N * r.last_N;
        if (d_W) MG_HIP(hipMemcpyAsync(d_W, r.W.p, 3 * NN * sizeof(T), hipMemcpyDeviceToDevice, lib_stream()));
        if (d_phi) launch_phi_interleave<T>(dptr<T>(r.phi), reinterpret_cast<T*>(d_phi), (i64)NN, lib_stream());
        MG_HIP(hipStreamSynchronize(lib_stream()));
    };
    if (c->prec == MG_F64) copy(*c->f64, double(0)); else copy(*c->f32, float(0));
    MG_API_END
}

int mg_knn(const float* T, int64_t nt, const float* Q, int64_t m, int d, int k, int64_t self_offset, int32_t* idx, float* dist2) {
    MG_API_BEGIN
    static thread_local DevBuf pd, pi;   // scratch per host thread: threads on their own streams run concurrently
    require(m >= 0 && nt >= 1 && (m == 0 || (T && Q && idx && dist2)), "bad k-NN arguments");
    if (m > 0) {
        // m * chunks stays below 2^19 lists of 32 while there is more than one chunk (about 1024 blocks of 256
        // queries): the first call takes room for all of them, so that no later call grows the scratch (a
        // hipFree waits for the whole device, and so for what other threads run beside this one)
        const size_t sz = std::max((size_t)m * knn_chunks(m, nt) * 32, (size_t)1 << 24);
        launch_knn(T, nt, Q, m, d, k, self_offset, reinterpret_cast<float*>(pd.ensure(sz * sizeof(float))),
                   reinterpret_cast<int*>(pi.ensure(sz * sizeof(int))), idx, dist2, lib_stream());
        MG_HIP(hipStreamSynchronize(lib_stream()));
    }
    MG_API_END
}

int mg_knn_noalloc(const float* T, int64_t nt, const float* Q, int64_t m, int d, int k, int64_t self_offset, int32_t* idx,
                   float* dist2) {
    MG_API_BEGIN
    require(m >= 0 && nt >= 1 && (m == 0 || (T && Q && idx && dist2)), "bad k-NN arguments");
    launch_knn_noalloc(T, nt, Q, m, d, k, self_offset, idx, dist2, lib_stream());
    MG_API_END
}

int mg_smooth_knn(const float* dist, int64_t m, int k, int n_neighbors, const float* mean_all, float* vals) {
    MG_API_BEGIN
    require(m == 0 || (dist && mean_all && vals), "null array");
    require(n_neighbors >= 2, "n_neighbors must be >= 2");
    launch_smooth_knn(dist, m, k, n_neighbors, mean_all, vals, lib_stream());
    MG_API_END
}

int mg_perplexity_rows(const float* d2, int64_t n, int k, double perplexity, double tol, int n_iter, double* P, double* beta) {
    MG_API_BEGIN
    require(n >= 0, "n must be >= 0");
    require(n == 0 || (d2 && P), "null array");
    launch_perplexity(d2, n, k, perplexity, tol, n_iter, P, beta, lib_stream());
    MG_API_END
}

int mg_symmetrize_knn(int mode, const int32_t* idx, const double* val, int64_t n, int k, double param, int32_t* indptr,
                      int32_t* out_idx, float* out_val, int64_t cap, int64_t* h_nnz) {
    MG_API_BEGIN
    require(idx && val && indptr && out_idx && out_val && h_nnz, "null array");
    require(n >= 1 && k >= 1, "n and k must be positive");
    DevBuf scratch;
    scratch.ensure((size_t)symmetrize_scratch_ints(n, k) * sizeof(int));
    const i64 nnz = symmetrize_knn(mode, idx, val, n, k, param, indptr, out_idx, out_val, cap, dptr<int>(scratch), lib_stream());
    *h_nnz = nnz < 0 ? -nnz : nnz;
    require(nnz >= 0, "output capacity " + std::to_string(cap) + " < " + std::to_string(-nnz) + " entries");
    MG_HIP(hipStreamSynchronize(lib_stream()));   // the scratch is freed on return
    MG_API_END
}

int mg_nn_descent(const float* X, int64_t n, int d, int k, int max_candidates, int n_iters, float delta, int n_trees,
                  int leaf_size, uint64_t seed, int verbose, int32_t* idx, float* dist2) {
    MG_API_BEGIN
    require(X && idx && dist2, "null array");
    nn_descent(X, n, d, k, max_candidates, n_iters, delta, n_trees, leaf_size, seed, idx, dist2, verbose, lib_stream());
    MG_API_END
}

int mg_mid_near_pairs(const float* X, int64_t n, int d, int n_pairs, int n_cand, uint64_t seed, int32_t* out) {
    MG_API_BEGIN
    require(X && out, "null array");
    mid_near_pairs(X, n, d, n_pairs, n_cand, seed, out, lib_stream());
    MG_API_END
}

int mg_knn_exact_graph(const float* X, int64_t n, int d, int k, int32_t* idx, float* dist2) {
    MG_API_BEGIN
    require(X && idx && dist2, "null array");
    knn_self_exact(X, n, d, k, idx, dist2, lib_stream());
    MG_API_END
}

int mg_topk_rows(const float* D, int64_t rows, int64_t cols, int k, int32_t* idx, float* val) {
    MG_API_BEGIN
    require(rows == 0 || (D && idx && val), "null array");
    launch_topk_rows(D, rows, cols, k, idx, val, lib_stream());
    MG_API_END
}

int mg_tsne_optimize_batch(int prec, const mg_tsne_params* p, int count, const mg_tsne_instance* inst, int n_streams,
                           int n_threads) {
    MG_API_BEGIN
    check_prec(prec);
    require(p != nullptr, "params must not be null");
    if (prec == MG_F64) optimize_batch<double>(p, count, inst, n_streams, n_threads);
    else optimize_batch<float>(p, count, inst, n_streams, n_threads);
    MG_API_END
}

int mg_tsne_batch_stats(int64_t* plans, int64_t* iterations, int64_t* replays, int64_t* launches, int64_t* checks,
                        int64_t* stalls, double* plan_seconds) {
    MG_API_BEGIN
    if (plans) *plans = g_bs_plans.load();
    if (iterations) *iterations = g_bs_iters.load();
    if (stalls) *stalls = g_bs_stalls.load();
    if (replays) *replays = g_bs_replays.load();
    if (launches) *launches = g_bs_launches.load();
    if (checks) *checks = g_bs_checks.load();
    if (plan_seconds) *plan_seconds = 1e-6 * (double)g_bs_plan_us.load();
    MG_API_END
}

int mg_tsne_transform(mg_fftconv* c, const void* Ytr, int64_t n_train, const int32_t* nbr, const void* w, int k, void* Yn,
                      int64_t m, int n_iter, double lr, double mom0, double mom1, int mom_switch, double exaggeration, double dof) {
    MG_API_BEGIN
    require(c != nullptr, "null object");
    require(n_train >= 1 && k >= 1 && m >= 0 && n_iter >= 0, "bad sizes");
    require(Ytr && (m == 0 || (nbr && w && Yn)), "null array");
    require(lr > 0 && dof > 0, "learning_rate and dof must be positive");
    auto run = [&](auto& r, auto zero) {
        typedef decltype(zero) T;
        require(r.last_N > 0, "the field has not been computed: call mg_fftconv_compute on the training layout first");
        double z[2];
        MG_HIP(hipMemcpyAsync(z, r.zout.p, sizeof(z), hipMemcpyDeviceToHost, lib_stream()));
        MG_HIP(hipStreamSynchronize(lib_stream()));
        launch_transform<T>((const T*)Ytr, n_train, nbr, (const T*)w, k, (T*)Yn, m, r.last_g, r.ni, r.centred, dptr<T>(r.phi), z[1],
                            n_iter, lr, mom0, mom1, mom_switch, exaggeration, dof, dptr<double>(r.centroid), lib_stream());
        MG_HIP(hipStreamSynchronize(lib_stream()));
    };
    if (c->prec == MG_F64) run(*c->f64, double(0)); else run(*c->f32, float(0));
    MG_API_END
}

// ---- fallback kernels
int mg_interp_scatter(int prec, const void* Y, int64_t n, const void* q, int C, double xmin, double ymin, double width,
                      int nb, int ni, void* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(n >= 0 && (n == 0 || (Y && q)) && out, "null array");
    require(nb >= 1 && ni >= 1 && ni <= 8, "bad grid parameters");
    DevBuf ws;
    int* w = reinterpret_cast<int*>(ws.ensure(shim_scratch_ints(n, nb) * sizeof(int)));
    if (prec == MG_F64) launch_interp_scatter<double>((const double*)Y, n, (const double*)q, C, xmin, ymin, width, nb, ni, (double*)out, w, lib_stream());
    else launch_interp_scatter<float>((const float*)Y, n, (const float*)q, C, xmin, ymin, width, nb, ni, (float*)out, w, lib_stream());
    MG_HIP(hipStreamSynchronize(lib_stream()));
    MG_API_END
}

int mg_interp_gather(int prec, const void* Y, int64_t n, const void* G, int C, double xmin, double ymin, double width,
                     int nb, int ni, void* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(n >= 0 && (n == 0 || (Y && G && out)), "null array");
    if (prec == MG_F64) launch_interp_gather<double>((const double*)Y, n, (const double*)G, C, xmin, ymin, width, nb, ni, (double*)out, lib_stream());
    else launch_interp_gather<float>((const float*)Y, n, (const float*)G, C, xmin, ymin, width, nb, ni, (float*)out, lib_stream());
    MG_API_END
}

int mg_scatter_padded_2d(int prec, const void* Y, int64_t n, double xmin, double ymin, double width, int nb, int ni, int M,
                         void* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(n >= 0 && (n == 0 || Y) && out, "null array");
    require(nb >= 1 && ni >= 1 && ni <= 8, "bad grid parameters");
    DevBuf ws;
    int* w = reinterpret_cast<int*>(ws.ensure(shim_scratch_ints(n, nb) * sizeof(int)));
    if (prec == MG_F64) launch_scatter_padded<double>((const double*)Y, n, xmin, ymin, width, nb, ni, M, (double*)out, w, lib_stream());
    else launch_scatter_padded<float>((const float*)Y, n, xmin, ymin, width, nb, ni, M, (float*)out, w, lib_stream());
    MG_HIP(hipStreamSynchronize(lib_stream()));
    MG_API_END
}

int mg_spectral_products(int prec, const void* K1, const void* K2, const void* W, int64_t M, int64_t Mh, void* out) {
    MG_API_BEGIN
    check_prec(prec);
    require(M >= 0 && Mh >= 0, "bad shapes");
    require(M * Mh == 0 || (K1 && K2 && W && out), "null array");
    if (prec == MG_F64) launch_spectral_products<double>((const double*)K1, (const double*)K2, (const double*)W, M, Mh, (double*)out, lib_stream());
    else launch_spectral_products<float>((const float*)K1, (const float*)K2, (const float*)W, M, Mh, (float*)out, lib_stream());
    MG_API_END
}

int mg_gather_forces_2d(int prec, const void* Y, int64_t n, const void* phi, int M, double xmin, double ymin, double width,
                        int nb, int ni, void* F, double* h_sumQ) {
    MG_API_BEGIN
    check_prec(prec);
    require(n >= 0 && (n == 0 || (Y && phi && F)), "null array");
    DevBuf part;
    double* p = reinterpret_cast<double*>(part.ensure((size_t)std::max<i64>(1, (n + 255) / 256) * sizeof(double)));
    int nparts;
    if (prec == MG_F64) nparts = launch_gather_forces<double>((const double*)Y, n, (const double*)phi, M, xmin, ymin, width, nb, ni, (double*)F, p, lib_stream());
    else nparts = launch_gather_forces<float>((const float*)Y, n, (const float*)phi, M, xmin, ymin, width, nb, ni, (float*)F, p, lib_stream());
    std::vector<double> h((size_t)std::max(1, nparts), 0.0);
    if (nparts > 0) MG_HIP(hipMemcpyAsync(h.data(), p, (size_t)nparts * sizeof(double), hipMemcpyDeviceToHost, lib_stream()));
    MG_HIP(hipStreamSynchronize(lib_stream()));
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += h[k];
    if (h_sumQ) *h_sumQ = s - (double)n;
    MG_API_END
}

int mg_tsne_optimize(int prec, const mg_tsne_params* p, const int32_t* ip, const int32_t* ix, const void* data, int64_t nnz,
                     int64_t n, void* Y, void* U, void* G, const int32_t* pairs, int m, int it0, int it1, double* h_sumQ,
                     mg_tsne_callback cb, void* user) {
    MG_API_BEGIN
    check_prec(prec);
    require(p != nullptr, "params must not be null");
    require(nnz == 0 || data != nullptr, "data must be a device array");
    check_schedule(p);
    if (prec == MG_F64) {
        Optimizer<double> o(p->n_interp);
        o.run(p, ip, ix, (const double*)data, nnz, n, (double*)Y, (double*)U, (double*)G, pairs, m, it0, it1, h_sumQ, cb, user, lib_stream());
    } else {
        Optimizer<float> o(p->n_interp);
        o.run(p, ip, ix, (const float*)data, nnz, n, (float*)Y, (float*)U, (float*)G, pairs, m, it0, it1, h_sumQ, cb, user, lib_stream());
    }
    MG_API_END
}

int mg_optimize_layout(const mg_layout_params* p, float* Y, int64_t n, int dim, const int32_t* head, const int32_t* tail,
                       const float* eps, float* eons, float* eonns, int64_t m, const int32_t* pairs, int n_mid,
                       const int32_t* rev_ptr, const int32_t* rev_idx) {
    MG_API_BEGIN
    require(p != nullptr, "params must not be null");
    require(dim == 2, "the device sampled layout is 2-D", MG_ERR_UNSUP);
    require(n >= 1 && n < (int64_t)INT32_MAX && Y, "Y must be (n, 2) with n >= 1");
    require(m >= 0 && m < (int64_t)INT32_MAX && (m == 0 || (head && tail && eps && eons && eonns)), "bad edge arrays");
    require(p->epoch_start >= 0 && p->epoch_end >= p->epoch_start, "invalid epoch range");
    require(p->negative_sample_rate > 0.f, "negative_sample_rate must be positive");
    require(p->a > 0.f && p->b > 0.f, "kernel a and b must be positive");
    require(p->n_fixed >= 0 && p->n_fixed <= n, "n_fixed out of range");
    require(n_mid == 0 || (pairs && n_mid >= 1), "mid_pairs must be (n, m)");
    require((rev_ptr == nullptr) == (rev_idx == nullptr), "reverse pairs need indptr and indices");
    require(p->clip_mode == 0 || p->clip_mode == 1, "clip_mode must be 0 (coord) or 1 (norm)");
    DevBuf scratch;
    launch_layout(*p, Y, n, head, tail, eps, eons, eonns, m, pairs, n_mid, rev_ptr, rev_idx,
                  scratch.ensure(layout_scratch_bytes(n)), lib_stream());
    MG_HIP(hipStreamSynchronize(lib_stream()));
    MG_API_END
}

int mg_tsne_session_create(mg_tsne_session** out, int prec, const mg_tsne_params* p, const int32_t* ip, const int32_t* ix,
                           const void* data, int64_t nnz, int64_t n, const int32_t* pairs, int m) {
    MG_API_BEGIN
    check_prec(prec);
    require(out != nullptr, "null output pointer");
    *out = nullptr;
    check_schedule(p);
    require(n >= 1, "n must be >= 1");
    require(nnz == 0 || data != nullptr, "data must be a device array");
    require(!pairs || m >= 1, "pairs must be (n, m) with m >= 1");
    validate_csr(ip, ix, n, nnz, lib_stream());
    std::unique_ptr<mg_tsne_session> s(new mg_tsne_session());
    s->prec = prec; s->p = *p; s->ip = ip; s->ix = ix; s->data = data; s->nnz = nnz; s->n = n; s->pairs = pairs; s->m = m;
    if (prec == MG_F64) s->f64.reset(new Optimizer<double>(p->n_interp));
    else s->f32.reset(new Optimizer<float>(p->n_interp));
    *out = s.release();
    MG_API_END
}

int mg_tsne_session_run(mg_tsne_session* s, void* Y, void* U, void* G, int it0, int it1, double* h_sumQ, mg_tsne_callback cb,
                        void* user) {
    MG_API_BEGIN
    require(s != nullptr, "null session");
    if (s->prec == MG_F64)
        s->f64->run(&s->p, s->ip, s->ix, (const double*)s->data, s->nnz, s->n, (double*)Y, (double*)U, (double*)G, s->pairs, s->m,
                    it0, it1, h_sumQ, cb, user, lib_stream(), false);
    else
        s->f32->run(&s->p, s->ip, s->ix, (const float*)s->data, s->nnz, s->n, (float*)Y, (float*)U, (float*)G, s->pairs, s->m,
                    it0, it1, h_sumQ, cb, user, lib_stream(), false);
    MG_API_END
}

int mg_tsne_session_trace(mg_tsne_session* s, double* h_kl, double* h_grad_norm) {
    MG_API_BEGIN
    require(s != nullptr, "null session");
    if (s->prec == MG_F64) { s->f64->h_kl = h_kl; s->f64->h_gnorm = h_grad_norm; }
    else { s->f32->h_kl = h_kl; s->f32->h_gnorm = h_grad_norm; }
    MG_API_END
}

int mg_tsne_session_destroy(mg_tsne_session* s) { MG_API_BEGIN delete s; MG_API_END }

}  // extern "C"
