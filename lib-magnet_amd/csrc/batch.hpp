// Batched launches: one launch of a stage covers the work of many independent embeddings (instances).
// A launch's blocks are split into consecutive ranges, one per instance (BMap, passed by value); a block finds
// its instance and its block index within that instance, and runs the same block body the single-instance
// launch runs for that block index. Each instance's arithmetic, block decomposition and summation order are
// those of its single-instance launch, so a batched run gives each instance the single run's result bit for bit.
//
// The batch is device-driven: at the start of every iteration a decision kernel derives each instance's grid
// from the bounding box its last update wrote, and every later kernel reads the grid from the table. Launch
// shapes depend only on a plan the host makes now and then: each instance's window of FFT sizes [Mlo, Mhi],
// whose largest grid sizes its buffers and its block ranges; blocks beyond the live grid exit at once. The
// launches of one iteration therefore do not change from one iteration to the next, and are captured once in
// a hipGraph and replayed.
#pragma once
#include "common.hpp"

namespace mg {

constexpr int BMAX = 64;   // instances per batched launch

struct BMap {
    int count;                 // ranges in use
    int start[BMAX + 1];       // blocks [start[k], start[k + 1]) belong to instance inst[k]
    int inst[BMAX];
};

// range of block b (uniform over the block): binary search over the kernel argument
__device__ __forceinline__ int bmap_find(const BMap& m, int b) {
    int lo = 0, hi = m.count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (m.start[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Per instance, in device memory (one table for the batch). Point arrays are concatenated: instance i owns rows
// [off, off + n) of Y, U, G, the attraction and the repulsion forces, and positions [off, off + n) of the sorted
// key / index arrays. The host writes the fixed part when it plans the launches; the decision kernel (kb_decide)
// writes the per-iteration part at the start of every iteration.
template <class T> struct BInst {
    // ---- fixed for a plan
    const int32_t* ip;          // CSR of the instance (local column indices)
    const int32_t* ix;
    const T* P;
    i64 n, off;
    T lr;
    int Mlo, Mhi;               // the plan's window of FFT sizes; the grid buffers hold the largest grid of Mhi
    int* item_key;              // spread chunk items (2 per chunk)
    int* item_valid;
    T* item_val;
    T* W;                       // (3, N, N)
    cplx<T>* Wh;                // (3, Mh, N)
    cplx<T>* Ph;                // (4, Mh, N)
    cplx<T>* Tr;                // (4, N, Mh)
    T* phi;                     // (4, N, N)
    int KM[4];                  // the window's sizes and their kernel spectra at the unclipped cell width (tables shared
    const T* KS[4][4];          // by every instance of the batch)
    T* Kw[4];                   // the instance's own spectra of a clipped grid (the width varies), at the largest
    cplx<T>* spec_tmp;          // size only, with their staging (4 Mh^2); null when its window cannot clip
    double* zpart;              // gather partials, sumQ and 1/sumQ
    double* zout;
    T* inv_out;
    double* upart;              // update partials, {mean, bbox}
    T* mb;
    double* sumq_base;          // sumQ of every iteration (n_iter)
    // ---- this iteration (kb_decide)
    GridT<T> g;                 // grid, FFT size M
    int M;
    const T* K[4];              // the spectra of (M, delta)
    int active;                 // 0: the instance has finished, or its grid needs a size outside [Mlo, Mhi]
    int fresh;                  // the spectra of (M, delta) are to be computed into Kw
    T rho, mom;                 // exaggeration and momentum of the iteration
    double delta;
    double* sumq;               // this iteration's sumQ slot
};

// Per instance, in device memory, carried from iteration to iteration and from plan to plan.
struct BState {
    int next_it;                // the iteration the instance runs next
    int M;                      // FFT size of the last decided grid
    int stall;                  // 1: that size lies outside the plan's window (the instance waits for a new plan);
                                // 2: the layout is not finite
    int sM;                     // (sM, sdelta): the spectra the instance's own tables (Kw) hold (sM 0: none)
    double sdelta;
};

// The schedule as the decision kernel needs it
struct BSched {
    int n_iter, early_iter, ni, min_boxes, max_boxes, count;
    double box_width, rho0, rho1, mom0, mom1;
    double delta0;              // box_width / ni: the cell width of every grid that does not clip
};

// ---- work lists: the live blocks of a grid stage, built on the device after the decision (kb_worklists). A list is
// the ints [total, n, cap, inst[cap], start[cap + 1]]: n instances with live work, instance inst[k] owning the
// virtual blocks [start[k], start[k + 1]), total = start[n]. The stage's launch has a fixed grid whose blocks stride
// over the virtual blocks, so its shape does not depend on the grids, and no block is dispatched for nothing.
enum WorkKind { WK_ROWS3 = 0, WK_COLS = 1, WK_ROWS4 = 2, WK_TR0 = 3, WK_TR1 = 4 };
struct WorkDesc { int kind, M, cap, off, cand; };   // cand: offset of the list's candidate instances
inline int wl_ints(int cap) { return 3 + 2 * cap + 1; }
__device__ __forceinline__ int wl_find(const int* wl, int v) {
    const int n = wl[1];
    const int* start = wl + 3 + wl[2];
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int wl_inst(const int* wl, int k) { return wl[3 + k]; }
__device__ __forceinline__ int wl_start(const int* wl, int k) { return wl[3 + wl[2] + k]; }
template <class T> void launch_worklists(const BInst<T>* I, const WorkDesc* d, int nd, const int* cand, int* wl, hipStream_t st);

// ---- batched launchers (all on `st`; I is the device table, m the block ranges)
int update_blocks(i64 n);
template <class T> void launch_attraction_batch(const BInst<T>* I, const BMap& m, int G, const T* Y, T inv_dof, T* out,
                                                hipStream_t st);
int attraction_lanes(i64 n, i64 nnz);                          // lanes per CSR row of the 2-D attraction
inline int attraction_blocks(i64 n, int G) { return (int)((n + 256 / G - 1) / (256 / G)); }
// the fused step and recentring of every active instance (rho, mom, lr from the table)
template <class T> void launch_update_batch(const BInst<T>* I, const BMap& m, T* Y, T* U, T* G, const T* A, const T* F,
                                            hipStream_t st);
// the grid of every instance from the bbox its last update wrote, the FFT size check against the plan's window, and
// the iteration's schedule values (one thread per instance)
template <class T> void launch_decide_batch(BInst<T>* I, BState* S, const BSched& q, hipStream_t st);
// clears each active instance's charge grid W (3 N^2 values, 4096 per virtual block of the list; grid blocks)
// keys (inst << box_bits | box key) and global point indices of every point (m: 256-point blocks); the points of
// an inactive instance all take box 0
template <class T> void launch_keys_batch(const BInst<T>* I, const BMap& m, const T* Y, int box_bits, int* keys, int* vals,
                                          hipStream_t st);
inline int key_blocks(i64 n) { return (int)((n + 255) / 256); }
// stable sort of keys (< 2^bits) with their values, both in A; the result's keys keep only their low
// `keep_bits` bits; returns the buffer with the result (0 = A, 1 = B)
int launch_key_sort_masked(int* keysA, int* valsA, int* keysB, int* valsB, int* hist, i64 n, int bits, int keep_bits,
                           hipStream_t st);
template <class T> void launch_spread_batch(const BInst<T>* I, const BMap& m, const int* keys, const int* vals, const T* Y,
                                            int ni, bool centred, hipStream_t st);
template <class T> void launch_gather_batch(const BInst<T>* I, const BMap& m, const int* keys, const int* vals, const T* Y,
                                            int ni, bool centred, T* F, hipStream_t st);
// FFT stages of the instances at FFT size M, over the work list wl (WK_ROWS3, WK_COLS, WK_ROWS4 of size M), grid
// blocks striding over its virtual blocks (each the single launch's block of the same index)
template <class T> void launch_rows_forward_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st);
template <class T> void launch_columns_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st);
template <class T> void launch_rows_inverse_wl(const BInst<T>* I, const int* wl, int M, int grid, hipStream_t st);
// which: 0 = Tr (3, N, Mh) -> Wh (3, Mh, N) (list WK_TR0); 1 = Ph (4, Mh, N) -> Tr (4, N, Mh) (WK_TR1)
template <class T> void launch_transpose_wl(const BInst<T>* I, const int* wl, int which, int grid, hipStream_t st);
// spectra of the instances at size M whose table is fresh: stage 1 (m1: fft_col_blocks(M)), transpose (mt:
// transpose_blocks(Mh, Mh, 1)), stage 2 (ntab 2: m1; ntab 4, by pairs of columns: m2, spec_pair_blocks(M), then the
// transposed table, mr: transpose_blocks(Mh, Mh, 1))
template <class T> void launch_spectra_batch(const BInst<T>* I, const BMap& m1, const BMap& mt, const BMap& m2, const BMap& mr,
                                             int M, double dof, bool round_kernel, int ntab, hipStream_t st);
int spec_pair_blocks(int M);
template <class T> void prepare_fft_kernels();   // every FFT kernel's LDS limit, set before threads start
int fft_row_blocks(int N, int M, int channels);
int fft_col_blocks(int M);
inline int transpose_blocks(int R, int C, int batch) { return ((C + 31) / 32) * ((R + 31) / 32) * batch; }
// the transposes of the grid passes over the largest grid of Mhi (which 0: 3 channels, 1: 4)
inline int transpose_cap_blocks(int Mhi, int which) {
    return which == 0 ? transpose_blocks(Mhi / 2, Mhi / 2 + 1, 3) : transpose_blocks(Mhi / 2 + 1, Mhi / 2, 4);
}

}  // namespace mg
