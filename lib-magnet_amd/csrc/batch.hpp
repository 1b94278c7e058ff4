// Batched launches: one launch of a stage covers the work of many independent embeddings (instances).
// A launch's blocks are split into consecutive ranges, one per instance (BMap, passed by value); a block finds
// its instance and its block index within that instance, and runs the same block body the single-instance
// launch runs for that block index. Each instance's arithmetic, block decomposition and summation order are
// those of its single-instance launch, so a batched run gives each instance the single run's result bit for bit.
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

// Per instance and iteration, in device memory (one table for the batch). Point arrays are concatenated:
// instance i owns rows [off, off + n) of Y, U, G, the attraction and the repulsion forces, and positions
// [off, off + n) of the sorted key / index arrays.
template <class T> struct BInst {
    const int32_t* ip;          // CSR of the instance (local column indices)
    const int32_t* ix;
    const T* P;
    i64 n, off;
    T lr;
    GridT<T> g;                 // this iteration's grid, FFT size M
    int M;
    int* item_key;              // spread chunk items (2 per chunk)
    int* item_valid;
    T* item_val;
    T* W;                       // (3, N, N)
    cplx<T>* Wh;                // (3, Mh, N)
    cplx<T>* Ph;                // (4, Mh, N)
    cplx<T>* Tr;                // (4, N, Mh)
    T* phi;                     // (4, N, N)
    const T* K[4];              // kernel spectra of (M, delta)
    double* zpart;              // gather partials, sumQ and 1/sumQ
    double* zout;
    T* inv_out;
    double* sumq;               // this iteration's sumQ slot
    double* upart;              // update partials, {mean, bbox}
    T* mb;
    double delta;               // spectra to (re)compute: cell width delta into Kw, through spec_tmp
    cplx<T>* spec_tmp;
    T* Kw[4];
};

// ---- batched launchers (all on `st`; I is the device table, m the block ranges)
int update_blocks(i64 n);
template <class T> void launch_attraction_batch(const BInst<T>* I, const BMap& m, int G, const T* Y, T inv_dof, T* out,
                                                hipStream_t st);
int attraction_lanes(i64 n, i64 nnz);                          // lanes per CSR row of the 2-D attraction
inline int attraction_blocks(i64 n, int G) { return (int)((n + 256 / G - 1) / (256 / G)); }
template <class T> void launch_update_batch(const BInst<T>* I, const BMap& m, T* Y, T* U, T* G, const T* A, const T* F,
                                            T rho, T mom, hipStream_t st);
// keys (inst << box_bits | box key) and global point indices of every point (m: 256-point blocks)
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
// FFT stages over instances of one FFT size M (rows: blocks per instance fft_row_blocks, columns fft_col_blocks)
template <class T> void launch_rows_forward_batch(const BInst<T>* I, const BMap& m, int M, hipStream_t st);
template <class T> void launch_columns_batch(const BInst<T>* I, const BMap& m, int M, hipStream_t st);
template <class T> void launch_rows_inverse_batch(const BInst<T>* I, const BMap& m, int M, hipStream_t st);
// which: 0 = Tr (3, N, Mh) -> Wh (3, Mh, N); 1 = Ph (4, Mh, N) -> Tr (4, N, Mh); 2 / 3 = spectra stage 1 -> 2
// (spec_tmp, ntab 2 / 4); blocks per instance from transpose_blocks
template <class T> void launch_transpose_batch(const BInst<T>* I, const BMap& m, int which, hipStream_t st);
template <class T> void launch_spectra_batch(const BInst<T>* I, const BMap& m1, const BMap& mt, int M, double dof,
                                             bool round_kernel, int ntab, hipStream_t st);
int fft_row_blocks(int N, int M, int channels);
int fft_col_blocks(int M);
inline int transpose_blocks(int R, int C, int batch) { return ((C + 31) / 32) * ((R + 31) / 32) * batch; }

}  // namespace mg
