// Host side of the FFT stages: twiddle tables per device and the supported transform sizes. The kernels and
// their launchers live in k_fft_impl.hpp, compiled once per precision (k_fft32.hip, k_fft64.hip).
#include "common.hpp"
#include "fft_core.hpp"
#include "batch.hpp"
#include <cmath>
#include <mutex>
#include <atomic>
#include <vector>

namespace mg {

namespace {

#define MG_ALL_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096) X(6144) X(8192)
constexpr int kSizes[] = {64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192};
constexpr int kNumSizes = sizeof(kSizes) / sizeof(kSizes[0]);
int size_index(int L) {
    for (int i = 0; i < kNumSizes; ++i) if (kSizes[i] == L) return i;
    throw Error(ERR_UNSUP, "FFT size not supported");
}

// exp(-2 pi i m / n), from long double with the angle reduced to the first octant
cplx<double> root(long long m, long long n) {
    m %= n;
    if (m < 0) m += n;
    const long double pi = 3.141592653589793238462643383279502884L;
    // angle 2 pi m / n = quarter turns q plus a remainder; cos / sin of the remainder, rotated
    const long long q = (4 * m) / n;
    const long long r4 = 4 * m - q * n;   // remainder in units of 1/(4n) turns: angle 2 pi r4 / (4n)
    const long double a = 2.0L * pi * (long double)r4 / (4.0L * (long double)n);
    long double c = cosl(a), s = sinl(a);
    for (long long i = 0; i < q; ++i) { const long double t = c; c = -s; s = t; }   // + quarter turn
    return {(double)c, -(double)s};
}

template <class Pl> void plan_tables(std::vector<cplx<double>>& out) {
    for (int I = 1; I < Pl::P; ++I) {
        const int Ns = plan_ns<Pl>(I), R = Pl::R[I], NB = log2c(R);
        for (int b = 0; b < NB; ++b)
            for (int k = 0; k < Ns; ++k) out.push_back(root((long long)k << b, (long long)Ns * R));
    }
}

struct Tables {
    bool ready = false;
    cplx<float>* f32 = nullptr;
    cplx<double>* f64 = nullptr;
    size_t off_fwd[kNumSizes], off_rev[kNumSizes], off_half[kNumSizes];
};
Tables g_tw[64];
std::mutex g_tw_mu;
int current_device() { int d = 0; MG_HIP(hipGetDevice(&d)); return d; }

}  // namespace

void fft_init_twiddles() {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    const int dev = current_device();
    require(dev >= 0 && dev < 64, "device index out of range");
    Tables& t = g_tw[dev];
    if (t.ready) return;
    std::vector<cplx<double>> h;
#define MG_T(L_) { \
        const int i = size_index(L_); \
        t.off_fwd[i] = h.size(); plan_tables<FwdPlan<double, L_>::type>(h); \
        t.off_rev[i] = h.size(); plan_tables<Rev<FwdPlan<double, L_>::type>>(h); \
        t.off_half[i] = h.size(); for (int k = 0; k <= L_; ++k) h.push_back(root(k, 2 * L_)); }
    MG_ALL_SIZES(MG_T)
#undef MG_T
    std::vector<cplx<float>> hf(h.size());
    for (size_t i = 0; i < h.size(); ++i) hf[i] = {(float)h[i].x, (float)h[i].y};
    MG_HIP(hipMalloc(&t.f64, h.size() * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&t.f32, h.size() * sizeof(cplx<float>)));
    MG_HIP(hipMemcpy(t.f64, h.data(), h.size() * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(t.f32, hf.data(), hf.size() * sizeof(cplx<float>), hipMemcpyHostToDevice));
    t.ready = true;
}

template <class T> static TwSet<T> set_of(const Tables& t, const cplx<T>* base, int L) {
    const int i = size_index(L);
    return {base + t.off_fwd[i], base + t.off_rev[i], base + t.off_half[i]};
}
template <> TwSet<float> fft_twiddle_set<float>(int L) { const Tables& t = g_tw[current_device()]; return set_of<float>(t, t.f32, L); }
template <> TwSet<double> fft_twiddle_set<double>(int L) { const Tables& t = g_tw[current_device()]; return set_of<double>(t, t.f64, L); }

int device_cus() {
    static std::atomic<int> cus[64] = {};   // per device, read once
    const int dev = current_device();
    if (dev < 0 || dev >= 64) return 256;
    int c = cus[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        int n = 0;
        MG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        c = n > 0 ? n : 256;
        cus[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

void kernel_lds_limit(const void* f, int bytes) {
    struct Set { int dev; const void* f; int bytes; };
    static std::mutex mu;
    static std::vector<Set> done;
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(mu);
    for (auto& d : done)
        if (d.dev == dev && d.f == f) {
            if (d.bytes >= bytes) return;
            MG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            d.bytes = bytes;
            return;
        }
    MG_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done.push_back({dev, f, bytes});
}

// blocks of the batched row and column passes (transforms per workgroup as in fft_batch)
static int batch_of(int L) { (void)size_index(L); return L / 16 >= 256 ? 1 : 256 / (L / 16); }
int fft_row_blocks(int N, int M, int channels) { const int B = batch_of(M / 2); return (N + B - 1) / B * channels; }
int fft_col_blocks(int M) { const int B = batch_of(M); return (M / 2 + 1 + B - 1) / B; }
int spec_pair_blocks(int M) { const int B = batch_of(M), np = (M / 2 + 2) / 2; return (np + B - 1) / B; }

}  // namespace mg
