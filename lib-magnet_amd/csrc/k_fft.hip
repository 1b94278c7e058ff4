// Host side of the FFT stages: twiddle tables per device and the supported transform sizes. The kernels and
// their launchers live in k_fft_impl.hpp, compiled once per precision (k_fft32.hip, k_fft64.hip).
#include "common.hpp"
#include "fft_core.hpp"
#include <cmath>
#include <mutex>
#include <vector>

namespace mg {

namespace {
struct Twiddles {
    cplx<float>* hi32 = nullptr; cplx<float>* lo32 = nullptr;
    cplx<double>* hi64 = nullptr; cplx<double>* lo64 = nullptr;
};
Twiddles g_tw[64];
std::mutex g_tw_mu;
int current_device() { int d = 0; MG_HIP(hipGetDevice(&d)); return d; }
}  // namespace

// exp(-2 pi i m / TW_PERIOD) for m = 64 a (hi) and m = b (lo), from long double arithmetic
void fft_init_twiddles() {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    const int dev = current_device();
    require(dev >= 0 && dev < 64, "device index out of range");
    if (g_tw[dev].hi64) return;
    const long double pi = 3.141592653589793238462643383279502884L;
    auto entry = [&](int m) {
        const long double ang = -2.0L * pi * (long double)m / (long double)TW_PERIOD;
        return cplx<double>{(double)cosl(ang), (double)sinl(ang)};
    };
    std::vector<cplx<double>> hi64(TW_HI), lo64(TW_LO);
    std::vector<cplx<float>> hi32(TW_HI), lo32(TW_LO);
    for (int a = 0; a < TW_HI; ++a) { hi64[a] = entry(64 * a); hi32[a] = {(float)hi64[a].x, (float)hi64[a].y}; }
    for (int b = 0; b < TW_LO; ++b) { lo64[b] = entry(b); lo32[b] = {(float)lo64[b].x, (float)lo64[b].y}; }
    Twiddles t;
    MG_HIP(hipMalloc(&t.hi64, TW_HI * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&t.lo64, TW_LO * sizeof(cplx<double>)));
    MG_HIP(hipMalloc(&t.hi32, TW_HI * sizeof(cplx<float>)));
    MG_HIP(hipMalloc(&t.lo32, TW_LO * sizeof(cplx<float>)));
    MG_HIP(hipMemcpy(t.hi64, hi64.data(), TW_HI * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(t.lo64, lo64.data(), TW_LO * sizeof(cplx<double>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(t.hi32, hi32.data(), TW_HI * sizeof(cplx<float>), hipMemcpyHostToDevice));
    MG_HIP(hipMemcpy(t.lo32, lo32.data(), TW_LO * sizeof(cplx<float>), hipMemcpyHostToDevice));
    g_tw[dev] = t;
}

template <> TwPtr<float> fft_twiddles<float>() { const Twiddles& t = g_tw[current_device()]; return {t.hi32, t.lo32}; }
template <> TwPtr<double> fft_twiddles<double>() { const Twiddles& t = g_tw[current_device()]; return {t.hi64, t.lo64}; }

// the transform sizes with compiled plans: 2^s and 3 * 2^s from 128 to 8192
int fft_size(int N) {
    for (int M = 128; M <= MAX_M; M *= 2) {
        if (3 * (M / 4) >= 2 * N && 3 * (M / 4) >= 128) return 3 * (M / 4);
        if (M >= 2 * N) return M;
    }
    return 0;
}

}  // namespace mg
