// FFT kernels and launchers instantiated for double.
#include "k_fft_impl.hpp"

namespace mg {
MG_FFT_INST(double)
}  // namespace mg
