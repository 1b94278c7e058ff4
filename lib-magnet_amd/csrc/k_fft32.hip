// FFT kernels and launchers instantiated for float.
#include "k_fft_impl.hpp"

namespace mg {
MG_FFT_INST(float)
}  // namespace mg
