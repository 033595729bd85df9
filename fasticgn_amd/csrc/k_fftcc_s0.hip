// Instantiations of the FFTCC kernels for window sizes 2, 10, 18, 26, 34, 42, 50, 58 (split across
// translation units so they compile in parallel).
#include "k_fftcc_impl.h"

namespace fi {
template void launch_fftcc_n<2>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<10>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<18>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<26>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<34>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<42>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<50>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<58>(const FftccParams&, hipStream_t);
}  // namespace fi
