// Instantiations of the FFTCC kernels for window sizes 4, 12, 20, 28, 36, 44, 52, 60 (split across
// translation units so they compile in parallel).
#include "k_fftcc_impl.h"

namespace fi {
template void launch_fftcc_n<4>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<12>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<20>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<28>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<36>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<44>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<52>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<60>(const FftccParams&, hipStream_t);
}  // namespace fi
