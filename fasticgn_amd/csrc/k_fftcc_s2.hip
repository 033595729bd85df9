// Instantiations of the FFTCC kernels for window sizes 6, 14, 22, 30, 38, 46, 54, 62 (split across
// translation units so they compile in parallel).
#include "k_fftcc_impl.h"

namespace fi {
template void launch_fftcc_n<6>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<14>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<22>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<30>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<38>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<46>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<54>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<62>(const FftccParams&, hipStream_t);
}  // namespace fi
