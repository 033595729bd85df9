// Instantiations of the FFTCC kernels for window sizes 8, 16, 24, 32, 40, 48, 56, 64 (split across
// translation units so they compile in parallel).
#include "k_fftcc_impl.h"

namespace fi {
template void launch_fftcc_n<8>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<16>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<24>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<32>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<40>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<48>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<56>(const FftccParams&, hipStream_t);
template void launch_fftcc_n<64>(const FftccParams&, hipStream_t);
}  // namespace fi
