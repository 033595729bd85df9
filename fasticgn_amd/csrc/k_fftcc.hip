// FFTCC launch dispatch over the window size n = 2r (kernels in
// k_fftcc_impl.h, instantiated in k_fftcc_s*.hip).
#include <hip/hip_runtime.h>

#include "fi_launch.h"

namespace fi {

template <int N>
void launch_fftcc_n(const FftccParams& p, hipStream_t s);

size_t fftcc_lds_bytes(int r, int mode)
{
    (void)mode;
    const size_t n = 2 * (size_t)r, hc = r + 1;
    const size_t G = hc <= 16 ? 16 : (hc <= 32 ? 32 : 64);
    const size_t ppb = 64 / G;
    return ppb * (n * hc * 2 * sizeof(float) + n * sizeof(float));
}

void launch_fftcc(const FftccParams& p, hipStream_t s)
{
    if (p.n_pois <= 0) return;
    switch (2 * p.r) {
#define FI_N(n) \
    case n: launch_fftcc_n<n>(p, s); break;
        FI_N(2) FI_N(4) FI_N(6) FI_N(8) FI_N(10) FI_N(12) FI_N(14) FI_N(16)
        FI_N(18) FI_N(20) FI_N(22) FI_N(24) FI_N(26) FI_N(28) FI_N(30) FI_N(32)
        FI_N(34) FI_N(36) FI_N(38) FI_N(40) FI_N(42) FI_N(44) FI_N(46) FI_N(48)
        FI_N(50) FI_N(52) FI_N(54) FI_N(56) FI_N(58) FI_N(60) FI_N(62) FI_N(64)
#undef FI_N
    default: break;
    }
}

}  // namespace fi
