// Two-entries-per-wave IC-GN kernels (FI_ICGN_PAIR); split from k_icgn.hip so
// the instantiations compile in parallel.
#include "k_icgn_impl.h"

namespace fi {

#define X(S) template void launch_icgn_pair_t<S>(const IcgnParams&, hipStream_t);
FI_ICGN_PAIR(X)
#undef X

bool icgn_pair_supported(int rx, int ry)
{
    const int R = pair_rows(rx, ry);
#define X(S) if (R == S) return true;
    FI_ICGN_PAIR(X)
#undef X
    return false;
}

}  // namespace fi
