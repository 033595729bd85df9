// Exact-shape IC-GN kernels (FI_ICGN_EXACT_B); split from k_icgn.hip so the
// instantiations compile in parallel.
#include "k_icgn_impl.h"

namespace fi {

#define X(W, S)                                                                 \
    template void launch_icgn_t<W, S, true>(const IcgnParams&, hipStream_t);     \
    template int icgn_queue_per_cu_t<W, S, true>(const IcgnParams&);
FI_ICGN_EXACT_B(X)
#undef X

}  // namespace fi
