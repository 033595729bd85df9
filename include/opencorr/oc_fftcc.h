// oc_fftcc.h -- the reference's header name for the OpenCorr-shaped API of this
// library (src/oc_fftcc.h): callers written against OpenCorr include it unchanged
// with -I include/opencorr.
#pragma once
#include "../fasticgn_opencorr.hpp"
