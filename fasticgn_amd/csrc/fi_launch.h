// Host-side launchers for the gfx950 kernels (defined in k_*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fi_params.h"

namespace fi {

void launch_to_float(const DevImage& img, float* out, hipStream_t s);
void launch_gradient(const float* img, int w, int h, float* gx, float* gy, hipStream_t s);
void launch_bspline(const DevImage& img, float* coef, hipStream_t s);

// LDS bytes of one FFTCC block for radius r (its static allocation).
size_t fftcc_lds_bytes(int r, int mode);
void launch_fftcc(const FftccParams& p, hipStream_t s);

// Largest radius the block-per-POI IC-GN kernels handle.
constexpr int kIcgnMaxRadius = 31;
void launch_icgn_prep(const IcgnParams& p, hipStream_t s);
// Lane shape: waves per POI and rows per band (k_icgn_impl.h).  The launchers
// return false when the shape does not fit the subset of p.rx, p.ry.
bool icgn_shape_supported(int waves, int spl);
bool icgn_shape_covers(int waves, int spl, int rx, int ry);
void icgn_default_shape(int rx, int ry, int& waves, int& spl);
// Every instantiated shape that covers the subset (at most cap); returns the count.
int icgn_shape_candidates(int rx, int ry, int* waves, int* spl, int cap);
bool launch_icgn_shape(const IcgnParams& p, int waves, int spl, hipStream_t s);
// Resident blocks per CU of the persistent-queue launch of shape (waves, spl)
// for p's radius, image size and sampling (p.tar null = table); 0 if the shape
// does not cover the radius.
int icgn_queue_per_cu(const IcgnParams& p, int waves, int spl);
bool launch_icgn(const IcgnParams& p, hipStream_t s);
// Whether the two-entries-per-wave kernel (IcgnParams::group == 2) runs
// uniform subsets of this radius; launches it does not run take the shared-
// tile kernel.
bool icgn_pair_supported(int rx, int ry);

// Frame-policy helpers (k_policy.hip).
struct FrameCounters {
    unsigned long long bad;          // isBadPoiForFallback
    unsigned long long valid;        // finite zncc >= 0
    unsigned long long iter_sum;     // sum of finite iteration counts
    unsigned long long iter_finite;  // POIs with a finite iteration count
    unsigned long long recovered;    // repaired POIs that are no longer bad
    unsigned int iter_max;
    unsigned int candidates;         // POI-level repair candidates written
    unsigned long long iterations;   // the frame's executed iterations, copied by k_qa
    // set by k_decide on the device (temporal frames):
    unsigned int repair_n;           // per class: candidates to repair (0 under a whole-frame fallback)
    unsigned int fallback_n;         // per class: POIs the whole-frame fallback re-runs (0 or the class size)
    unsigned int fallback;           // frame entry: 1 when the whole-frame fallback ran
    unsigned int temporal;           // frame entry: seeds set by temporal extrapolation
};

}  // namespace fi
