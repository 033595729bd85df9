// Device-side frame policy of the multi-frame driver: temporal seeding, the
// bad-POI scan with repair-candidate collection, the per-POI merges and the
// frame QA counters.  The predicates restate opencorr_cli_cuda_mvp.cpp:483-558
// and seedPoiQueueFromPrevious (:560-624, temporal extrapolation off); they
// run where the POIs already live so a frame needs no host scan.
#include <hip/hip_runtime.h>

#include "fi_host.h"

namespace fi {

constexpr float kGoodZncc = 0.8f;  // kSkipFftccZnccThreshold, :479

__device__ __forceinline__ bool finite_f(float v) { return isfinite(v); }

// isBadPoiForFallback (:493-497)
__device__ __forceinline__ bool is_bad(const fi_poi& p)
{
    const float z = p.result[2];
    return z != -3.f && (!finite_f(z) || z < kGoodZncc);
}

// shouldReplacePoiWithFallback (:505-517)
__device__ __forceinline__ bool should_replace(const fi_poi& cur, const fi_poi& fb)
{
    const float zf = fb.result[2], zc = cur.result[2];
    if (!finite_f(zf) || zf == -3.f) return false;
    if (!finite_f(zc)) return true;
    return zf >= zc;
}

// shouldTakeFrameFallbackResult (:527-539)
__device__ __forceinline__ bool take_frame_fallback(const fi_poi& skip, const fi_poi& scratch)
{
    const float zs = scratch.result[2];
    if (skip.result[2] == -3.f) return finite_f(zs) && zs != -3.f && zs >= kGoodZncc;
    if (is_bad(skip)) return should_replace(skip, scratch);
    return false;
}

// isPoiLevelFallbackCandidate (:541-558)
__device__ __forceinline__ bool is_candidate(const fi_poi& cur, const fi_poi* prev, bool bad)
{
    if (!bad) return false;
    const float z = cur.result[2];
    if (!finite_f(z) || z < 0.f) return true;
    if (prev == nullptr || !(cur.x == prev->x && cur.y == prev->y) || !finite_f(prev->result[2]))
        return true;
    return prev->result[2] >= kGoodZncc;
}

__global__ void k_seed(const fi_poi* __restrict__ tmpl, const fi_poi* __restrict__ prev,
                       fi_poi* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fi_poi q = tmpl[i];
    const fi_poi& pv = prev[i];
    if (finite_f(pv.result[2]) && pv.result[2] >= 0.f) {
#pragma unroll
        for (int k = 0; k < 12; k++) q.deformation[k] = pv.deformation[k];
#pragma unroll
        for (int k = 0; k < 6; k++) q.result[k] = 0.f;
#pragma unroll
        for (int k = 0; k < 3; k++) q.strain[k] = 0.f;
    }
    out[i] = q;
}

__global__ void k_scan(const fi_poi* __restrict__ cur, const fi_poi* __restrict__ prev, int n,
                       int collect, FrameCounters* counters, uint32_t* candidates)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool bad = is_bad(cur[i]);
    if (bad) atomicAdd(&counters->bad, 1ull);
    if (collect && is_candidate(cur[i], prev ? prev + i : nullptr, bad)) {
        const unsigned slot = atomicAdd(&counters->candidates, 1u);
        candidates[slot] = (uint32_t)i;
    }
}

__global__ void k_gather(const fi_poi* __restrict__ src, const uint32_t* __restrict__ idx, int n,
                         fi_poi* __restrict__ dst)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[idx[j]];
}

__global__ void k_merge_frame(const fi_poi* __restrict__ skip, fi_poi* __restrict__ scratch, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!take_frame_fallback(skip[i], scratch[i])) scratch[i] = skip[i];
}

__global__ void k_merge_repair(fi_poi* __restrict__ cur, const fi_poi* __restrict__ rep,
                               const uint32_t* __restrict__ idx, int n, FrameCounters* counters)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = idx[j];
    if (should_replace(cur[i], rep[j])) {
        const bool recovered = !is_bad(rep[j]);
        cur[i] = rep[j];
        if (recovered) atomicAdd(&counters->recovered, 1ull);
    }
}

// computeFrameQa (:629-700): valid = finite zncc >= 0; iteration statistics
// over finite iteration counts; bad per isBadPoiForFallback.
__global__ void k_qa(const fi_poi* __restrict__ cur, int n, FrameCounters* counters)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fi_poi& p = cur[i];
    const float z = p.result[2];
    if (finite_f(z) && z >= 0.f) atomicAdd(&counters->valid, 1ull);
    if (is_bad(p)) atomicAdd(&counters->bad, 1ull);
    const float it = p.result[3];
    if (finite_f(it)) {
        atomicAdd(&counters->iter_finite, 1ull);
        atomicAdd(&counters->iter_sum, (unsigned long long)(long long)it);
        atomicMax(&counters->iter_max, (unsigned)(int)(it + 0.5f));
    }
}

static dim3 grid_for(int n) { return dim3((n + 255) / 256); }

void launch_seed_from_previous(const fi_poi* tmpl, const fi_poi* prev, fi_poi* out, int n,
                               hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_seed, grid_for(n), dim3(256), 0, s, tmpl, prev, out, n);
}

void launch_scan(const fi_poi* cur, const fi_poi* prev, int n, int collect,
                 FrameCounters* counters, uint32_t* candidates, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_scan, grid_for(n), dim3(256), 0, s, cur, prev, n, collect, counters,
                           candidates);
}

void launch_gather(const fi_poi* src, const uint32_t* idx, int n, fi_poi* dst, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_gather, grid_for(n), dim3(256), 0, s, src, idx, n, dst);
}

void launch_merge_frame_fallback(const fi_poi* skip, fi_poi* scratch_inout, int n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_merge_frame, grid_for(n), dim3(256), 0, s, skip, scratch_inout, n);
}

void launch_merge_repair(fi_poi* cur, const fi_poi* repaired, const uint32_t* idx, int n,
                         FrameCounters* counters, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_merge_repair, grid_for(n), dim3(256), 0, s, cur, repaired, idx, n,
                           counters);
}

void launch_qa(const fi_poi* cur, int n, FrameCounters* counters, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_qa, grid_for(n), dim3(256), 0, s, cur, n, counters);
}

}  // namespace fi
