// Device-side frame policy of the multi-frame driver: temporal seeding (with
// the optional two-frame extrapolation), the bad-POI scan with repair-candidate
// collection, the whole-frame fallback decision, the per-POI merges, the frame
// QA counters and the raw output records.  The predicates restate
// opencorr_cli_cuda_mvp.cpp:483-558 and seedPoiQueueFromPrevious (:560-624);
// they run where the POIs already live so a frame needs no host scan.
#include <hip/hip_runtime.h>

#include "fi_device.h"
#include "fi_host.h"

namespace fi {

constexpr float kGoodZncc = 0.8f;  // kSkipFftccZnccThreshold, :479

__device__ __forceinline__ bool finite_f(float v) { return isfinite(v); }

// isBadPoiForFallback (:493-497)
__device__ __forceinline__ bool is_bad_zncc(float z) { return z != -3.f && (!finite_f(z) || z < kGoodZncc); }

__device__ __forceinline__ bool is_bad(const fi_poi& p) { return is_bad_zncc(p.result[2]); }

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

// Block sums of per-thread counts: one atomic per block and counter instead
// of one per POI (same-address atomics serialise in L2).
constexpr int kPolicyThreads = 256;

template <int N, int THREADS = kPolicyThreads>
__device__ __forceinline__ void block_reduce(unsigned (&v)[N], unsigned (&mx)[1])
{
    __shared__ unsigned red[N + 1][THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; q++) {
        unsigned x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        v[q] = x;
    }
    unsigned m = mx[0];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor(m, o));
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) red[q][wave] = v[q];
        red[N][wave] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            unsigned t = 0;
            for (int u = 0; u < THREADS / 64; u++) t += red[q][u];
            v[q] = t;
        }
        unsigned t = 0;
        for (int u = 0; u < THREADS / 64; u++) t = max(t, red[N][u]);
        mx[0] = t;
    }
}

constexpr float kTemporalMinStep = 0.05f;  // kTemporalSeedMinStepPx, :481

// isEligibleTemporalSeedHistory (:499-504)
__device__ __forceinline__ bool eligible_history(const fi_poi& p)
{
    return finite_f(p.deformation[0]) && finite_f(p.deformation[6]) && finite_f(p.result[2]) &&
           p.result[2] >= kGoodZncc;
}

// seedPoiQueueFromPrevious (:560-624): the template, with the twelve shape
// parameters of the previous result where it is valid (finite zncc >= 0) and
// its result and strain cleared; with prev2 (the frame before), a POI whose two
// previous results are good at the template location moves on by their
// difference in u, v when that step is at least kTemporalMinStep px and each
// component at most max_delta.  Extrapolated seeds are counted into
// frame->temporal.
__global__ __launch_bounds__(kPolicyThreads) void k_seed(const fi_poi* __restrict__ tmpl,
                                                          const fi_poi* __restrict__ prev,
                                                          const fi_poi* __restrict__ prev2, float max_delta,
                                                          fi_poi* __restrict__ out, int n, FrameCounters* frame)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool extrapolated = false;
    if (i < n) {
        fi_poi q = tmpl[i];
        const fi_poi& pv = prev[i];
        if (finite_f(pv.result[2]) && pv.result[2] >= 0.f) {
#pragma unroll
            for (int k = 0; k < 12; k++) q.deformation[k] = pv.deformation[k];
            if (prev2) {
                const fi_poi& p2 = prev2[i];
                if (pv.x == p2.x && pv.y == p2.y && pv.x == q.x && pv.y == q.y && eligible_history(pv) &&
                    eligible_history(p2)) {
                    const float du = pv.deformation[0] - p2.deformation[0];
                    const float dv = pv.deformation[6] - p2.deformation[6];
                    const float pu = pv.deformation[0] + du, pw = pv.deformation[6] + dv;
                    const float step2 = du * du + dv * dv;  // two roundings and a sum (no contraction)
                    if (finite_f(pu) && finite_f(pw) && step2 >= kTemporalMinStep * kTemporalMinStep &&
                        fabsf(du) <= max_delta && fabsf(dv) <= max_delta) {
                        q.deformation[0] = pu;
                        q.deformation[6] = pw;
                        extrapolated = true;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 6; k++) q.result[k] = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++) q.strain[k] = 0.f;
        }
        out[i] = q;
    }
    if (prev2) {
        unsigned v[1] = { extrapolated ? 1u : 0u }, mx[1] = { 0u };
        block_reduce<1>(v, mx);
        if (threadIdx.x == 0 && v[0]) atomicAdd(&frame->temporal, v[0]);
    }
}

// processFrame's whole-frame fallback decision (:3086-3115) on the device:
// bad ratio = (float)bad / (float)n over the classes' scans, compared with the
// threshold (baseline + margin) kept in device memory.  Sets the frame flag and,
// per class, how many repair candidates the POI-level repair takes (none under
// a fallback) and how many POIs the fallback re-runs (the class, or none).
__global__ void k_decide(FrameCounters* __restrict__ counters, int nclass, const unsigned* __restrict__ class_n,
                         int n, const float* __restrict__ threshold, int collect)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    unsigned long long bad = 0;
    for (int c = 0; c < nclass; c++) bad += counters[c].bad;
    const float ratio = n ? (float)bad / (float)n : 0.f;
    const bool fb = ratio > *threshold;
    counters[nclass].fallback = fb ? 1u : 0u;
    for (int c = 0; c < nclass; c++) {
        counters[c].repair_n = (fb || !collect) ? 0u : counters[c].candidates;
        counters[c].fallback_n = fb ? class_n[c] : 0u;
    }
}

__global__ __launch_bounds__(kPolicyThreads) void k_scan(const fi_poi* __restrict__ cur,
                                                          const fi_poi* __restrict__ prev, int n,
                                                          int collect, FrameCounters* counters,
                                                          uint32_t* candidates)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    const bool bad = in && is_bad(cur[i]);
    const bool cand = in && collect && is_candidate(cur[i], prev ? prev + i : nullptr, bad);
    // candidates: one slot range per wave, in lane order
    const unsigned long long mask = __builtin_amdgcn_ballot_w64(cand);
    if (mask) {
        const int lane = threadIdx.x & 63;
        unsigned base = 0;
        if (lane == __builtin_ctzll(mask)) base = atomicAdd(&counters->candidates, (unsigned)__popcll(mask));
        base = __shfl(base, __builtin_ctzll(mask));
        if (cand) {
            const unsigned below = (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
            candidates[base + below] = (uint32_t)i;
        }
    }
    unsigned v[1] = { bad ? 1u : 0u }, mx[1] = { 0u };
    block_reduce<1>(v, mx);
    if (threadIdx.x == 0 && v[0]) atomicAdd(&counters->bad, (unsigned long long)v[0]);
}

// Learned-seed fallback candidates: not seeded, or an IC-GN result that is not
// a finite ZNCC at or above the threshold (failures -3/-4/-5 included).
__global__ __launch_bounds__(kPolicyThreads) void k_seed_scan(const fi_poi* __restrict__ cur,
                                                               const uint32_t* __restrict__ ok, int n,
                                                               float threshold, FrameCounters* counters,
                                                               uint32_t* candidates)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool cand = false;
    if (i < n) {
        const float z = cur[i].result[2];
        cand = ok[i] == 0u || !(finite_f(z) && z >= threshold);
    }
    const unsigned long long mask = __builtin_amdgcn_ballot_w64(cand);
    if (mask) {
        const int lane = threadIdx.x & 63;
        unsigned base = 0;
        if (lane == __builtin_ctzll(mask)) base = atomicAdd(&counters->candidates, (unsigned)__popcll(mask));
        base = __shfl(base, __builtin_ctzll(mask));
        if (cand) candidates[base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)i;
    }
}

// Entries a launch covers: n, or the device count when one is given (at most
// n, the capacity the grid was sized for).
__device__ __forceinline__ int bounded(int n, const unsigned* count)
{
    return count ? min((int)*count, n) : n;
}

__global__ void k_gather(const fi_poi* __restrict__ src, const uint32_t* __restrict__ idx, int n,
                         fi_poi* __restrict__ dst, const unsigned* count, uint8_t* __restrict__ flags)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < bounded(n, count)) {
        const uint32_t i = idx[j];
        dst[j] = src[i];
        if (flags) flags[i] = 1u;
    }
}

// The whole-frame fallback's per-POI merge; flags (optional): every POI went
// through FFTCC.
__global__ void k_merge_frame(const fi_poi* __restrict__ skip, fi_poi* __restrict__ scratch, int n,
                              const unsigned* gate, uint8_t* __restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (gate && *gate == 0u)) return;
    if (flags) flags[i] = 1u;
    if (!take_frame_fallback(skip[i], scratch[i])) scratch[i] = skip[i];
}

// seeded (optional): the learned seeder's per-POI flags; a POI it did not
// seed takes the repair unconditionally, so its record is the FFTCC-seeded
// path's whatever the IC-GN from its stale guess reached
__global__ void k_merge_repair(fi_poi* __restrict__ cur, const fi_poi* __restrict__ rep,
                               const uint32_t* __restrict__ idx, int n, FrameCounters* counters,
                               const unsigned* count, const uint32_t* __restrict__ seeded)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= bounded(n, count)) return;
    const uint32_t i = idx[j];
    if ((seeded && seeded[i] == 0u) || should_replace(cur[i], rep[j])) {
        const bool recovered = !is_bad(rep[j]);
        cur[i] = rep[j];
        if (recovered) atomicAdd(&counters->recovered, 1ull);
    }
}

// computeFrameQa (:629-700): valid = finite zncc >= 0; iteration statistics
// over finite iteration counts; bad per isBadPoiForFallback.  Also the frame's
// executed IC-GN iterations: the sum of iters[0..n_iters), the per-entry
// counts every IC-GN launch of the frame wrote (zeroed at the frame's start).
constexpr int kQaThreads = 1024;

__global__ __launch_bounds__(kQaThreads) void k_qa(const fi_poi* __restrict__ cur, int n,
                                                   FrameCounters* counters,
                                                   const uint32_t* __restrict__ iters, size_t n_iters)
{
    // grid-stride over a few hundred blocks, so the six counters see one
    // atomic per block (same-address atomics serialise); blocks of 1024
    // threads and loads batched four deep keep enough reads in flight
    unsigned v[5] = { 0u, 0u, 0u, 0u, 0u }, mx[1] = { 0u };
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = t; j < n_iters; j += 4 * stride) {
        uint32_t a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = j + u * stride < n_iters ? iters[j + u * stride] : 0u;
#pragma unroll
        for (int u = 0; u < 4; u++) v[4] += a[u];
    }
    for (size_t i0 = t; i0 < (size_t)n; i0 += 4 * stride) {
        float2 r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const size_t i = i0 + u * stride;
            // result[2] and result[3]: zncc and the iteration count
            r[u] = i < (size_t)n ? make_float2(cur[i].result[2], cur[i].result[3]) : make_float2(-1.f, NAN);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i0 + u * stride >= (size_t)n) continue;
            const float z = r[u].x, it = r[u].y;
            v[0] += (finite_f(z) && z >= 0.f) ? 1u : 0u;
            v[1] += is_bad_zncc(z) ? 1u : 0u;
            if (finite_f(it)) {
                v[2] += 1u;
                v[3] += (unsigned)(long long)it;
                mx[0] = max(mx[0], (unsigned)(int)(it + 0.5f));
            }
        }
    }
    block_reduce<5, kQaThreads>(v, mx);
    if (threadIdx.x == 0) {
        if (v[4]) atomicAdd(&counters->iterations, (unsigned long long)v[4]);
        if (v[0]) atomicAdd(&counters->valid, (unsigned long long)v[0]);
        if (v[1]) atomicAdd(&counters->bad, (unsigned long long)v[1]);
        if (v[2]) atomicAdd(&counters->iter_finite, (unsigned long long)v[2]);
        if (v[3]) atomicAdd(&counters->iter_sum, (unsigned long long)v[3]);
        if (mx[0]) atomicMax(&counters->iter_max, mx[0]);
    }
}

static dim3 grid_for(int n) { return dim3((n + kPolicyThreads - 1) / kPolicyThreads); }
constexpr int kReduceBlocks = 256;  // grid of the counting kernels

// Zero n 32-bit words (a kernel rather than hipMemsetAsync, so that frame
// graphs capture no memset node).
__global__ void k_zero_words(uint32_t* __restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_fill_bytes(uint8_t* __restrict__ p, size_t n, uint8_t v, const unsigned* gate)
{
    if (gate && *gate == 0u) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Start of a frame in one launch: zero the counters and the per-entry
// iteration counts, set every per-POI FFTCC-redo flag to v and, when src is
// given, copy bytes (a multiple of 4; both buffers 16-byte aligned) from src to
// dst: the POI template into the frame's queue, without a copy node of its own.
__global__ void k_frame_reset(uint32_t* __restrict__ words, size_t n_words, uint32_t* __restrict__ words2,
                              size_t n_words2, uint8_t* __restrict__ flags, size_t n, uint8_t v,
                              const void* __restrict__ src, void* __restrict__ dst, size_t bytes)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = t; i < n_words; i += stride) words[i] = 0u;
    for (size_t i = t; i < n_words2; i += stride) words2[i] = 0u;
    for (size_t i = t; i < n; i += stride) flags[i] = v;
    if (src) {
        const size_t n16 = bytes / 16;
        const uint4* s4 = static_cast<const uint4*>(src);
        uint4* d4 = static_cast<uint4*>(dst);
        for (size_t i = t; i < n16; i += stride) d4[i] = s4[i];
        const size_t tail = (bytes - 16 * n16) / 4;
        if (t < tail)
            static_cast<uint32_t*>(dst)[4 * n16 + t] = static_cast<const uint32_t*>(src)[4 * n16 + t];
    }
}

void launch_frame_reset(uint32_t* words, size_t n_words, uint32_t* words2, size_t n_words2, uint8_t* flags,
                        size_t n, uint8_t v, hipStream_t s, const void* src, void* dst, size_t bytes)
{
    const size_t m = std::max(std::max(std::max(n_words, n_words2), n), src ? bytes / 16 : 0);
    hipLaunchKernelGGL(k_frame_reset, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((m + 255) / 256, 1024))),
                       dim3(256), 0, s, words, n_words, words2, n_words2, flags, n, v, src, dst, bytes);
}

void launch_fill_bytes(uint8_t* p, size_t n, uint8_t v, hipStream_t s, const unsigned* gate)
{
    if (n > 0)
        hipLaunchKernelGGL(k_fill_bytes, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, s,
                           p, n, v, gate);
}

void launch_zero_words(uint32_t* p, size_t n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, s,
                           p, n);
}

void launch_seed_from_previous(const fi_poi* tmpl, const fi_poi* prev, const fi_poi* prev2, float max_delta,
                               fi_poi* out, int n, FrameCounters* frame, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_seed, grid_for(n), dim3(kPolicyThreads), 0, s, tmpl, prev, prev2, max_delta, out, n,
                           frame);
}

void launch_decide(FrameCounters* counters, int nclass, const unsigned* class_n, int n, const float* threshold,
                   int collect, hipStream_t s)
{
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, s, counters, nclass, class_n, n, threshold, collect);
}

// dst = src (bytes, a multiple of 4; both 16-byte aligned) when *gate != 0.
__global__ void k_copy_if(const void* __restrict__ src, void* __restrict__ dst, size_t bytes,
                          const unsigned* __restrict__ gate)
{
    if (*gate == 0u) return;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    const size_t n16 = bytes / 16;
    const uint4* s4 = static_cast<const uint4*>(src);
    uint4* d4 = static_cast<uint4*>(dst);
    for (size_t i = t; i < n16; i += stride) d4[i] = s4[i];
    const size_t tail = (bytes - 16 * n16) / 4;
    if (t < tail) static_cast<uint32_t*>(dst)[4 * n16 + t] = static_cast<const uint32_t*>(src)[4 * n16 + t];
}

void launch_copy_if(const void* src, void* dst, size_t bytes, const unsigned* gate, hipStream_t s)
{
    if (bytes > 0)
        hipLaunchKernelGGL(k_copy_if, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((bytes / 16 + 255) / 256, 1024))),
                           dim3(256), 0, s, src, dst, bytes, gate);
}

// (int32_t) of a float coordinate as the x86 host converts it (truncation;
// NaN and values outside the int32 range give INT32_MIN), so the device records
// equal fi_format_raw's byte for byte.
__device__ __forceinline__ int32_t coord_i32(float v)
{
    return (v >= -2147483648.f && v < 2147483648.f) ? (int32_t)v : INT32_MIN;
}

// formatRawBinary (:2498-2536, no output filter): the 36-byte record x, y
// (int32), u, v, ux, uy, vx, vy, zncc (float32) of layout entry i at record
// order[i] (the caller's template order; identity when order is null).
__global__ void k_pack_raw(const fi_poi* __restrict__ src, const uint32_t* __restrict__ order, int n,
                           uint32_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fi_poi& p = src[i];
    uint32_t* r = out + 9 * (size_t)(order ? order[i] : (uint32_t)i);
    r[0] = (uint32_t)coord_i32(p.x);
    r[1] = (uint32_t)coord_i32(p.y);
    r[2] = __float_as_uint(p.deformation[0]);
    r[3] = __float_as_uint(p.deformation[6]);
    r[4] = __float_as_uint(p.deformation[1]);
    r[5] = __float_as_uint(p.deformation[2]);
    r[6] = __float_as_uint(p.deformation[7]);
    r[7] = __float_as_uint(p.deformation[8]);
    r[8] = __float_as_uint(p.result[2]);
}

void launch_pack_raw(const fi_poi* src, const uint32_t* order, int n, void* out, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_pack_raw, grid_for(n), dim3(256), 0, s, src, order, n, static_cast<uint32_t*>(out));
}

void launch_scan(const fi_poi* cur, const fi_poi* prev, int n, int collect,
                 FrameCounters* counters, uint32_t* candidates, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_scan, grid_for(n), dim3(kPolicyThreads), 0, s, cur, prev, n, collect, counters,
                           candidates);
}

void launch_seed_scan(const fi_poi* cur, const uint32_t* ok, int n, float threshold, FrameCounters* counters,
                      uint32_t* candidates, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_seed_scan, grid_for(n), dim3(kPolicyThreads), 0, s, cur, ok, n, threshold, counters,
                           candidates);
}

void launch_gather(const fi_poi* src, const uint32_t* idx, int n, fi_poi* dst, hipStream_t s,
                   const unsigned* count, uint8_t* flags)
{
    if (n > 0) hipLaunchKernelGGL(k_gather, grid_for(n), dim3(256), 0, s, src, idx, n, dst, count, flags);
}

void launch_merge_frame_fallback(const fi_poi* skip, fi_poi* scratch_inout, int n, hipStream_t s,
                                 const unsigned* gate, uint8_t* flags)
{
    if (n > 0) hipLaunchKernelGGL(k_merge_frame, grid_for(n), dim3(256), 0, s, skip, scratch_inout, n, gate, flags);
}

void launch_merge_repair(fi_poi* cur, const fi_poi* repaired, const uint32_t* idx, int n,
                         FrameCounters* counters, hipStream_t s, const unsigned* count, const uint32_t* seeded)
{
    if (n > 0)
        hipLaunchKernelGGL(k_merge_repair, grid_for(n), dim3(256), 0, s, cur, repaired, idx, n,
                           counters, count, seeded);
}

void launch_qa(const fi_poi* cur, int n, FrameCounters* counters, hipStream_t s, const uint32_t* iters,
               size_t n_iters)
{
    const size_t m = std::max((size_t)n, n_iters);
    hipLaunchKernelGGL(k_qa, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((m + kQaThreads - 1) / kQaThreads,
                                                                                    kReduceBlocks))),
                       dim3(kQaThreads), 0, s, cur, n, counters, iters, n_iters);
}


// ---------------------------------------------------------------------------
// Neighbourhood consensus of learned seeds.  The network predicts each POI's
// (u, v) and gradients from its own crops, and its errors differ from POI to
// POI, while the displacement field varies little over a subset's footprint
// (the first-order warp IC-GN fits assumes as much).  Each seeded POI takes
// the mean of its seeded neighbours' predictions within +-radius, each
// neighbour's (u, v) transported to the POI with that neighbour's gradients,
// u_j - (ux_j dx + uy_j dy) with dx = x_j - x_i, over the neighbours whose
// transported (u, v) agree with the POI's own prediction within tau (a
// neighbour across a discontinuity does not pull the seed).  The mean
// gradients then carry the mean (u, v) from the neighbours' centroid to the
// POI, so a linear field is reproduced whatever the neighbourhood's shape.
// Plain FP32 in a fixed order (cells row by row, a cell's POIs in position
// order), no fused multiply-adds: a host restatement reproduces it exactly.
// ---------------------------------------------------------------------------
// raw[k] = the prediction of POI c.idx[k] (positions grouped by cell), with x
// at 3e38 (and the rest zero) when the POI is not seeded or a value of its
// prediction is not finite: every neighbour test rejects it and every term it
// adds (w = 0 times a finite value) is an exact zero
__global__ void k_seed_raw(const fi_poi* __restrict__ pois, const uint32_t* __restrict__ ok, int m,
                           SeedCells c, float4* __restrict__ raw)
{
    // XCD-aware: each XCD packs a contiguous run of cells, the run its own
    // consensus blocks read (k_seed_consensus), so the reads hit its L2
    const int k = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t i = c.idx[k];
    const fi_poi& q = pois[i];
    const float* d = q.deformation;
    const bool use = (!c.sample || c.sample[i]) && ok[i] && finite_f(d[0]) && finite_f(d[6]) && finite_f(d[1]) &&
                     finite_f(d[2]) && finite_f(d[7]) && finite_f(d[8]);
    raw[2 * k] = use ? make_float4(q.x, q.y, d[0], d[6]) : make_float4(3e38f, 0.f, 0.f, 0.f);
    raw[2 * k + 1] = use ? make_float4(d[1], d[2], d[7], d[8]) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// One thread per POI: its (2K+1)^2 cells are 2K+1 runs of raw (a cell row's
// cells are consecutive), read without indirection.  An entry adds w * value
// with w = 1 when it passes the tests, else 0: a rejected entry adds +-0,
// which leaves each sum's bits as they are (the sums start at +0 and never
// become -0), the loads of a run are independent and each sum is one FMA.
// A POI the network did not seed (c.sample) has no prediction of its own:
// the tau test is measured from the mean of the carried predictions within
// +-radius (a first pass), and its ok is set by whether any survive.
__global__ void k_seed_consensus(fi_poi* __restrict__ pois, uint32_t* __restrict__ ok,
                                 const float4* __restrict__ raw, int n, SeedCells c)
{
    const int i = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool own = !c.sample || c.sample[i];
    if (own && !ok[i]) return;
    const int ci = c.cell[i];
    if (ci < 0) {
        if (!own) ok[i] = 0;
        return;
    }
    fi_poi& q = pois[i];
    const float ax = q.x, ay = q.y;
    float au = q.deformation[0], av = q.deformation[6];
    const int cx = ci % c.ncx, cy = ci / c.ncx;
    const int xa = max(cx - c.k, 0), xb = min(cx + c.k, c.ncx - 1);
    const int ya = max(cy - c.k, 0), yb = min(cy + c.k, c.ncy - 1);
    if (!own) {
        float n1 = 0.f, s1u = 0.f, s1v = 0.f;
        for (int yy = ya; yy <= yb; yy++) {
            const uint32_t k0 = c.start[yy * c.ncx + xa], k1 = c.start[yy * c.ncx + xb + 1];
#pragma unroll 8
            for (uint32_t k = k0; k < k1; k++) {
                const float4 b = raw[2 * k], g = raw[2 * k + 1];
                const float dx = b.x - ax, dy = b.y - ay;
                const float tu = b.z - (g.x * dx + g.y * dy);
                const float tv = b.w - (g.z * dx + g.w * dy);
                const float w = (fabsf(dx) <= c.radius && fabsf(dy) <= c.radius) ? 1.f : 0.f;
                n1 += w;
                s1u = __builtin_fmaf(w, tu, s1u);
                s1v = __builtin_fmaf(w, tv, s1v);
            }
        }
        if (n1 == 0.f) {
            ok[i] = 0;
            return;
        }
        const float inv1 = 1.f / n1;
        au = s1u * inv1;
        av = s1v * inv1;
    }
    float cnt = 0.f, su = 0.f, sv = 0.f, sdx = 0.f, sdy = 0.f, sux = 0.f, suy = 0.f, svx = 0.f, svy = 0.f;
    for (int yy = ya; yy <= yb; yy++) {
        const uint32_t k0 = c.start[yy * c.ncx + xa], k1 = c.start[yy * c.ncx + xb + 1];
#pragma unroll 8
        for (uint32_t k = k0; k < k1; k++) {
            const float4 b = raw[2 * k], g = raw[2 * k + 1];  // x, y, u, v; ux, uy, vx, vy
            const float dx = b.x - ax, dy = b.y - ay;
            const float tu = b.z - (g.x * dx + g.y * dy);
            const float tv = b.w - (g.z * dx + g.w * dy);
            const float w = (fabsf(dx) <= c.radius && fabsf(dy) <= c.radius && fabsf(tu - au) <= c.tau &&
                             fabsf(tv - av) <= c.tau) ? 1.f : 0.f;
            cnt += w;
            su = __builtin_fmaf(w, b.z, su);
            sv = __builtin_fmaf(w, b.w, sv);
            sdx = __builtin_fmaf(w, dx, sdx);
            sdy = __builtin_fmaf(w, dy, sdy);
            sux = __builtin_fmaf(w, g.x, sux);
            suy = __builtin_fmaf(w, g.y, suy);
            svx = __builtin_fmaf(w, g.z, svx);
            svy = __builtin_fmaf(w, g.w, svy);
        }
    }
    if (cnt == 0.f) {  // own prediction not finite, or no neighbour agrees with the mean
        if (!own) ok[i] = 0;
        return;
    }
    const float inv = 1.f / cnt;
    const float ux = sux * inv, uy = suy * inv, vx = svx * inv, vy = svy * inv;
    const float mdx = sdx * inv, mdy = sdy * inv;
    q.deformation[0] = su * inv - (ux * mdx + uy * mdy);
    q.deformation[6] = sv * inv - (vx * mdx + vy * mdy);
    q.deformation[1] = ux;
    q.deformation[2] = uy;
    q.deformation[7] = vx;
    q.deformation[8] = vy;
    if (!own) ok[i] = 1;
}

void launch_seed_consensus(fi_poi* pois, uint32_t* ok, float4* raw, int n, const SeedCells& c,
                           hipStream_t s)
{
    if (n <= 0) return;
    // every position with a cell is in idx once: its length is at most n
    hipLaunchKernelGGL(k_seed_raw, grid_for(c.m), dim3(kPolicyThreads), 0, s, pois, ok, c.m, c, raw);
    hipLaunchKernelGGL(k_seed_consensus, grid_for(n), dim3(kPolicyThreads), 0, s, pois, ok, raw, n, c);
}

}  // namespace fi
