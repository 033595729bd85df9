// Device-side helpers shared by the gfx950 kernels: wave64 reductions built on
// DPP row operations plus lane reads (every lane ends with the same,
// wave-uniform value), compensated FP32 accumulation, XCD-aware block
// remapping, and pixel loads from raw u8/u16/f32 images.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fi_params.h"

namespace fi {

// DPP controls (GCN/CDNA encoding): quad_perm xor 1, quad_perm xor 2,
// row_ror 4, row_ror 8, row_half_mirror.
constexpr int kDppQuadXor1 = 0xB1;
constexpr int kDppQuadXor2 = 0x4E;
constexpr int kDppRowRor4 = 0x124;
constexpr int kDppRowRor8 = 0x128;
constexpr int kDppRowHalfMirror = 0x141;

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ float lane_read(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Sum over the 64 lanes.  Rows of 16 lanes reduce with DPP, the four row
// totals are read from lanes 0/16/32/48 and added in a fixed order, so the
// result is bit-identical on every lane (callers branch on it).
__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_f<kDppQuadXor1>(v);
    v += dpp_f<kDppQuadXor2>(v);
    v += dpp_f<kDppRowRor4>(v);
    v += dpp_f<kDppRowRor8>(v);
    float r0 = lane_read(v, 0), r1 = lane_read(v, 16), r2 = lane_read(v, 32), r3 = lane_read(v, 48);
    return (r0 + r1) + (r2 + r3);
}

// Knuth TwoSum: s + e == a + b exactly (no magnitude precondition).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    float bp = s - a;
    float ap = s - bp;
    e = (a - ap) + (b - bp);
}

// Compensated accumulator: running sum plus the exact rounding errors.
struct CSum {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x)
    {
        float t, e;
        two_sum(s, x, t, e);
        s = t;
        c += e;
    }
};

// Wave sum of a compensated pair: each DPP combine is a TwoSum whose error
// joins the carried compensation; the four row totals are folded the same way.
__device__ __forceinline__ float wave_sum_comp(CSum a)
{
    float s = a.s, c = a.c, t, e;
#define FI_COMP_STEP(CTRL)                      \
    {                                           \
        float os = dpp_f<CTRL>(s);              \
        float oc = dpp_f<CTRL>(c);              \
        two_sum(s, os, t, e);                   \
        s = t;                                  \
        c = (c + oc) + e;                       \
    }
    FI_COMP_STEP(kDppQuadXor1)
    FI_COMP_STEP(kDppQuadXor2)
    FI_COMP_STEP(kDppRowRor4)
    FI_COMP_STEP(kDppRowRor8)
#undef FI_COMP_STEP
    float s0 = lane_read(s, 0), s1 = lane_read(s, 16), s2 = lane_read(s, 32), s3 = lane_read(s, 48);
    float c0 = lane_read(c, 0), c1 = lane_read(c, 16), c2 = lane_read(c, 32), c3 = lane_read(c, 48);
    float t01, e01, t23, e23, tt, et;
    two_sum(s0, s1, t01, e01);
    two_sum(s2, s3, t23, e23);
    two_sum(t01, t23, tt, et);
    float comp = ((c0 + c1) + (c2 + c3)) + ((e01 + e23) + et);
    return tt + comp;
}

// Compensated wave sum of two per-lane partial sums: they are combined
// exactly (TwoSum), then reduced across lanes with the TwoSum tree.
__device__ __forceinline__ float wave_sum_pair(float a, float b)
{
    CSum c;
    two_sum(a, b, c.s, c.c);
    return wave_sum_comp(c);
}

__device__ __forceinline__ float wave_max(float v)
{
    v = fmaxf(v, dpp_f<kDppQuadXor1>(v));
    v = fmaxf(v, dpp_f<kDppQuadXor2>(v));
    v = fmaxf(v, dpp_f<kDppRowRor4>(v));
    v = fmaxf(v, dpp_f<kDppRowRor8>(v));
    float r0 = lane_read(v, 0), r1 = lane_read(v, 16), r2 = lane_read(v, 32), r3 = lane_read(v, 48);
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

__device__ __forceinline__ int wave_min_i(int v)
{
    v = min(v, dpp_i<kDppQuadXor1>(v));
    v = min(v, dpp_i<kDppQuadXor2>(v));
    v = min(v, dpp_i<kDppRowRor4>(v));
    v = min(v, dpp_i<kDppRowRor8>(v));
    int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return min(min(r0, r1), min(r2, r3));
}

__device__ __forceinline__ bool wave_any(bool p)
{
    return __builtin_amdgcn_ballot_w64(p) != 0;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8),
// each XCD with its own L2.  Give every XCD one contiguous run of work items so
// neighbouring POIs, which share image texels, hit the same L2.  Bijective on
// [0, nblocks).
__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    const int nx = 8;
    const int q = nblocks / nx, rem = nblocks % nx;
    const int xcd = b % nx, slot = b / nx;
    return xcd * q + min(xcd, rem) + slot;
}

// BC = B * C of the bicubic B-spline, src/oc_cubic_bspline.h:52-58, as the
// float quotients the reference's initialisers produce.  Row i holds the
// coefficients of d^(3-i) of the four neighbour weights.
__device__ __forceinline__ float bc(int i, int j)
{
    constexpr float m[4][4] = {
        { -144.0f / 336.0f, 384.0f / 336.0f, -384.0f / 336.0f, 144.0f / 336.0f },
        { 342.0f / 336.0f, -702.0f / 336.0f, 450.0f / 336.0f, -90.0f / 336.0f },
        { -198.0f / 336.0f, -18.0f / 336.0f, 270.0f / 336.0f, -54.0f / 336.0f },
        { 0.0f / 336.0f, 336.0f / 336.0f, 0.0f / 336.0f, 0.0f / 336.0f },
    };
    return m[i][j];
}

template <typename T>
__device__ __forceinline__ float load_px(const T* img, size_t idx, float scale);

template <>
__device__ __forceinline__ float load_px<uint8_t>(const uint8_t* img, size_t idx, float scale)
{
    return (float)img[idx] * scale;
}

template <>
__device__ __forceinline__ float load_px<uint16_t>(const uint16_t* img, size_t idx, float scale)
{
    return (float)img[idx] * scale;
}

template <>
__device__ __forceinline__ float load_px<float>(const float* img, size_t idx, float scale)
{
    return img[idx] * scale;
}

}  // namespace fi
