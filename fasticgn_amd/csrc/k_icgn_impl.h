// ICGN2D1 iteration kernel for gfx950 (template; instantiated per lane shape
// in k_icgn*.hip).
//
// Semantics follow the CPU reference ICGN2D1::compute(POI2D*),
// src/oc_icgn.cpp:168-387: the entry gate (-3), the geometric out-of-range
// predicate of :39-43 (-3, deformation keeps the incoming guess), a NaN sample
// coordinate (-5), not converged within the iteration cap (-4), NaN result
// (-5 with u, v reset to the initial guess).  The reference GPU's extra gates
// (margin-2 reference bound, zero-texture -3, |dp| > 100) are deliberately not
// applied: the CPU path is the one this is judged against.
//
// An iteration reduces twice.  With t' = t - ref_mean for the target samples
// (kept per lane) and r = ref - ref_mean, the first round sums t', t'^2 and
// t' r, which give the target mean and norm and the ZNCC
//     sum (t - tbar)^2 = sum t'^2 - (sum t')^2 / n,
//     ZNCC = (sum t' r - tbar' * sum r) / (nr * nt);
// the second sums SD_j e with the reference's per-sample error
// e = (t - tbar) * (nr / nt) - r, so the numerator, small near convergence, is
// never the difference of two large sums.  Shifting by ref_mean keeps the
// one-pass variance free of cancellation.  Per-lane sums are FMA chains of at
// most SPL terms; the cross-lane and cross-wave trees are compensated
// (TwoSum); the convergence norm is accumulated in FP64.
//
// Lane layout ("bands"): the subset's sw columns map to lanes, G = NT / sw
// bands of R = ceil(sh / G) rows stack vertically, the last band ending at the
// bottom row (its first ovl rows repeat rows of the band above and are masked
// out of the sums).  A lane's local x is fixed and its y steps by one per
// slot, so the warped coordinate of slot k is base + k * W01 and the y-weighted
// sums come from a running suffix sum.  Lanes beyond the bands mirror band 0
// (their samples are real ones) and contribute nothing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "fi_device.h"
#include "fi_launch.h"

namespace fi {

// entry gate of src/oc_icgn.cpp:184-186 (reference side; the deformation and
// zncc parts are checked by the caller) plus finite coordinates, which keep
// (int) conversions in bounds
__device__ __forceinline__ bool ref_gate_ok(float x, float y, int rx, int ry, int w, int h)
{
    if (!(isfinite(x) && isfinite(y))) return false;
    return !(y - ry < 0 || x - rx < 0 || y + ry > h - 1 || x + rx > w - 1);
}

struct Bands {
    int R;        // rows per band
    int ovl;      // leading rows of the last band already covered above
    int col;      // this lane's subset column
    int row0;     // this lane's first subset row
    bool active;  // the lane owns a column of a band
    bool last;    // the lane is in the last band and that band overlaps
};

// a / b for 0 <= a < 2^21 and b > 0: (a + 0.5) / b lies at least 0.5 / b from
// an integer, and the float product is within (a / b) 2^-22 of it, so the
// truncation is the quotient (a reciprocal and a multiply where the integer
// division takes four quarter-rate multiplies).
__device__ __forceinline__ int div_small(int a, int b)
{
    return (int)(((float)a + 0.5f) * __builtin_amdgcn_rcpf((float)b));
}

__device__ __forceinline__ Bands make_bands(int rx, int ry, int nt, int tid)
{
    Bands b;
    const int sw = 2 * rx + 1, sh = 2 * ry + 1;
    const int G = div_small(nt, sw);
    b.R = div_small(sh + G - 1, G);
    const int gu = div_small(sh + b.R - 1, b.R);
    b.ovl = gu * b.R - sh;
    int g = div_small(tid, sw);
    b.col = tid - g * sw;
    b.active = g < gu;
    if (!b.active) g = 0;
    b.row0 = (g == gu - 1) ? sh - b.R : g * b.R;
    b.last = b.active && g == gu - 1 && b.ovl > 0;
    return b;
}

// Bicubic B-spline sample from the 16-float texel (coef[i][j] multiplies
// dy^i dx^j, src/oc_cubic_bspline.cpp:154-177), Horner in dx then dy.
__device__ __forceinline__ float bspline_eval(const float4 c0, const float4 c1, const float4 c2,
                                              const float4 c3, float dx, float dy)
{
    const float r0 = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c0.w, dx, c0.z), dx, c0.y), dx, c0.x);
    const float r1 = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c1.w, dx, c1.z), dx, c1.y), dx, c1.x);
    const float r2 = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c2.w, dx, c2.z), dx, c2.y), dx, c2.x);
    const float r3 = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c3.w, dx, c3.z), dx, c3.y), dx, c3.x);
    return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(r3, dy, r2), dy, r1), dy, r0);
}

// The coefficient table is read through a raw buffer resource whose range is
// the texel rows (64 * w * h bytes).  A texel offset is range checked by the
// hardware (zeros past the end, no access); the plane is added as the scalar
// offset, at most 3 * 16 * w bytes, which the three zero planes after the last
// row cover.  So a sample outside the image (whose value is discarded) needs
// no clamp and can never read outside the allocation.
constexpr uint32_t kBufferWord3 = 0x00020000;  // gfx9-family raw buffer, dword data

__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const float* coef, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(coef), (short)0, (int)bytes, (int)kBufferWord3);
}

// One float of an f32 image read through a raw buffer resource: the lane's
// byte offset in a VGPR, a wave-uniform row offset as the scalar offset, so a
// run of rows costs no per-load address arithmetic.
__device__ __forceinline__ float ld1(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
}

__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
}

// Sample at (gx, gy).  For coordinates in 1 <= x < w-2, 1 <= y < h-2 the
// truncating conversion is the floor and fract is exact.
__device__ __forceinline__ float bspline_sample(__amdgpu_buffer_rsrc_t r, uint32_t row_bytes,
                                                uint32_t plane_bytes, float gx, float gy)
{
    const float dx = __builtin_amdgcn_fractf(gx), dy = __builtin_amdgcn_fractf(gy);
    const uint32_t off = __umul24((uint32_t)(int)gy, row_bytes) + ((uint32_t)(int)gx << 4);
    return bspline_eval(ld4(r, off, 0), ld4(r, off, plane_bytes), ld4(r, off, 2 * plane_bytes),
                        ld4(r, off, 3 * plane_bytes), dx, dy);
}

// ---------------------------------------------------------------------------
// Sampling from the raw target staged in LDS.  The B-spline value
// sum_{k,l} coef[k][l] dy^k dx^l with coef = BC Q BC^T (flipped) is, term by
// term, sum_n wy_n sum_m wx_m Q[n][m] over the 4x4 neighbourhood Q (rows
// iy-1..iy+2, columns ix-1..ix+2) with w_m(d) = BC[0][m] d^3 + BC[1][m] d^2 +
// BC[2][m] d + BC[3][m]: the same polynomial, evaluated from 16 raw pixels
// instead of 16 stored coefficients.
// ---------------------------------------------------------------------------
// The four cubics, factored: with t = 1 - d and P = d t (BC's columns are
// mirror images, w3(d) = w0(t), w2(d) = w1(t), and every weight but the
// interpolating part of w1, w2 vanishes at d = 0 and d = 1)
//     w0 = P (24 d - 33) / 56           w1 = t + P (318 - 384 d) / 336
//     w3 = -P (24 d + 9) / 56           w2 = d + P (384 d - 66) / 336
// which equal BC's columns as polynomials: ten operations instead of twelve.
__device__ __forceinline__ void spline_weights(float d, float& w0, float& w1, float& w2, float& w3)
{
    const float t = 1.f - d, P = d * t;
    w0 = P * __builtin_fmaf(24.f / 56.f, d, -33.f / 56.f);
    w3 = P * __builtin_fmaf(-24.f / 56.f, d, -9.f / 56.f);
    w1 = __builtin_fmaf(P, __builtin_fmaf(-384.f / 336.f, d, 318.f / 336.f), t);
    w2 = __builtin_fmaf(P, __builtin_fmaf(384.f / 336.f, d, -66.f / 336.f), d);
}

__device__ __forceinline__ float row4(const float* q, float w0, float w1, float w2, float w3)
{
    return __builtin_fmaf(w3, q[3], __builtin_fmaf(w2, q[2], __builtin_fmaf(w1, q[1], w0 * q[0])));
}

// bias + sum_n w_n(d) r_n for the four neighbourhood rows, from the factored
// weights above: with P = d (1 - d), d21 = r2 - r1 and w_n = [lerp part] +
// P (a_n d + b_n),
//     sum = r1 + d d21 + P (d A + B),   A = (24/56)(r0 - r3) + (384/336) d21,
//     B = -(33/56) r0 + (318/336) r1 - (66/336) r2 - (9/56) r3
// (a_1 = -a_2 and a_3 = -a_0 give A; the b_n are the constant terms): thirteen
// operations with the bias, where the four weights and their combination take
// fourteen and the bias one more.
__device__ __forceinline__ float ycombine(float d, float r0, float r1, float r2, float r3, float bias)
{
    const float P = __builtin_fmaf(-d, d, d);
    const float d21 = r2 - r1;
    const float A = __builtin_fmaf(24.f / 56.f, r0 - r3, (384.f / 336.f) * d21);
    const float B = __builtin_fmaf(-9.f / 56.f, r3,
                                   __builtin_fmaf(-66.f / 336.f, r2, __builtin_fmaf(318.f / 336.f, r1, (-33.f / 56.f) * r0)));
    return __builtin_fmaf(P, __builtin_fmaf(d, A, B), __builtin_fmaf(d, d21, r1 + bias));
}

// LDS row stride of the per-POI tile (tiles up to 64 pixels wide, radius <=
// 28): a neighbourhood's four rows lie 0, 64, 128 and 192 words from one
// address, so its 16 pixels are eight ds_read2_b32 with immediate offsets.
constexpr int kTileStride = 64;

// bias + sample at (gx, gy) from a tile whose element (0, 0) is image pixel
// (tx0, ty0); cbase = (ty0 + 1) * kTileStride + tx0 + 1 places the
// neighbourhood's first row and column.  The caller guarantees the
// neighbourhood lies in the tile.
__device__ __forceinline__ float tile_sample(const float* tile, int cbase, float gx, float gy, float bias)
{
    const float dx = __builtin_amdgcn_fractf(gx), dy = __builtin_amdgcn_fractf(gy);
    const int base = ((int)gy << 6) + (int)gx - cbase;
    static_assert(kTileStride == 64, "base uses a shift by 6");
    const float* q0 = tile + base;
    const float* q1 = q0 + kTileStride;
    const float* q2 = q1 + kTileStride;
    const float* q3 = q2 + kTileStride;
    float x0, x1, x2, x3;
    spline_weights(dx, x0, x1, x2, x3);
    const float r0 = row4(q0, x0, x1, x2, x3), r1 = row4(q1, x0, x1, x2, x3);
    const float r2 = row4(q2, x0, x1, x2, x3), r3 = row4(q3, x0, x1, x2, x3);
    return ycombine(dy, r0, r1, r2, r3, bias);
}

// bias + sample at (gx, gy) from the f32 target in global memory (rows w floats
// apart) with the tile path's weights: an iteration whose neighbourhoods
// outgrow the LDS tile, when no coefficient table was built.  The caller's
// range test (1 <= gx < w-2, 1 <= gy < h-2) keeps the 4x4 neighbourhood in the
// image; a row's four pixels are one unaligned 16-byte load.
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ float target_sample(const float* tar, uint32_t w, float gx, float gy, float bias)
{
    const float dx = __builtin_amdgcn_fractf(gx), dy = __builtin_amdgcn_fractf(gy);
    const float* q0 = tar + ((size_t)((uint32_t)(int)gy - 1u) * w + ((uint32_t)(int)gx - 1u));
    const f32x4u a = *reinterpret_cast<const f32x4u*>(q0);
    const f32x4u b = *reinterpret_cast<const f32x4u*>(q0 + w);
    const f32x4u c = *reinterpret_cast<const f32x4u*>(q0 + 2 * (size_t)w);
    const f32x4u d = *reinterpret_cast<const f32x4u*>(q0 + 3 * (size_t)w);
    float x0, x1, x2, x3;
    spline_weights(dx, x0, x1, x2, x3);
    auto row = [&](const f32x4u q) {
        return __builtin_fmaf(x3, q[3], __builtin_fmaf(x2, q[2], __builtin_fmaf(x1, q[1], x0 * q[0])));
    };
    return ycombine(dy, row(a), row(b), row(c), row(d), bias);
}

// ---------------------------------------------------------------------------
// Shared coefficient tile (sampling mode FI_SAMPLING_GROUP): kGroupPois
// neighbouring POIs, one wave each, share one LDS tile holding, per pixel
// (row n, column ix), the x-polynomial of the row segment [ix, ix+1]:
//     r_n(dx) = sum_m w_m(dx) Q[n][ix-1+m] = c0 + c1 dx + c2 dx^2 + c3 dx^3,
//     c3 = sum_m BC[0][m] Q[n][ix-1+m], c2 = sum_m BC[1][m] .., c1 = sum_m BC[2][m] ..,
//     c0 = Q[n][ix] (BC's last row is (0, 1, 0, 0)),
// so a sample is four Horner rows and the y weights: about 14 VALU fewer than
// the raw tile's weights and 4x4 combination.  16 bytes per pixel is too much
// LDS for one POI per block at 12 blocks per CU; a block of four neighbouring
// POIs (5 px apart on a grid) shares one.
// ---------------------------------------------------------------------------
constexpr int kGroupPois = 4;
constexpr int kGroupStride = 64;  // entries per LDS row: four rows are immediate offsets of one address
constexpr int kGroupRowSlack = 8;  // rows of vertical spread the four neighbourhoods may add
// pixels of slack around the union of the four incoming neighbourhoods: motion
// the entries' iterations may make before they leave the tile
#ifndef FI_GROUP_MARGIN
#define FI_GROUP_MARGIN 2
#endif
constexpr int kGroupMargin = FI_GROUP_MARGIN;

// bias + sample at (gx, gy).  Entry (row, col) of a tile whose pixel (0, 0)
// is image pixel (tx0, ty0) holds the polynomial of pixel column tx0 + 1 +
// col; cb = -(ty0 + 1) * kGroupStride - (tx0 + 1) places the sample's first
// row.  The caller guarantees the neighbourhood lies in the tile.
template <int STRIDE = kGroupStride>
__device__ __forceinline__ float coef_sample(const float4* ct, int cb, float gx, float gy, float bias)
{
    static_assert(kGroupStride * sizeof(float4) == 1024, "a row of entries is 1 KiB: a shift by 10");
    const float dx = __builtin_amdgcn_fractf(gx), dy = __builtin_amdgcn_fractf(gy);
    // rows of 1 KiB: a shift; other strides: one 24-bit multiply-add
    const uint32_t rowb = STRIDE == kGroupStride
                              ? ((uint32_t)(int)gy << 10) + (uint32_t)(cb * 16)
                              : __umul24((uint32_t)(int)gy, (uint32_t)(STRIDE * 16)) + (uint32_t)(cb * 16);
    const float4* q = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(ct) +
                                                      (((uint32_t)(int)gx << 4) + rowb));
    auto horner = [&](const float4 c) {
        return __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(c.w, dx, c.z), dx, c.y), dx, c.x);
    };
    const float r0 = horner(q[0]), r1 = horner(q[STRIDE]);
    const float r2 = horner(q[2 * STRIDE]), r3 = horner(q[3 * STRIDE]);
    return ycombine(dy, r0, r1, r2, r3, bias);
}

// Coefficients of pixels tx0 + 1 .. tx0 + tw - 3 of rows ty0 .. ty0 + th - 1
// (the image pixels tx0 .. tx0 + tw - 1 they read lie in the image).  Lane =
// column (tw - 3 <= 64), wave = row modulo the block's waves; three rows per
// pass, so twelve loads are in flight before the first store.
__device__ __forceinline__ float4 row_coef(float q0, float q1, float q2, float q3)
{
    float4 v;
    v.x = q1;
    v.y = __builtin_fmaf(bc(2, 3), q3, __builtin_fmaf(bc(2, 2), q2, __builtin_fmaf(bc(2, 1), q1, bc(2, 0) * q0)));
    v.z = __builtin_fmaf(bc(1, 3), q3, __builtin_fmaf(bc(1, 2), q2, __builtin_fmaf(bc(1, 1), q1, bc(1, 0) * q0)));
    v.w = __builtin_fmaf(bc(0, 3), q3, __builtin_fmaf(bc(0, 2), q2, __builtin_fmaf(bc(0, 1), q1, bc(0, 0) * q0)));
    return v;
}

template <int WAVES, int STRIDE = kGroupStride>
__device__ __forceinline__ void stage_coef(float4* ct, const float* tar, int w, int tx0, int ty0, int tw, int th,
                                           int lane, int wave)
{
    if (lane >= tw - 3) return;
    constexpr int kRows = 3;
    const float* q = tar + (size_t)(ty0 + wave) * w + (tx0 + lane);
    float4* d = ct + wave * STRIDE + lane;
    const size_t step = (size_t)WAVES * w;
    int r = wave;
    for (; r + (kRows - 1) * WAVES < th; r += kRows * WAVES) {
        float v[kRows][4];
#pragma unroll
        for (int j = 0; j < kRows; j++)
#pragma unroll
            for (int m = 0; m < 4; m++) v[j][m] = q[j * step + m];
#pragma unroll
        for (int j = 0; j < kRows; j++) d[j * WAVES * STRIDE] = row_coef(v[j][0], v[j][1], v[j][2], v[j][3]);
        q += kRows * step;
        d += kRows * WAVES * STRIDE;
    }
    for (; r < th; r += WAVES) {
        *d = row_coef(q[0], q[1], q[2], q[3]);
        q += step;
        d += WAVES * STRIDE;
    }
}

// Copy image rows ty0.. and columns tx0.. (tw x th pixels) into the tile,
// twelve loads per lane in flight before their stores.
__device__ __forceinline__ void stage_tile(float* tile, const float* tar, int w, int tx0, int ty0,
                                           int tw, int th, int stride, int tid, int nt)
{
    constexpr int kBatch = 12;
    const uint32_t magic = (uint32_t)((0x100000000ull + (unsigned long long)tw - 1) / (unsigned long long)tw);
    const int n = tw * th;
    for (int base = tid; base < n; base += kBatch * nt) {
        float v[kBatch];
        int dst[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; j++) {
            const int idx = min(base + j * nt, n - 1);  // a clamped read is not stored
            const int r = (int)__umulhi((uint32_t)idx, magic);
            const int c = idx - r * tw;
            v[j] = tar[(size_t)(ty0 + r) * w + (tx0 + c)];
            dst[j] = r * stride + c;
        }
#pragma unroll
        for (int j = 0; j < kBatch; j++)
            if (base + j * nt < n) tile[dst[j]] = v[j];
    }
}

// Per-lane partial sums -> wave totals, identical on every lane.  Transposing
// exchanges (permlane32_swap, permlane16_swap, and a rotate by 8 inside the
// rows with selects) fold pairs of values so each register carries up to
// eight values, one per 8-lane half row; DPP then reduces inside the groups
// and lane reads collect the totals.
__device__ __forceinline__ float2 swap32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return make_float2(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

__device__ __forceinline__ float2 swap16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return make_float2(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

// Compensated pair: value s plus the accumulated rounding errors c.
struct CPair {
    float s, c;
};

// s + t exactly as a rounded sum and its error (Knuth TwoSum), the errors of
// both operands carried along
__device__ __forceinline__ CPair cadd(CPair a, CPair b)
{
    float s, e;
    two_sum(a.s, b.s, s, e);
    return { s, (a.c + b.c) + e };
}

// fold32(a, b): rows 0-1 carry a summed over the two lane halves, rows 2-3 b
__device__ __forceinline__ CPair fold32(float a, float b)
{
    const float2 v = swap32(a, b);
    CPair r;
    two_sum(v.x, v.y, r.s, r.c);
    return r;
}

// fold16(p, q) of two fold32 results: rows carry p.lo, q.lo, p.hi, q.hi, each
// summed over four lanes 16 apart
__device__ __forceinline__ CPair fold16(CPair p, CPair q)
{
    const float2 s = swap16(p.s, q.s);
    const float2 c = swap16(p.c, q.c);
    return cadd({ s.x, c.x }, { s.y, c.y });
}

template <int CTRL>
__device__ __forceinline__ CPair cdpp(CPair v)
{
    return cadd(v, { dpp_f<CTRL>(v.s), dpp_f<CTRL>(v.c) });
}

// sum of the 16 lanes of each row (every lane of the row holds it)
__device__ __forceinline__ CPair row_sum16(CPair v)
{
    v = cdpp<kDppQuadXor1>(v);
    v = cdpp<kDppQuadXor2>(v);
    v = cdpp<kDppRowRor4>(v);
    v = cdpp<kDppRowRor8>(v);
    return v;
}

// fold8(p, q) of two fold16 results: in every row, lanes 0-7 carry p and
// lanes 8-15 q, each summed over two lanes 8 apart
__device__ __forceinline__ CPair fold8(CPair p, CPair q)
{
    const bool lo = (__lane_id() & 8) == 0;
    const CPair a = { lo ? p.s : q.s, lo ? p.c : q.c };
    const CPair x = { lo ? q.s : p.s, lo ? q.c : p.c };
    return cadd(a, { dpp_f<kDppRowRor8>(x.s), dpp_f<kDppRowRor8>(x.c) });
}

// sum of the 8 lanes of each half row (every lane of the half holds it)
__device__ __forceinline__ CPair half_sum8(CPair v)
{
    v = cdpp<kDppQuadXor1>(v);
    v = cdpp<kDppQuadXor2>(v);
    v = cdpp<kDppRowHalfMirror>(v);
    return v;
}

// Per-lane partial sums -> wave totals with a compensated (TwoSum) tree.
__device__ __forceinline__ void wave_sum3(float (&v)[3])
{
    const CPair w0 = fold32(v[0], v[1]);
    const CPair w1 = fold32(v[2], v[2]);
    const CPair z = row_sum16(fold16(w0, w1));  // rows: 0, 2, 1, 2
    const float t = z.s + z.c;
    v[0] = lane_read(t, 0);
    v[2] = lane_read(t, 16);
    v[1] = lane_read(t, 32);
}

__device__ __forceinline__ void wave_sum6(float (&v)[6])
{
    const CPair w0 = fold32(v[0], v[1]);
    const CPair w1 = fold32(v[2], v[3]);
    const CPair w2 = fold32(v[4], v[5]);
    // rows 0, 2, 1, 3 beside rows 4, 4, 5, 5: half rows 0|4, 2|4, 1|5, 3|5
    const CPair z = half_sum8(fold8(fold16(w0, w1), fold16(w2, w2)));
    const float t = z.s + z.c;
    v[0] = lane_read(t, 0);
    v[4] = lane_read(t, 8);
    v[2] = lane_read(t, 16);
    v[1] = lane_read(t, 32);
    v[5] = lane_read(t, 40);
    v[3] = lane_read(t, 48);
}

__device__ __forceinline__ void wave_sum(float (&v)[3]) { wave_sum3(v); }
__device__ __forceinline__ void wave_sum(float (&v)[6]) { wave_sum6(v); }

// Block totals: wave totals combined through LDS in wave order with TwoSum,
// so every wave holds the same bits.
// lds holds WAVES * kBlockSumMax floats.
constexpr int kBlockSumMax = 6;

template <int WAVES, int N>
__device__ __forceinline__ void block_sum(float (&v)[N], float* lds, int wave, int lane)
{
    static_assert(N <= kBlockSumMax, "the block's reduction buffer holds kBlockSumMax values per wave");
    wave_sum(v);
    if (WAVES == 1) return;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) lds[wave * N + q] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; q++) {
        CPair acc = { lds[q], 0.f };
#pragma unroll
        for (int u = 1; u < WAVES; u++) acc = cadd(acc, { lds[u * N + q], 0.f });
        v[q] = acc.s + acc.c;
    }
    __syncthreads();
}

// Any lane of the block: bit 0 = a, bit 1 = b.
template <int WAVES>
__device__ __forceinline__ int block_any2(bool a, bool b, int* lds, int wave, int lane)
{
    const int m = (__builtin_amdgcn_ballot_w64(a) != 0 ? 1 : 0) | (__builtin_amdgcn_ballot_w64(b) != 0 ? 2 : 0);
    if (WAVES == 1) return m;
    if (lane == 0) lds[wave] = m;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int u = 0; u < WAVES; u++) r |= lds[u];
    __syncthreads();
    return r;
}

// A wave-uniform pointer into the constant address space that the optimiser
// cannot prove loop-invariant: loads through it are scalar loads at their
// point of use instead of values hoisted into registers held across the
// sampling pass.  For data the kernel never writes.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* reread(const T* ptr)
{
    uint64_t v = reinterpret_cast<uint64_t>(ptr);
    asm volatile("" : "+v"(v));
    // readfirstlane returns int: widen through uint32_t (no sign extension)
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    const uint64_t u = ((uint64_t)hi << 32) | (uint64_t)lo;
    return (const __attribute__((address_space(4))) T*)u;
}

// Pixels of slack around a subset's neighbourhood in the LDS tile, so the
// tile survives the first iterations' motion without restaging.
constexpr int kTileMargin = 2;

// Corners of a warped subset closer than this to the interpolatable range's
// edge send the iteration through the exact per-sample test; a sample and the
// box bound it is tested against differ by at most kTileSlack in rounding.
constexpr float kEdgeMargin = 1.f / 64.f;
constexpr float kTileSlack = 0.125f;

// Largest band overlap a launched shape may have (checked on the host).
constexpr int kMaxOverlap = 4;

// Minimum resident waves per SIMD the register allocation must allow (the
// compiler otherwise trades occupancy for deeper load pipelining).
#ifndef FI_ICGN_MIN_WAVES
#define FI_ICGN_MIN_WAVES 3
#endif

// One block of WAVES wavefronts per POI; each lane owns SPL rows of one subset
// column.  EXACT: the host guarantees R == SPL, so no slot is skipped.  Every
// early return is taken by the whole block.
// GROUP (WAVES == 1, k_icgn_group): tile is the block's shared coefficient
// tile, pixel (0, 0) = image (gx0, gy0), gw x gh pixels (gw = 0: none); it is
// never restaged, and an iteration whose neighbourhoods leave it samples the
// target in device memory.  No block barrier is reached on this path, so the
// block's waves run their entries independently.
// OFFSET: p.offset holds the entries' centre offsets (compiled only where a
// launch needs them: the offsets are uniform values live across the whole
// iteration, which the register budget of the plain kernels cannot afford).
template <int WAVES, int SPL, bool EXACT, bool GROUP = false, bool OFFSET = false>
__device__ __forceinline__ void icgn_entry(const IcgnParams& p, const int b, float* red, int* flags, float* tile,
                                           float* hl, int gx0 = 0, int gy0 = 0, int gw = 0, int gh = 0)
{
    static_assert(!GROUP || WAVES == 1, "a shared coefficient tile serves one-wave entries");
    constexpr int NT = 64 * WAVES;
    // the thread's index within the entry: in a group block each wave is an
    // entry of its own
    const int tid = GROUP ? (int)(threadIdx.x & 63) : (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = p.index ? (int)p.index[b] : b;
    fi_poi* poi = p.pois + i;
    const float x = poi->x, y = poi->y;
    // compute(POI2D*, Point2D& center_offset), src/oc_icgn.cpp:399-615: the
    // warp acts about (x + ox, y + oy) on local coordinates minus the offset
    float ox = 0.f, oy = 0.f;
    if constexpr (OFFSET) {
        ox = p.offset[2 * (size_t)i];
        oy = p.offset[2 * (size_t)i + 1];
    }
    const float xc = OFFSET ? x + ox : x, yc = OFFSET ? y + oy : y;
    // the warped image of the local box's centre (-ox, -oy), relative to (xc, yc)
    auto box_dx = [&](float W00_, float W01_, float W02_) {
        return OFFSET ? __builtin_fmaf(-W00_, ox, __builtin_fmaf(-W01_, oy, W02_)) : W02_;
    };
    const float u_in = poi->deformation[0], ux_in = poi->deformation[1], uy_in = poi->deformation[2];
    const float v_in = poi->deformation[6], vx_in = poi->deformation[7], vy_in = poi->deformation[8];
    const float z_in = poi->result[2];
    int rx = p.rx, ry = p.ry;
    if (p.self_adaptive) {
        rx = (int)poi->subset_radius[0];
        ry = (int)poi->subset_radius[1];
    }
    const int w = p.w, h = p.h;
    // entry gate, src/oc_icgn.cpp:184-191
    if (!ref_gate_ok(x, y, rx, ry, w, h) || fabsf(u_in) >= (float)w || fabsf(v_in) >= (float)h
        || z_in < 0.f || isnan(u_in) || isnan(v_in)) {
        if (tid == 0) {
            poi->result[2] = (z_in >= 0.f) ? -3.f : z_in;
            if (p.iters) p.iters[b] = 0;
        }
        return;
    }
    const PoiPrep* pr = p.prep + (p.prep_index ? p.prep_index[i] : (uint32_t)i);
    const float ref_mean = pr->mean, ref_norm = pr->norm;
    const float inv_ref_norm = 1.f / ref_norm;
    // the wave's copy of the inverse Hessian (hl: 36 floats of LDS per wave),
    // from which lane a < 6 forms dp[a] each iteration; written and read by
    // this wave only
    hl += 36 * wave;
    if (lane < 36) hl[lane] = pr->hinv[lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const Bands bd = make_bands(rx, ry, NT, tid);
    const int R = bd.R, ovl = bd.ovl;
    // Per-slot reference data; zero where the slot repeats a row of the band
    // above or the lane mirrors band 0, which removes those samples from every
    // sum but the two of t' itself (masked with m_all / m_dup below).
    const int ulx = (int)(x - rx), uly = (int)(y - ry);
    const int px = (int)x, py = (int)y;
    // warp matrix, Deformation2D1::setWarp (src/oc_deformation.cpp:117-128)
    float W00 = 1.f + ux_in, W01 = uy_in, W02 = u_in;
    float W10 = vx_in, W11 = 1.f + vy_in, W12 = v_in;
    float W20 = 0.f, W21 = 0.f, W22 = 1.f;
    const float wm2 = (float)(w - 2), hm2 = (float)(h - 2);
    const float rxf = (float)rx, ryf = (float)ry;
    // small integers: the float products are exact
    const float rx2 = (float)rx * (float)rx, ry2 = (float)ry * (float)ry;
    const float inv_ns = 1.f / ((float)(2 * rx + 1) * (float)(2 * ry + 1));
    const float conv_crit = p.conv, stop = p.stop;

    // The LDS tile holds the raw target around the warped subset; it is
    // (re)staged when an iteration's neighbourhoods leave it, and an iteration
    // whose neighbourhoods cannot fit in it samples the table instead.
    const int tile_w = GROUP ? gw : p.tile_w, tile_h = GROUP ? gh : p.tile_h, tile_stride = p.tile_stride;
    int tx0 = GROUP ? gx0 : -(1 << 30), ty0 = GROUP ? gy0 : -(1 << 30);

    // Every lane loads (idle lanes and repeated rows address real subset
    // samples; a generic shape's slots past R repeat row R-1), then masks: no
    // divergent branch, so all loads are in flight together.
    // The images are read as buffers: the host caps w * h * 4 below 2^31, and
    // every slot addresses a subset sample inside the image (entry gate).
    float rz[SPL], gxv[SPL], gyv[SPL];
    const uint32_t pitch = (uint32_t)w * 4u, img_bytes = pitch * (uint32_t)h;
    const __amdgpu_buffer_rsrc_t rs_ref = table_rsrc(p.ref, img_bytes);
    const __amdgpu_buffer_rsrc_t rs_gx = table_rsrc(p.gx, img_bytes), rs_gy = table_rsrc(p.gy, img_bytes);
    // rows and widths below 2^24: 24-bit multiplies
    const uint32_t ro = (__umul24((uint32_t)(uly + bd.row0), (uint32_t)w) + (uint32_t)(ulx + bd.col)) * 4u;
    const uint32_t go = (__umul24((uint32_t)(py - ry + bd.row0), (uint32_t)w) + (uint32_t)(px - rx + bd.col)) * 4u;
#pragma unroll
    for (int k = 0; k < SPL; k++) {
        const uint32_t so = (uint32_t)(EXACT ? k : min(k, R - 1)) * pitch;
        const float rv = ld1(rs_ref, ro, so);
        const float gxl = ld1(rs_gx, go, so), gyl = ld1(rs_gy, go, so);
        const bool valid = bd.active && !(bd.last && k < ovl) && (EXACT || k < R);
        rz[k] = valid ? rv - ref_mean : 0.f;
        gxv[k] = valid ? gxl : 0.f;
        gyv[k] = valid ? gyl : 0.f;
    }
    // Stage the tile for the incoming warp now, so its loads are in flight
    // together with the reference loads above (one round trip, not two).
    if constexpr (!GROUP) {
        const float ex = __builtin_fmaf(fabsf(W00), rxf, fabsf(W01) * ryf);
        const float ey = __builtin_fmaf(fabsf(W10), rxf, fabsf(W11) * ryf);
        const float cx = xc + box_dx(W00, W01, W02);
        const float cy = yc + box_dx(W10, W11, W12);
        if (tile_w > 0 && cx - ex >= 1.f + kEdgeMargin && cx + ex < wm2 - kEdgeMargin &&
            cy - ey >= 1.f + kEdgeMargin && cy + ey < hm2 - kEdgeMargin) {
            const int nx0 = max((int)(cx - ex - kTileSlack) - 1, 0);
            const int nx1 = min((int)(cx + ex + kTileSlack) + 2, w - 1);
            const int ny0 = max((int)(cy - ey - kTileSlack) - 1, 0);
            const int ny1 = min((int)(cy + ey + kTileSlack) + 2, h - 1);
            if (nx1 - nx0 < tile_w && ny1 - ny0 < tile_h) {
                tx0 = min(max(nx0 - (tile_w - (nx1 - nx0 + 1)) / 2, 0), w - tile_w);
                ty0 = min(max(ny0 - (tile_h - (ny1 - ny0 + 1)) / 2, 0), h - tile_h);
                stage_tile(tile, p.tar, w, tx0, ty0, tile_w, tile_h, tile_stride, tid, NT);
                __syncthreads();
            }
        }
    }
    const float m_all = bd.active ? 1.f : 0.f;
    const float m_dup = bd.last ? 0.f : 1.f;
    const float xl = OFFSET ? (float)(bd.col - rx) - ox : (float)(bd.col - rx);
    const float yl0 = OFFSET ? (float)(bd.row0 - ry) - oy : (float)(bd.row0 - ry);
    const __amdgpu_buffer_rsrc_t table = table_rsrc(p.coef, (uint32_t)w * (uint32_t)h * 64u);
    const uint32_t row_bytes = (uint32_t)w * 64u, plane_bytes = (uint32_t)w * 16u;

    // Shared tile (GROUP): an iteration samples it when its neighbourhood,
    // columns max((int)(cx - ex - kTileSlack) - 1, 0) .. min((int)(cx + ex +
    // kTileSlack) + 2, w - 1) and the rows alike, lies in the tile's columns
    // tx0 .. tx0 + tile_w - 1 and rows.  With c = cx - ex >= 0 (samples in
    // range) the truncation is the floor and c - kTileSlack is exact (c <
    // 2^20), so the first column test is c >= tx0 + 1 + kTileSlack, or always
    // for tx0 = 0; the last, c' = cx + ex < tx0 + tile_w - 2 - kTileSlack, or
    // always for a tile reaching the image's last column.
    float g_lox = INFINITY, g_hix = -INFINITY, g_loy = INFINITY, g_hiy = -INFINITY;
    if constexpr (GROUP) {
        if (tile_w > 0) {
            g_lox = tx0 > 0 ? (float)tx0 + (1.f + kTileSlack) : -INFINITY;
            g_loy = ty0 > 0 ? (float)ty0 + (1.f + kTileSlack) : -INFINITY;
            g_hix = tx0 + tile_w >= w ? INFINITY : (float)(tx0 + tile_w) - (2.f + kTileSlack);
            g_hiy = ty0 + tile_h >= h ? INFINITY : (float)(ty0 + tile_h) - (2.f + kTileSlack);
        }
    }

    int it = 0;
    float zncc = 0.f;
    double cn;
    do {
        it++;
        // sample k of the lane sits at (xb + k W01, yb + k W11): the POI
        // centre, the warp of the lane's column and first row, then its steps
        const float xb = xc + __builtin_fmaf(W00, xl, __builtin_fmaf(W01, yl0, W02));
        const float yb = yc + __builtin_fmaf(W10, xl, __builtin_fmaf(W11, yl0, W12));
        // The samples are an affine image of the grid: they lie within the
        // warped bounding box x + W02 +- (|W00| rx + |W01| ry), so all are in
        // the interpolatable range (:39-43) when the box is, by a margin.  A NaN
        // anywhere in the warp fails the test.
        const float ex = __builtin_fmaf(fabsf(W00), rxf, fabsf(W01) * ryf);
        const float ey = __builtin_fmaf(fabsf(W10), rxf, fabsf(W11) * ryf);
        const float cx = xc + box_dx(W00, W01, W02);
        const float cy = yc + box_dx(W10, W11, W12);
        const bool box_in = (cx - ex >= 1.f + kEdgeMargin) & (cx + ex < wm2 - kEdgeMargin) &
                            (cy - ey >= 1.f + kEdgeMargin) & (cy + ey < hm2 - kEdgeMargin);
        if (!box_in) {
            // Per-sample test: out of range (-3) takes precedence over a NaN
            // coordinate (-5), :290-309; neither means the box was only near
            // the edge.
            bool oob = false, nan = false;
#pragma unroll
            for (int k = 0; k < SPL; k++) {
                if (EXACT || k < R) {
                    const float gxp = __builtin_fmaf((float)k, W01, xb);
                    const float gyp = __builtin_fmaf((float)k, W11, yb);
                    oob |= (gxp < 1.f) | (gyp < 1.f) | (gxp >= wm2) | (gyp >= hm2);
                    nan |= (gxp != gxp) | (gyp != gyp);
                }
            }
            const int any = block_any2<WAVES>(oob, nan, flags, wave, lane);
            if (any) {
                if (tid == 0) {
                    poi->result[2] = (any & 1) ? -3.f : -5.f;
                    if (p.iters) p.iters[b] = (uint32_t)it;
                }
                return;
            }
        }

        bool use_tile;
        if constexpr (GROUP) {
            // the neighbourhood test below against the fixed shared tile, as
            // four compares with the bounds formed before the loop
            use_tile = (cx - ex >= g_lox) & (cx + ex < g_hix) & (cy - ey >= g_loy) & (cy + ey < g_hiy);
        } else {
            // neighbourhood columns/rows this iteration needs (within the
            // image: every sample passed the range test)
            const int nx0 = max((int)(cx - ex - kTileSlack) - 1, 0);
            const int nx1 = min((int)(cx + ex + kTileSlack) + 2, w - 1);
            const int ny0 = max((int)(cy - ey - kTileSlack) - 1, 0);
            const int ny1 = min((int)(cy + ey + kTileSlack) + 2, h - 1);
            use_tile = tile_w > 0 && nx1 - nx0 < tile_w && ny1 - ny0 < tile_h;
            if (use_tile && !(nx0 >= tx0 && nx1 < tx0 + tile_w && ny0 >= ty0 && ny1 < ty0 + tile_h)) {
                // centre the tile on the neighbourhood, inside the image
                tx0 = min(max(nx0 - (tile_w - (nx1 - nx0 + 1)) / 2, 0), w - tile_w);
                ty0 = min(max(ny0 - (tile_h - (ny1 - ny0 + 1)) / 2, 0), h - tile_h);
                if (WAVES > 1) __syncthreads();  // every wave is done with the old tile
                stage_tile(tile, p.tar, w, tx0, ty0, tile_w, tile_h, tile_stride, tid, NT);
                __syncthreads();
            }
        }
        const int cbase = GROUP ? -(ty0 + 1) * kGroupStride - (tx0 + 1) : (ty0 + 1) * tile_stride + tx0 + 1;

        // pass 1: sample; t' = t - ref_mean per slot (the samplers subtract
        // it), and the sums of t', t'^2, t' r (a repeated row counts once: its
        // t' is masked, its r is zero)
        float tpv[SPL];
        float st = 0.f, stt = 0.f, str = 0.f;
        auto pass1 = [&](auto sample) {
#pragma unroll
            for (int k = 0; k < SPL; k++) {
                tpv[k] = 0.f;
                if (EXACT || k < R) {
                    const float gxp = __builtin_fmaf((float)k, W01, xb);
                    const float gyp = __builtin_fmaf((float)k, W11, yb);
                    const float tp = sample(gxp, gyp, k);
                    const float tpm = (k < kMaxOverlap) ? tp * ((k < ovl) ? m_dup : 1.f) : tp;
                    tpv[k] = tp;
                    // (a sum's first term is assigned: 0 + x is not folded)
                    st = (EXACT && k == 0) ? tpm : st + tpm;
                    stt = __builtin_fmaf(tpm, tp, stt);
                    str = __builtin_fmaf(tp, rz[k], str);
                }
            }
        };
        if (use_tile) {
            if constexpr (GROUP)
                pass1([&](float gx, float gy, int) {
                    return coef_sample(reinterpret_cast<const float4*>(tile), cbase, gx, gy, -ref_mean);
                });
            else
                pass1([&](float gx, float gy, int) { return tile_sample(tile, cbase, gx, gy, -ref_mean); });
        } else if (p.coef)
            pass1([&](float gx, float gy, int) {
                return bspline_sample(table, row_bytes, plane_bytes, gx, gy) - ref_mean;
            });
        else
            pass1([&](float gx, float gy, int) { return target_sample(p.tar, (uint32_t)w, gx, gy, -ref_mean); });
        // lanes mirroring band 0 drop their t' sums (their r is zero)
        float s1[3] = { st * m_all, stt * m_all, str };
        block_sum<WAVES>(s1, red, wave, lane);

        // ZNSSD terms (Subset2D::zeroMeanNorm, :311-322); the per-POI
        // constant is re-read (scalar load) rather than held in a register
        const auto pc = reread(pr);
        const float tm = s1[0] * inv_ns;                      // target mean - reference mean
        const float szz = __builtin_fmaf(-s1[0], tm, s1[1]);  // sum of squared centred target
        const float inv_tar_norm = __builtin_amdgcn_rsqf(szz);
        // e = 0 * inf when the target subset is flat, as in the reference
        const float scale = __int_as_float(__builtin_amdgcn_readfirstlane(
            __float_as_int((szz > 0.f) ? ref_norm * inv_tar_norm : __builtin_nanf(""))));
        zncc = __builtin_fmaf(-tm, pc->ref_sum, s1[2]) * inv_tar_norm * inv_ref_norm;

        // pass 2: e = (t - tbar) * (nr / nt) - r and the numerator sum SD_j e
        // (:311-322); slots in descending order so AK = sum_k k * a_k
        // accumulates as a running suffix sum (repeated rows have SD = 0)
        float A = 0.f, AK = 0.f, B = 0.f, BK = 0.f;
#pragma unroll
        for (int k = SPL - 1; k >= 0; k--) {
            if (EXACT || k < R) {
                const float e = __builtin_fmaf(scale, tpv[k] - tm, -rz[k]);
                AK = (EXACT && k == SPL - 2) ? A : AK + A;
                BK = (EXACT && k == SPL - 2) ? B : BK + B;
                A = __builtin_fmaf(gxv[k], e, A);
                B = __builtin_fmaf(gyv[k], e, B);
            }
        }
        float bnum[6] = { A, xl * A, __builtin_fmaf(yl0, A, AK), B, xl * B, __builtin_fmaf(yl0, B, BK) };
        block_sum<WAVES>(bnum, red, wave, lane);

        // dp = H^-1 b (src/oc_icgn.cpp:324-331): lane a < 6 forms dp[a] from
        // its row (lanes above repeat row 5), the wave reads the six back
        float dpl = 0.f;
        {
            const float* hr = hl + 6 * min(lane, 5);
#pragma unroll
            for (int c = 0; c < 6; c++) dpl = __builtin_fmaf(hr[c], bnum[c], dpl);
        }
        float dp[6];
#pragma unroll
        for (int a = 0; a < 6; a++) dp[a] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dpl), a));
        // convergence norm (:343-352), accumulated in FP64; the loop compares
        // its square with the launcher's threshold, the root is taken once.
        // It is finite exactly when every dp is (a float squared cannot
        // overflow FP64).
        cn = (double)dp[0] * dp[0];
        cn += (double)dp[1] * dp[1] * rx2;
        cn += (double)dp[2] * dp[2] * ry2;
        cn += (double)dp[3] * dp[3];
        cn += (double)dp[4] * dp[4] * rx2;
        cn += (double)dp[5] * dp[5] * ry2;

        // W <- W * W(dp)^-1 (:333-340); the 3x3 inverse by cofactors down column 0
        const float D00 = 1.f + dp[1], D01 = dp[2], D02 = dp[0];
        const float D10 = dp[4], D11 = 1.f + dp[5], D12 = dp[3];
        if (cn <= __DBL_MAX__) {
            // Finite dp: the terms of the 3x3 product that hold a zero of the
            // affine last rows (0 0 1) are exact zeros and drop out, leaving
            // every other product and sum with the rounding of the full form.
            const float c20 = D01 * D12 - D02 * D11, c21 = D02 * D10 - D00 * D12;
            const float det = D11 * D00 + -(D01 * D10);
            const float invdet = __builtin_amdgcn_rcpf(det);
            const float I00 = D11 * invdet, I01 = -D01 * invdet, I02 = c20 * invdet;
            const float I10 = -D10 * invdet, I11 = D00 * invdet, I12 = c21 * invdet;
            const float I22 = (D00 * D11 - D01 * D10) * invdet;
            const float N00 = W00 * I00 + W01 * I10, N01 = W00 * I01 + W01 * I11;
            const float N02 = W00 * I02 + (W01 * I12 + W02 * I22);
            const float N10 = W10 * I00 + W11 * I10, N11 = W10 * I01 + W11 * I11;
            const float N12 = W10 * I02 + (W11 * I12 + W12 * I22);
            W00 = N00; W01 = N01; W02 = N02; W10 = N10; W11 = N11; W12 = N12;
            W22 = W22 * I22;
        } else {
            const float D20 = 0.f, D21 = 0.f, D22 = 1.f;
            const float c00 = D11 * D22 - D12 * D21, c01 = D12 * D20 - D10 * D22, c02 = D10 * D21 - D11 * D20;
            const float c10 = D21 * D02 - D22 * D01, c11 = D22 * D00 - D20 * D02, c12 = D20 * D01 - D21 * D00;
            const float c20 = D01 * D12 - D02 * D11, c21 = D02 * D10 - D00 * D12, c22 = D00 * D11 - D01 * D10;
            const float det = c00 * D00 + (c10 * D10 + c20 * D20);
            const float invdet = __builtin_amdgcn_rcpf(det);
            const float I00 = c00 * invdet, I01 = c10 * invdet, I02 = c20 * invdet;
            const float I10 = c01 * invdet, I11 = c11 * invdet, I12 = c21 * invdet;
            const float I20 = c02 * invdet, I21 = c12 * invdet, I22 = c22 * invdet;
            const float N00 = W00 * I00 + (W01 * I10 + W02 * I20), N01 = W00 * I01 + (W01 * I11 + W02 * I21);
            const float N02 = W00 * I02 + (W01 * I12 + W02 * I22), N10 = W10 * I00 + (W11 * I10 + W12 * I20);
            const float N11 = W10 * I01 + (W11 * I11 + W12 * I21), N12 = W10 * I02 + (W11 * I12 + W12 * I22);
            const float N20 = W20 * I00 + (W21 * I10 + W22 * I20), N21 = W20 * I01 + (W21 * I11 + W22 * I21);
            const float N22 = W20 * I02 + (W21 * I12 + W22 * I22);
            W00 = N00; W01 = N01; W02 = N02; W10 = N10; W11 = N11; W12 = N12; W20 = N20; W21 = N21; W22 = N22;
        }
    } while ((float)it < stop && cn >= p.conv2);
    const float conv = (float)sqrt(cn);

    if (tid == 0) {
        if (p.iters) p.iters[b] = (uint32_t)it;
        float u = W02, ux = W00 - 1.f, uy = W01, v = W12, vx = W10, vy = W11 - 1.f;
        if (conv >= conv_crit && (float)it >= stop) zncc = -4.f;
        if (isnan(zncc) || isnan(u) || isnan(v)) {
            u = u_in;
            v = v_in;
            zncc = -5.f;
        }
        poi->deformation[0] = u;
        poi->deformation[1] = ux;
        poi->deformation[2] = uy;
        poi->deformation[6] = v;
        poi->deformation[7] = vx;
        poi->deformation[8] = vy;
        poi->result[0] = u_in;
        poi->result[1] = v_in;
        poi->result[2] = zncc;
        poi->result[3] = (float)it;
        poi->result[4] = conv;
        poi->subset_radius[0] = (float)rx;
        poi->subset_radius[1] = (float)ry;
    }
}

// Persistent work queue: entries [0, n) split into kQueueParts contiguous
// ranges, one per XCD (workgroups are dealt to the XCDs round-robin, so block
// b runs on XCD b % 8 and neighbouring POIs share that XCD's L2); a block
// claims kQueueChunk entries at a time from its own range, then from the
// others' until every range is drained.  p.queue holds the kQueueParts
// counters, kQueueStride ints apart, zeroed.
constexpr int kQueueChunk = 4;

template <int WAVES, int SPL, bool EXACT, bool OFFSET = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(FI_ICGN_MIN_WAVES)))
void k_icgn(IcgnParams p)
{
    __shared__ float red[WAVES > 1 ? WAVES * kBlockSumMax : 1];
    __shared__ int flags[WAVES];
    __shared__ float hl[WAVES * 36];
    extern __shared__ float tile[];
    // one call site: the entry is inlined once
    const int b = p.count ? (int)blockIdx.x : xcd_remap(blockIdx.x, gridDim.x);
    const int n = p.count ? min((int)*p.count, p.n) : p.n;
    if (b < n) icgn_entry<WAVES, SPL, EXACT, false, OFFSET>(p, b, red, flags, tile, hl);
}

template <int WAVES, int SPL, bool EXACT>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(FI_ICGN_MIN_WAVES)))
void k_icgn_queue(IcgnParams p)
{
    __shared__ float red[WAVES > 1 ? WAVES * kBlockSumMax : 1];
    __shared__ int flags[WAVES];
    __shared__ float hl[WAVES * 36];
    // claim[0]: the entry the block runs next, -1 once every range is
    // drained; claim[1]: end of the chunk it came from; claim[2]: ranges
    // thread 0 has passed since its own.  The walk state lives here rather
    // than in registers, so that only the entry index stays live across an
    // entry, as in k_icgn (the entry needs every scalar register it can get).
    __shared__ int claim[3];
    extern __shared__ float tile[];
    if (threadIdx.x == 0) {
        claim[0] = -1;
        claim[1] = 0;
        claim[2] = 0;
    }
    for (;;) {
        if (threadIdx.x == 0) {
            int b = claim[0] + 1;
            if (claim[0] < 0 || b >= claim[1]) {
                const int home = (int)(blockIdx.x % kQueueParts);
                int k = claim[2];
                b = -1;
                for (; k < kQueueParts; k++) {
                    const int part = (home + k) % kQueueParts;
                    const int lo = (int)((long long)p.n * part / kQueueParts);
                    const int hi = (int)((long long)p.n * (part + 1) / kQueueParts);
                    // a drained range is seen with a load, not another atomic
                    if (lo + __hip_atomic_load(p.queue + part * kQueueStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi)
                        continue;
                    const int c = lo + atomicAdd(p.queue + part * kQueueStride, kQueueChunk);
                    if (c >= hi) continue;
                    b = c;
                    claim[1] = min(c + kQueueChunk, hi);
                    break;
                }
                claim[2] = k;
            }
            claim[0] = b;
        }
        __syncthreads();
        const int b = __builtin_amdgcn_readfirstlane(claim[0]);
        __syncthreads();  // every thread has read the entry before it is rewritten
        if (b < 0) break;
        icgn_entry<WAVES, SPL, EXACT>(p, b, red, flags, tile, hl);
        __syncthreads();  // shared state is reused by the next entry
    }
}

// Pixel neighbourhood (columns nx0..nx1, rows ny0..ny1) of entry b's incoming
// warp, as icgn_entry computes it; false when the entry fails the gate or its
// warped box is not inside the interpolatable range by kEdgeMargin.
__device__ __forceinline__ bool entry_box(const IcgnParams& p, int b, int& nx0, int& nx1, int& ny0, int& ny1)
{
    const int i = p.index ? (int)p.index[b] : b;
    const fi_poi* poi = p.pois + i;
    const float x = poi->x, y = poi->y;
    float ox = 0.f, oy = 0.f;
    if (p.offset) {
        ox = p.offset[2 * (size_t)i];
        oy = p.offset[2 * (size_t)i + 1];
    }
    const float u = poi->deformation[0], ux = poi->deformation[1], uy = poi->deformation[2];
    const float v = poi->deformation[6], vx = poi->deformation[7], vy = poi->deformation[8];
    int rx = p.rx, ry = p.ry;
    if (p.self_adaptive) {
        rx = (int)poi->subset_radius[0];
        ry = (int)poi->subset_radius[1];
    }
    const int w = p.w, h = p.h;
    if (!ref_gate_ok(x, y, rx, ry, w, h) || fabsf(u) >= (float)w || fabsf(v) >= (float)h || poi->result[2] < 0.f ||
        isnan(u) || isnan(v))
        return false;
    const float W00 = 1.f + ux, W01 = uy, W10 = vx, W11 = 1.f + vy;
    const float ex = __builtin_fmaf(fabsf(W00), (float)rx, fabsf(W01) * (float)ry);
    const float ey = __builtin_fmaf(fabsf(W10), (float)rx, fabsf(W11) * (float)ry);
    const float cx = (x + ox) + __builtin_fmaf(-W00, ox, __builtin_fmaf(-W01, oy, u));
    const float cy = (y + oy) + __builtin_fmaf(-W10, ox, __builtin_fmaf(-W11, oy, v));
    if (!(cx - ex >= 1.f + kEdgeMargin && cx + ex < (float)(w - 2) - kEdgeMargin && cy - ey >= 1.f + kEdgeMargin &&
          cy + ey < (float)(h - 2) - kEdgeMargin))
        return false;
    nx0 = max((int)(cx - ex - kTileSlack) - 1, 0);
    nx1 = min((int)(cx + ex + kTileSlack) + 2, w - 1);
    ny0 = max((int)(cy - ey - kTileSlack) - 1, 0);
    ny1 = min((int)(cy + ey + kTileSlack) + 2, h - 1);
    return true;
}

// kGroupPois consecutive entries per block, one wave each, sharing one
// coefficient tile that covers their incoming neighbourhoods plus
// kGroupMargin (p.tile_h = rows allocated; a union that does not fit leaves the
// block without a tile).  The two barriers come before any wave starts its
// entry.
template <int SPL, bool EXACT>
__global__ __launch_bounds__(64 * kGroupPois) __attribute__((amdgpu_waves_per_eu(FI_ICGN_MIN_WAVES)))
void k_icgn_group(IcgnParams p)
{
    __shared__ int gbox[kGroupPois][4];
    __shared__ float red[1];
    __shared__ int flags[kGroupPois];
    __shared__ float hl[kGroupPois * 36];
    extern __shared__ float4 ctile[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = xcd_remap(blockIdx.x, gridDim.x) * kGroupPois + wave;
    int bx0 = 0, bx1 = -1, by0 = 0, by1 = -1;
    const bool has = b < p.n && entry_box(p, b, bx0, bx1, by0, by1);
    if (lane == 0) {
        gbox[wave][0] = has ? bx0 : INT_MAX;
        gbox[wave][1] = bx1;
        gbox[wave][2] = by0;
        gbox[wave][3] = by1;
    }
    __syncthreads();
    int X0 = INT_MAX, X1 = INT_MIN, Y0 = INT_MAX, Y1 = INT_MIN;
#pragma unroll
    for (int v = 0; v < kGroupPois; v++)
        if (gbox[v][0] != INT_MAX) {
            X0 = min(X0, gbox[v][0]);
            X1 = max(X1, gbox[v][1]);
            Y0 = min(Y0, gbox[v][2]);
            Y1 = max(Y1, gbox[v][3]);
        }
    int gw = 0, gh = 0;
    if (X0 <= X1) {
        X0 = max(X0 - kGroupMargin, 0);
        X1 = min(X1 + kGroupMargin, p.w - 1);
        Y0 = max(Y0 - kGroupMargin, 0);
        Y1 = min(Y1 + kGroupMargin, p.h - 1);
        if (X1 - X0 - 2 <= kGroupStride && Y1 - Y0 + 1 <= p.tile_h) {
            gw = X1 - X0 + 1;
            gh = Y1 - Y0 + 1;
        }
    }
    if (gw > 0) stage_coef<kGroupPois>(ctile, p.tar, p.w, X0, Y0, gw, gh, lane, wave);
    __syncthreads();
    if (b < p.n)
        icgn_entry<1, SPL, EXACT, true>(p, b, red, flags + wave, reinterpret_cast<float*>(ctile), hl + 36 * wave, X0,
                                        Y0, gw, gh);
}

// ---------------------------------------------------------------------------
// Two entries per wave (sampling mode FI_SAMPLING_PAIR, k_icgn_pair): lanes
// 0-31 run one entry and lanes 32-63 the next, each with the band layout of a
// 32-lane entry (G = 32 / sw bands of R = SPL rows).  Both halves execute every
// instruction, so the work of an iteration that does not grow with the subset
// (the cross-lane sums, the 6x6 solve, the warp update, the convergence test)
// serves two entries, and the per-entry set-up is shared the same way.  The
// cross-lane trees never cross the halves (permlane16 swaps, rotations inside
// 16-lane rows, DPP inside 8-lane half rows); each lane reads its half's totals
// back with ds_bpermute, which occupies the LDS crossbar instead of VALU issue.
// The wave iterates while either entry does: a finished entry has written its
// record, and its lanes then sample at the other entry's coordinates (always a
// valid path) and their sums are discarded.  Eight neighbouring entries (four
// waves) share one coefficient tile, kPairStride entries per row.
// ---------------------------------------------------------------------------
constexpr int kPairEntries = 8;
constexpr int kPairStride = 80;  // 1280 B rows: eight POIs 5 px apart at r = 15 need 74 entries
constexpr int kPairLanes = 32;

// the value of v in the lane at byte address addr (lane * 4)
__device__ __forceinline__ float lane_fetch(float v, int addr)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// the same lane's v in the other half of the wave
__device__ __forceinline__ float other_half(float v, bool upper)
{
    const float2 s = swap32(v, v);  // x: lower half of v in both halves, y: upper half
    return upper ? s.x : s.y;
}

// fold16 of raw per-lane values inside each 32-lane half: in each half, the
// first row carries a and the second b, each summed over two lanes 16 apart
__device__ __forceinline__ CPair hfold16(float a, float b)
{
    const float2 v = swap16(a, b);
    CPair r;
    two_sum(v.x, v.y, r.s, r.c);
    return r;
}

// Per-lane partials -> totals of each 32-lane half (compensated), read back by
// every lane of the half; hb = byte address of the half's lane 0.
__device__ __forceinline__ void half_sum3(float (&v)[3], int hb)
{
    const CPair w0 = hfold16(v[0], v[1]);      // rows per half: v0, v1
    const CPair w1 = hfold16(v[2], v[2]);      // rows per half: v2, v2
    const CPair z = half_sum8(fold8(w0, w1));  // half rows: v0 | v2, v1 | v2
    const float t = z.s + z.c;
    v[0] = lane_fetch(t, hb);
    v[2] = lane_fetch(t, hb + 32);
    v[1] = lane_fetch(t, hb + 64);
}

__device__ __forceinline__ void half_sum6(float (&v)[6], int hb)
{
    const CPair w0 = hfold16(v[0], v[1]);
    const CPair w1 = hfold16(v[2], v[3]);
    const CPair w2 = hfold16(v[4], v[5]);
    const CPair z0 = half_sum8(fold8(w0, w1));          // half rows: v0 | v2, v1 | v3
    const CPair z1 = half_sum8(cdpp<kDppRowRor8>(w2));  // rows: v4, v5
    const float t0 = z0.s + z0.c, t1 = z1.s + z1.c;
    v[0] = lane_fetch(t0, hb);
    v[2] = lane_fetch(t0, hb + 32);
    v[1] = lane_fetch(t0, hb + 64);
    v[3] = lane_fetch(t0, hb + 96);
    v[4] = lane_fetch(t1, hb);
    v[5] = lane_fetch(t1, hb + 64);
}

// any lane of this lane's half
__device__ __forceinline__ bool half_any(bool v, bool upper)
{
    const uint64_t m = __builtin_amdgcn_ballot_w64(v);
    return (upper ? (m >> 32) : (m & 0xffffffffull)) != 0;
}

// Band layout of a 32-lane entry: the shape k_icgn_pair<SPL> runs (0 = none).
inline int pair_rows(int rx, int ry)
{
    const int sw = 2 * rx + 1, sh = 2 * ry + 1;
    if (rx < 1 || ry < 1 || sw > kPairLanes) return 0;
    const int G = kPairLanes / sw;
    const int R = (sh + G - 1) / G;
    const int gu = (sh + R - 1) / R;
    return gu * R - sh <= kMaxOverlap ? R : 0;
}

// Entry b of the launch on this lane's half (b >= p.n: no entry), slot e of
// the block.  The entry's set-up loads are issued first, so they are in
// flight while the block agrees on its coefficient tile: the union of the
// entries' incoming neighbourhoods plus kGroupMargin (gbox: one box per slot),
// staged once (p.tile_h rows of kPairStride entries; a union that does not fit
// leaves the block without a tile).  The tile is never restaged; an iteration
// whose neighbourhoods leave it samples the target in device memory.  Uniform
// radius p.rx, p.ry (the host launches no self-adaptive or centre-offset work
// here).  Both barriers are reached by every thread before any iteration.
template <int SPL>
__device__ __forceinline__ void icgn_pair_entry(const IcgnParams& p, const int b, const int e, float* hl,
                                                float4* ct, int (*gbox)[4])
{
    const int lane = threadIdx.x & 63, hlane = lane & 31;
    const bool upper = lane >= 32;
    const int hb = (lane & 32) * 4;
    const bool in = b < p.n;
    const int i = in ? (p.index ? (int)p.index[b] : b) : 0;
    fi_poi* poi = p.pois + i;
    float xc = poi->x, yc = poi->y;
    const float u_in = poi->deformation[0], ux_in = poi->deformation[1], uy_in = poi->deformation[2];
    const float v_in = poi->deformation[6], vx_in = poi->deformation[7], vy_in = poi->deformation[8];
    const float z_in = poi->result[2];
    const int rx = p.rx, ry = p.ry, w = p.w, h = p.h;
    // entry gate, src/oc_icgn.cpp:184-191
    // (a NaN zncc passes: the gate rejects zncc < 0 only)
    bool live = in && ref_gate_ok(xc, yc, rx, ry, w, h) && fabsf(u_in) < (float)w && fabsf(v_in) < (float)h &&
                !(z_in < 0.f) && !isnan(u_in) && !isnan(v_in);
    if (in && !live && hlane == 0) {
        poi->result[2] = (z_in >= 0.f) ? -3.f : z_in;
        if (p.iters) p.iters[b] = 0;
    }
    const PoiPrep* pr = p.prep + (p.prep_index ? p.prep_index[i] : (uint32_t)i);
    const float ref_mean = pr->mean, ref_norm = pr->norm, ref_sum = pr->ref_sum;
    const float inv_ref_norm = 1.f / ref_norm;
    // the entry's inverse Hessian in LDS (36 floats), for lanes a < 6
    hl[hlane] = pr->hinv[hlane];
    if (hlane < 4) hl[32 + hlane] = pr->hinv[32 + hlane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const Bands bd = make_bands(rx, ry, kPairLanes, hlane);
    const int ovl = bd.ovl;
    float W00 = 1.f + ux_in, W01 = uy_in, W02 = u_in;
    float W10 = vx_in, W11 = 1.f + vy_in, W12 = v_in;
    float W20 = 0.f, W21 = 0.f, W22 = 1.f;
    const float wm2 = (float)(w - 2), hm2 = (float)(h - 2);
    const float rxf = (float)rx, ryf = (float)ry;
    const float rx2 = (float)rx * (float)rx, ry2 = (float)ry * (float)ry;
    const float inv_ns = 1.f / ((float)(2 * rx + 1) * (float)(2 * ry + 1));

    // Reference data per slot (zero on masked slots).  A half without an
    // entry reads nothing: its offsets lie past the buffers' ranges, which the
    // hardware answers with zeros.
    float rz[SPL], gxv[SPL], gyv[SPL];
    const uint32_t pitch = (uint32_t)w * 4u, img_bytes = pitch * (uint32_t)h;
    const __amdgpu_buffer_rsrc_t rs_ref = table_rsrc(p.ref, img_bytes);
    const __amdgpu_buffer_rsrc_t rs_gx = table_rsrc(p.gx, img_bytes), rs_gy = table_rsrc(p.gy, img_bytes);
    uint32_t ro = 0xfffffff0u;
    if (live) {
        const int ulx = (int)(xc - rx), uly = (int)(yc - ry);
        ro = (__umul24((uint32_t)(uly + bd.row0), (uint32_t)w) + (uint32_t)(ulx + bd.col)) * 4u;
    }
#pragma unroll
    for (int k = 0; k < SPL; k++) {
        const uint32_t so = (uint32_t)k * pitch;
        const float rv = ld1(rs_ref, ro, so);
        const float gxl = ld1(rs_gx, ro, so), gyl = ld1(rs_gy, ro, so);
        const bool valid = bd.active && !(bd.last && k < ovl);
        rz[k] = valid ? rv - ref_mean : 0.f;
        gxv[k] = valid ? gxl : 0.f;
        gyv[k] = valid ? gyl : 0.f;
    }

    // the block's tile: this entry's incoming neighbourhood (entry_box)
    {
        const float ex = __builtin_fmaf(fabsf(W00), rxf, fabsf(W01) * ryf);
        const float ey = __builtin_fmaf(fabsf(W10), rxf, fabsf(W11) * ryf);
        const float cx = xc + W02, cy = yc + W12;
        const bool has = live && cx - ex >= 1.f + kEdgeMargin && cx + ex < wm2 - kEdgeMargin &&
                         cy - ey >= 1.f + kEdgeMargin && cy + ey < hm2 - kEdgeMargin;
        if (hlane == 0) {
            gbox[e][0] = has ? max((int)(cx - ex - kTileSlack) - 1, 0) : INT_MAX;
            gbox[e][1] = min((int)(cx + ex + kTileSlack) + 2, w - 1);
            gbox[e][2] = max((int)(cy - ey - kTileSlack) - 1, 0);
            gbox[e][3] = min((int)(cy + ey + kTileSlack) + 2, h - 1);
        }
    }
    __syncthreads();
    int gx0 = INT_MAX, gx1 = INT_MIN, gy0 = INT_MAX, gy1 = INT_MIN;
#pragma unroll
    for (int v = 0; v < kPairEntries; v++)
        if (gbox[v][0] != INT_MAX) {
            gx0 = min(gx0, gbox[v][0]);
            gx1 = max(gx1, gbox[v][1]);
            gy0 = min(gy0, gbox[v][2]);
            gy1 = max(gy1, gbox[v][3]);
        }
    int gw = 0, gh = 0;
    if (gx0 <= gx1) {
        gx0 = max(gx0 - kGroupMargin, 0);
        gx1 = min(gx1 + kGroupMargin, w - 1);
        gy0 = max(gy0 - kGroupMargin, 0);
        gy1 = min(gy1 + kGroupMargin, h - 1);
        if (gx1 - gx0 - 2 <= kPairStride && gy1 - gy0 + 1 <= p.tile_h) {
            gw = gx1 - gx0 + 1;
            gh = gy1 - gy0 + 1;
        }
    }
    if (gw > 0) {
        // columns lane and lane + 64 (a tile row holds up to kPairStride entries)
        const int wave = threadIdx.x >> 6;
        stage_coef<kPairEntries / 2, kPairStride>(ct, p.tar, w, gx0, gy0, gw, gh, lane, wave);
        if (gw - 3 > 64)
            stage_coef<kPairEntries / 2, kPairStride>(ct + 64, p.tar, w, gx0 + 64, gy0, gw - 64, gh, lane, wave);
    }
    __syncthreads();

    const float m_all = bd.active ? 1.f : 0.f;
    const float m_dup = bd.last ? 0.f : 1.f;
    const float xl = (float)(bd.col - rx), yl0 = (float)(bd.row0 - ry);
    const int cbase = -(gy0 + 1) * kPairStride - (gx0 + 1);
    const float4* tile = ct;

    int it = 0;
    while (__builtin_amdgcn_ballot_w64(live) != 0) {
        // a finished (or empty) half takes the other half's centre and warp
        if (__builtin_amdgcn_ballot_w64(!live) != 0) {
            const float oxc = other_half(xc, upper), oyc = other_half(yc, upper);
            const float o00 = other_half(W00, upper), o01 = other_half(W01, upper), o02 = other_half(W02, upper);
            const float o10 = other_half(W10, upper), o11 = other_half(W11, upper), o12 = other_half(W12, upper);
            if (!live) {
                xc = oxc; yc = oyc;
                W00 = o00; W01 = o01; W02 = o02; W10 = o10; W11 = o11; W12 = o12;
            }
        }
        it += live ? 1 : 0;
        float xb = xc + __builtin_fmaf(W00, xl, __builtin_fmaf(W01, yl0, W02));
        float yb = yc + __builtin_fmaf(W10, xl, __builtin_fmaf(W11, yl0, W12));
        float ex = __builtin_fmaf(fabsf(W00), rxf, fabsf(W01) * ryf);
        float ey = __builtin_fmaf(fabsf(W10), rxf, fabsf(W11) * ryf);
        float cx = xc + W02, cy = yc + W12;
        const bool box_in = (cx - ex >= 1.f + kEdgeMargin) & (cx + ex < wm2 - kEdgeMargin) &
                            (cy - ey >= 1.f + kEdgeMargin) & (cy + ey < hm2 - kEdgeMargin);
        if (__builtin_amdgcn_ballot_w64(!box_in) != 0) {
            // per-sample test (:290-309): out of range (-3) before NaN (-5)
            bool oob = false, nan = false;
#pragma unroll
            for (int k = 0; k < SPL; k++) {
                const float gxp = __builtin_fmaf((float)k, W01, xb);
                const float gyp = __builtin_fmaf((float)k, W11, yb);
                oob |= (gxp < 1.f) | (gyp < 1.f) | (gxp >= wm2) | (gyp >= hm2);
                nan |= (gxp != gxp) | (gyp != gyp);
            }
            const bool any_oob = half_any(oob, upper), any_nan = half_any(nan, upper);
            const bool fail = live && (any_oob || any_nan);
            if (fail && hlane == 0) {
                poi->result[2] = any_oob ? -3.f : -5.f;
                if (p.iters) p.iters[b] = (uint32_t)it;
            }
            live = live && !fail;
            if (__builtin_amdgcn_ballot_w64(live) == 0) break;
            if (__builtin_amdgcn_ballot_w64(fail) != 0) {
                // the failed half samples at the other half's coordinates
                const float oxc = other_half(xc, upper), oyc = other_half(yc, upper);
                const float o00 = other_half(W00, upper), o01 = other_half(W01, upper), o02 = other_half(W02, upper);
                const float o10 = other_half(W10, upper), o11 = other_half(W11, upper), o12 = other_half(W12, upper);
                if (fail) {
                    xc = oxc; yc = oyc;
                    W00 = o00; W01 = o01; W02 = o02; W10 = o10; W11 = o11; W12 = o12;
                    xb = xc + __builtin_fmaf(W00, xl, __builtin_fmaf(W01, yl0, W02));
                    yb = yc + __builtin_fmaf(W10, xl, __builtin_fmaf(W11, yl0, W12));
                    ex = __builtin_fmaf(fabsf(W00), rxf, fabsf(W01) * ryf);
                    ey = __builtin_fmaf(fabsf(W10), rxf, fabsf(W11) * ryf);
                    cx = xc + W02;
                    cy = yc + W12;
                }
            }
        }
        const int nx0 = max((int)(cx - ex - kTileSlack) - 1, 0);
        const int nx1 = min((int)(cx + ex + kTileSlack) + 2, w - 1);
        const int ny0 = max((int)(cy - ey - kTileSlack) - 1, 0);
        const int ny1 = min((int)(cy + ey + kTileSlack) + 2, h - 1);
        const bool use_tile = gw > 0 && nx0 >= gx0 && nx1 < gx0 + gw && ny0 >= gy0 && ny1 < gy0 + gh;

        // pass 1: t' = t - ref_mean per slot, sums of t', t'^2, t' r
        float tpv[SPL];
        float st = 0.f, stt = 0.f, str = 0.f;
        auto pass1 = [&](auto sample) {
#pragma unroll
            for (int k = 0; k < SPL; k++) {
                const float gxp = __builtin_fmaf((float)k, W01, xb);
                const float gyp = __builtin_fmaf((float)k, W11, yb);
                const float tp = sample(gxp, gyp);
                const float tpm = (k < kMaxOverlap) ? tp * ((k < ovl) ? m_dup : 1.f) : tp;
                tpv[k] = tp;
                st += tpm;
                stt = __builtin_fmaf(tpm, tp, stt);
                str = __builtin_fmaf(tp, rz[k], str);
            }
        };
        if (use_tile)
            pass1([&](float gx, float gy) { return coef_sample<kPairStride>(tile, cbase, gx, gy, -ref_mean); });
        else
            pass1([&](float gx, float gy) { return target_sample(p.tar, (uint32_t)w, gx, gy, -ref_mean); });
        float s1[3] = { st * m_all, stt * m_all, str };
        half_sum3(s1, hb);

        // ZNSSD terms (Subset2D::zeroMeanNorm, :311-322)
        const float tm = s1[0] * inv_ns;
        const float szz = __builtin_fmaf(-s1[0], tm, s1[1]);
        const float inv_tar_norm = __builtin_amdgcn_rsqf(szz);
        const float scale = (szz > 0.f) ? ref_norm * inv_tar_norm : __builtin_nanf("");
        const float zncc = __builtin_fmaf(-tm, ref_sum, s1[2]) * inv_tar_norm * inv_ref_norm;

        // pass 2: e = (t - tbar) * (nr / nt) - r, sums SD_j e; slots in
        // descending order so AK = sum_k k a_k is a running suffix sum
        float A = 0.f, AK = 0.f, B = 0.f, BK = 0.f;
#pragma unroll
        for (int k = SPL - 1; k >= 0; k--) {
            const float e = __builtin_fmaf(scale, tpv[k] - tm, -rz[k]);
            AK += A;
            BK += B;
            A = __builtin_fmaf(gxv[k], e, A);
            B = __builtin_fmaf(gyv[k], e, B);
        }
        float bnum[6] = { A, xl * A, __builtin_fmaf(yl0, A, AK), B, xl * B, __builtin_fmaf(yl0, B, BK) };
        half_sum6(bnum, hb);

        // dp = H^-1 b (:324-331): lane a < 6 of the half forms dp[a]
        float dpl = 0.f;
        {
            const float* hr = hl + 6 * min(hlane, 5);
#pragma unroll
            for (int c = 0; c < 6; c++) dpl = __builtin_fmaf(hr[c], bnum[c], dpl);
        }
        float dp[6];
#pragma unroll
        for (int a = 0; a < 6; a++) dp[a] = lane_fetch(dpl, hb + 4 * a);
        // convergence norm (:343-352) in FP64, against the launcher's threshold
        double cn = (double)dp[0] * dp[0];
        cn += (double)dp[1] * dp[1] * rx2;
        cn += (double)dp[2] * dp[2] * ry2;
        cn += (double)dp[3] * dp[3];
        cn += (double)dp[4] * dp[4] * rx2;
        cn += (double)dp[5] * dp[5] * ry2;

        // W <- W * W(dp)^-1 (:333-340), as in icgn_entry
        const float D00 = 1.f + dp[1], D01 = dp[2], D02 = dp[0];
        const float D10 = dp[4], D11 = 1.f + dp[5], D12 = dp[3];
        if (cn <= __DBL_MAX__) {
            const float c20 = D01 * D12 - D02 * D11, c21 = D02 * D10 - D00 * D12;
            const float det = D11 * D00 + -(D01 * D10);
            const float invdet = __builtin_amdgcn_rcpf(det);
            const float I00 = D11 * invdet, I01 = -D01 * invdet, I02 = c20 * invdet;
            const float I10 = -D10 * invdet, I11 = D00 * invdet, I12 = c21 * invdet;
            const float I22 = (D00 * D11 - D01 * D10) * invdet;
            const float N00 = W00 * I00 + W01 * I10, N01 = W00 * I01 + W01 * I11;
            const float N02 = W00 * I02 + (W01 * I12 + W02 * I22);
            const float N10 = W10 * I00 + W11 * I10, N11 = W10 * I01 + W11 * I11;
            const float N12 = W10 * I02 + (W11 * I12 + W12 * I22);
            W00 = N00; W01 = N01; W02 = N02; W10 = N10; W11 = N11; W12 = N12;
            W22 = W22 * I22;
        } else {
            const float D20 = 0.f, D21 = 0.f, D22 = 1.f;
            const float c00 = D11 * D22 - D12 * D21, c01 = D12 * D20 - D10 * D22, c02 = D10 * D21 - D11 * D20;
            const float c10 = D21 * D02 - D22 * D01, c11 = D22 * D00 - D20 * D02, c12 = D20 * D01 - D21 * D00;
            const float c20 = D01 * D12 - D02 * D11, c21 = D02 * D10 - D00 * D12, c22 = D00 * D11 - D01 * D10;
            const float det = c00 * D00 + (c10 * D10 + c20 * D20);
            const float invdet = __builtin_amdgcn_rcpf(det);
            const float I00 = c00 * invdet, I01 = c10 * invdet, I02 = c20 * invdet;
            const float I10 = c01 * invdet, I11 = c11 * invdet, I12 = c21 * invdet;
            const float I20 = c02 * invdet, I21 = c12 * invdet, I22 = c22 * invdet;
            const float N00 = W00 * I00 + (W01 * I10 + W02 * I20), N01 = W00 * I01 + (W01 * I11 + W02 * I21);
            const float N02 = W00 * I02 + (W01 * I12 + W02 * I22), N10 = W10 * I00 + (W11 * I10 + W12 * I20);
            const float N11 = W10 * I01 + (W11 * I11 + W12 * I21), N12 = W10 * I02 + (W11 * I12 + W12 * I22);
            const float N20 = W20 * I00 + (W21 * I10 + W22 * I20), N21 = W20 * I01 + (W21 * I11 + W22 * I21);
            const float N22 = W20 * I02 + (W21 * I12 + W22 * I22);
            W00 = N00; W01 = N01; W02 = N02; W10 = N10; W11 = N11; W12 = N12; W20 = N20; W21 = N21; W22 = N22;
        }
        const bool fin = live && !((float)it < p.stop && cn >= p.conv2);
        if (__builtin_amdgcn_ballot_w64(fin) != 0) {
            if (fin && hlane == 0) {
                const float conv = (float)sqrt(cn);
                if (p.iters) p.iters[b] = (uint32_t)it;
                float u = W02, ux = W00 - 1.f, uy = W01, v = W12, vx = W10, vy = W11 - 1.f;
                float z = zncc;
                if (conv >= p.conv && (float)it >= p.stop) z = -4.f;
                if (isnan(z) || isnan(u) || isnan(v)) {
                    u = u_in;
                    v = v_in;
                    z = -5.f;
                }
                poi->deformation[0] = u;
                poi->deformation[1] = ux;
                poi->deformation[2] = uy;
                poi->deformation[6] = v;
                poi->deformation[7] = vx;
                poi->deformation[8] = vy;
                poi->result[0] = u_in;
                poi->result[1] = v_in;
                poi->result[2] = z;
                poi->result[3] = (float)it;
                poi->result[4] = conv;
                poi->subset_radius[0] = (float)rx;
                poi->subset_radius[1] = (float)ry;
            }
            live = live && !fin;
        }
    }
}

// kPairEntries consecutive entries per block, two per wave, sharing one
// coefficient tile (icgn_pair_entry).
template <int SPL>
__global__ __launch_bounds__(64 * kPairEntries / 2) __attribute__((amdgpu_waves_per_eu(2)))
void k_icgn_pair(IcgnParams p)
{
    __shared__ int gbox[kPairEntries][4];
    __shared__ float hl[kPairEntries * 36];
    extern __shared__ float4 ctile[];
    const int e = 2 * (int)(threadIdx.x >> 6) + (int)((threadIdx.x & 63) >> 5);
    const int b = xcd_remap(blockIdx.x, gridDim.x) * kPairEntries + e;
    icgn_pair_entry<SPL>(p, b, e, hl + 36 * e, ctile, gbox);
}

template <int SPL>
void launch_icgn_pair_t(const IcgnParams& p, hipStream_t s)
{
    const size_t lds = (size_t)p.tile_h * kPairStride * sizeof(float4);
    hipLaunchKernelGGL((k_icgn_pair<SPL>), dim3((p.n + kPairEntries - 1) / kPairEntries),
                       dim3(64 * kPairEntries / 2), lds, s, p);
}

// Instantiated pair shapes (rows per band of a 32-lane entry).
#define FI_ICGN_PAIR(X) X(31)

// Instantiated lane shapes (waves, rows per band).  EXACT_A/B are compiled in
// k_icgn_xa.hip / k_icgn_xb.hip, GENERIC in k_icgn.hip.
#define FI_ICGN_EXACT_A(X) \
    X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 6) X(1, 7) X(1, 12) X(1, 13) X(1, 14) X(1, 15) X(1, 16)
#define FI_ICGN_EXACT_B(X)                                                      \
    X(2, 4) X(2, 8) X(2, 11) X(2, 12) X(2, 13) X(2, 14)                          \
    X(4, 2) X(4, 4) X(4, 7) X(4, 9) X(4, 10) X(4, 11) X(4, 14) X(4, 15) X(4, 16)
#define FI_ICGN_GENERIC(X) X(1, 16) X(2, 16) X(4, 16)

template <int WAVES, int SPL, bool EXACT>
void launch_icgn_t(const IcgnParams& p, hipStream_t s)
{
    if constexpr (WAVES == 1) {
        if (p.group) {
            // p.tile_h: coefficient rows per block
            const size_t glds = (size_t)p.tile_h * kGroupStride * sizeof(float4);
            hipLaunchKernelGGL((k_icgn_group<SPL, EXACT>), dim3((p.n + kGroupPois - 1) / kGroupPois),
                               dim3(64 * kGroupPois), glds, s, p);
            return;
        }
    }
    const size_t lds = p.tile_w > 0 ? (size_t)p.tile_stride * p.tile_h * sizeof(float) : 0;
    if (p.queue && !p.count)
        hipLaunchKernelGGL((k_icgn_queue<WAVES, SPL, EXACT>), dim3(std::max(1, std::min(p.n, p.queue_blocks))),
                           dim3(64 * WAVES), lds, s, p);
    else
        hipLaunchKernelGGL((k_icgn<WAVES, SPL, EXACT>), dim3(p.n), dim3(64 * WAVES), lds, s, p);
}

// Centre offsets (p.offset): the generic shape of the same width, one block
// per entry.
template <int WAVES>
void launch_icgn_offset_t(const IcgnParams& p, hipStream_t s)
{
    const size_t lds = p.tile_w > 0 ? (size_t)p.tile_stride * p.tile_h * sizeof(float) : 0;
    hipLaunchKernelGGL((k_icgn<WAVES, 16, false, true>), dim3(p.n), dim3(64 * WAVES), lds, s, p);
}

// Blocks of k_icgn_queue<WAVES, SPL, EXACT> resident on one CU at p's tile
// size (register, LDS and wave limits, as the runtime computes them); 0 on
// error.
template <int WAVES, int SPL, bool EXACT>
int icgn_queue_per_cu_t(const IcgnParams& p)
{
    const size_t lds = p.tile_w > 0 ? (size_t)p.tile_stride * p.tile_h * sizeof(float) : 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void*>(&k_icgn_queue<WAVES, SPL, EXACT>), 64 * WAVES, lds) != hipSuccess)
        return 0;
    return nb;
}

}  // namespace fi
