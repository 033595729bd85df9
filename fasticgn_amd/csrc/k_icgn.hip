// ICGN2D1 on gfx950: one wavefront (64 lanes) per POI.
//
//   k_icgn_prep  once per reference layout: reference subset mean and norm,
//                steepest-descent Hessian (compensated FP32 sums), FP64
//                partial-pivot inversion, stored as a 160-byte PoiPrep record.
//   k_icgn       once per frame: the inverse-compositional Gauss-Newton loop
//                (warp, bicubic B-spline sample, ZNSSD, numerator, dp, warp
//                update, convergence) with compensated FP32 reductions and the
//                convergence norm in FP64.
//
// Semantics follow the CPU reference ICGN2D1::compute(POI2D*),
// src/oc_icgn.cpp:168-387: the entry gate (-3), the geometric out-of-range
// predicate of :39-43 (-3, deformation keeps the incoming guess), a NaN sample
// coordinate (-5), not converged within the iteration cap (-4), NaN result
// (-5 with u, v reset to the initial guess).  The reference GPU's extra gates
// (margin-2 reference bound, zero-texture -3, |dp| > 100) are deliberately not
// applied: the CPU path is the one this is judged against.
//
// Sample layout: the (2rx+1) x (2ry+1) subset is a flat row-major list; lane l
// owns samples l, l+64, l+128, ...  With CACHE the reference value, gradients
// and local coordinates of those samples live in VGPRs for the whole solve;
// without it (large radii) they are reloaded from L2 in the passes that need
// them.
#include <hip/hip_runtime.h>

#include "fi_device.h"
#include "fi_launch.h"

namespace fi {

struct SubsetGeom {
    int rx, ry, sw, sh, ns;
    uint32_t div_magic;  // ceil(2^32 / sw): row = umulhi(s, magic)
};

__device__ __forceinline__ SubsetGeom make_geom(int rx, int ry)
{
    SubsetGeom g;
    g.rx = rx;
    g.ry = ry;
    g.sw = 2 * rx + 1;
    g.sh = 2 * ry + 1;
    g.ns = g.sw * g.sh;
    g.div_magic = (uint32_t)((0x100000000ull + (unsigned long long)g.sw - 1) / (unsigned long long)g.sw);
    return g;
}

__device__ __forceinline__ void sample_coords(const SubsetGeom& g, int s, int& c, int& r)
{
    r = (int)__umulhi((uint32_t)s, g.div_magic);
    c = s - r * g.sw;
}

// entry gate of src/oc_icgn.cpp:184-186 (reference side; the deformation and
// zncc parts are checked by the caller) plus finite coordinates, which keep
// (int) conversions in bounds
__device__ __forceinline__ bool ref_gate_ok(float x, float y, int rx, int ry, int w, int h)
{
    if (!(isfinite(x) && isfinite(y))) return false;
    return !(y - ry < 0 || x - rx < 0 || y + ry > h - 1 || x + rx > w - 1);
}

// ---------------------------------------------------------------------------
// FP64 6x6 inverse, partial-pivot LU with the same structure as the CPU
// reference's Eigen PartialPivLU (largest |a_ik|, first on ties; a zero pivot
// column is skipped so a singular Hessian yields inf/NaN exactly as there).
// ---------------------------------------------------------------------------
__device__ void inverse6_f64(const double* a, double* inv)
{
    double m[6][6];
    int perm[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        perm[i] = i;
#pragma unroll
        for (int j = 0; j < 6; j++) m[i][j] = a[i * 6 + j];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int piv = k;
        double best = fabs(m[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const double s = fabs(m[i][k]);
            if (s > best) { best = s; piv = i; }
        }
        if (best != 0.0) {
            if (piv != k) {
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    // select-based swap keeps the matrix in registers
#pragma unroll
                    for (int i = k + 1; i < 6; i++)
                        if (i == piv) { const double t = m[k][j]; m[k][j] = m[i][j]; m[i][j] = t; }
                }
#pragma unroll
                for (int i = k + 1; i < 6; i++)
                    if (i == piv) { const int t = perm[k]; perm[k] = perm[i]; perm[i] = t; }
            }
#pragma unroll
            for (int i = k + 1; i < 6; i++) m[i][k] /= m[k][k];
        }
#pragma unroll
        for (int i = k + 1; i < 6; i++)
#pragma unroll
            for (int j = k + 1; j < 6; j++) m[i][j] -= m[i][k] * m[k][j];
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        double x[6];
#pragma unroll
        for (int i = 0; i < 6; i++) x[i] = (perm[i] == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 6; k++)
#pragma unroll
            for (int i = k + 1; i < 6; i++) x[i] -= m[i][k] * x[k];
#pragma unroll
        for (int k = 5; k >= 0; k--) {
            x[k] /= m[k][k];
#pragma unroll
            for (int i = 0; i < k; i++) x[i] -= m[i][k] * x[k];
        }
#pragma unroll
        for (int i = 0; i < 6; i++) inv[i * 6 + c] = x[i];
    }
}

// ---------------------------------------------------------------------------
// Reference-side prep: mean, norm, Hessian, inverse.
// ---------------------------------------------------------------------------
// Three streaming passes over the subset (sum, centred sum of squares,
// Hessian); samples are re-read from L1/L2 instead of held in VGPRs, so one
// kernel serves every radius.
__global__ __launch_bounds__(64) void k_icgn_prep(IcgnParams p)
{
    const int lane = threadIdx.x;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= p.n) return;
    const int i = p.index ? (int)p.index[b] : b;
    const fi_poi* poi = p.pois + i;
    const float x = poi->x, y = poi->y;
    int rx = p.rx, ry = p.ry;
    if (p.self_adaptive) {
        rx = (int)poi->subset_radius[0];
        ry = (int)poi->subset_radius[1];
    }
    PoiPrep* out = p.prep + i;
    const int w = p.w, h = p.h;
    if (!ref_gate_ok(x, y, rx, ry, w, h)) {
        if (lane == 0) { out->status = 0; out->rx = rx; }
        return;
    }
    const SubsetGeom g = make_geom(rx, ry);
    const int ulx = (int)(x - rx), uly = (int)(y - ry);  // Subset2D::fill, src/oc_subset.cpp:39-44
    const int px = (int)x, py = (int)y;                   // gradients at (int)x + local, :206-210

    CSum sum_r;
    for (int s = lane; s < g.ns; s += 64) {
        int c, r;
        sample_coords(g, s, c, r);
        sum_r.add(p.ref[(size_t)(uly + r) * w + (ulx + c)]);
    }
    const float mean = wave_sum_comp(sum_r) / (float)g.ns;
    CSum sq;
    for (int s = lane; s < g.ns; s += 64) {
        int c, r;
        sample_coords(g, s, c, r);
        const float d = p.ref[(size_t)(uly + r) * w + (ulx + c)] - mean;
        sq.add(d * d);
    }
    const float norm = sqrtf(wave_sum_comp(sq));

    CSum hs[21];
    for (int s = lane; s < g.ns; s += 64) {
        int c, r;
        sample_coords(g, s, c, r);
        const size_t gi = (size_t)(py + r - ry) * w + (px + c - rx);
        const float gxv = p.gx[gi], gyv = p.gy[gi];
        const float xl = (float)(c - rx), yl = (float)(r - ry);
        const float sd[6] = { gxv, gxv * xl, gxv * yl, gyv, gyv * xl, gyv * yl };
        int q = 0;
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int e = 0; e <= a; e++) hs[q++].add(sd[a] * sd[e]);
    }
    double H[36];
    {
        int q = 0;
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int e = 0; e <= a; e++) {
                const double v = (double)wave_sum_comp(hs[q++]);
                H[a * 6 + e] = v;
                H[e * 6 + a] = v;
            }
    }
    double Hi[36];
    inverse6_f64(H, Hi);
    if (lane == 0) {
        out->mean = mean;
        out->norm = norm;
        out->status = 1;
        out->rx = rx;
        float4* dst = reinterpret_cast<float4*>(out->hinv);
#pragma unroll
        for (int q = 0; q < 9; q++)
            dst[q] = make_float4((float)Hi[4 * q], (float)Hi[4 * q + 1], (float)Hi[4 * q + 2],
                                 (float)Hi[4 * q + 3]);
    }
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

// The coefficient table is read through a raw buffer resource: one 32-bit
// byte offset per sample, and the hardware range check returns zeros for any
// offset past the table, so a sample outside the image (whose value is
// discarded) needs no clamp and can never fault.
constexpr uint32_t kBufferWord3 = 0x00020000;  // gfx9-family raw buffer, dword data

__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const float* coef, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(coef), (short)0, (int)bytes, (int)kBufferWord3);
}

__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
}

// The table is four planes of float4 (plane k = coef[k][0..3] of every texel);
// plane_bytes = 16 * w * h.  An offset past a plane's end reads the next plane
// or, past the table, zeros: either way a value that is discarded.
__device__ __forceinline__ float bspline_sample(__amdgpu_buffer_rsrc_t r, int w, uint32_t plane_bytes,
                                                float gx, float gy)
{
    const float fx = floorf(gx), fy = floorf(gy);
    const float dx = gx - fx, dy = gy - fy;
    const uint32_t off = ((uint32_t)((int)fy * w + (int)fx)) << 4;
    return bspline_eval(ld4(r, off, 0), ld4(r, off, plane_bytes), ld4(r, off, 2 * plane_bytes),
                        ld4(r, off, 3 * plane_bytes), dx, dy);
}

// ---------------------------------------------------------------------------
// IC-GN iteration.  One block of WAVES wavefronts per POI; thread t owns
// subset samples t, t + 64*WAVES, ... (SPL of them), with the reference value,
// gradients and local coordinates held in VGPRs for the whole solve.  Sums are
// per-lane FMA chains, a compensated (TwoSum) DPP tree inside each wave and,
// for WAVES > 1, a compensated combine of the per-wave results through LDS in a
// fixed order, so every wave holds bit-identical totals and takes the same
// branches.
// ---------------------------------------------------------------------------

// Minimum resident waves per SIMD the register allocation must allow (the
// compiler otherwise trades occupancy for deeper load pipelining).
#ifndef FI_ICGN_MIN_WAVES
#define FI_ICGN_MIN_WAVES 3
#endif

template <int WAVES>
struct BlockSum {
    // per wave: NV compensated pairs
    static constexpr int NV = 7;
    float* lds;
    __device__ __forceinline__ void reduce(float* vals, int nv, int wave, int lane)
    {
        // vals[i] arrive as per-lane partial sums
        CSum part[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) {
            if (i < nv) {
                // wave_sum_comp returns s + c; keep both for the cross-wave step
                CSum c;
                c.s = vals[i];
                c.c = 0.f;
                part[i].s = wave_sum_comp(c);
                part[i].c = 0.f;
            }
        }
        if (WAVES == 1) {
#pragma unroll
            for (int i = 0; i < NV; i++)
                if (i < nv) vals[i] = part[i].s;
            return;
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++)
                if (i < nv) lds[wave * NV + i] = part[i].s;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; i++) {
            if (i < nv) {
                CSum acc;
#pragma unroll
                for (int w = 0; w < WAVES; w++) acc.add(lds[w * NV + i]);
                vals[i] = acc.s + acc.c;
            }
        }
        __syncthreads();
    }
};

template <int WAVES, int SPL>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(FI_ICGN_MIN_WAVES)))
void k_icgn(IcgnParams p)
{
    constexpr int NT = 64 * WAVES;
    __shared__ float red[WAVES > 1 ? WAVES * BlockSum<WAVES>::NV : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    BlockSum<WAVES> bs{ red };
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= p.n) return;
    const int i = p.index ? (int)p.index[b] : b;
    fi_poi* poi = p.pois + i;
    const float x = poi->x, y = poi->y;
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
        if (tid == 0) poi->result[2] = (z_in >= 0.f) ? -3.f : z_in;
        return;
    }
    const PoiPrep* pr = p.prep + (p.prep_index ? p.prep_index[i] : (uint32_t)i);
    const float ref_mean = pr->mean, ref_norm = pr->norm;
    float hinv[36];
#pragma unroll
    for (int q = 0; q < 36; q++) hinv[q] = pr->hinv[q];

    const SubsetGeom g = make_geom(rx, ry);
    const int ulx = (int)(x - rx), uly = (int)(y - ry);
    const int px = (int)x, py = (int)y;

    // Slots past the subset (s >= ns) sample the centre (xl = yl = 0) with zero
    // reference data and are masked out of the sums: no per-slot branch.
    float rz[SPL], gxv[SPL], gyv[SPL], xlv[SPL], ylv[SPL];
#pragma unroll
    for (int k = 0; k < SPL; k++) {
        const int s = tid + NT * k;
        rz[k] = gxv[k] = gyv[k] = xlv[k] = ylv[k] = 0.f;
        if (s < g.ns) {
            int c, r;
            sample_coords(g, s, c, r);
            rz[k] = p.ref[(size_t)(uly + r) * w + (ulx + c)] - ref_mean;
            const size_t gi = (size_t)(py + r - ry) * w + (px + c - rx);
            gxv[k] = p.gx[gi];
            gyv[k] = p.gy[gi];
            xlv[k] = (float)(c - rx);
            ylv[k] = (float)(r - ry);
        }
    }
    const __amdgpu_buffer_rsrc_t table = table_rsrc(p.coef, (uint32_t)((size_t)w * h * 64));
    const uint32_t plane_bytes = (uint32_t)((size_t)w * h * 16);

    // warp matrix, Deformation2D1::setWarp (src/oc_deformation.cpp:117-128)
    float W00 = 1.f + ux_in, W01 = uy_in, W02 = u_in;
    float W10 = vx_in, W11 = 1.f + vy_in, W12 = v_in;
    float W20 = 0.f, W21 = 0.f, W22 = 1.f;
    const float wm2 = (float)(w - 2), hm2 = (float)(h - 2);
    const float rx2 = (float)(rx * rx), ry2 = (float)(ry * ry);
    const float conv_crit = p.conv, stop = p.stop;

    float t[SPL];
    int it = 0;
    float conv = 0.f, znssd = 0.f;
    do {
        it++;
        // pass 1: warp, range test, B-spline sample, sum of t (two
        // interleaved per-lane accumulators keep the FMA chains short)
        bool outside = false;
        float sa = 0.f, sb = 0.f;
#pragma unroll
        for (int k = 0; k < SPL; k++) {
            const bool slot = tid + NT * k < g.ns;
            const float gxp = x + __builtin_fmaf(W00, xlv[k], __builtin_fmaf(W01, ylv[k], W02));
            const float gyp = y + __builtin_fmaf(W10, xlv[k], __builtin_fmaf(W11, ylv[k], W12));
            // false for a NaN coordinate as well as for an out-of-range one
            const bool in = (gxp >= 1.f) & (gyp >= 1.f) & (gxp < wm2) & (gyp < hm2);
            outside |= !in;
            const float v = bspline_sample(table, w, plane_bytes, gxp, gyp);
            t[k] = (in & slot) ? v : 0.f;
            if (k & 1) sb += t[k]; else sa += t[k];
        }
        // first reduction: the sum of t and the number of lanes holding an
        // unusable sample (a count, so the compensated sum is exact)
        float v2[2];
        {
            float a, e;
            two_sum(sa, sb, a, e);
            v2[0] = a + e;
            v2[1] = outside ? 1.f : 0.f;
        }
        bs.reduce(v2, 2, wave, lane);
        if (v2[1] > 0.f) {
            // Some sample is unusable: out of the interpolatable range (-3,
            // :39-43) takes precedence over a NaN coordinate (-5), :290-309.
            bool oob = false;
#pragma unroll
            for (int k = 0; k < SPL; k++) {
                const float gxp = x + __builtin_fmaf(W00, xlv[k], __builtin_fmaf(W01, ylv[k], W02));
                const float gyp = y + __builtin_fmaf(W10, xlv[k], __builtin_fmaf(W11, ylv[k], W12));
                oob |= (gxp < 1.f) || (gyp < 1.f) || (gxp >= wm2) || (gyp >= hm2);
            }
            float v1[1] = { oob ? 1.f : 0.f };
            bs.reduce(v1, 1, wave, lane);
            if (tid == 0) {
                poi->result[2] = (v1[0] > 0.f) ? -3.f : -5.f;
                if (p.iter_total) atomicAdd(p.iter_total, (unsigned long long)it);
            }
            return;
        }
        float v1[1] = { v2[0] };
        const float tar_mean = v1[0] / (float)g.ns;
        // pass 2: centre, sum of squares (Subset2D::zeroMeanNorm, two-pass)
        float qa = 0.f, qb = 0.f;
#pragma unroll
        for (int k = 0; k < SPL; k++) {
            t[k] = (tid + NT * k < g.ns) ? t[k] - tar_mean : 0.f;
            if (k & 1) qb = __builtin_fmaf(t[k], t[k], qb); else qa = __builtin_fmaf(t[k], t[k], qa);
        }
        {
            float a, e;
            two_sum(qa, qb, a, e);
            v1[0] = a + e;
        }
        bs.reduce(v1, 1, wave, lane);
        const float tar_norm = sqrtf(v1[0]);
        const float scale = ref_norm / tar_norm;

        // pass 3: error e = t * (nr / nt) - r, ZNSSD numerator and the six
        // steepest-descent products (:311-322); padding slots give e = 0
        float acc[7] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int k = 0; k < SPL; k++) {
            const float e = __builtin_fmaf(t[k], scale, -rz[k]);
            acc[0] = __builtin_fmaf(e, e, acc[0]);
            const float ge = gxv[k] * e, he = gyv[k] * e;
            acc[1] += ge;
            acc[2] = __builtin_fmaf(ge, xlv[k], acc[2]);
            acc[3] = __builtin_fmaf(ge, ylv[k], acc[3]);
            acc[4] += he;
            acc[5] = __builtin_fmaf(he, xlv[k], acc[5]);
            acc[6] = __builtin_fmaf(he, ylv[k], acc[6]);
        }
        bs.reduce(acc, 7, wave, lane);
        znssd = acc[0] / (ref_norm * ref_norm);
        const float* bnum = acc + 1;

        // dp = H^-1 b (src/oc_icgn.cpp:324-331)
        float dp[6];
#pragma unroll
        for (int a = 0; a < 6; a++) {
            float d = 0.f;
#pragma unroll
            for (int c = 0; c < 6; c++) d = __builtin_fmaf(hinv[a * 6 + c], bnum[c], d);
            dp[a] = d;
        }
        // W <- W * W(dp)^-1 (:333-340); the 3x3 inverse by cofactors down column 0
        const float D00 = 1.f + dp[1], D01 = dp[2], D02 = dp[0];
        const float D10 = dp[4], D11 = 1.f + dp[5], D12 = dp[3];
        const float D20 = 0.f, D21 = 0.f, D22 = 1.f;
        const float c00 = D11 * D22 - D12 * D21, c01 = D12 * D20 - D10 * D22, c02 = D10 * D21 - D11 * D20;
        const float c10 = D21 * D02 - D22 * D01, c11 = D22 * D00 - D20 * D02, c12 = D20 * D01 - D21 * D00;
        const float c20 = D01 * D12 - D02 * D11, c21 = D02 * D10 - D00 * D12, c22 = D00 * D11 - D01 * D10;
        const float det = c00 * D00 + (c10 * D10 + c20 * D20);
        const float invdet = 1.f / det;
        const float I00 = c00 * invdet, I01 = c10 * invdet, I02 = c20 * invdet;
        const float I10 = c01 * invdet, I11 = c11 * invdet, I12 = c21 * invdet;
        const float I20 = c02 * invdet, I21 = c12 * invdet, I22 = c22 * invdet;
        const float N00 = W00 * I00 + (W01 * I10 + W02 * I20), N01 = W00 * I01 + (W01 * I11 + W02 * I21);
        const float N02 = W00 * I02 + (W01 * I12 + W02 * I22), N10 = W10 * I00 + (W11 * I10 + W12 * I20);
        const float N11 = W10 * I01 + (W11 * I11 + W12 * I21), N12 = W10 * I02 + (W11 * I12 + W12 * I22);
        const float N20 = W20 * I00 + (W21 * I10 + W22 * I20), N21 = W20 * I01 + (W21 * I11 + W22 * I21);
        const float N22 = W20 * I02 + (W21 * I12 + W22 * I22);
        W00 = N00; W01 = N01; W02 = N02; W10 = N10; W11 = N11; W12 = N12; W20 = N20; W21 = N21; W22 = N22;

        // convergence norm (:343-352), accumulated in FP64
        double cn = (double)dp[0] * dp[0];
        cn += (double)dp[1] * dp[1] * rx2;
        cn += (double)dp[2] * dp[2] * ry2;
        cn += (double)dp[3] * dp[3];
        cn += (double)dp[4] * dp[4] * rx2;
        cn += (double)dp[5] * dp[5] * ry2;
        conv = (float)sqrt(cn);
    } while ((float)it < stop && conv >= conv_crit);

    if (tid == 0) {
        if (p.iter_total) atomicAdd(p.iter_total, (unsigned long long)it);
        float u = W02, ux = W00 - 1.f, uy = W01, v = W12, vx = W10, vy = W11 - 1.f;
        float zncc = 0.5f * (2.f - znssd);
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

// Lane-to-POI shapes: (waves per POI, samples per lane).  The default picks
// the fewest waves that keep SPL <= 16; an explicit shape (autotuner) must
// cover the subset.
#define FI_ICGN_SHAPES(X)                                                        \
    X(1, 1) X(1, 2) X(1, 4) X(1, 7) X(1, 10) X(1, 16) X(2, 8) X(2, 14) X(2, 16)   \
    X(4, 4) X(4, 7) X(4, 10) X(4, 16)

bool icgn_shape_supported(int waves, int spl)
{
#define X(W, S) if (waves == W && spl == S) return true;
    FI_ICGN_SHAPES(X)
#undef X
    return false;
}

void icgn_default_shape(int ns, int& waves, int& spl)
{
    const int cand[][2] = { { 1, 1 }, { 1, 2 }, { 1, 4 }, { 1, 7 }, { 1, 10 }, { 1, 16 },
                            { 2, 14 }, { 2, 16 }, { 4, 10 }, { 4, 16 } };
    for (auto& c : cand)
        if (64 * c[0] * c[1] >= ns) { waves = c[0]; spl = c[1]; return; }
    waves = 4;
    spl = 16;
}

void launch_icgn_prep(const IcgnParams& p, hipStream_t s)
{
    if (p.n <= 0) return;
    hipLaunchKernelGGL(k_icgn_prep, dim3(p.n), dim3(64), 0, s, p);
}

void launch_icgn_shape(const IcgnParams& p, int waves, int spl, hipStream_t s)
{
    if (p.n <= 0) return;
    const dim3 grid(p.n);
#define X(W, S)                                                                 \
    if (waves == W && spl == S) {                                               \
        hipLaunchKernelGGL((k_icgn<W, S>), grid, dim3(64 * W), 0, s, p);        \
        return;                                                                 \
    }
    FI_ICGN_SHAPES(X)
#undef X
}

void launch_icgn(const IcgnParams& p, hipStream_t s)
{
    int waves, spl;
    icgn_default_shape((2 * p.rx + 1) * (2 * p.ry + 1), waves, spl);
    launch_icgn_shape(p, waves, spl, s);
}

}  // namespace fi
