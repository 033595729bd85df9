// ICGN2D1 on gfx950: the per-POI reference records and the IC-GN launch
// dispatch over lane shapes.
//
//   k_icgn_prep  once per reference layout: reference subset mean and norm,
//                steepest-descent Hessian (compensated FP32 sums), FP64
//                partial-pivot inversion and sum (ref - mean); stored as a
//                176-byte PoiPrep record.
//   k_icgn       once per frame (k_icgn_impl.h): the inverse-compositional
//                Gauss-Newton loop, one block per POI.
//
// Semantics follow src/oc_icgn.cpp:168-387 (see k_icgn_impl.h).
#include <hip/hip_runtime.h>

#include "k_icgn_impl.h"

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
    // centre offset (src/oc_icgn.cpp:446-447): steepest-descent local coordinates
    const float ox = p.offset ? p.offset[2 * (size_t)i] : 0.f, oy = p.offset ? p.offset[2 * (size_t)i + 1] : 0.f;
    if (!ref_gate_ok(x, y, rx, ry, w, h)) {
        if (lane == 0) { out->status = 0; out->rx = rx; out->ry = ry; }
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
    CSum sq, rs;
    for (int s = lane; s < g.ns; s += 64) {
        int c, r;
        sample_coords(g, s, c, r);
        const float d = p.ref[(size_t)(uly + r) * w + (ulx + c)] - mean;
        sq.add(d * d);
        rs.add(d);
    }
    const float norm = sqrtf(wave_sum_comp(sq));
    const float ref_sum = wave_sum_comp(rs);

    CSum hs[21];
    for (int s = lane; s < g.ns; s += 64) {
        int c, r;
        sample_coords(g, s, c, r);
        const size_t gi = (size_t)(py + r - ry) * w + (px + c - rx);
        const float gxv = p.gx[gi], gyv = p.gy[gi];
        const float xl = (float)(c - rx) - ox, yl = (float)(r - ry) - oy;
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
        out->ry = ry;
        float4* dst = reinterpret_cast<float4*>(out->hinv);
#pragma unroll
        for (int q = 0; q < 9; q++)
            dst[q] = make_float4((float)Hi[4 * q], (float)Hi[4 * q + 1], (float)Hi[4 * q + 2],
                                 (float)Hi[4 * q + 3]);
        out->ref_sum = ref_sum;
    }
}

// ---------------------------------------------------------------------------
// Lane shapes.  A shape (waves, spl) runs a (2rx+1) x (2ry+1) subset when
// G = 64 * waves / (2rx+1) >= 1 bands of R = ceil((2ry+1) / G) <= spl rows fit.
// Exact shapes (R == spl, every slot a sample) exist for every square radius
// 1..31 at the fewest waves, plus alternatives for radii 10, 15 and 20; the
// generic ones (spl = 16, slots past R skipped) cover the rest.
// ---------------------------------------------------------------------------
#define X(W, S)                                                                 \
    extern template void launch_icgn_t<W, S, true>(const IcgnParams&, hipStream_t); \
    extern template int icgn_queue_per_cu_t<W, S, true>(const IcgnParams&);
FI_ICGN_EXACT_A(X)
FI_ICGN_EXACT_B(X)
#undef X

static bool exact_shape(int waves, int spl)
{
#define X(W, S) if (waves == W && spl == S) return true;
    FI_ICGN_EXACT_A(X)
    FI_ICGN_EXACT_B(X)
#undef X
    return false;
}

static bool generic_shape(int waves, int spl)
{
#define X(W, S) if (waves == W && spl == S) return true;
    FI_ICGN_GENERIC(X)
#undef X
    return false;
}

// Rows per band of the shape for the subset, 0 if it does not fit (no band,
// or a last-band overlap beyond what the kernel masks).
static int band_rows(int waves, int rx, int ry)
{
    const int sh = 2 * ry + 1;
    const int G = 64 * waves / (2 * rx + 1);
    if (G < 1) return 0;
    const int R = (sh + G - 1) / G;
    const int gu = (sh + R - 1) / R;
    return gu * R - sh <= kMaxOverlap ? R : 0;
}

bool icgn_shape_supported(int waves, int spl) { return exact_shape(waves, spl) || generic_shape(waves, spl); }

bool icgn_shape_covers(int waves, int spl, int rx, int ry)
{
    const int R = band_rows(waves, rx, ry);
    if (R == 0 || R > spl) return false;
    return (R == spl && exact_shape(waves, spl)) || generic_shape(waves, spl);
}

int icgn_shape_candidates(int rx, int ry, int* waves, int* spl, int cap)
{
    int m = 0;
    auto add = [&](int W, int S) {
        if (!icgn_shape_covers(W, S, rx, ry)) return;
        for (int i = 0; i < m; i++)
            if (waves[i] == W && spl[i] == S) return;
        if (m < cap) {
            waves[m] = W;
            spl[m] = S;
            m++;
        }
    };
#define X(W, S) add(W, S);
    FI_ICGN_EXACT_A(X)
    FI_ICGN_EXACT_B(X)
    FI_ICGN_GENERIC(X)
#undef X
    return m;
}

void icgn_default_shape(int rx, int ry, int& waves, int& spl)
{
    for (int W : { 1, 2, 4 }) {
        const int R = band_rows(W, rx, ry);
        if (R == 0 || R > 16) continue;
        waves = W;
        spl = exact_shape(W, R) ? R : 16;
        return;
    }
    waves = spl = 0;
}

void launch_icgn_prep(const IcgnParams& p, hipStream_t s)
{
    if (p.n <= 0) return;
    hipLaunchKernelGGL(k_icgn_prep, dim3(p.n), dim3(64), 0, s, p);
}

// The per-POI LDS tile: the subset's neighbourhood (2r+1 + 3 pixels) plus
// kTileMargin on each side, rows kTileStride words apart, when the image is
// that large, the tile at most kTileStride wide (radius <= 28) and the f32
// target is given; otherwise the kernel samples the B-spline table.
// (float)sqrt(s2) >= c, the reference's loop test on the FP64 squared norm
// s2 (src/oc_icgn.cpp:343-352), as one FP64 compare s2 >= conv_threshold2(c):
// for c > 0, (float)sqrt(s2) >= c exactly when sqrt(s2) reaches the midpoint
// m of c and the float below it (an exact tie aside), and m, with at most 25
// significant bits, squares exactly in FP64.  c = 0 passes every norm, c < 0
// every norm but NaN, c = NaN none.
double conv_threshold2(float c)
{
    if (c != c) return (double)c;
    if (c < 0.f) return -HUGE_VAL;
    if (c == 0.f) return 0.0;
    const double m = 0.5 * ((double)nextafterf(c, 0.f) + (double)c);
    return m * m;
}

static IcgnParams with_tile(const IcgnParams& p)
{
    IcgnParams q = p;
    q.conv2 = conv_threshold2(p.conv);
    q.tile_w = q.tile_h = q.tile_stride = 0;
    if (p.tar == nullptr) return q;
    const int tw = 2 * p.rx + 4 + 2 * kTileMargin, th = 2 * p.ry + 4 + 2 * kTileMargin;
    if (tw > p.w || th > p.h || tw > kTileStride) return q;
    q.tile_w = tw;
    q.tile_h = th;
    q.tile_stride = kTileStride;
    return q;
}

bool launch_icgn_shape(const IcgnParams& p0, int waves, int spl, hipStream_t s)
{
    IcgnParams p = with_tile(p0);
    if (p.offset) {
        // centre offsets: the generic shape of this width covers the subset
        const int R = band_rows(waves, p.rx, p.ry);
        if (R == 0 || R > spl || spl > 16) return false;
        if (p.n <= 0) return true;
        p.group = 0;
        p.queue = nullptr;
        if (waves == 1) launch_icgn_offset_t<1>(p, s);
        else if (waves == 2) launch_icgn_offset_t<2>(p, s);
        else if (waves == 4) launch_icgn_offset_t<4>(p, s);
        else return false;
        return true;
    }
    if (p.group) {
        // shared coefficient tiles serve one-wave shapes on neighbouring
        // entries; anything else runs with a tile per entry
        if (waves != 1 || p.tar == nullptr || p.count)
            p.group = 0;
        else
            p.tile_h = 2 * p.ry + 4 + 2 * kGroupMargin + kGroupRowSlack;
    }
    const int R = band_rows(waves, p.rx, p.ry);
    if (R == 0 || R > spl) return false;
    if (p.n <= 0) return true;
    if (R == spl && exact_shape(waves, spl)) {
#define X(W, S)                                                                 \
        if (waves == W && spl == S) {                                           \
            launch_icgn_t<W, S, true>(p, s);                                    \
            return true;                                                        \
        }
        FI_ICGN_EXACT_A(X)
        FI_ICGN_EXACT_B(X)
#undef X
    }
#define X(W, S)                                                                 \
    if (waves == W && spl == S) {                                               \
        launch_icgn_t<W, S, false>(p, s);                                       \
        return true;                                                            \
    }
    FI_ICGN_GENERIC(X)
#undef X
    return false;
}

int icgn_queue_per_cu(const IcgnParams& p0, int waves, int spl)
{
    const IcgnParams p = with_tile(p0);
    const int R = band_rows(waves, p.rx, p.ry);
    if (R == 0 || R > spl) return 0;
    if (R == spl && exact_shape(waves, spl)) {
#define X(W, S)                                                                 \
        if (waves == W && spl == S) return icgn_queue_per_cu_t<W, S, true>(p);
        FI_ICGN_EXACT_A(X)
        FI_ICGN_EXACT_B(X)
#undef X
    }
#define X(W, S)                                                                 \
    if (waves == W && spl == S) return icgn_queue_per_cu_t<W, S, false>(p);
    FI_ICGN_GENERIC(X)
#undef X
    return 0;
}

bool launch_icgn(const IcgnParams& p, hipStream_t s)
{
    int waves, spl;
    icgn_default_shape(p.rx, p.ry, waves, spl);
    return waves != 0 && launch_icgn_shape(p, waves, spl, s);
}

}  // namespace fi
