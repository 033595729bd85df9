// ICGN2D1 on gfx950: the per-POI reference records and the IC-GN launch
// dispatch over lane shapes.
//
//   k_icgn_prep  once per reference layout, one wave per POI: reference
//                subset mean and norm, the 21 steepest-descent Hessian sums
//                (compensated FP32) and sum (ref - mean), into a 176-byte
//                PoiPrep record;
//   k_prep_inverse then one lane per POI: FP64 partial-pivot inversion of
//                the Hessian, stored FP32 in the record.
//   k_icgn       once per frame (k_icgn_impl.h): the inverse-compositional
//                Gauss-Newton loop, one block per POI.
//
// Semantics follow src/oc_icgn.cpp:168-387 (see k_icgn_impl.h).
#include <hip/hip_runtime.h>

#include "k_icgn_impl.h"

namespace fi {

// ---------------------------------------------------------------------------
// FP64 6x6 inverse, partial-pivot LU with the same structure as the CPU
// reference's Eigen PartialPivLU (largest |a_ik|, first on ties; a zero pivot
// column is skipped so a singular Hessian yields inf/NaN exactly as there).
// One lane per POI: every pivot row differs per lane, so row swaps are
// selects over the candidate rows (registers are never indexed at run time).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lu6_f64(double (&m)[6][6], int (&perm)[6])
{
#pragma unroll
    for (int i = 0; i < 6; i++) perm[i] = i;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        int piv = k;
        double best = fabs(m[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const double s = fabs(m[i][k]);
            if (s > best) { best = s; piv = i; }
        }
        const bool nz = best != 0.0;
        if (!nz) piv = k;
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const bool sw = piv == i;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const double a = m[k][j], c = m[i][j];
                m[k][j] = sw ? c : a;
                m[i][j] = sw ? a : c;
            }
            const int pa = perm[k], pc = perm[i];
            perm[k] = sw ? pc : pa;
            perm[i] = sw ? pa : pc;
        }
#pragma unroll
        for (int i = k + 1; i < 6; i++) {
            const double l = m[i][k] / m[k][k];
            m[i][k] = nz ? l : m[i][k];
        }
#pragma unroll
        for (int i = k + 1; i < 6; i++)
#pragma unroll
            for (int j = k + 1; j < 6; j++) m[i][j] -= m[i][k] * m[k][j];
    }
}

// Column c of the inverse from the factors: forward then back substitution on
// the permuted unit vector.
__device__ __forceinline__ void solve6_col_f64(const double (&m)[6][6], const int (&perm)[6], int c, double (&x)[6])
{
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
}

// ---------------------------------------------------------------------------
// Reference-side prep: mean, norm, Hessian, inverse.
// ---------------------------------------------------------------------------
// One wave per POI.  The subset's columns map to lanes and, when the subset is
// at most 32 wide, the wave's two (or more) lane groups take alternate rows;
// rows are streamed in batches of kPrepRows with every load of a batch in
// flight at once.  Pass 1 accumulates the reference sum and the 21 Hessian
// entries as per-lane FMA chains (at most 63 terms), reduced across lanes by
// the transposing TwoSum tree six sums at a time; pass 2 (the rows again, from
// cache) the centred sums of squares, reduced the same way.
constexpr int kPrepRows = 8;

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
    const int sw = 2 * rx + 1, sh = 2 * ry + 1, ns = sw * sh;
    const int ulx = (int)(x - rx), uly = (int)(y - ry);  // Subset2D::fill, src/oc_subset.cpp:39-44
    const int px = (int)x, py = (int)y;                   // gradients at (int)x + local, :206-210
    const int groups = 64 / sw;                           // lane groups, one row each per step
    const int grp = lane / sw, col = lane - grp * sw;
    const bool on = grp < groups;
    const int c0 = on ? col : 0, g0 = on ? grp : 0;
    const float* ref_col = p.ref + (size_t)(uly + g0) * w + (ulx + c0);
    const size_t go = (size_t)(py - ry + g0) * w + (px - rx + c0);
    const float* gx_col = p.gx + go;
    const float* gy_col = p.gy + go;
    const float xl = (float)(c0 - rx) - ox;
    const size_t step = (size_t)groups * w;  // rows between a lane's samples
    const int rows = on ? (sh - g0 + groups - 1) / groups : 0;  // this lane's rows
    const int last = rows > 0 ? rows - 1 : 0;

    float sum_r = 0.f;
    float hs[21];
#pragma unroll
    for (int q = 0; q < 21; q++) hs[q] = 0.f;
    for (int k0 = 0; k0 < rows; k0 += kPrepRows) {
        float rv[kPrepRows], gxv[kPrepRows], gyv[kPrepRows];
        // every load is issued (rows past the lane's last re-read its last
        // row, in bounds) and the values are masked afterwards: no branch
        // separates the loads of a batch
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) {
            const size_t o = (size_t)min(k0 + k, last) * step;
            rv[k] = ref_col[o];
            gxv[k] = gx_col[o];
            gyv[k] = gy_col[o];
        }
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) {
            const bool ok = k0 + k < rows;
            rv[k] = ok ? rv[k] : 0.f;
            gxv[k] = ok ? gxv[k] : 0.f;
            gyv[k] = ok ? gyv[k] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) {
            // rows past the lane's last give zero terms (all three loads are 0)
            sum_r += rv[k];
            const float yl = (float)(g0 + (k0 + k) * groups - ry) - oy;
            const float sd[6] = { gxv[k], gxv[k] * xl, gxv[k] * yl, gyv[k], gyv[k] * xl, gyv[k] * yl };
            // entry (a, e <= a) at a (a + 1) / 2 + e
#pragma unroll
            for (int a = 0; a < 6; a++)
#pragma unroll
                for (int e = 0; e <= a; e++)
                    hs[a * (a + 1) / 2 + e] = __builtin_fmaf(sd[a], sd[e], hs[a * (a + 1) / 2 + e]);
        }
    }
    // 21 Hessian entries + the reference sum: four transposing reductions
    float t[4][6];
#pragma unroll
    for (int q = 0; q < 21; q++) t[q / 6][q % 6] = hs[q];
    t[3][3] = sum_r;
    t[3][4] = 0.f;
    t[3][5] = 0.f;
#pragma unroll
    for (int g = 0; g < 4; g++) wave_sum6(t[g]);
    const float mean = t[3][3] / (float)ns;

    float sq = 0.f, rs = 0.f;
    for (int k0 = 0; k0 < rows; k0 += kPrepRows) {
        float rv[kPrepRows];
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) rv[k] = ref_col[(size_t)min(k0 + k, last) * step];
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) rv[k] = k0 + k < rows ? rv[k] : mean;  // past the end: d = 0
#pragma unroll
        for (int k = 0; k < kPrepRows; k++) {
            const float d = rv[k] - mean;
            sq = __builtin_fmaf(d, d, sq);
            rs += d;
        }
    }
    float u[3] = { sq, rs, 0.f };
    wave_sum3(u);
    const float norm = sqrtf(u[0]);
    const float ref_sum = u[1];

    // the 21 Hessian sums (packed lower triangle, entry (a, e <= a) at
    // a (a + 1) / 2 + e) wait in the record's hinv for k_prep_inverse
    if (lane == 0) {
        float4* hv = reinterpret_cast<float4*>(out->hinv);
#pragma unroll
        for (int v = 0; v < 5; v++)
            hv[v] = make_float4(t[(4 * v) / 6][(4 * v) % 6], t[(4 * v + 1) / 6][(4 * v + 1) % 6],
                                t[(4 * v + 2) / 6][(4 * v + 2) % 6], t[(4 * v + 3) / 6][(4 * v + 3) % 6]);
        out->hinv[20] = t[3][2];
    }
    if (lane == 0) {
        out->mean = mean;
        out->norm = norm;
        out->status = 1;
        out->rx = rx;
        out->ry = ry;
        out->ref_sum = ref_sum;
    }
}

// The inverse Hessian of every prepared record, one lane per POI: the FP64
// LU of the 21 sums k_icgn_prep left in hinv, then the six columns of the
// inverse, stored FP32 over the sums.
__global__ __launch_bounds__(64) void k_prep_inverse(IcgnParams p)
{
    const int b = (int)(blockIdx.x * 64 + threadIdx.x);
    if (b >= p.n) return;
    const int i = p.index ? (int)p.index[b] : b;
    PoiPrep* rec = p.prep + i;
    if (rec->status == 0) return;
    const float4* hv = reinterpret_cast<const float4*>(rec->hinv);
    float t[24];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const float4 q = hv[v];
        t[4 * v] = q.x;
        t[4 * v + 1] = q.y;
        t[4 * v + 2] = q.z;
        t[4 * v + 3] = q.w;
    }
    t[20] = rec->hinv[20];
    double m[6][6];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int e = 0; e <= a; e++) {
            const double v = (double)t[a * (a + 1) / 2 + e];
            m[a][e] = v;
            m[e][a] = v;
        }
    int perm[6];
    lu6_f64(m, perm);
    float inv[36];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        double xc[6];
        solve6_col_f64(m, perm, c, xc);
#pragma unroll
        for (int r = 0; r < 6; r++) inv[r * 6 + c] = (float)xc[r];
    }
    float4* ho = reinterpret_cast<float4*>(rec->hinv);
#pragma unroll
    for (int v = 0; v < 9; v++) ho[v] = make_float4(inv[4 * v], inv[4 * v + 1], inv[4 * v + 2], inv[4 * v + 3]);
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
#define X(S) extern template void launch_icgn_pair_t<S>(const IcgnParams&, hipStream_t);
FI_ICGN_PAIR(X)
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
    hipLaunchKernelGGL(k_prep_inverse, dim3((p.n + 63) / 64), dim3(64), 0, s, p);
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
// every norm but NaN, c = NaN none.  For c = +inf the float below is FLT_MAX
// and the midpoint is 2^128 - 2^103, where (float)sqrt rounds up to inf (the
// tie goes to the even significand, which is inf's).
double conv_threshold2(float c)
{
    if (c != c) return (double)c;
    if (c < 0.f) return -HUGE_VAL;
    if (c == 0.f) return 0.0;
    const double below = (double)nextafterf(c, 0.f);
    const double m = c == HUGE_VALF ? below + 0x1p103 : 0.5 * (below + (double)c);
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
    if (p.group == 2) {
        // two entries per wave (the lane shape does not apply): uniform
        // subsets whose 32-lane band layout is instantiated, in one block per
        // eight entries; anything else runs as the shared-tile kernel
        const int R = pair_rows(p.rx, p.ry);
        if (!p.count && !p.self_adaptive && p.tar != nullptr) {
            IcgnParams q = p;
            q.tile_h = 2 * p.ry + 4 + 2 * kGroupMargin + kGroupRowSlack;
            q.tile_stride = kPairStride;
#define X(S)                                                                    \
            if (R == S) {                                                       \
                if (q.n > 0) launch_icgn_pair_t<S>(q, s);                       \
                return true;                                                    \
            }
            FI_ICGN_PAIR(X)
#undef X
        }
        p.group = 1;
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
