// FFTCC2D integer-pixel seeding on gfx950: one wavefront per POI, the whole
// (2r)x(2r) correlation done in LDS -- window load, zero mean, R2C DFT,
// conj(R)*T against the resident reference spectrum, C2R DFT and the argmax
// with the reference's first-maximum rule -- with no vendor FFT.
//
// Reference behaviour: src/oc_fftcc.cpp:246-313 (prepareRef), :317-394
// (computePreparedRef), :396-496 (compute(POI2D*)).
//
// DFT specification.  The reference delegates the transforms to FFTW, whose
// rounding is not reproducible here.  This kernel and the CPU oracle both
// follow one specification of the separable DFT (one twiddle table
// tw[k] = (cos 2pi k/n, -sin 2pi k/n) rounded from double; fixed summation
// order; the fused multiply-adds written out below; everything else unfused,
// this file is compiled with -ffp-contract=off), so integer peaks, zncc and
// every intermediate are bit-identical between the two:
//   A[y][kx]  = sum_x   fma(w, tw)                         (real rows, kx <= n/2)
//   X[ky][kx] = sum_y   A (*) tw,   (*) = 4 fmas: re=fma(ar,wr,re); re=fma(-ai,wi,re);
//                                                 im=fma(ar,wi,im); im=fma(ai,wr,im)
//   Z         = (Rr*Tr + Ri*Ti, Rr*Ti - Ri*Tr)             (conj(R) * T)
//   G[y][kx]  = sum_ky  Z (*) conj(tw): re=fma(zr,wr,re); re=fma(zi,wi,re);
//                                       im=fma(zi,wr,im); im=fma(-zr,wi,im)
//   out[y][x] = G[y][0].re + sum_{kx=1}^{n/2-1} fma(2, fma(gr,wr, gi*wi), .)
//               +/- G[y][n/2].re
// Window mean and squared norm: row sums left to right, then the row sums top
// to bottom.
#include <hip/hip_runtime.h>

#include <climits>

#include "fi_device.h"
#include "fi_launch.h"

namespace fi {

struct FftccLds {
    float* win;   // n*n real window, later the correlation surface
    float2* A;    // n*hc scratch spectrum
    float2* B;    // n*hc target spectrum / cross power
    float2* C;    // n*hc reference spectrum (uncached mode only)
    float2* tw;   // n twiddles
    float* rows;  // n row partial sums
};

__device__ __forceinline__ FftccLds fftcc_carve(float* lds, int n, bool with_ref)
{
    const int hc = n / 2 + 1, nn = n * n, nh = n * hc;
    FftccLds L;
    L.win = lds;
    L.A = reinterpret_cast<float2*>(lds + nn);
    L.B = L.A + nh;
    L.C = L.B + nh;
    L.tw = with_ref ? L.C + nh : L.C;
    L.rows = reinterpret_cast<float*>(L.tw + n);
    return L;
}

// mean by row partials; returns it (wave-uniform)
__device__ __forceinline__ float window_mean(const FftccLds& L, int n, int lane)
{
    if (lane < n) {
        float rs = 0.f;
        for (int x = 0; x < n; x++) rs += L.win[lane * n + x];
        L.rows[lane] = rs;
    }
    __syncthreads();
    float tot = 0.f;
    for (int y = 0; y < n; y++) tot += L.rows[y];
    __syncthreads();
    return tot / (float)(n * n);
}

// subtract the mean in place; return the sum of squares (row partials)
__device__ __forceinline__ float window_zero_mean_sq(const FftccLds& L, int n, int lane, float mean)
{
    if (lane < n) {
        float rs = 0.f;
        for (int x = 0; x < n; x++) {
            float d = L.win[lane * n + x] - mean;
            L.win[lane * n + x] = d;
            rs += d * d;
        }
        L.rows[lane] = rs;
    }
    __syncthreads();
    float tot = 0.f;
    for (int y = 0; y < n; y++) tot += L.rows[y];
    __syncthreads();
    return tot;
}

// forward R2C of L.win into dst (n rows ky by hc columns kx)
__device__ __forceinline__ void dft_r2c(const FftccLds& L, int n, int lane, float2* dst)
{
    const int hc = n / 2 + 1, nh = n * hc;
    for (int o = lane; o < nh; o += 64) {
        const int y = o / hc, kx = o - y * hc;
        float re = 0.f, im = 0.f;
        int t = 0;
        const float* row = L.win + y * n;
        for (int x = 0; x < n; x++) {
            const float v = row[x];
            const float2 w = L.tw[t];
            re = __builtin_fmaf(v, w.x, re);
            im = __builtin_fmaf(v, w.y, im);
            t += kx;
            t -= (t >= n) ? n : 0;
        }
        L.A[o] = make_float2(re, im);
    }
    __syncthreads();
    for (int o = lane; o < nh; o += 64) {
        const int ky = o / hc, kx = o - ky * hc;
        float re = 0.f, im = 0.f;
        int t = 0;
        for (int y = 0; y < n; y++) {
            const float2 a = L.A[y * hc + kx];
            const float2 w = L.tw[t];
            re = __builtin_fmaf(a.x, w.x, re);
            re = __builtin_fmaf(-a.y, w.y, re);
            im = __builtin_fmaf(a.x, w.y, im);
            im = __builtin_fmaf(a.y, w.x, im);
            t += ky;
            t -= (t >= n) ? n : 0;
        }
        dst[o] = make_float2(re, im);
    }
    __syncthreads();
}

// inverse C2R of L.B (cross power) into L.win
__device__ __forceinline__ void dft_c2r(const FftccLds& L, int n, int lane)
{
    const int hc = n / 2 + 1, nh = n * hc, nn = n * n;
    for (int o = lane; o < nh; o += 64) {
        const int y = o / hc, kx = o - y * hc;
        float re = 0.f, im = 0.f;
        int t = 0;
        for (int ky = 0; ky < n; ky++) {
            const float2 z = L.B[ky * hc + kx];
            const float2 w = L.tw[t];
            re = __builtin_fmaf(z.x, w.x, re);
            re = __builtin_fmaf(z.y, w.y, re);
            im = __builtin_fmaf(z.y, w.x, im);
            im = __builtin_fmaf(-z.x, w.y, im);
            t += y;
            t -= (t >= n) ? n : 0;
        }
        L.A[o] = make_float2(re, im);
    }
    __syncthreads();
    for (int o = lane; o < nn; o += 64) {
        const int y = o / n, x = o - y * n;
        const float2* g = L.A + y * hc;
        float acc = g[0].x;
        int t = x;
        for (int kx = 1; kx < hc - 1; kx++) {
            const float2 w = L.tw[t];
            float s = g[kx].y * w.y;
            s = __builtin_fmaf(g[kx].x, w.x, s);
            acc = __builtin_fmaf(2.f, s, acc);
            t += x;
            t -= (t >= n) ? n : 0;
        }
        const float nyq = g[hc - 1].x;
        acc += (x & 1) ? -nyq : nyq;
        L.win[o] = acc;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ void load_window_int(const FftccLds& L, const DevImage& img, int ix,
                                                int iy, int r, int lane)
{
    const int n = 2 * r, nn = n * n;
    const T* px = static_cast<const T*>(img.data);
    for (int o = lane; o < nn; o += 64) {
        const int rr = o / n, c = o - rr * n;
        L.win[o] = load_px<T>(px, (size_t)(iy + rr - r) * img.pitch + (ix + c - r), img.scale);
    }
}

// pixel (int)(y + rr - r + d), (int)(x + c - r + d) in float arithmetic,
// src/oc_fftcc.cpp:341-343 and :418-430
template <typename T>
__device__ __forceinline__ void load_window_f(const FftccLds& L, const DevImage& img, float x,
                                              float y, float du, float dv, int r, int lane)
{
    const int n = 2 * r, nn = n * n;
    const T* px = static_cast<const T*>(img.data);
    for (int o = lane; o < nn; o += 64) {
        const int rr = o / n, c = o - rr * n;
        const float xp = x + (float)c - (float)r;
        const float yp = y + (float)rr - (float)r;
        const float tx = xp + du, ty = yp + dv;
        L.win[o] = load_px<T>(px, (size_t)(int)ty * img.pitch + (int)tx, img.scale);
    }
}

__device__ __forceinline__ bool window_outside(float x, float y, float du, float dv, int r, int w,
                                               int h)
{
    // int(x + du) outside [r, w - r): src/oc_fftcc.cpp:328-330 and :409-411
    if (!(isfinite(x + du) && isfinite(y + dv))) return true;
    const int tx = (int)(x + du), ty = (int)(y + dv);
    return tx < r || tx >= w - r || ty < r || ty >= h - r;
}

template <typename TR, typename TT, int MODE>
__global__ __launch_bounds__(64) void k_fftcc(FftccParams p)
{
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    if (i >= p.n_pois) return;
    const int r = p.r, n = 2 * r, hc = r + 1, nh = n * hc;
    const int w = p.ref.width, h = p.ref.height;
    const FftccLds L = fftcc_carve(lds, n, MODE == kFftccUncached);
    for (int k = lane; k < n; k += 64)
        L.tw[k] = make_float2(p.twiddles[2 * k], p.twiddles[2 * k + 1]);

    fi_poi* poi = p.pois + i;
    const float x = poi->x, y = poi->y;

    if (MODE == kFftccPrepare) {
        // prepareRef(queue): integer centre, validity r <= c < size - r
        const bool finite = isfinite(x) && isfinite(y);
        const int ix = finite ? (int)x : 0, iy = finite ? (int)y : 0;
        float2* spec = reinterpret_cast<float2*>(p.spectra) + (size_t)i * nh;
        const bool ok = finite && ix >= r && ix < w - r && iy >= r && iy < h - r;
        if (!ok) {
            for (int o = lane; o < nh; o += 64) spec[o] = make_float2(0.f, 0.f);
            if (lane == 0) p.slots[i] = FftccSlot{ ix, iy, 0.f, 0 };
            return;
        }
        load_window_int<TR>(L, p.ref, ix, iy, r, lane);
        __syncthreads();
        const float mean = window_mean(L, n, lane);
        const float norm = window_zero_mean_sq(L, n, lane, mean);
        dft_r2c(L, n, lane, L.B);
        for (int o = lane; o < nh; o += 64) spec[o] = L.B[o];
        if (lane == 0) p.slots[i] = FftccSlot{ ix, iy, norm, 1 };
        return;
    }

    const float u0 = poi->deformation[0], v0 = poi->deformation[6];
    float ref_norm;
    const float2* R;
    if (MODE == kFftccCached) {
        const uint32_t s = p.slot ? p.slot[i] : (uint32_t)i;
        const FftccSlot meta = p.slots[s];
        if (!meta.valid || window_outside(x, y, u0, v0, r, w, h)) return;
        ref_norm = meta.norm;
        R = reinterpret_cast<const float2*>(p.spectra) + (size_t)s * nh;
    } else {
        // compute(POI2D*): reference window inside, target window inside
        if (!(isfinite(x) && isfinite(y))) return;
        const int ix = (int)x, iy = (int)y;
        if (ix < r || ix >= w - r || iy < r || iy >= h - r || window_outside(x, y, u0, v0, r, w, h))
            return;
        load_window_f<TR>(L, p.ref, x, y, 0.f, 0.f, r, lane);
        __syncthreads();
        const float mean = window_mean(L, n, lane);
        ref_norm = window_zero_mean_sq(L, n, lane, mean);
        dft_r2c(L, n, lane, L.C);
        R = L.C;
    }

    load_window_f<TT>(L, p.tar, x, y, u0, v0, r, lane);
    __syncthreads();
    const float tmean = window_mean(L, n, lane);
    const float tar_norm = window_zero_mean_sq(L, n, lane, tmean);
    dft_r2c(L, n, lane, L.B);
    for (int o = lane; o < nh; o += 64) {
        const float2 rv = R[o], tv = L.B[o];
        L.B[o] = make_float2((rv.x * tv.x) + (rv.y * tv.y), (rv.x * tv.y) - (rv.y * tv.x));
    }
    __syncthreads();
    dft_c2r(L, n, lane);

    // first maximum above -2 (strict >), src/oc_fftcc.cpp:367-376
    float m = -2.f;
    int idx = INT_MAX;
    const int nn = n * n;
    for (int o = lane; o < nn; o += 64) {
        const float v = L.win[o];
        if (v > m) { m = v; idx = o; }
    }
    const float mall = wave_max(m);
    int iall = wave_min_i(m == mall ? idx : INT_MAX);
    if (iall == INT_MAX) iall = 0;
    if (lane == 0) {
        int lu = iall % n, lv = iall / n;
        if (lu > r) lu -= n;
        if (lv > r) lv -= n;
        poi->deformation[0] = (float)lu + u0;
        poi->deformation[6] = (float)lv + v0;
        poi->result[0] = u0;
        poi->result[1] = v0;
        poi->result[2] = (float)((double)mall / (sqrt((double)(ref_norm * tar_norm)) * (double)nn));
    }
}

size_t fftcc_lds_bytes(int r, int mode)
{
    const size_t n = 2 * (size_t)r, hc = r + 1, nh = n * hc;
    const size_t bufs = (mode == kFftccUncached) ? 3 : 2;
    return sizeof(float) * (n * n + 2 * nh * bufs + 2 * n + n);
}

template <typename TR, typename TT>
static void launch_mode(const FftccParams& p, hipStream_t s)
{
    const size_t lds = fftcc_lds_bytes(p.r, p.mode);
    const dim3 grid(p.n_pois), block(64);
    if (p.mode == kFftccPrepare)
        hipLaunchKernelGGL((k_fftcc<TR, TT, kFftccPrepare>), grid, block, lds, s, p);
    else if (p.mode == kFftccCached)
        hipLaunchKernelGGL((k_fftcc<TR, TT, kFftccCached>), grid, block, lds, s, p);
    else
        hipLaunchKernelGGL((k_fftcc<TR, TT, kFftccUncached>), grid, block, lds, s, p);
}

template <typename TR>
static void launch_tar(const FftccParams& p, hipStream_t s)
{
    switch (p.tar.pixel) {
    case kU8: launch_mode<TR, uint8_t>(p, s); break;
    case kU16: launch_mode<TR, uint16_t>(p, s); break;
    default: launch_mode<TR, float>(p, s); break;
    }
}

void launch_fftcc(const FftccParams& p, hipStream_t s)
{
    if (p.n_pois <= 0) return;
    switch (p.ref.pixel) {
    case kU8: launch_tar<uint8_t>(p, s); break;
    case kU16: launch_tar<uint16_t>(p, s); break;
    default: launch_tar<float>(p, s); break;
    }
}

}  // namespace fi
