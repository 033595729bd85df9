// FFTCC2D integer-pixel seeding on gfx950, no vendor FFT.
//
// Reference behaviour: src/oc_fftcc.cpp:246-313 (prepareRef), :317-394
// (computePreparedRef), :396-496 (compute(POI2D*)): zero-mean (2r)x(2r)
// windows, R2C, conj(R)*T, C2R, first maximum above -2, wrap, zncc.
//
// Layout: one 64-lane wavefront serves 64/G POIs, G = 16, 32 or 64 lanes per
// POI (enough for the n/2 + 1 spectrum columns, n = 2r).  Each lane holds whole
// 1-D transforms in VGPRs: a pair of window rows packed into one complex
// length-n FFT for the row pass, one spectrum column for the column passes.
// LDS carries only the row <-> column transposes and the row sums.
//
// Transform specification.  The reference delegates the transforms to FFTW,
// whose rounding is not reproducible here.  This kernel and the CPU oracle
// (oracle/oracle.c) perform one specification operation for operation:
// recursive decimation-in-time FFT with radix 4, 2, 3, 5 (direct sum for a
// prime > 5), the butterflies and complex product written out below with
// their fused multiply-adds, twiddles tw[j] = (cos, -sin)(2 pi j / n) rounded
// from double, two real rows per complex FFT, inverse = conj(fft(conj(.))),
// imaginary parts of the kx = 0 and kx = n/2 bins ignored by the inverse.
// Everything not written as an fma is unfused (-ffp-contract=off), so integer
// peaks, zncc and every stored float are bit-identical to the oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "fi_device.h"
#include "fi_launch.h"

namespace fi {

template <int N>
struct CV {
    float2 v[N];
};

template <int I, int E, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}

constexpr float kK3 = 0.866025403784438646763723170753f;    // sin(2 pi / 3)
constexpr float kC51 = 0.309016994374947424102293417183f;   // cos(2 pi / 5)
constexpr float kC52 = -0.809016994374947424102293417183f;  // cos(4 pi / 5)
constexpr float kS51 = 0.951056516295153572116439333379f;   // sin(2 pi / 5)
constexpr float kS52 = 0.587785252292473129168705954639f;   // sin(4 pi / 5)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
    return make_float2(__builtin_fmaf(a.x, w.x, -(a.y * w.y)), __builtin_fmaf(a.x, w.y, a.y * w.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

constexpr int fft_radix(int n)
{
    return n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : n;
}

// X[OFF + M*t] = butterfly_P(T)[t]; STEP (prime P > 5 only) = table size / P
template <int P, int M, int OFF, int NX, int STEP>
__device__ __forceinline__ void butterfly(const float2 (&T)[P], CV<NX>& X, const float2* __restrict__ tw)
{
    if constexpr (P == 2) {
        X.v[OFF] = cadd(T[0], T[1]);
        X.v[OFF + M] = csub(T[0], T[1]);
    } else if constexpr (P == 3) {
        const float2 s1 = cadd(T[1], T[2]), s2 = csub(T[1], T[2]);
        X.v[OFF] = cadd(T[0], s1);
        const float2 m = make_float2(T[0].x - 0.5f * s1.x, T[0].y - 0.5f * s1.y);
        const float2 q = make_float2(kK3 * s2.x, kK3 * s2.y);
        X.v[OFF + M] = make_float2(m.x + q.y, m.y - q.x);
        X.v[OFF + 2 * M] = make_float2(m.x - q.y, m.y + q.x);
    } else if constexpr (P == 4) {
        const float2 s0 = cadd(T[0], T[2]), s1 = csub(T[0], T[2]);
        const float2 s2 = cadd(T[1], T[3]), s3 = csub(T[1], T[3]);
        X.v[OFF] = cadd(s0, s2);
        X.v[OFF + 2 * M] = csub(s0, s2);
        X.v[OFF + M] = make_float2(s1.x + s3.y, s1.y - s3.x);
        X.v[OFF + 3 * M] = make_float2(s1.x - s3.y, s1.y + s3.x);
    } else if constexpr (P == 5) {
        const float2 t1 = cadd(T[1], T[4]), t2 = cadd(T[2], T[3]);
        const float2 t3 = csub(T[1], T[4]), t4 = csub(T[2], T[3]);
        X.v[OFF] = cadd(T[0], cadd(t1, t2));
        const float2 a1 = make_float2(T[0].x + __builtin_fmaf(kC51, t1.x, kC52 * t2.x),
                                      T[0].y + __builtin_fmaf(kC51, t1.y, kC52 * t2.y));
        const float2 a2 = make_float2(T[0].x + __builtin_fmaf(kC52, t1.x, kC51 * t2.x),
                                      T[0].y + __builtin_fmaf(kC52, t1.y, kC51 * t2.y));
        const float2 b1 = make_float2(__builtin_fmaf(kS51, t3.x, kS52 * t4.x),
                                      __builtin_fmaf(kS51, t3.y, kS52 * t4.y));
        const float2 b2 = make_float2(__builtin_fmaf(kS52, t3.x, -(kS51 * t4.x)),
                                      __builtin_fmaf(kS52, t3.y, -(kS51 * t4.y)));
        X.v[OFF + M] = make_float2(a1.x + b1.y, a1.y - b1.x);
        X.v[OFF + 4 * M] = make_float2(a1.x - b1.y, a1.y + b1.x);
        X.v[OFF + 2 * M] = make_float2(a2.x + b2.y, a2.y - b2.x);
        X.v[OFF + 3 * M] = make_float2(a2.x - b2.y, a2.y + b2.x);
    } else {
        static_for<0, P>([&](auto t) {
            float2 acc = T[0];
            static_for<1, P>([&](auto r) {
                constexpr int j = (r * t) % P;
                if constexpr (j == 0)
                    acc = cadd(acc, T[r]);
                else
                    acc = cadd(acc, cmul(T[r], tw[j * STEP]));
            });
            X.v[OFF + t * M] = acc;
        });
    }
}

// FFT of length N over x.v[IOFF], x.v[IOFF + S], ...; STEP = table size / N
template <int N, int S, int IOFF, int NIN, int STEP>
__device__ __forceinline__ CV<N> fft_rec(const CV<NIN>& x, const float2* __restrict__ tw)
{
    CV<N> X;
    if constexpr (N == 1) {
        X.v[0] = x.v[IOFF];
    } else {
        constexpr int P = N <= 5 ? N : fft_radix(N);
        constexpr int M = N / P;
        if constexpr (M == 1) {
            float2 T[P];
            static_for<0, P>([&](auto r) { T[r] = x.v[IOFF + r * S]; });
            butterfly<P, 1, 0, N, STEP>(T, X, tw);
        } else {
            CV<M> Y[P];
            static_for<0, P>([&](auto r) { Y[r] = fft_rec<M, S * P, IOFF + r * S, NIN, STEP * P>(x, tw); });
            static_for<0, M>([&](auto k) {
                float2 T[P];
                static_for<0, P>([&](auto r) {
                    if constexpr (r * k == 0)
                        T[r] = Y[r].v[k];
                    else
                        T[r] = cmul(Y[r].v[k], tw[r * k * STEP]);
                });
                butterfly<P, M, k, N, STEP * M>(T, X, tw);
            });
        }
    }
    return X;
}

template <int N>
__device__ __forceinline__ CV<N> fft(const CV<N>& x, const float2* __restrict__ tw)
{
    return fft_rec<N, 1, 0, N, 1>(x, tw);
}

template <int N>
__device__ __forceinline__ CV<N> ifft(const CV<N>& z, const float2* __restrict__ tw)
{
    CV<N> c;
#pragma unroll
    for (int k = 0; k < N; k++) c.v[k] = cconj(z.v[k]);
    CV<N> r = fft<N>(c, tw);
#pragma unroll
    for (int k = 0; k < N; k++) r.v[k] = cconj(r.v[k]);
    return r;
}

// int(x + du) outside [r, size - r): src/oc_fftcc.cpp:328-330 and :409-411
__device__ __forceinline__ bool window_outside(float x, float y, float du, float dv, int r, int w, int h)
{
    if (!(isfinite(x + du) && isfinite(y + dv))) return true;
    const int tx = (int)(x + du), ty = (int)(y + dv);
    return tx < r || tx >= w - r || ty < r || ty >= h - r;
}

// One POI group's state inside a block.
template <int N>
struct FftccGroup {
    static constexpr int R = N / 2, HC = N / 2 + 1, PAIRS = N / 2;
    static constexpr int G = HC <= 16 ? 16 : (HC <= 32 ? 32 : 64);
    static constexpr int PPB = 64 / G;
};

// Loads the two window rows of pair j into a[] and b[]: integer centre
// (prepareRef) or float arithmetic (int)(c + k - r + d) per pixel (compute).
template <int N, bool INT_CENTRE>
__device__ __forceinline__ void load_rows(const float* __restrict__ img, int w, float x, float y, float du,
                                          float dv, int j, float (&a)[N], float (&b)[N])
{
    constexpr int R = N / 2;
    if (INT_CENTRE) {
        const int ix = (int)x, iy = (int)y;
        const float* ra = img + (size_t)(iy + 2 * j - R) * w + (ix - R);
        const float* rb = ra + w;
#pragma unroll
        for (int c = 0; c < N; c++) {
            a[c] = ra[c];
            b[c] = rb[c];
        }
    } else {
        const float ya = y + (float)(2 * j) - (float)R, yb = y + (float)(2 * j + 1) - (float)R;
        const size_t rowa = (size_t)(int)(ya + dv) * w, rowb = (size_t)(int)(yb + dv) * w;
#pragma unroll
        for (int c = 0; c < N; c++) {
            const float xp = x + (float)c - (float)R;
            const int tx = (int)(xp + du);
            a[c] = img[rowa + tx];
            b[c] = img[rowb + tx];
        }
    }
}

// Window statistics by row partial sums (left to right, then top to bottom)
// and the packed forward row FFT of pair j into spec[y][kx].
template <int N>
__device__ __forceinline__ void rows_forward(float (&a)[N], float (&b)[N], int j, bool act, float* rows,
                                             float2* spec, const float2* __restrict__ tw, float& norm_out)
{
    constexpr int HC = N / 2 + 1, PAIRS = N / 2;
    if (act && j < PAIRS) {
        float ra = 0.f, rb = 0.f;
#pragma unroll
        for (int c = 0; c < N; c++) ra += a[c];
#pragma unroll
        for (int c = 0; c < N; c++) rb += b[c];
        rows[2 * j] = ra;
        rows[2 * j + 1] = rb;
    }
    __syncthreads();
    float tot = 0.f;
    if (act)
        for (int y = 0; y < N; y++) tot += rows[y];
    const float mean = tot / (float)(N * N);
    __syncthreads();
    if (act && j < PAIRS) {
        float ra = 0.f, rb = 0.f;
#pragma unroll
        for (int c = 0; c < N; c++) {
            a[c] = a[c] - mean;
            ra += a[c] * a[c];
        }
#pragma unroll
        for (int c = 0; c < N; c++) {
            b[c] = b[c] - mean;
            rb += b[c] * b[c];
        }
        rows[2 * j] = ra;
        rows[2 * j + 1] = rb;
    }
    __syncthreads();
    float nrm = 0.f;
    if (act)
        for (int y = 0; y < N; y++) nrm += rows[y];
    norm_out = nrm;
    if (act && j < PAIRS) {
        CV<N> z;
#pragma unroll
        for (int c = 0; c < N; c++) z.v[c] = make_float2(a[c], b[c]);
        const CV<N> Z = fft<N>(z, tw);
#pragma unroll
        for (int kx = 0; kx < HC; kx++) {
            const int nk = (N - kx) % N;
            spec[(2 * j) * HC + kx] =
                make_float2(0.5f * (Z.v[kx].x + Z.v[nk].x), 0.5f * (Z.v[kx].y - Z.v[nk].y));
            spec[(2 * j + 1) * HC + kx] =
                make_float2(0.5f * (Z.v[kx].y + Z.v[nk].y), 0.5f * (Z.v[nk].x - Z.v[kx].x));
        }
    }
    __syncthreads();
}

template <int N, int MODE>
__global__ __launch_bounds__(64) void k_fftcc(FftccParams p)
{
    using Gp = FftccGroup<N>;
    constexpr int R = Gp::R, HC = Gp::HC, PAIRS = Gp::PAIRS, G = Gp::G, PPB = Gp::PPB, NH = N * HC;
    __shared__ float2 spec_all[PPB][NH];
    __shared__ float rows_all[PPB][N];
    const int lane = threadIdx.x, g = lane / G, j = lane % G;
    const int i = (p.count ? (int)blockIdx.x : xcd_remap(blockIdx.x, gridDim.x)) * PPB + g;
    const int n_pois = p.count ? min((int)*p.count, p.n_pois) : p.n_pois;
    float2* spec = spec_all[g];
    float* rows = rows_all[g];
    const float2* tw = reinterpret_cast<const float2*>(p.twiddles);
    const int w = p.w, h = p.h;

    bool act = i < n_pois;
    float x = 0.f, y = 0.f, u0 = 0.f, v0 = 0.f;
    fi_poi* poi = nullptr;
    if (act) {
        poi = p.pois + i;
        x = poi->x;
        y = poi->y;
        u0 = poi->deformation[0];
        v0 = poi->deformation[6];
    }
    float a[N], b[N];

    if (MODE == kFftccPrepare) {
        // prepareRef(queue): integer centre, valid iff r <= c < size - r
        const bool finite = isfinite(x) && isfinite(y);
        const int ix = finite ? (int)x : 0, iy = finite ? (int)y : 0;
        const bool ok = act && finite && ix >= R && ix < w - R && iy >= R && iy < h - R;
        if (ok && j < PAIRS) load_rows<N, true>(p.ref, w, x, y, 0.f, 0.f, j, a, b);
        float norm;
        rows_forward<N>(a, b, j, ok, rows, spec, tw, norm);
        if (act && j < HC) {
            float2* out = reinterpret_cast<float2*>(p.spectra) + (size_t)i * NH;
            if (ok) {
                CV<N> col;
#pragma unroll
                for (int yy = 0; yy < N; yy++) col.v[yy] = spec[yy * HC + j];
                const CV<N> C = fft<N>(col, tw);
#pragma unroll
                for (int ky = 0; ky < N; ky++) out[ky * HC + j] = C.v[ky];
            } else {
#pragma unroll
                for (int ky = 0; ky < N; ky++) out[ky * HC + j] = make_float2(0.f, 0.f);
            }
            if (j == 0) p.slots[i] = FftccSlot{ ix, iy, ok ? norm : 0.f, ok ? 1 : 0 };
        }
        return;
    }

    // compute: validity (a POI failing it is left untouched)
    bool ok = act;
    uint32_t slot = 0;
    float ref_norm = 0.f;
    if (MODE == kFftccCached) {
        if (act) {
            slot = p.slot ? p.slot[i] : (uint32_t)i;
            const FftccSlot meta = p.slots[slot];
            ok = meta.valid && !window_outside(x, y, u0, v0, R, w, h);
            ref_norm = meta.norm;
        }
    } else if (act) {
        const bool fin = isfinite(x) && isfinite(y);
        const int ix = fin ? (int)x : 0, iy = fin ? (int)y : 0;
        ok = fin && ix >= R && ix < w - R && iy >= R && iy < h - R && !window_outside(x, y, u0, v0, R, w, h);
    }

    CV<N> Rcol;  // uncached: this lane's column of the reference spectrum
    if (MODE == kFftccUncached) {
        if (ok && j < PAIRS) load_rows<N, false>(p.ref, w, x, y, 0.f, 0.f, j, a, b);
        rows_forward<N>(a, b, j, ok, rows, spec, tw, ref_norm);
        if (ok && j < HC) {
            CV<N> col;
#pragma unroll
            for (int yy = 0; yy < N; yy++) col.v[yy] = spec[yy * HC + j];
            Rcol = fft<N>(col, tw);
        }
        __syncthreads();
    }

    if (ok && j < PAIRS) load_rows<N, false>(p.tar, w, x, y, u0, v0, j, a, b);
    float tar_norm;
    rows_forward<N>(a, b, j, ok, rows, spec, tw, tar_norm);
    if (ok && j < HC) {
        CV<N> col;
#pragma unroll
        for (int yy = 0; yy < N; yy++) col.v[yy] = spec[yy * HC + j];
        const CV<N> T = fft<N>(col, tw);
        const float2* Rg = reinterpret_cast<const float2*>(p.spectra) + (size_t)slot * NH;
        CV<N> Z;
#pragma unroll
        for (int ky = 0; ky < N; ky++) {
            const float2 rv = (MODE == kFftccCached) ? Rg[ky * HC + j] : Rcol.v[ky];
            const float2 tv = T.v[ky];
            Z.v[ky] = make_float2((rv.x * tv.x) + (rv.y * tv.y), (rv.x * tv.y) - (rv.y * tv.x));
        }
        const CV<N> Gc = ifft<N>(Z, tw);
#pragma unroll
        for (int yy = 0; yy < N; yy++) spec[yy * HC + j] = Gc.v[yy];
    }
    __syncthreads();

    // inverse rows, then the first maximum above -2 (strict >), :367-376
    float m = -2.f;
    int idx = INT_MAX;
    if (ok && j < PAIRS) {
        float2 A[HC], B[HC];
#pragma unroll
        for (int k = 0; k < HC; k++) {
            A[k] = spec[(2 * j) * HC + k];
            B[k] = spec[(2 * j + 1) * HC + k];
        }
        A[0].y = 0.f;
        B[0].y = 0.f;
        A[HC - 1].y = 0.f;
        B[HC - 1].y = 0.f;
        CV<N> Zr;
#pragma unroll
        for (int k = 0; k < N; k++) {
            if (k <= N / 2) {
                Zr.v[k] = make_float2(A[k].x - B[k].y, A[k].y + B[k].x);
            } else {
                const int q = N - k;
                Zr.v[k] = make_float2(A[q].x + B[q].y, B[q].x - A[q].y);
            }
        }
        const CV<N> z = ifft<N>(Zr, tw);
#pragma unroll
        for (int c = 0; c < N; c++)
            if (z.v[c].x > m) { m = z.v[c].x; idx = (2 * j) * N + c; }
#pragma unroll
        for (int c = 0; c < N; c++)
            if (z.v[c].y > m) { m = z.v[c].y; idx = (2 * j + 1) * N + c; }
    }
    // reduce over the group's G lanes: larger value, then smaller index
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
        const float om = __shfl_xor(m, off, G);
        const int oi = __shfl_xor(idx, off, G);
        if (om > m || (om == m && oi < idx)) { m = om; idx = oi; }
    }
    if (ok && j == 0) {
        const int best = (idx == INT_MAX) ? 0 : idx;
        int lu = best % N, lv = best / N;
        if (lu > R) lu -= N;
        if (lv > R) lv -= N;
        poi->deformation[0] = (float)lu + u0;
        poi->deformation[6] = (float)lv + v0;
        poi->result[0] = u0;
        poi->result[1] = v0;
        poi->result[2] = (float)((double)m / (sqrt((double)(ref_norm * tar_norm)) * (double)(N * N)));
    }
}

template <int N>
void launch_fftcc_n(const FftccParams& p, hipStream_t s)
{
    const int ppb = FftccGroup<N>::PPB;
    const dim3 grid((p.n_pois + ppb - 1) / ppb), block(64);
    if (p.mode == kFftccPrepare)
        hipLaunchKernelGGL((k_fftcc<N, kFftccPrepare>), grid, block, 0, s, p);
    else if (p.mode == kFftccCached)
        hipLaunchKernelGGL((k_fftcc<N, kFftccCached>), grid, block, 0, s, p);
    else
        hipLaunchKernelGGL((k_fftcc<N, kFftccUncached>), grid, block, 0, s, p);
}

}  // namespace fi
