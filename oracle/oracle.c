/*
 * CPU oracle: plain-C restatement of the reference's 2D-DIC path.
 *
 * TEST INFRASTRUCTURE ONLY -- see oracle.h.  Nothing in the product links
 * or loads this library.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   src/oc_gradient.cpp:37-79            Gradient2D4                or_gradient_*
 *   src/oc_cubic_bspline.h:52-58,
 *   src/oc_cubic_bspline.cpp:84-181      BicubicBspline             or_bspline_*
 *   src/oc_fftcc.cpp:246-496             FFTCC2D (cached, uncached) or_fftcc_*
 *   src/oc_icgn.cpp:39-43,168-387        ICGN2D1::compute(POI2D*)   or_icgn_compute_*
 *   src/oc_subset.cpp:39-53              Subset2D fill/zeroMeanNorm
 *   src/oc_deformation.cpp:94-128        Deformation2D1 warp/update
 *   src/oc_image.cpp:163-185             12-bit alignment           or_image_to_float
 *
 * Third-party arithmetic the reference delegates (absent here, restated):
 *   FFTW3f >= 3.3 (README.md:51): fftwf_plan_dft_r2c_2d / c2r_2d, unnormalised.
 *     Restated as the DFT specification below.  FFTW's own rounding is not
 *     reproduced: bit-level parity with FFTW is unpinned; integer peaks are
 *     pinned by the reference's LCG known-answer test and by FP64 brute-force
 *     circular correlation (tests/test_oracle.py).
 *   Eigen >= 3.3 (README.md:48): Matrix6f::inverse (PartialPivLU), Matrix3f
 *     inverse and products, mean()/squaredNorm().  Restated in oracle_core.inc;
 *     bit-level parity with Eigen's reduction order is unpinned, results are
 *     pinned by FP64 brute force and closed forms.
 *
 * Pinning (each function has a check in tests/test_oracle.py that a plausible
 * mistake fails):
 *   or_gradient       exact on cubic ramps; hand-computed taps
 *   or_bspline_*      constant image, node interpolation, brute-force BC*Q*BC^T
 *   or_inverse6       H * H^-1 = I on SPD and pivoting cases; numpy.linalg.inv
 *   or_fftcc_*        LCG known answer (smoke/library_boundary_probe.cpp:41-62),
 *                     FP64 brute-force circular correlation, wrap rule,
 *                     zncc = 1 on identical windows, boundary returns
 *   or_icgn_*         integer shift converges to 0.01 px in one step
 *                     (smoke/run_cache_regression.py:22-28), analytic affine
 *                     speckle truth, OOB band closed form
 *                     (smoke/run_oob_rejection_gate.py:53-56), zero texture -5
 *                     (smoke/run_zero_texture_gate.py:9-35), FP32 vs FP64 twin
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* BC = B * C, src/oc_cubic_bspline.h:52-58; the FP32 entries are the float
 * quotients the reference's initialisers produce. */
static void or_bc_matrix_fill(int f64, void* out)
{
    static const double num[4][4] = {
        { -144.0, 384.0, -384.0, 144.0 },
        { 342.0, -702.0, 450.0, -90.0 },
        { -198.0, -18.0, 270.0, -54.0 },
        { 0.0, 336.0, 0.0, 0.0 },
    };
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            if (f64) ((double*)out)[i * 4 + j] = num[i][j] / 336.0;
            else ((float*)out)[i * 4 + j] = (float)num[i][j] / 336.0f;
        }
}

#define REAL float
#define SFX(n) n##_f32
#include "oracle_core.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX(n) n##_f64
#include "oracle_core.inc"
#undef REAL
#undef SFX

/* ---------------------------------------------------------------------------
 * Image ingest.  u8 as is; u16 as is (right aligned 12-bit must stay <= 4095)
 * or divided by 16 when left aligned (src/oc_image.cpp:163-185).
 * ------------------------------------------------------------------------- */
int or_image_to_float(const void* px, int pixel, int width, int height,
                      size_t stride_bytes, int depth12, float* out)
{
    for (int y = 0; y < height; y++) {
        const uint8_t* row = (const uint8_t*)px + (size_t)y * stride_bytes;
        for (int x = 0; x < width; x++) {
            float v;
            if (pixel == 0) v = (float)row[x];
            else if (pixel == 1) {
                uint16_t s;
                memcpy(&s, row + 2 * (size_t)x, 2);
                if (depth12 == 1 && s > 4095) return -1;
                v = (float)s;
                if (depth12 == 2) v = (float)((double)v * (1.0 / 16.0));
            } else {
                memcpy(&v, row + 4 * (size_t)x, 4);
            }
            out[(size_t)y * width + x] = v;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * FFTCC DFT specification.
 *
 * FFTW's r2c_2d(n0 = n, n1 = n) on the row-major window window[y*n + x] keeps
 * the half spectrum along x: X[ky][kx], ky < n, kx <= n/2.  The c2r inverse is
 * unnormalised.  The restatement computes both as separable DFTs with one
 * twiddle table and a fixed operation order, so that any implementation that
 * performs the same float operations in the same order reproduces every bit:
 *
 *   forward, along x:  A[y][kx]  = sum_{x asc} w[y][x] * tw[(x*kx) % n]
 *                        re = fma(v, wr, re); im = fma(v, wi, im)
 *   forward, along y:  X[ky][kx] = sum_{y asc} A[y][kx] * tw[(y*ky) % n]
 *                        re = fma(ar, wr, re); re = fma(-ai, wi, re);
 *                        im = fma(ar, wi, im); im = fma(ai, wr, im)
 *   product:           Z = conj(R) * T = (Rr*Tr + Ri*Ti, Rr*Ti - Ri*Tr), unfused
 *   inverse, along y:  G[y][kx]  = sum_{ky asc} Z[ky][kx] * conj(tw[(y*ky) % n])
 *                        re = fma(zr, wr, re); re = fma(zi, wi, re);
 *                        im = fma(zi, wr, im); im = fma(-zr, wi, im)
 *   inverse, along x:  out[y][x] = G[y][0].re
 *                        + sum_{kx=1..n/2-1} [s = gi*wi; s = fma(gr, wr, s);
 *                                             acc = fma(2, s, acc)]
 *                        + (x odd ? -1 : 1) * G[y][n/2].re
 * Each sum starts from 0.0f and fma is the correctly rounded fused
 * multiply-add (C99 fmaf).  Imaginary parts of the kx = 0 and kx = n/2 bins
 * are ignored by the inverse, as FFTW's c2r does.
 *
 * Window mean and squared norm use row partial sums: for each row the values
 * are summed left to right from 0.0f, and the row sums are added top to
 * bottom from 0.0f (the reference sums serially, src/oc_fftcc.cpp:415-437).
 * ------------------------------------------------------------------------- */
void or_dft_twiddles(int n, float* tw)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < n; k++) {
        double a = two_pi * (double)k / (double)n;
        tw[2 * k] = (float)cos(a);
        tw[2 * k + 1] = (float)(-sin(a));
    }
}

static void dft_r2c_2d(const float* x, int n, const float* tw, float* tmp, float* X)
{
    const int hc = n / 2 + 1;
    for (int y = 0; y < n; y++)
        for (int kx = 0; kx < hc; kx++) {
            float re = 0.f, im = 0.f;
            for (int xx = 0; xx < n; xx++) {
                int t = (xx * kx) % n;
                float v = x[y * n + xx];
                re = fmaf(v, tw[2 * t], re);
                im = fmaf(v, tw[2 * t + 1], im);
            }
            tmp[2 * (y * hc + kx)] = re;
            tmp[2 * (y * hc + kx) + 1] = im;
        }
    for (int kx = 0; kx < hc; kx++)
        for (int ky = 0; ky < n; ky++) {
            float re = 0.f, im = 0.f;
            for (int y = 0; y < n; y++) {
                int t = (y * ky) % n;
                float ar = tmp[2 * (y * hc + kx)], ai = tmp[2 * (y * hc + kx) + 1];
                float wr = tw[2 * t], wi = tw[2 * t + 1];
                re = fmaf(ar, wr, re);
                re = fmaf(-ai, wi, re);
                im = fmaf(ar, wi, im);
                im = fmaf(ai, wr, im);
            }
            X[2 * (ky * hc + kx)] = re;
            X[2 * (ky * hc + kx) + 1] = im;
        }
}

static void dft_c2r_2d(const float* Z, int n, const float* tw, float* tmp, float* out)
{
    const int hc = n / 2 + 1;
    for (int kx = 0; kx < hc; kx++)
        for (int y = 0; y < n; y++) {
            float re = 0.f, im = 0.f;
            for (int ky = 0; ky < n; ky++) {
                int t = (y * ky) % n;
                float zr = Z[2 * (ky * hc + kx)], zi = Z[2 * (ky * hc + kx) + 1];
                float wr = tw[2 * t], wi = tw[2 * t + 1];
                re = fmaf(zr, wr, re);
                re = fmaf(zi, wi, re);
                im = fmaf(zi, wr, im);
                im = fmaf(-zr, wi, im);
            }
            tmp[2 * (y * hc + kx)] = re;
            tmp[2 * (y * hc + kx) + 1] = im;
        }
    for (int y = 0; y < n; y++)
        for (int xx = 0; xx < n; xx++) {
            float acc = tmp[2 * (y * hc)];
            for (int kx = 1; kx < hc - 1; kx++) {
                int t = (xx * kx) % n;
                float gr = tmp[2 * (y * hc + kx)], gi = tmp[2 * (y * hc + kx) + 1];
                float s = gi * tw[2 * t + 1];
                s = fmaf(gr, tw[2 * t], s);
                acc = fmaf(2.f, s, acc);
            }
            float nyq = tmp[2 * (y * hc + hc - 1)];
            acc += (xx & 1) ? -nyq : nyq;
            out[y * n + xx] = acc;
        }
}

static void cross_power(const float* R, const float* T, int len, float* Z)
{
    for (int i = 0; i < len; i++) {
        float rr = R[2 * i], ri = R[2 * i + 1], tr = T[2 * i], ti = T[2 * i + 1];
        Z[2 * i] = (rr * tr) + (ri * ti);
        Z[2 * i + 1] = (rr * ti) - (ri * tr);
    }
}

static float window_mean(const float* v, int n)
{
    float tot = 0.f;
    for (int y = 0; y < n; y++) {
        float rs = 0.f;
        for (int x = 0; x < n; x++) rs += v[y * n + x];
        tot += rs;
    }
    return tot / (float)(n * n);
}

/* subtracts the mean in place and returns the sum of squares (not sqrt) */
static float window_zero_mean_sq(float* v, int n, float mean)
{
    float tot = 0.f;
    for (int y = 0; y < n; y++) {
        float rs = 0.f;
        for (int x = 0; x < n; x++) {
            float d = v[y * n + x] - mean;
            v[y * n + x] = d;
            rs += d * d;
        }
        tot += rs;
    }
    return tot;
}

void or_fftcc_correlate(const float* ref_win, const float* tar_win, int n, float* out)
{
    const int hc = n / 2 + 1;
    float* tw = (float*)malloc(sizeof(float) * 2 * n);
    float* tmp = (float*)malloc(sizeof(float) * 2 * n * hc);
    float* R = (float*)malloc(sizeof(float) * 2 * n * hc);
    float* T = (float*)malloc(sizeof(float) * 2 * n * hc);
    float* Z = (float*)malloc(sizeof(float) * 2 * n * hc);
    or_dft_twiddles(n, tw);
    dft_r2c_2d(ref_win, n, tw, tmp, R);
    dft_r2c_2d(tar_win, n, tw, tmp, T);
    cross_power(R, T, n * hc, Z);
    dft_c2r_2d(Z, n, tw, tmp, out);
    free(tw); free(tmp); free(R); free(T); free(Z);
}

typedef struct fftcc_ws {
    float *tw, *refw, *tarw, *tmp, *R, *T, *Z, *out;
} fftcc_ws;

static void fftcc_ws_alloc(fftcc_ws* ws, int n)
{
    const int hc = n / 2 + 1;
    ws->tw = (float*)malloc(sizeof(float) * 2 * n);
    ws->refw = (float*)malloc(sizeof(float) * n * n);
    ws->tarw = (float*)malloc(sizeof(float) * n * n);
    ws->out = (float*)malloc(sizeof(float) * n * n);
    ws->tmp = (float*)malloc(sizeof(float) * 2 * n * hc);
    ws->R = (float*)malloc(sizeof(float) * 2 * n * hc);
    ws->T = (float*)malloc(sizeof(float) * 2 * n * hc);
    ws->Z = (float*)malloc(sizeof(float) * 2 * n * hc);
    or_dft_twiddles(n, ws->tw);
}

static void fftcc_ws_free(fftcc_ws* ws)
{
    free(ws->tw); free(ws->refw); free(ws->tarw); free(ws->out);
    free(ws->tmp); free(ws->R); free(ws->T); free(ws->Z);
}

/* argmax (first maximum above -2, strict >), wrap, store: :368-393 / :467-495 */
static void fftcc_finish(or_poi* poi, const float* corr, int r, float u0, float v0,
                         float ref_norm, float tar_norm)
{
    const int n = 2 * r, size = n * n;
    float mx = -2.f;
    int mi = 0;
    for (int i = 0; i < size; i++)
        if (corr[i] > mx) { mx = corr[i]; mi = i; }
    int lu = mi % n, lv = mi / n;
    if (lu > r) lu -= n;
    if (lv > r) lv -= n;
    poi->p[0] = (float)lu + u0;
    poi->p[6] = (float)lv + v0;
    poi->r[0] = u0;
    poi->r[1] = v0;
    poi->r[2] = (float)((double)mx / (sqrt((double)(ref_norm * tar_norm)) * (double)size));
}

/* target window of computePreparedRef / compute(POI2D*): pixel
 * (int)(y + r' - r + v0), (int)(x + c - r + u0) in float arithmetic (:341-343) */
static void fill_target_window(const float* tar, int w, const or_poi* poi, int r, float u0,
                               float v0, float* dst)
{
    const int n = 2 * r;
    for (int rr = 0; rr < n; rr++)
        for (int c = 0; c < n; c++) {
            float rxp = poi->x + c - r;
            float ryp = poi->y + rr - r;
            float txp = rxp + u0, typ = ryp + v0;
            dst[rr * n + c] = tar[(size_t)(int)typ * w + (int)txp];
        }
}

static int target_window_outside(const or_poi* poi, int w, int h, int r, float u0, float v0)
{
    return (int)(poi->x + u0) < r || (int)(poi->x + u0) >= w - r
        || (int)(poi->y + v0) < r || (int)(poi->y + v0) >= h - r;
}

static void fftcc_one(const float* ref, const float* tar, int w, int h, int r, or_poi* poi,
                      fftcc_ws* ws)
{
    const int n = 2 * r, hc = r + 1;
    const float u0 = poi->p[0], v0 = poi->p[6];
    if ((int)poi->x < r || (int)poi->x >= w - r || (int)poi->y < r || (int)poi->y >= h - r
        || target_window_outside(poi, w, h, r, u0, v0))
        return;
    for (int rr = 0; rr < n; rr++)
        for (int c = 0; c < n; c++) {
            float rxp = poi->x + c - r;
            float ryp = poi->y + rr - r;
            ws->refw[rr * n + c] = ref[(size_t)(int)ryp * w + (int)rxp];
        }
    fill_target_window(tar, w, poi, r, u0, v0, ws->tarw);
    float ref_mean = window_mean(ws->refw, n);
    float tar_mean = window_mean(ws->tarw, n);
    float ref_norm = window_zero_mean_sq(ws->refw, n, ref_mean);
    float tar_norm = window_zero_mean_sq(ws->tarw, n, tar_mean);
    dft_r2c_2d(ws->refw, n, ws->tw, ws->tmp, ws->R);
    dft_r2c_2d(ws->tarw, n, ws->tw, ws->tmp, ws->T);
    cross_power(ws->R, ws->T, n * hc, ws->Z);
    dft_c2r_2d(ws->Z, n, ws->tw, ws->tmp, ws->out);
    fftcc_finish(poi, ws->out, r, u0, v0, ref_norm, tar_norm);
}

void or_fftcc_compute(const float* ref, const float* tar, int w, int h, int r,
                      or_poi* pois, size_t n, int threads)
{
#pragma omp parallel num_threads(threads)
    {
        fftcc_ws ws;
        fftcc_ws_alloc(&ws, 2 * r);
#pragma omp for schedule(dynamic, 16)
        for (long i = 0; i < (long)n; i++) fftcc_one(ref, tar, w, h, r, &pois[i], &ws);
        fftcc_ws_free(&ws);
    }
}

void or_fftcc_prepare_ref(const float* ref, int w, int h, int r, const or_poi* layout,
                          size_t n, float* spectra, float* norm, int* valid,
                          int* xs, int* ys, int threads)
{
    const int nn = 2 * r, hc = r + 1, len = nn * hc;
#pragma omp parallel num_threads(threads)
    {
        fftcc_ws ws;
        fftcc_ws_alloc(&ws, nn);
#pragma omp for schedule(dynamic, 16)
        for (long i = 0; i < (long)n; i++) {
            const int px = (int)layout[i].x, py = (int)layout[i].y;
            float* spec = spectra + 2 * (size_t)len * i;
            xs[i] = px;
            ys[i] = py;
            memset(spec, 0, sizeof(float) * 2 * len);
            norm[i] = 0.f;
            valid[i] = 0;
            if (px < r || px >= w - r || py < r || py >= h - r) continue;
            for (int rr = 0; rr < nn; rr++)
                for (int c = 0; c < nn; c++)
                    ws.refw[rr * nn + c] = ref[(size_t)(py + rr - r) * w + (px + c - r)];
            float mean = window_mean(ws.refw, nn);
            norm[i] = window_zero_mean_sq(ws.refw, nn, mean);
            dft_r2c_2d(ws.refw, nn, ws.tw, ws.tmp, spec);
            valid[i] = 1;
        }
        fftcc_ws_free(&ws);
    }
}

void or_fftcc_compute_prepared(const float* tar, int w, int h, int r,
                               const float* spectra, const float* norm, const int* valid,
                               or_poi* pois, size_t n, const size_t* idx, int threads)
{
    const int nn = 2 * r, hc = r + 1, len = nn * hc;
#pragma omp parallel num_threads(threads)
    {
        fftcc_ws ws;
        fftcc_ws_alloc(&ws, nn);
#pragma omp for schedule(dynamic, 16)
        for (long i = 0; i < (long)n; i++) {
            or_poi* poi = &pois[i];
            const size_t slot = idx ? idx[i] : (size_t)i;
            const float u0 = poi->p[0], v0 = poi->p[6];
            if (!valid[slot] || target_window_outside(poi, w, h, r, u0, v0)) continue;
            fill_target_window(tar, w, poi, r, u0, v0, ws.tarw);
            float mean = window_mean(ws.tarw, nn);
            float tar_norm = window_zero_mean_sq(ws.tarw, nn, mean);
            dft_r2c_2d(ws.tarw, nn, ws.tw, ws.tmp, ws.T);
            cross_power(spectra + 2 * (size_t)len * slot, ws.T, len, ws.Z);
            dft_c2r_2d(ws.Z, nn, ws.tw, ws.tmp, ws.out);
            fftcc_finish(poi, ws.out, r, u0, v0, norm[slot], tar_norm);
        }
        fftcc_ws_free(&ws);
    }
}
