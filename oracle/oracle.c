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
 *   src/oc_image.cpp:125-217             depth, alignment, normalise or_image_to_float
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
/* loadImage2DMat (src/oc_image.cpp:125-217) on decoded pixels.  mode = depth
 * (bits 0-3: 0 as stored, 1 12-bit right, 2 12-bit left, 3 12-bit with the
 * alignment inferred (left when the maximum exceeds 4095), 4 depth inferred:
 * 16-bit data up to 4095 is 12-bit right-aligned, else 16-bit, 5 16-bit) | normalise
 * (bits 4-7: 0 none, 1 to 8 bits, 2 to 16 bits, 3 to [0, 1], each relative to
 * the logical maximum 255 / 4095 / 65535 of the effective depth).  The scale
 * is formed in FP64 as there and applied as float(pixel) * float(scale).
 * Returns 0, -1 for right-aligned 12-bit data above 4095, -2 for a mode the
 * pixel type does not take (float pixels are taken as stored). */
int or_image_to_float(const void* px, int pixel, int width, int height,
                      size_t stride_bytes, int mode, float* out)
{
    const int depth = mode & 15, norm = (mode >> 4) & 15;
    if (depth > 5 || norm > 3 || (mode & ~0xff)) return -2;
    if (pixel == 2 && (depth != 0 || norm != 0)) return -2;
    unsigned mx = 0;
    for (int y = 0; y < height; y++) {
        const uint8_t* row = (const uint8_t*)px + (size_t)y * stride_bytes;
        for (int x = 0; x < width; x++) {
            unsigned v = 0;
            if (pixel == 0) v = row[x];
            else if (pixel == 1) { uint16_t s; memcpy(&s, row + 2 * (size_t)x, 2); v = s; }
            if (v > mx) mx = v;
        }
    }
    /* effective depth (8, 12, 16, or 0 for float) and alignment */
    int eff = pixel == 0 ? 8 : pixel == 2 ? 0 : 16, left = 0;
    if (pixel != 2) {
        if (depth == 1) eff = 12;
        else if (depth == 2) { eff = 12; left = 1; }
        else if (depth == 3) { eff = 12; left = mx > 4095; }
        else if (depth == 5) eff = 16;
        else if (depth == 4 && pixel == 1) eff = mx <= 4095 ? 12 : 16;
        if (eff == 12 && !left && mx > 4095) return -1;
    }
    double scale = left ? 1.0 / 16.0 : 1.0;
    const double logical = eff == 8 ? 255.0 : eff == 12 ? 4095.0 : 65535.0;
    if (norm == 1) scale *= 255.0 / logical;
    else if (norm == 2) scale *= 65535.0 / logical;
    else if (norm == 3) scale *= 1.0 / logical;
    const float s32 = (float)scale;
    for (int y = 0; y < height; y++) {
        const uint8_t* row = (const uint8_t*)px + (size_t)y * stride_bytes;
        for (int x = 0; x < width; x++) {
            float v;
            if (pixel == 0) v = (float)row[x];
            else if (pixel == 1) { uint16_t s; memcpy(&s, row + 2 * (size_t)x, 2); v = (float)s; }
            else memcpy(&v, row + 4 * (size_t)x, 4);
            out[(size_t)y * width + x] = v * s32;
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * FFTCC transform specification.
 *
 * FFTW's r2c_2d(n0 = n, n1 = n) on the row-major window w[y*n + x] keeps the
 * half spectrum along x: X[ky][kx], ky < n, kx <= n/2; its c2r inverse is
 * unnormalised and ignores the imaginary parts of the kx = 0 and kx = n/2
 * bins.  FFTW's own rounding is not reproducible, so this restatement fixes one
 * algorithm, written out to the individual float operation; any
 * implementation performing the same operations in the same order (the gfx950
 * kernel does) reproduces every bit.
 *
 * fft(n): length-n complex DFT, X[k] = sum_j x[j] e^{-2 pi i jk/n}, by
 * recursive decimation in time.  Base cases n = 1..5 are the butterflies
 * below.  Otherwise p = 4 if 4 | n, else 2 if 2 | n, else 3 if 3 | n, else 5
 * if 5 | n, else n (a prime > 5: direct sum).  With m = n/p:
 *   Y_r = fft(m) of x[r], x[r+p], ...              (r = 0..p-1)
 *   T_r = Y_r[k] if r*k == 0 else cmul(Y_r[k], tw[r*k*step])
 *   X[k + m*t], t = 0..p-1 = butterfly_p(T_0..T_{p-1})   (k = 0..m-1)
 * tw[j] = (cos 2pi j/N, -sin 2pi j/N) rounded from double, N the top-level
 * size, step = N/n.  cmul(a, w) = (fma(ax, wx, -(ay*wy)), fma(ax, wy, ay*wx)).
 * Butterflies (componentwise on (re, im)):
 *   2: X0 = a + b; X1 = a - b
 *   3: s1 = b + c; s2 = b - c; X0 = a + s1; m = a - 0.5*s1; q = K3*s2;
 *      X1 = (m.x + q.y, m.y - q.x); X2 = (m.x - q.y, m.y + q.x)
 *   4: s0 = a + c; s1 = a - c; s2 = b + d; s3 = b - d; X0 = s0 + s2;
 *      X2 = s0 - s2; X1 = (s1.x + s3.y, s1.y - s3.x); X3 = (s1.x - s3.y, s1.y + s3.x)
 *   5: t1 = x1 + x4; t2 = x2 + x3; t3 = x1 - x4; t4 = x2 - x3;
 *      X0 = x0 + (t1 + t2); a1 = x0 + fma(C1, t1, C2*t2); a2 = x0 + fma(C2, t1, C1*t2);
 *      b1 = fma(S1, t3, S2*t4); b2 = fma(S2, t3, -(S1*t4));
 *      X1 = (a1.x + b1.y, a1.y - b1.x); X4 = (a1.x - b1.y, a1.y + b1.x);
 *      X2 = (a2.x + b2.y, a2.y - b2.x); X3 = (a2.x - b2.y, a2.y + b2.x)
 *   prime p > 5: X[t] = x[0] + sum_{r=1}^{p-1} (j = r*t mod p; j == 0 ? x[r] : cmul(x[r], tw[j*step]))
 *   with K3 = sin(2pi/3), C1 = cos(2pi/5), C2 = cos(4pi/5), S1 = sin(2pi/5),
 *   S2 = sin(4pi/5) as float literals.
 * ifft(z) = conj(fft(conj(z))), unnormalised.
 *
 * 2-D forward (real window w, n even): rows in pairs (2j, 2j+1):
 *   z[x] = (w[2j][x], w[2j+1][x]); Z = fft(z); for kx <= n/2, nk = (n - kx) % n:
 *   A[kx] = (0.5*(Z[kx].x + Z[nk].x), 0.5*(Z[kx].y - Z[nk].y))  -> row 2j
 *   B[kx] = (0.5*(Z[kx].y + Z[nk].y), 0.5*(Z[nk].x - Z[kx].x))  -> row 2j+1
 * then fft along y of every column kx.
 * Cross power: Z = conj(R) * T = (Rr*Tr + Ri*Ti, Rr*Ti - Ri*Tr), unfused.
 * 2-D inverse: ifft along y of every column, then rows in pairs: the
 * imaginary parts of kx = 0 and kx = n/2 of both rows are zeroed, then for
 * k <= n/2: Zr[k] = (A.x - B.y, A.y + B.x); for k > n/2, j = n - k:
 * Zr[k] = (A[j].x + B[j].y, B[j].x - A[j].y); z = ifft(Zr); row 2j = z.x,
 * row 2j+1 = z.y.
 *
 * Window mean and squared norm use row partial sums by default: for each row
 * the values are summed left to right from 0.0f, and the row sums are added
 * top to bottom from 0.0f (the device's order).  or_fftcc_set_serial_sums(1)
 * switches to the reference's order, one running sum over the window in row
 * order (src/oc_fftcc.cpp:280-299, 415-437), so the two orders can be
 * compared.
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

typedef struct cpx { float x, y; } cpx;

static const float K3 = 0.866025403784438646763723170753f;
static const float C51 = 0.309016994374947424102293417183f;
static const float C52 = -0.809016994374947424102293417183f;
static const float S51 = 0.951056516295153572116439333379f;
static const float S52 = 0.587785252292473129168705954639f;

static cpx cadd(cpx a, cpx b) { cpx r = { a.x + b.x, a.y + b.y }; return r; }
static cpx csub(cpx a, cpx b) { cpx r = { a.x - b.x, a.y - b.y }; return r; }
static cpx cmul(cpx a, cpx w)
{
    cpx r;
    r.x = fmaf(a.x, w.x, -(a.y * w.y));
    r.y = fmaf(a.x, w.y, a.y * w.x);
    return r;
}
static cpx cconj(cpx a) { cpx r = { a.x, -a.y }; return r; }

static int fft_radix(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    return n;
}

static void butterfly(int p, const cpx* T, cpx* X, int m, const cpx* tw, int step)
{
    /* X[m*t] for t < p from T[0..p-1] */
    if (p == 2) {
        X[0] = cadd(T[0], T[1]);
        X[m] = csub(T[0], T[1]);
    } else if (p == 3) {
        cpx s1 = cadd(T[1], T[2]), s2 = csub(T[1], T[2]);
        X[0] = cadd(T[0], s1);
        cpx mm = { T[0].x - 0.5f * s1.x, T[0].y - 0.5f * s1.y };
        cpx q = { K3 * s2.x, K3 * s2.y };
        cpx x1 = { mm.x + q.y, mm.y - q.x }, x2 = { mm.x - q.y, mm.y + q.x };
        X[m] = x1;
        X[2 * m] = x2;
    } else if (p == 4) {
        cpx s0 = cadd(T[0], T[2]), s1 = csub(T[0], T[2]), s2 = cadd(T[1], T[3]), s3 = csub(T[1], T[3]);
        X[0] = cadd(s0, s2);
        X[2 * m] = csub(s0, s2);
        cpx x1 = { s1.x + s3.y, s1.y - s3.x }, x3 = { s1.x - s3.y, s1.y + s3.x };
        X[m] = x1;
        X[3 * m] = x3;
    } else if (p == 5) {
        cpx t1 = cadd(T[1], T[4]), t2 = cadd(T[2], T[3]), t3 = csub(T[1], T[4]), t4 = csub(T[2], T[3]);
        X[0] = cadd(T[0], cadd(t1, t2));
        cpx a1 = { T[0].x + fmaf(C51, t1.x, C52 * t2.x), T[0].y + fmaf(C51, t1.y, C52 * t2.y) };
        cpx a2 = { T[0].x + fmaf(C52, t1.x, C51 * t2.x), T[0].y + fmaf(C52, t1.y, C51 * t2.y) };
        cpx b1 = { fmaf(S51, t3.x, S52 * t4.x), fmaf(S51, t3.y, S52 * t4.y) };
        cpx b2 = { fmaf(S52, t3.x, -(S51 * t4.x)), fmaf(S52, t3.y, -(S51 * t4.y)) };
        cpx x1 = { a1.x + b1.y, a1.y - b1.x }, x4 = { a1.x - b1.y, a1.y + b1.x };
        cpx x2 = { a2.x + b2.y, a2.y - b2.x }, x3 = { a2.x - b2.y, a2.y + b2.x };
        X[m] = x1;
        X[2 * m] = x2;
        X[3 * m] = x3;
        X[4 * m] = x4;
    } else {
        /* prime p > 5: direct sum; step here is N/p */
        for (int t = 0; t < p; t++) {
            cpx acc = T[0];
            for (int r = 1; r < p; r++) {
                int j = (r * t) % p;
                acc = (j == 0) ? cadd(acc, T[r]) : cadd(acc, cmul(T[r], tw[j * step]));
            }
            X[t * m] = acc;
        }
    }
}

/* X[0..n-1] = fft of x[0], x[s], ..., x[(n-1)s]; step = N / n */
static void fft_rec(const cpx* x, int s, cpx* X, int n, const cpx* tw, int step)
{
    if (n == 1) {
        X[0] = x[0];
        return;
    }
    int p = (n <= 5) ? n : fft_radix(n);
    int m = n / p;
    if (m == 1) {
        cpx T[64];
        for (int r = 0; r < p; r++) T[r] = x[r * s];
        butterfly(p, T, X, 1, tw, step);
        return;
    }
    cpx Y[64];
    for (int r = 0; r < p; r++) fft_rec(x + r * s, s * p, Y + r * m, m, tw, step * p);
    for (int k = 0; k < m; k++) {
        cpx T[64];
        for (int r = 0; r < p; r++)
            T[r] = (r * k == 0) ? Y[r * m + k] : cmul(Y[r * m + k], tw[r * k * step]);
        butterfly(p, T, X + k, m, tw, step * m);
    }
}

static void fft_n(const cpx* x, cpx* X, int n, const cpx* tw) { fft_rec(x, 1, X, n, tw, 1); }

static void ifft_n(const cpx* z, cpx* out, int n, const cpx* tw)
{
    cpx a[64] = { { 0.f, 0.f } }, b[64];
    for (int i = 0; i < n; i++) a[i] = cconj(z[i]);
    fft_rec(a, 1, b, n, tw, 1);
    for (int i = 0; i < n; i++) out[i] = cconj(b[i]);
}

/* forward 2-D of the real window x (n x n) into X (n rows ky, hc columns kx) */
static void dft_r2c_2d(const float* x, int n, const float* twf, float* tmp, float* Xf)
{
    const int hc = n / 2 + 1;
    const cpx* tw = (const cpx*)twf;
    cpx* S = (cpx*)tmp;
    cpx* X = (cpx*)Xf;
    for (int j = 0; j < n / 2; j++) {
        cpx z[64], Z[64];
        for (int xx = 0; xx < n; xx++) {
            z[xx].x = x[(2 * j) * n + xx];
            z[xx].y = x[(2 * j + 1) * n + xx];
        }
        fft_n(z, Z, n, tw);
        for (int kx = 0; kx < hc; kx++) {
            int nk = (n - kx) % n;
            cpx A = { 0.5f * (Z[kx].x + Z[nk].x), 0.5f * (Z[kx].y - Z[nk].y) };
            cpx B = { 0.5f * (Z[kx].y + Z[nk].y), 0.5f * (Z[nk].x - Z[kx].x) };
            S[(2 * j) * hc + kx] = A;
            S[(2 * j + 1) * hc + kx] = B;
        }
    }
    for (int kx = 0; kx < hc; kx++) {
        cpx col[64], C[64];
        for (int y = 0; y < n; y++) col[y] = S[y * hc + kx];
        fft_n(col, C, n, tw);
        for (int ky = 0; ky < n; ky++) X[ky * hc + kx] = C[ky];
    }
}

/* inverse 2-D of the half spectrum Z into the real n x n out */
static void dft_c2r_2d(const float* Zf, int n, const float* twf, float* tmp, float* out)
{
    const int hc = n / 2 + 1;
    const cpx* tw = (const cpx*)twf;
    const cpx* Z = (const cpx*)Zf;
    cpx* G = (cpx*)tmp;
    for (int kx = 0; kx < hc; kx++) {
        cpx col[64], g[64];
        for (int ky = 0; ky < n; ky++) col[ky] = Z[ky * hc + kx];
        ifft_n(col, g, n, tw);
        for (int y = 0; y < n; y++) G[y * hc + kx] = g[y];
    }
    for (int j = 0; j < n / 2; j++) {
        cpx A[33], B[33], Zr[64], z[64];
        for (int k = 0; k < hc; k++) {
            A[k] = G[(2 * j) * hc + k];
            B[k] = G[(2 * j + 1) * hc + k];
        }
        A[0].y = 0.f; B[0].y = 0.f; A[hc - 1].y = 0.f; B[hc - 1].y = 0.f;
        for (int k = 0; k < n; k++) {
            if (k <= n / 2) {
                Zr[k].x = A[k].x - B[k].y;
                Zr[k].y = A[k].y + B[k].x;
            } else {
                int q = n - k;
                Zr[k].x = A[q].x + B[q].y;
                Zr[k].y = B[q].x - A[q].y;
            }
        }
        ifft_n(Zr, z, n, tw);
        for (int xx = 0; xx < n; xx++) {
            out[(2 * j) * n + xx] = z[xx].x;
            out[(2 * j + 1) * n + xx] = z[xx].y;
        }
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

static int fftcc_serial_sums = 0;

void or_fftcc_set_serial_sums(int on) { fftcc_serial_sums = on != 0; }

static float window_mean(const float* v, int n)
{
    float tot = 0.f;
    if (fftcc_serial_sums) {
        for (int i = 0; i < n * n; i++) tot += v[i];
    } else {
        for (int y = 0; y < n; y++) {
            float rs = 0.f;
            for (int x = 0; x < n; x++) rs += v[y * n + x];
            tot += rs;
        }
    }
    return tot / (float)(n * n);
}

/* subtracts the mean in place and returns the sum of squares (not sqrt) */
static float window_zero_mean_sq(float* v, int n, float mean)
{
    float tot = 0.f;
    if (fftcc_serial_sums) {
        for (int i = 0; i < n * n; i++) {
            v[i] -= mean;
            tot += v[i] * v[i];
        }
        return tot;
    }
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
