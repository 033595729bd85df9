/*
 * CPU oracle for the 2D-DIC hot path: a plain-C restatement of OpenCorr's
 * FFTCC2D, Gradient2D4, BicubicBspline and ICGN2D1 (first order), written from
 * the reference sources cited beside each function in oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fasticgn_amd/, include/)
 * links, loads or calls this library.  It is used by tests/, by the smoke()
 * check in __graft_entry__.py and by the cpu_baseline leg of bench.py, and
 * only as the checker or the timed CPU baseline, never as a compute path.
 *
 * Two precisions are built from one source:
 *   *_f32  faithful FP32 restatement (the parity oracle),
 *   *_f64  the same algorithm carried out in FP64 (the "FP64 oracle").
 *
 * Pinning status per function is listed in oracle.c's header comment.
 */
#ifndef FI_ORACLE_H
#define FI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout-identical to opencorr::POI2D (src/oc_poi.h:102-108): Point2D x,y,
 * DeformationVector2D p[12], Result2D r[6], StrainVector2D e[3],
 * subset_radius (x,y).  25 floats = 100 bytes. */
typedef struct or_poi {
    float x, y;
    float p[12]; /* u ux uy uxx uxy uyy v vx vy vxx vxy vyy */
    float r[6];  /* u0 v0 zncc iteration convergence feature */
    float e[3];  /* exx eyy exy */
    float sr[2]; /* subset_radius x, y */
} or_poi;

/* ---- image ingest (src/oc_image.cpp:125-217) ---------------------------- */
/* pixel: 0 = u8, 1 = u16, 2 = f32.  mode: depth (0 as stored, 1 12-bit
 * right, 2 12-bit left, 3 12-bit inferred alignment, 4 inferred depth) |
 * normalise << 4 (0 none, 1 to 8 bits, 2 to 16 bits, 3 unit).  Output
 * row-major float, width*height.  0, -1 (right-aligned 12-bit above 4095) or
 * -2 (mode not valid for the pixel type). */
int or_image_to_float(const void* px, int pixel, int width, int height,
                      size_t stride_bytes, int mode, float* out);

/* ---- Gradient2D4 (src/oc_gradient.cpp:37-79) ----------------------------- */
void or_gradient_f32(const float* img, int w, int h, float* gx, float* gy, int threads);
void or_gradient_f64(const float* img, int w, int h, double* gx, double* gy, int threads);

/* ---- BicubicBspline (src/oc_cubic_bspline.cpp:84-181) -------------------- */
void or_bspline_prepare_f32(const float* img, int w, int h, float* coef, int threads);
void or_bspline_prepare_f64(const float* img, int w, int h, double* coef, int threads);
float or_bspline_eval_f32(const float* coef, int w, int h, float x, float y);
double or_bspline_eval_f64(const double* coef, int w, int h, double x, double y);

/* ---- FFTCC2D (src/oc_fftcc.cpp:246-557) ---------------------------------- */
/* Twiddle table shared by the FFTCC DFT specification: tw[2k] = cos(2*pi*k/n),
 * tw[2k+1] = -sin(2*pi*k/n), each rounded from double. */
void or_dft_twiddles(int n, float* tw);
/* Circular cross-correlation of two zero-mean n x n windows through the
 * DFT specification (R2C, conj(R)*T, C2R), unnormalised like FFTW. */
void or_fftcc_correlate(const float* ref_win, const float* tar_win, int n, float* out);
/* window mean / norm summation order: 0 row partials (default, the device's),
 * 1 one serial sum in row order (the reference's) */
void or_fftcc_set_serial_sums(int on);
/* Uncached compute over a queue (FFTCC2D::compute(POI2D*), :396-496). */
void or_fftcc_compute(const float* ref, const float* tar, int w, int h, int r,
                      or_poi* pois, size_t n, int threads);
/* Cached reference spectra (FFTCC2D::prepareRef(queue), :246-313).
 * spectra: n * (2r) * (r+1) complex (float pairs); norm, valid, xs, ys: n. */
void or_fftcc_prepare_ref(const float* ref, int w, int h, int r, const or_poi* layout,
                          size_t n, float* spectra, float* norm, int* valid,
                          int* xs, int* ys, int threads);
/* computePreparedRef (:317-394) for POI i against cache slot idx[i]
 * (idx == NULL means slot i). */
void or_fftcc_compute_prepared(const float* tar, int w, int h, int r,
                               const float* spectra, const float* norm, const int* valid,
                               or_poi* pois, size_t n, const size_t* idx, int threads);

/* ---- ICGN2D1 (src/oc_icgn.cpp:39-43, 168-387) ----------------------------- */
void or_icgn_compute_f32(const float* ref, int w, int h, const float* gx, const float* gy,
                         const float* coef, int rx, int ry, float conv, float stop,
                         int self_adaptive, or_poi* pois, size_t n, int threads);
void or_icgn_compute_f64(const float* ref, int w, int h, const double* gx, const double* gy,
                         const double* coef, int rx, int ry, double conv, double stop,
                         int self_adaptive, or_poi* pois, size_t n, int threads);
/* compute(POI2D*, Point2D& center_offset) / compute(queue, center_offset_queue)
 * (src/oc_icgn.cpp:399-625): offsets holds n (x, y) pairs. */
void or_icgn_compute_offset_f32(const float* ref, int w, int h, const float* gx, const float* gy,
                                const float* coef, int rx, int ry, float conv, float stop,
                                int self_adaptive, or_poi* pois, size_t n, const float* offsets,
                                int threads);
void or_icgn_compute_offset_f64(const float* ref, int w, int h, const double* gx, const double* gy,
                                const double* coef, int rx, int ry, double conv, double stop,
                                int self_adaptive, or_poi* pois, size_t n, const float* offsets,
                                int threads);

/* ---- linear algebra pieces, exported for the oracle's own tests ---------- */
/* 6x6 inverse by partial-pivot LU (Eigen PartialPivLU semantics), row-major. */
void or_inverse6_f32(const float* a, float* inv);
void or_inverse6_f64(const double* a, double* inv);

#ifdef __cplusplus
}
#endif

#endif
