/*
 * fasticgn.h -- C ABI of the MI355X (gfx950) 2D-DIC hot path.
 *
 * The reference exposes this path only as C++ classes (namespace opencorr):
 *   FFTCC2D        src/oc_fftcc.h:56-100      (CPU)   FFTCC2DCuda  src/fftcc2d_cuda.h:57-114
 *   ICGN2D1        src/oc_icgn.h:45-77        (CPU)   ICGN2D1CudaMvp src/icgn2d1_cuda_mvp.h:62-143
 *   POI2D          src/oc_poi.h:102-136
 * and the multi-frame driver inside its CLI (opencorr_cli_cuda_mvp.cpp:3024-3210,
 * 3834-4509).  This header is the thin C boundary those classes sit on in this
 * project (include/fasticgn_opencorr.hpp) and that Python reaches via ctypes.
 *
 * Conventions (mirroring the reference's library contract):
 *   - Every function returns an int status: FI_OK (0) or a negative FI_E_* code.
 *     Nothing throws across the ABI.  fi_last_error() returns the message of
 *     the last failure on the calling host thread.
 *   - Per-POI numerical failures are not errors: they are the reference's zncc
 *     sentinels -3 (rejected at entry / out of interpolatable range), -4 (not
 *     converged within the iteration cap), -5 (NaN), src/oc_dic.h:29-35.
 *   - fi_poi is layout-identical to opencorr::POI2D (25 floats, 100 bytes), so
 *     std::vector<POI2D>::data() and a numpy structured array pass through.
 *   - Calls are synchronous: results are in caller memory when a call returns.
 *     The caller owns every buffer it passes.  One handle is never used from
 *     two host threads at once; distinct handles may be.
 *   - Images are row-major.  pixel: FI_PIXEL_U8 / U16 / F32.  The image
 *     argument named depth12 is the input mode of loadImage2DMat
 *     (src/oc_image.cpp:125-217): a depth FI_DEPTH_* or-ed with a normalise
 *     FI_NORMALIZE_* (below); the pixels stay as uploaded in HBM and the
 *     kernels convert them as they load them.
 */
#ifndef FASTICGN_H
#define FASTICGN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define FI_OK 0
#define FI_E_INVALID_ARGUMENT (-1) /* std::invalid_argument in the reference */
#define FI_E_LOGIC (-2)            /* std::logic_error */
#define FI_E_OUT_OF_RANGE (-3)     /* std::out_of_range */
#define FI_E_DEVICE (-4)           /* HIP runtime failure (std::runtime_error) */
#define FI_E_UNSUPPORTED (-5)      /* valid for the reference, not for this device path */
#define FI_E_NO_MEMORY (-6)

/* ---- pixel formats ------------------------------------------------------ */
#define FI_PIXEL_U8 0
#define FI_PIXEL_U16 1
#define FI_PIXEL_F32 2
/* Input mode (ImageLoadOptions, src/oc_image.h:26-53; loadImage2DMat,
 * src/oc_image.cpp:125-217) = depth | normalise.
 *   FI_DEPTH_NATIVE   the stored type's depth: 8-bit data 8-bit, 16-bit data 16-bit
 *   FI_DEPTH12_RIGHT  12-bit right-aligned (values above 4095 are an error)
 *   FI_DEPTH12_LEFT   12-bit left-aligned (divided by 16)
 *   FI_DEPTH12_AUTO   12-bit, left-aligned when the maximum exceeds 4095
 *   FI_DEPTH16        16-bit (ImageInputDepthMode::Depth16)
 *   FI_DEPTH_AUTO     inferred: 8-bit data 8-bit; 16-bit data up to 4095
 *                     12-bit right-aligned, else 16-bit (ImageInputDepthMode::Auto)
 *   FI_NORMALIZE_TO8 / TO16 / UNIT scale to 0..255 / 0..65535 / 0..1 from the
 *   logical maximum of the effective depth (255, 4095 or 65535).  The scale is
 *   formed in FP64 and applied as float(pixel) * float(scale).  Depth modes
 *   take 8- or 16-bit pixels; float pixels take FI_DEPTH_NATIVE only. */
#define FI_DEPTH_NATIVE 0
#define FI_DEPTH12_RIGHT 1
#define FI_DEPTH12_LEFT 2
#define FI_DEPTH12_AUTO 3
#define FI_DEPTH_AUTO 4
#define FI_DEPTH16 5
#define FI_DEPTH_MASK 0x0f
#define FI_NORMALIZE_NONE 0x00
#define FI_NORMALIZE_TO8 0x10
#define FI_NORMALIZE_TO16 0x20
#define FI_NORMALIZE_UNIT 0x30
#define FI_NORMALIZE_MASK 0xf0

/* How an image is converted (ImageLoadInfo, src/oc_image.h:55-73). */
typedef struct fi_ingest_info {
    int32_t effective_bits;   /* 8, 12, 16, or 32 for float pixels */
    int32_t left_aligned;     /* 12-bit data divided by 16 */
    double source_max;        /* maximum pixel value (16-bit data, when the mode needs it), else 0 */
    double scale;             /* the FP64 scale; pixels become float(p) * float(scale) */
} fi_ingest_info;

/* opencorr::POI2D (src/oc_poi.h:102-108): Point2D x,y; DeformationVector2D
 * p[12] = u ux uy uxx uxy uyy v vx vy vxx vxy vyy; Result2D r[6] = u0 v0 zncc
 * iteration convergence feature; StrainVector2D e[3]; subset_radius x,y. */
typedef struct fi_poi {
    float x, y;
    float deformation[12];
    float result[6];
    float strain[3];
    float subset_radius[2];
} fi_poi;

/* Per-frame quality record of the raw manifest (opencorr_cli_cuda_mvp.cpp:654-723). */
typedef struct fi_frame_stats {
    int32_t frame_index;
    int32_t skip_fftcc_used;      /* frame seeded from the previous result */
    int32_t fallback_used;        /* whole-frame FFTCC re-run and per-POI merge */
    int32_t poi_fallback_count;   /* POI-level repair candidates */
    int32_t poi_fallback_recovered;
    int32_t iteration_max;
    float bad_ratio;              /* after repair / merge */
    float pre_repair_bad_ratio;
    float baseline_bad_ratio;     /* NaN until established */
    float fallback_threshold;     /* NaN until established */
    float valid_ratio;
    float iteration_mean;
    int64_t valid_count;
    int64_t iteration_sum;        /* sum of executed IC-GN iterations */
    int64_t fftcc_pois;           /* POIs run through FFTCC this frame */
    double device_ms;             /* device time of the frame (HIP events) */
    double bspline_ms;            /* target B-spline table (on its own stream, beside the seeding;
                                     built only when a launch samples it, else 0) */
    double fftcc_ms;              /* FFTCC seeding (or temporal seed copy) */
    double icgn_ms;               /* first IC-GN pass over the frame */
    int64_t seed_pois;            /* POIs seeded by the learned seeder */
    int64_t seed_fallback_count;  /* of those, re-run from FFTCC (low ZNCC or failure) */
    double seed_ms;               /* learned-seeder inference */
    int32_t temporal_seed_count;  /* seeds set by the two-frame extrapolation (cli:3063-3067) */
    int32_t temporal_seed_used;   /* temporal_seed_count > 0 */
} fi_frame_stats;

typedef struct fi_fftcc fi_fftcc;
typedef struct fi_icgn fi_icgn;
typedef struct fi_frame fi_frame;
typedef struct fi_seeder fi_seeder;

/* ---- library ------------------------------------------------------------ */
int fi_version(int* major, int* minor, int* patch);
int fi_device_count(int* count);
/* hipDeviceSynchronize on the given device (timing brackets). */
int fi_device_synchronize(int device);
/* Copies the message of the calling thread's last failure (NUL-terminated,
 * truncated to len).  Returns the length of the full message. */
int fi_last_error(char* buf, size_t len);
/* Resolves an input mode for the given pixels (the effective depth and
 * alignment, inferred from the maximum where the mode says so, and the scale)
 * without uploading them; FI_E_INVALID_ARGUMENT as an upload would fail. */
int fi_image_resolve(const void* px, int width, int height, size_t stride_bytes, int pixel, int mode,
                     fi_ingest_info* out);
/* The FFTCC DFT twiddle table for length n (2n floats: cos, -sin), as the
 * device kernels use it.  For tests of the DFT specification. */
int fi_dft_twiddles(int n, float* out);
/* The FP64 threshold the IC-GN kernels compare their squared convergence norm
 * with (the loop test of src/oc_icgn.cpp:343-352 without a root per
 * iteration): (float)sqrt(s2) >= conv exactly when s2 >= *out, an exact tie
 * of sqrt(s2) with the rounding midpoint below conv aside.  Host only; for
 * tests of the loop test. */
int fi_conv_threshold2(float conv, double* out);

/* ---- FFTCC2D -------------------------------------------------------------
 * FFTCC2D(int rx, int ry, int threads)    src/oc_fftcc.cpp:154-170
 * FFTCC2DCuda(int rx, int ry, int device) src/fftcc2d_cuda.h:85-109
 * Square radii 1..23170 are valid (else FI_E_INVALID_ARGUMENT, as the
 * reference throws invalid_argument); this device path runs 1..32 and returns
 * FI_E_UNSUPPORTED above that. */
int fi_fftcc_create(int rx, int ry, int device, fi_fftcc** out);
int fi_fftcc_destroy(fi_fftcc* h);
/* setRefImage / setTarImage (:188-202): content-change notifications. */
int fi_fftcc_set_ref(fi_fftcc* h, const void* px, int width, int height,
                     size_t stride_bytes, int pixel, int depth12);
int fi_fftcc_set_tar(fi_fftcc* h, const void* px, int width, int height,
                     size_t stride_bytes, int pixel, int depth12);
/* prepareRef() (:231-236) clears the layout; prepareRef(queue) (:246-313)
 * caches the reference spectra of that layout on the device. */
int fi_fftcc_prepare_ref_clear(fi_fftcc* h);
int fi_fftcc_prepare_ref(fi_fftcc* h, const fi_poi* layout, size_t n);
int fi_fftcc_has_prepared_ref(fi_fftcc* h, int* out);
/* compute(vector<POI2D>&) (:529-557): cached when the queue matches the
 * prepared layout, uncached otherwise. */
int fi_fftcc_compute(fi_fftcc* h, fi_poi* pois, size_t n);
/* compute(queue, cache_indices) (:498-527): indexed against the cached layout;
 * FI_E_INVALID_ARGUMENT / FI_E_LOGIC / FI_E_OUT_OF_RANGE as the reference throws. */
int fi_fftcc_compute_indexed(fi_fftcc* h, fi_poi* pois, size_t n, const size_t* cache_indices);

/* ---- ICGN2D1 -------------------------------------------------------------
 * ICGN2D1(int rx, int ry, float conv, float stop, int threads) src/oc_icgn.cpp:95-112
 * ICGN2D1CudaMvp(rx, ry, conv, int stop, device, dual_gpu)    src/icgn2d1_cuda_mvp.h:98-136
 * This device path runs radii 1..31. */
int fi_icgn_create(int rx, int ry, float conv, float stop, int device, fi_icgn** out);
int fi_icgn_destroy(fi_icgn* h);
int fi_icgn_set_iteration(fi_icgn* h, float conv, float stop);       /* :120-124 */
int fi_icgn_set_self_adaptive(fi_icgn* h, int on);                   /* src/oc_dic.h:62 */
/* DIC::setSubset (src/oc_dic.cpp:28-32): the subset radii of later computes
 * (1..31 on the device, FI_E_UNSUPPORTED above). */
int fi_icgn_set_subset(fi_icgn* h, int rx, int ry);
/* Lane-to-POI shape of the IC-GN kernel: waves per POI (1, 2, 4) and rows per
 * band (the subset's columns map to lanes; 64 * waves / (2rx+1) bands of rows
 * stack vertically and must cover the 2ry+1 rows); 0, 0 restores the default.
 * A shape that does not fit the subset fails at compute (FI_E_INVALID_ARGUMENT).
 * No reference counterpart (the CUDA engine picks its block size by occupancy,
 * cu:331-359).  Results depend on the shape only through FP32 summation order. */
int fi_icgn_set_shape(fi_icgn* h, int waves, int rows_per_band);
/* Where the IC-GN kernel reads the target: FI_SAMPLING_AUTO stages each
 * subset's raw neighbourhood in LDS and evaluates the B-spline polynomial from
 * the 16 pixels around a sample (iterations whose warp outgrows the staged
 * tile read the B-spline table here, and the same 16 pixels from the target
 * in device memory in the frame driver, which builds no table for this mode);
 * FI_SAMPLING_TABLE reads every sample's 16 coefficients from the table
 * (src/oc_cubic_bspline.cpp:134-181).  All evaluate the same polynomial and
 * differ only in FP32 rounding.  No reference counterpart. */
#define FI_SAMPLING_AUTO 0
#define FI_SAMPLING_TABLE 1
/* FI_SAMPLING_GROUP: four neighbouring POIs (consecutive in the queue), one
 * wave each, share one LDS tile of per-pixel x-polynomial coefficients (Horner
 * rows instead of weights x pixels); an iteration that leaves it reads the
 * target in device memory.  One-wave lane shapes only; others run as AUTO. */
#define FI_SAMPLING_GROUP 2
/* FI_SAMPLING_PAIR: two POIs per wave (lanes 0-31 and 32-63, each with the
 * band layout of a 32-lane subset), eight neighbouring POIs per block sharing
 * one coefficient tile as in FI_SAMPLING_GROUP: the per-iteration sums,
 * solve and warp update serve two POIs.  Uniform subsets whose 32-lane layout
 * the build instantiates (radius 15); others run as FI_SAMPLING_GROUP.  Results
 * differ from the other modes only in FP32 summation order. */
#define FI_SAMPLING_PAIR 3
int fi_icgn_set_sampling(fi_icgn* h, int mode);
/* setRefImage / setTarImage.  The IC-GN kernels address the f32 images with
 * 32-bit offsets: a reference whose width * height * 4 bytes exceeds 2 GiB
 * returns FI_E_UNSUPPORTED (the reference itself has no such limit). */
int fi_icgn_set_ref(fi_icgn* h, const void* px, int width, int height,
                    size_t stride_bytes, int pixel, int depth12);
int fi_icgn_set_tar(fi_icgn* h, const void* px, int width, int height,
                    size_t stride_bytes, int pixel, int depth12);
int fi_icgn_prepare_ref(fi_icgn* h);   /* Gradient2D4 of the reference (:132-142) */
int fi_icgn_prepare_tar(fi_icgn* h);   /* BicubicBspline table of the target (:144-153) */
int fi_icgn_prepare(fi_icgn* h);       /* both (:155-159) */
/* compute(vector<POI2D>&) (:389-397, per POI :168-387). */
int fi_icgn_compute(fi_icgn* h, fi_poi* pois, size_t n);
/* compute(POI2D*, Point2D& center_offset) / compute(vector<POI2D>&,
 * vector<Point2D>& center_offset_queue) (src/oc_icgn.cpp:399-625; Zhou et al.,
 * Opt. Lasers Eng. 55 (2014) 5-11): center_offsets holds n (x, y) pairs; POI i
 * is matched with its warp centred at (x + ox, y + oy), so u, v are the
 * displacement of that point.  A non-finite offset gives the reference's
 * per-POI statuses (-5 for NaN, -3 out of range). */
int fi_icgn_compute_offset(fi_icgn* h, fi_poi* pois, size_t n, const float* center_offsets);

/* compute(vector<POI2D>&) on n POIs the caller keeps in device memory (updated
 * in place), allocating nothing: no device or pinned memory, no host staging.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-POI reference records live in a static device array of
 * the library that holds at most 65,536 POIs (11.5 MB per GPU, held until
 * the process ends); a larger n is FI_E_INVALID_ARGUMENT.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU race.  Uniform
 * subsets only (self-adaptive subsets are FI_E_UNSUPPORTED); prepareRef and
 * prepareTar must have run.  Synchronous on the handle's stream. */
int fi_icgn_compute_noalloc(fi_icgn* h, fi_poi* device_pois, size_t n);

/* Copies one of the engine's device-side arrays to host memory, for tests of
 * the preparation stages against the oracle; bytes must be the array's exact
 * size.  No reference counterpart (the reference keeps them private):
 *   FI_EXPORT_REF_F32   reference as f32 (w * h floats)         after prepareRef
 *   FI_EXPORT_GRAD_X/Y  Gradient2D4 of the reference (w * h)    after prepareRef
 *   FI_EXPORT_TAR_F32   target as f32 (w * h)                   after prepareTar
 *   FI_EXPORT_BSPLINE   BicubicBspline table, texel-major, 16 floats coef[k][l]
 *                       (k * 4 + l) per texel (w * h * 16)      after prepareTar
 *   FI_EXPORT_PREP      the last compute's per-POI reference records, 176 bytes
 *                       each: mean, norm (f32), status, rx (int32), inverse
 *                       Hessian (36 f32, row-major), sum (ref - mean) (f32),
 *                       ry (int32), 8 bytes of padding */
#define FI_EXPORT_REF_F32 0
#define FI_EXPORT_GRAD_X 1
#define FI_EXPORT_GRAD_Y 2
#define FI_EXPORT_TAR_F32 3
#define FI_EXPORT_BSPLINE 4
#define FI_EXPORT_PREP 5
int fi_icgn_export(fi_icgn* h, int what, void* out, size_t bytes);

/* ---- multi-frame driver --------------------------------------------------
 * The fixed-reference sequence of opencorr_cli_cuda_mvp.cpp:3834-4509 with the
 * frame policy of processFrame (:3024-3210): the reference, its gradients,
 * per-POI Hessians and FFTCC reference spectra stay resident; each frame
 * uploads one target.  policy flags: */
#define FI_POLICY_SKIP_FFTCC 1u   /* seed frames > 0 from the previous result (:560-624) */
#define FI_POLICY_POI_FALLBACK 2u /* POI-level repair of bad POIs (:2955-3022) */
#define FI_POLICY_DEFAULT (FI_POLICY_SKIP_FFTCC | FI_POLICY_POI_FALLBACK)
int fi_frame_create(int radius, float conv, float stop, int device, fi_frame** out);
int fi_frame_destroy(fi_frame* h);
/* Reference image of the sequence; as fi_icgn_set_ref, a reference above
 * 2 GiB as f32 returns FI_E_UNSUPPORTED. */
int fi_frame_set_ref(fi_frame* h, const void* px, int width, int height,
                     size_t stride_bytes, int pixel, int depth12);
/* POI template (buildPoiQueue, :2439-2472, or any explicit list). */
int fi_frame_set_pois(fi_frame* h, const fi_poi* pois, size_t n);
/* Queue one target (host memory) for the next fi_frame_run: a staging thread
 * copies it to pinned memory and the device while a running frame computes
 * (two slots; a third queued target is FI_E_LOGIC).  The pixels must stay
 * valid until the next fi_frame_run / fi_frame_autotune returns.  A run with
 * no new target processes the last one again.  The 12-bit range of a
 * FI_DEPTH12_RIGHT target is checked as the staging thread copies it: the
 * fi_frame_run (or autotune) that takes a target above 4095 fails with
 * FI_E_INVALID_ARGUMENT (src/oc_image.cpp:172-178) and the driver stays usable. */
int fi_frame_upload_target(fi_frame* h, const void* px, int width, int height,
                           size_t stride_bytes, int pixel, int depth12);
/* Process the uploaded target with the given policy; results stay on the
 * device (fi_frame_download copies them out).  stats may be NULL.  On a frame
 * seeded from the previous result, the whole-frame fallback (bad ratio above
 * baseline + 0.03: FFTCC + IC-GN from the template, merged per POI by
 * shouldTakeFrameFallbackResult, cli:3086-3152) is decided and run on the
 * device, inside the frame's graph when graphs are on. */
int fi_frame_run(fi_frame* h, unsigned policy, fi_frame_stats* stats);
/* fi_frame_run in two halves, for a caller that keeps the device busy across
 * the frame loop (processFrame per target, opencorr_cli_cuda_mvp.cpp:3024-3210,
 * which the reference runs one frame after another): fi_frame_run_async
 * launches the frame and returns; fi_frame_wait returns the oldest launched
 * frame's statistics once it has finished (FI_E_LOGIC when none is waiting).
 * A frame is queued on the device behind one still running when both are
 * replayed as graphs (fi_frame_set_graph), are not seeded from the previous
 * frame (no FI_POLICY_SKIP_FFTCC) and need no B-spline table; otherwise the
 * launch first completes the frames in flight (their statistics wait for
 * fi_frame_wait).  At most two frames are in flight.  Results, records and
 * statistics are those of fi_frame_run frame for frame.  Every other frame
 * call but fi_frame_upload_target and fi_frame_records completes the frames
 * in flight first; fi_frame_run fails with FI_E_LOGIC while any launched frame
 * has not been waited for.  Typical loop: run_async(k), wait(k - 1) and read
 * its records (age 1), upload_target(k + 1). */
int fi_frame_run_async(fi_frame* h, unsigned policy);
int fi_frame_wait(fi_frame* h, fi_frame_stats* stats);
int fi_frame_download(fi_frame* h, fi_poi* out, size_t n);
/* Per POI of the last frame (template order): 1 if it was run through FFTCC
 * (an FFTCC-seeded frame, the whole-frame fallback, POI-level repair or the
 * learned seed's fallback), else 0. */
int fi_frame_download_redo(fi_frame* h, uint8_t* out, size_t n);
/* Output records of a frame (formatRawBinary, cli:2498-2536, no output
 * filter): n 36-byte records x, y (int32), u, v, ux, uy, vx, vy, zncc
 * (float32) in template order, packed on the device.  age 0 = the last frame
 * run, 1 = the one before it.  *records points into pinned memory the driver
 * owns; it stays valid until the second fi_frame_run after the frame it
 * holds.  With fi_frame_set_records on, every run packs its records and
 * copies them to host memory on a copy stream that overlaps the next frame,
 * and this call only waits for that copy (so the records of age 1, read
 * after the next run, cost the host no wait); otherwise the last frame's
 * (age 0 only) are packed and copied now.  *n_records (may be NULL) = n. */
int fi_frame_records(fi_frame* h, int age, const void** records, size_t* n_records);
int fi_frame_set_records(fi_frame* h, int on);
/* Two-frame temporal extrapolation of the seeds (seedPoiQueueFromPrevious
 * with --temporal-extrapolation, cli:560-624, 3364-3371): on a frame seeded
 * from the previous result, a POI whose last two results are good (finite,
 * zncc >= 0.8) at its template location starts from u + (u - u_before),
 * v + (v - v_before) when that step is at least 0.05 px and each component at
 * most max_delta px (max_delta 0 = the driver's radius, as the CLI passes its
 * subset radius).  Off by default, as in the reference.  fi_frame_stats
 * counts the extrapolated seeds. */
int fi_frame_set_temporal_extrapolation(fi_frame* h, int on, float max_delta);
/* Reference-side preparation of the driver (prepareRef of both engines,
 * cli:4095-4104): device buffers, reference f32 + Gradient2D4, FFTCC reference
 * spectra and the per-POI IC-GN records (mean, norm, inverse Hessian).  The
 * first fi_frame_run after a new reference or template does it by itself;
 * fi_frame_prepare does it now (if needed) and reports the timing of the last
 * preparation. */
typedef struct fi_prepare_stats {
    double total_ms;      /* host wall time of the preparation */
    double alloc_ms;      /* device buffer allocation (host wall time) */
    double reference_ms;  /* reference f32 + gradients (device, HIP events) */
    double fftcc_ms;      /* FFTCC reference spectra (device) */
    double hessian_ms;    /* per-POI records: mean, norm, FP64 inverse Hessian (device) */
} fi_prepare_stats;
int fi_frame_prepare(fi_frame* h, fi_prepare_stats* stats);
/* Forget the previous frames (next run seeds by FFTCC again). */
int fi_frame_reset_sequence(fi_frame* h);
/* IC-GN lane-to-POI shape for this driver (see fi_icgn_set_shape). */
int fi_frame_set_shape(fi_frame* h, int waves, int rows_per_band);
/* IC-GN target sampling for this driver (see fi_icgn_set_sampling). */
int fi_frame_set_sampling(fi_frame* h, int mode);
/* Mixed subset radii (DIC::setSelfAdaptive, src/oc_dic.h:44-65, with the
 * per-POI subset_radius of ICGN2D1, src/oc_icgn.cpp:176-182): each template
 * POI carries a square subset_radius (1..31) used by both FFTCC and IC-GN;
 * POIs are grouped into radius classes on the device (downloads keep the
 * template order).  Off: every POI uses the driver's radius. */
int fi_frame_set_self_adaptive(fi_frame* h, int on);
/* Capture each frame's device work (target prefilter, seeding, IC-GN of every
 * radius class, fallback scan and decision, whole-frame fallback, POI-level
 * repair, QA) as a hipGraph and replay it; only repair candidates past the
 * in-graph capacity (n / 32 per radius class, at least 1,024) run eagerly
 * after it.  Results are byte-identical to
 * the eager frame; per-stage times in fi_frame_stats are NaN. */
int fi_frame_set_graph(fi_frame* h, int on);
/* IC-GN launch of every radius class: blocks > 0 runs the persistent
 * work-queue kernel with that many resident blocks claiming POIs from
 * per-XCD queues; 0 (default) launches one block per POI.  Replaces tuned
 * launch modes. */
int fi_frame_set_persistent(fi_frame* h, int blocks);
/* With several radius classes: on = one IC-GN launch over all of them (each
 * POI at its own radius, the lane shape and LDS tile of the largest radius),
 * off (default) = one launch per class.  The launch mode follows
 * fi_frame_set_persistent.  The autotuner chooses between the two. */
int fi_frame_set_fused(fi_frame* h, int on);
/* Autotuner: on the uploaded target (FFTCC-seeded), time every IC-GN lane
 * shape, sampling mode and launch mode (one block per POI, or the persistent
 * work queue with 12 waves per CU or with as many blocks as every CU holds)
 * for each radius class and keep the fastest; with several
 * classes, also one launch over all of them (report rows with radius 0),
 * kept when faster than the best per-class launches together.  Takes the
 * queued target as a frame does.  report (may be
 * NULL) receives up to report_len timed configurations; *report_n (may be
 * NULL) the number timed.  Does not advance the frame sequence. */
typedef struct fi_tune_entry {
    int32_t radius;
    int32_t waves;
    int32_t rows_per_band;
    int32_t sampling;     /* FI_SAMPLING_* */
    int32_t persistent;   /* resident blocks of the work-queue launch, 0 = one block per POI */
    int32_t chosen;       /* 1: kept for the radius class */
    double ms;            /* median IC-GN time of the class */
} fi_tune_entry;
int fi_frame_autotune(fi_frame* h, fi_tune_entry* report, size_t report_len, size_t* report_n);

/* ---- learned seeder --------------------------------------------------------
 * No reference counterpart: the seeder that replaces FFTCC2D as the IC-GN
 * initial guess.  A four-layer patchify CNN maps a reference/target crop pair
 * (32x32 each around the POI, the target crop shifted by the POI's rounded
 * u, v) to (u, v, ux, uy, vx, vy); BF16 MFMA with FP32 accumulation, trained on
 * the device by fi_seeder_train_epoch.  The parameter vector (FP32) is
 * W1 (32x16), b1, W2 (64x32), b2, W3 (128x64), b3, W4 (256x128), b4,
 * W5 (128x6), b5; outputs are (u, v) in pixels and gradients x 100.
 * Layer-1 input rows (BF16 bit patterns, 64 rows of 32 per sample) are
 * patch (py, px) of the normalised crops, column c*16 + dy*4 + dx. */
int fi_seeder_create(int device, fi_seeder** out);
int fi_seeder_destroy(fi_seeder* h);
int fi_seeder_param_count(size_t* n);
int fi_seeder_set_params(fi_seeder* h, const float* params, size_t n);
int fi_seeder_get_params(fi_seeder* h, float* params, size_t n);
/* Training set: n samples of layer-1 rows (n*64*32 BF16) and scaled targets
 * (n*6 FP32), kept on the device. */
int fi_seeder_set_training_data(fi_seeder* h, const uint16_t* x1, const float* targets, size_t n);
/* One epoch of Adam over a seeded shuffle in batches (multiple of 16; the
 * remainder of the shuffle is dropped).  mean_loss may be NULL. */
int fi_seeder_train_epoch(fi_seeder* h, int batch, float lr, uint64_t seed, float* mean_loss);
/* The training step's loss (mean squared error over n*6 scaled outputs) and
 * its gradient with respect to the parameter vector (n = batch, a multiple of
 * 16), computed by the kernels fi_seeder_train_epoch runs before its Adam
 * update; the parameters are not changed.  loss may be NULL. */
int fi_seeder_gradient(fi_seeder* h, const uint16_t* x1, const float* targets, size_t n, float* grads,
                       float* loss);
/* Network outputs (n*6, scaled) for layer-1 rows, through the training
 * kernels. */
int fi_seeder_forward(fi_seeder* h, const uint16_t* x1, size_t n, float* out);
/* Seed POIs in place from f32 images (the fused inference kernel); ok[i] = 1
 * where both crops lie inside the images (others untouched).  ok may be NULL. */
int fi_seeder_seed(fi_seeder* h, const float* ref, const float* tar, int width, int height, fi_poi* pois,
                   size_t n, uint32_t* ok);
/* fi_seeder_seed, then the neighbourhood consensus of the predictions: each
 * seeded POI takes the mean of its seeded neighbours' predictions within
 * +-radius px in x and y, each neighbour's u, v carried to the POI with that
 * neighbour's predicted gradients, over the neighbours whose carried u, v
 * agree with the POI's own within tau px; the mean gradients carry the mean
 * u, v from the neighbours' centroid to the POI.  passes (1..4): further seeder
 * passes as in fi_frame_set_seed_passes, each followed by the consensus over
 * the first pass's ok (the frame driver's learned seeding, run alone).
 * sample: FI_SEED_SAMPLE_* (below).  No reference counterpart. */
int fi_seeder_seed_consensus(fi_seeder* h, const float* ref, const float* tar, int width, int height,
                             fi_poi* pois, size_t n, uint32_t* ok, float radius, float tau, int passes,
                             int sample);

/* Frame driver with the learned seeder (policy FI_POLICY_LEARNED_SEED): the
 * FFTCC seeding of a frame is replaced by the seeder; POIs whose IC-GN result
 * has ZNCC below zncc_threshold (or failed) are re-run from FFTCC and merged
 * (better ZNCC wins); a POI the seeder could not seed takes the FFTCC re-run
 * as it is, so its record is the FFTCC-seeded frame's.  The seeder must
 * outlive its use. */
#define FI_POLICY_LEARNED_SEED 4u
int fi_frame_set_seeder(fi_frame* h, fi_seeder* seeder, float zncc_threshold);
/* Seeder passes per frame (1..4, default 1): pass k crops the target at the
 * rounded u, v of pass k-1, so a POI that starts far from its displacement
 * (the seeder's +-8 px range) is seeded from a near crop by the second pass. */
int fi_frame_set_seed_passes(fi_frame* h, int passes);
/* Neighbourhood consensus of the seeder's predictions after each pass (as
 * fi_seeder_seed_consensus, radius = the driver's subset radius; default on,
 * tau 0.5 px, the network on a checkerboard of the cells).  The seeds of a smooth field then lose most of the network's
 * per-POI error, and IC-GN needs fewer iterations; a neighbour across a
 * discontinuity disagrees by more than tau and is left out.  No reference
 * counterpart. */
int fi_frame_set_seed_consensus(fi_frame* h, int on, float tau, int sample);
/* Which POIs the network seeds when the consensus is on: every POI, a
 * checkerboard of the consensus cells (about half), or the cells with both
 * coordinates even (about a quarter); every other POI takes the consensus of
 * its network-seeded neighbours, the tau test then measured from their mean
 * (none left within tau: not seeded, FFTCC fallback). */
#define FI_SEED_SAMPLE_ALL 0
#define FI_SEED_SAMPLE_CHECKER 1
#define FI_SEED_SAMPLE_QUARTER 2

/* ---- raw output record (opencorr_cli_cuda_mvp.cpp:2498-2536) --------------
 * 36-byte little-endian record x,y (int32), u,v,ux,uy,vx,vy,zncc (float32). */
int fi_format_raw(const fi_poi* pois, size_t n, void* out, size_t out_len, size_t* written);
/* One raw frame file (frame_NNNNNN.bin of the opencorr-gpu-raw-v1 tree,
 * cli:3700-3720): the n records of fi_format_raw written to path (created or
 * truncated); *bytes (may be NULL) = bytes written.  The manifest
 * (writeRawManifest, cli:2615-2732) is written by the host layer
 * (fasticgn_amd/rawio.py). */
int fi_write_raw(const fi_poi* pois, size_t n, const char* path, size_t* bytes);

#ifdef __cplusplus
}
#endif

#endif
