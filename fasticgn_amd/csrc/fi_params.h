// Kernel parameter blocks and device-resident record layouts, shared by the
// host engines and the gfx950 kernels.  Plain structs, passed by value.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/fasticgn.h"

namespace fi {

enum PixelKind { kU8 = FI_PIXEL_U8, kU16 = FI_PIXEL_U16, kF32 = FI_PIXEL_F32 };

// A raw image resident on the device (row-major, pitch in elements).
struct DevImage {
    const void* data = nullptr;
    int width = 0, height = 0;
    int pixel = kF32;
    size_t pitch = 0;    // elements per row
    float scale = 1.f;   // 1, or 1/16 for left-aligned 12-bit input
};

// Per-POI reference-side data kept resident across frames (the reference
// recomputes it per POI per call on the CPU, src/oc_icgn.cpp:198-234, and
// caches it per layout on its GPU).  176 bytes, float4-aligned.
struct alignas(16) PoiPrep {
    float mean;       // reference subset mean
    float norm;       // sqrt(sum r^2), r = ref - mean
    int32_t status;   // 1 prepared, 0 rejected by the entry gate
    int32_t rx;       // radius the record was prepared for
    float hinv[36];   // inverse Hessian, row-major (FP64 inversion, stored FP32)
    float ref_sum;    // sum r (zero up to rounding), for the ZNCC
    int32_t ry;
    float pad[2];
};
static_assert(sizeof(PoiPrep) == 176, "PoiPrep layout");

// Target B-spline table: for texel row y, four planes of float4, plane k at
// float4 index (4 * y + k) * w + x holding coef[k][0..3] (the dy^k row), then
// three planes of zero padding so a plane offset added to any in-table texel
// offset stays inside the allocation.
inline size_t bspline_table_floats(int w, int h) { return (size_t)16 * w * h + (size_t)12 * w; }

// FFTCC reference-spectrum cache slot metadata.
struct FftccSlot {
    int32_t x, y;    // integer centre the spectrum was taken at
    float norm;      // sum of squares of the zero-mean window (not sqrt)
    int32_t valid;
};

enum FftccMode { kFftccPrepare = 0, kFftccCached = 1, kFftccUncached = 2 };

struct FftccParams {
    const float* ref;         // reference, f32 row-major, pitch = w
    const float* tar;         // target, f32 row-major, pitch = w
    int w, h;
    int r;                    // radius; window n = 2r
    int mode;
    fi_poi* pois;             // queue (prepare: the layout, read only)
    int n_pois;
    const uint32_t* slot;     // cached mode: slot per POI, nullptr = identity
    float* spectra;           // slots * n * (r+1) complex, interleaved
    FftccSlot* slots;
    const float* twiddles;    // 2n floats
    const unsigned* count;    // device POI count, at most n_pois (nullptr: n_pois); set, the
                              // blocks map to POIs in launch order
};

// Claim counters of one persistent-queue IC-GN launch: kQueueParts counters
// (one per XCD range), each on its own 256-byte line, since device-scope
// atomics from every XCD to one line serialize.
constexpr int kQueueParts = 8;
constexpr int kQueueStride = 64;  // ints between counters
constexpr int kQueueWords = kQueueParts * kQueueStride;

struct IcgnParams {
    const float* ref;         // reference, f32 row-major
    const float* gx;
    const float* gy;
    const float* coef;        // target B-spline table (bspline_table_floats); nullptr: an
                              // iteration that outgrows the LDS tile samples tar instead
    const float* tar;         // target, f32 row-major (nullptr: table sampling only)
    int w, h;
    fi_poi* pois;
    const uint32_t* index;    // POIs this launch processes (nullptr = all)
    int n;                    // number of POIs this launch processes
    PoiPrep* prep;            // per-POI reference records
    const uint32_t* prep_index;  // record of POI i (nullptr = record i)
    int rx, ry;               // uniform radius (self_adaptive = 0)
    int self_adaptive;
    float conv, stop;
    uint32_t* iters;          // optional: executed iterations of launch entry b
    // per-POI LDS tile of the target (set by the launcher; tile_w = 0: none)
    int tile_w, tile_h, tile_stride;
    // persistent work queue (nullptr: one block per entry): 8 zeroed claim
    // counters, and the number of resident blocks to launch
    int* queue;
    const unsigned* count;    // device entry count, at most n (nullptr: n); set, block b runs
                              // entry b (one block per entry, no queue)
    int group;                // 1: one-wave shapes run as k_icgn_group (four neighbouring
                              // entries share a coefficient tile of tile_h rows)
    const float* offset;      // per POI (indexed like pois) an (x, y) pair: subset centre
                              // offset of compute(POI2D*, Point2D&), nullptr = none
    int queue_blocks;
    double conv2;             // set by the launcher: the iteration continues while the FP64
                              // squared norm is >= conv2, i.e. (float)sqrt(norm^2) >= conv
};

// Neighbourhood consensus of learned seeds: POIs binned into square cells
// (row-major, ncx x ncy) of side >= radius / k, so the POIs within +-radius of
// a POI lie in the (2k+1) x (2k+1) cells around its own.
struct SeedCells {
    const uint32_t* start;  // ncx * ncy + 1 offsets into idx
    const uint32_t* idx;    // POI positions, grouped by cell
    const int32_t* cell;    // per POI position: its cell, -1 = none (non-finite position)
    int ncx, ncy;
    int k;                  // cells searched on each side
    int m;                  // length of idx (the POIs with a cell)
    const uint8_t* sample;  // per POI position: 1 = seeded by the network, 0 = by its neighbours
                            // alone; nullptr = every POI by the network
    float radius;           // neighbours within +-radius in x and in y
    float tau;              // largest |u, v| disagreement of a neighbour's transported seed
};

}  // namespace fi
