// Learned seeder: a four-layer patchify CNN mapping a reference/target subset
// pair (two 32x32 crops, each normalised to zero mean and unit RMS) to the six
// first-order shape parameters (u, v, ux, uy, vx, vy).  Every layer is
// "2x2 (first layer 4x4) non-overlapping patches -> linear -> ReLU", so each is
// one GEMM on a permuted copy of the previous activations:
//
//   layer  grid  rows/sample  K    N     (weights K x N)
//   1      8x8   64           32   16    input 2 channels x 4x4 pixels
//   2      4x4   16           64   32
//   3      2x2   4            128  64
//   4      1x1   1            256  128
//   head                      128  6     (FP32, no activation)
//
// GEMMs run on BF16 MFMA with FP32 accumulation (training:
// v_mfma_f32_16x16x16_bf16; fused inference: v_mfma_f32_16x16x32_bf16, head
// on v_mfma_f32_16x16x4_f32); activations and weights are rounded to BF16
// (round to nearest even) before each GEMM.  Outputs are scaled: (u, v) in
// pixels, gradients x kGradScale.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fi_params.h"

namespace fi {

constexpr int kSeedCrop = 32;       // crop side; the subset centre is pixel 15
constexpr int kSeedCropOrigin = 15; // crop column 0 = x - 15
constexpr float kGradScale = 100.f; // network outputs gradients x 100
constexpr int kSeedLayers = 4;
constexpr int kSeedRows[kSeedLayers] = { 64, 16, 4, 1 };
constexpr int kSeedK[kSeedLayers] = { 32, 64, 128, 256 };
constexpr int kSeedN[kSeedLayers] = { 16, 32, 64, 128 };
constexpr int kSeedOut = 6;

// Parameter vector (FP32) in this order: W1 (K1 x N1, row-major), b1, W2, b2,
// W3, b3, W4, b4, W5 (128 x 6), b5.
constexpr size_t seed_param_count()
{
    size_t n = 0;
    for (int l = 0; l < kSeedLayers; l++) n += (size_t)kSeedK[l] * kSeedN[l] + kSeedN[l];
    return n + 128 * kSeedOut + kSeedOut;
}

// Device weights for the kernels: BF16 copies of W1..W4 transposed to
// N x K (training GEMMs), the packed MFMA fragments of the fused inference
// kernel (k_pack_infer), FP32 biases, FP32 head.
struct SeedWeights {
    const uint16_t* wt[kSeedLayers];
    const uint16_t* frag;  // seed_frag_elems() BF16
    const float* w5f;      // seed_w5f_elems() FP32
    const float* b[kSeedLayers];
    const float* w5;  // 128 x 6
    const float* b5;  // 6
};

size_t seed_frag_elems();
size_t seed_w5f_elems();
void launch_pack_infer(const float* const w[kSeedLayers], const float* w5, uint16_t* frag, float* w5f,
                       hipStream_t s);

struct SeedInferParams {
    const float* ref;  // f32 row-major w x h
    const float* tar;
    int w, h;
    fi_poi* pois;      // seeds written into deformation[0..2, 6..8]
    const uint32_t* index;  // POIs to seed (nullptr = all)
    int n;
    SeedWeights wts;
    uint32_t* ok;      // per POI 1 = seeded, 0 = crop outside the image (untouched)
    int cus;           // compute units of the device (persistent grid size)
};

void launch_seed_infer(const SeedInferParams& p, hipStream_t s);

}  // namespace fi
