// Learned-seeder kernels for gfx950 (network in fi_seeder.h):
//
//   k_seed_infer     fused inference, 16 POIs per 256-thread block: crops are
//                    normalised in registers and fed straight into layer 1's
//                    MFMA fragments; layers 2-4 run from LDS; the FP32 head
//                    writes the seed into the POI record.
//   k_gemm_bf16      C = A B for training (any strides, BF16 in, FP32 out,
//                    optional split over K into partial sums).
//   small kernels    bias + ReLU + patchify, ReLU' + un-patchify, column sums,
//                    head forward/backward with the MSE loss, Adam.
//
// MFMA fragment layout of v_mfma_f32_16x16x16_bf16 (lane l of 64):
//   A (16 x 16, rows i, depth k): lane holds A[l % 16][4 (l / 16) + j], j < 4
//   B (16 x 16, depth k, cols n): lane holds B[4 (l / 16) + j][l % 16]
//   D (16 x 16):                  lane holds D[4 (l / 16) + j][l % 16]
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fi_device.h"
#include "fi_seeder.h"

namespace fi {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 load4(const uint16_t* p)
{
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFF);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFF);
    r[3] = (short)(v.y >> 16);
    return r;
}

// Row of a layer's activations -> (row, column) of the next layer's input:
// grid g x g of patches, 2 x 2 patches merge, channel blocks of width n in
// the order (dy, dx).
__device__ __forceinline__ void patchify(int s, int p, int g, int n, int c, int& row, int& col)
{
    const int py = p / g, px = p - py * g;
    const int g2 = g >> 1;
    row = s * g2 * g2 + (py >> 1) * g2 + (px >> 1);
    col = ((py & 1) * 2 + (px & 1)) * n + c;
}

// ---------------------------------------------------------------------------
// Fused inference
// ---------------------------------------------------------------------------
// One wave seeds 16 POIs per pass with every GEMM transposed: D (channels x
// POIs) = W^T (channels x depth) X^T (depth x POIs), POI s in MFMA column s
// (lane l works on POI l % 16).  The crop quadtree is walked depth first
// (layer-3 position q3, layer-2 position q2, layer-1 patch q1), so a layer's
// D fragments, after bias, ReLU and BF16 rounding, are the next layer's B
// fragment as they stand: in v_mfma_f32_16x16x32_bf16 lane l holds D rows
// 4g..4g+3 (g = l / 16) of a 16-row tile and B depth positions 8g..8g+7 of a
// 32-deep step, so two D tiles make one B fragment.  The packed weights
// (k_pack_infer) carry the matching depth order; nothing moves between lanes
// or through LDS from one layer to the next.  W1, W2 sit in VGPRs, W3 and W4
// (80 KiB) in LDS, staged once per persistent block.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// four consecutive pixels at any 4-byte boundary (global_load_dwordx4 takes
// dword-aligned addresses)
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

constexpr int kInferWaves = 4;
constexpr int kInferThreads = 64 * kInferWaves;
constexpr int kInferGroup = 16;  // POIs per wave pass
constexpr int kInferBlocksPerCu = 2;
constexpr int kXcds = 8;
// packed fragments: 64 lanes x 8 BF16 (1 KiB) each
constexpr int kFragOff2 = 1, kFragOff3 = kFragOff2 + 4, kFragOff4 = kFragOff3 + 16, kFragTotal = kFragOff4 + 64;
constexpr int kLdsFrags = kFragTotal - kFragOff3;  // W3 (16) + W4 (64)
constexpr int kW5Frags = 32;                       // head: 8 channel tiles x 4 rows, 64 FP32 each

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ __bf16 relu_bf16(float z) { return (__bf16)fmaxf(z, 0.f); }

__global__ __launch_bounds__(kInferThreads, 2) void k_seed_infer(SeedInferParams p)
{
    __shared__ bf16x8 wl[kLdsFrags * 64];  // W3 fragments, then W4
    const bf16x8* gf = reinterpret_cast<const bf16x8*>(p.wts.frag);
    for (int i = threadIdx.x; i < kLdsFrags * 64; i += kInferThreads) wl[i] = gf[kFragOff3 * 64 + i];
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, s = lane & 15;
    const int crop_c = g >> 1;       // this lane's pixels: 0 reference, 1 target
    const int rsub = 2 * (g & 1);    // and patch rows rsub, rsub + 1
    const int w = p.w;
    const bf16x8 w1 = gf[lane];
    bf16x8 w2[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w2[i] = gf[(kFragOff2 + i) * 64 + lane];
    float b1[4], b2[2][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        b1[i] = p.wts.b[0][4 * g + i];
        b2[0][i] = p.wts.b[1][4 * g + i];
        b2[1][i] = p.wts.b[1][16 + 4 * g + i];
    }
    const int groups = (p.n + kInferGroup - 1) / kInferGroup;
    if (p.w < kSeedCrop || p.h < kSeedCrop) {
        // no crop fits: nothing is seeded
        for (int b = (blockIdx.x * kInferWaves + wave) * 64 + lane; p.ok && b < p.n; b += gridDim.x * kInferThreads)
            p.ok[p.index ? p.index[b] : (uint32_t)b] = 0;
        return;
    }

    // XCD-aware: workgroups are dealt to the 8 XCDs round-robin, so with a grid
    // that is a multiple of 8 XCD x = blockIdx % 8 takes the contiguous
    // eighth x of the groups, and neighbouring POIs (overlapping crops) share
    // that XCD's L2
    int g0 = 0, g1 = groups, first = blockIdx.x * kInferWaves + wave, stride = gridDim.x * kInferWaves;
    if (gridDim.x % kXcds == 0) {
        const int xcd = blockIdx.x % kXcds, per = gridDim.x / kXcds;
        g0 = (int)((long long)groups * xcd / kXcds);
        g1 = (int)((long long)groups * (xcd + 1) / kXcds);
        first = g0 + (int)(blockIdx.x / kXcds) * kInferWaves + wave;
        stride = per * kInferWaves;
    }
    for (int grp = first; grp < g1; grp += stride) {
        const int b = grp * kInferGroup + s;
        const int qi = b < p.n ? (p.index ? (int)p.index[b] : b) : 0;
        int ok = 0, x0 = 0, y0 = 0;
        float gu = 0.f, gv = 0.f;
        if (b < p.n) {
            const fi_poi& q = p.pois[qi];
            gu = rintf(q.deformation[0]);
            gv = rintf(q.deformation[6]);
            const float cx = floorf(q.x), cy = floorf(q.y);
            const float rx0 = cx - kSeedCropOrigin, ry0 = cy - kSeedCropOrigin;
            const float tx0 = rx0 + gu, ty0 = ry0 + gv;
            const float hx = (float)(p.w - kSeedCrop), hy = (float)(p.h - kSeedCrop);
            // finite and inside both images (NaN compares false)
            ok = (rx0 >= 0.f && rx0 <= hx && ry0 >= 0.f && ry0 <= hy && tx0 >= 0.f && tx0 <= hx && ty0 >= 0.f &&
                  ty0 <= hy) ? 1 : 0;
            if (ok) {
                x0 = (int)(crop_c ? tx0 : rx0);
                y0 = (int)(crop_c ? ty0 : ry0);
            }
        }
        // crop origin; (0, 0) for a lane without a valid crop, so that every
        // load below stays inside the image (width, height >= 32)
        const float* base = (crop_c ? p.tar : p.ref) + (size_t)y0 * w + x0;

        // crop statistics: this lane sums rows r with (r >> 1) & 1 == g & 1,
        // the partner lane l ^ 16 the others
        float sum = 0.f, sq = 0.f;
        for (int rr = 0; rr < 16; rr++) {
            const float* row = base + (size_t)(4 * (rr >> 1) + rsub + (rr & 1)) * w;
#pragma unroll
            for (int x = 0; x < kSeedCrop; x++) {
                const float v = row[x];
                sum += v;
                sq = __builtin_fmaf(v, v, sq);
            }
        }
        sum += __shfl_xor(sum, 16);
        sq += __shfl_xor(sq, 16);
        const float mean = sum * (1.f / 1024.f);
        const float var = fmaxf(sq * (1.f / 1024.f) - mean * mean, 0.f);
        // a lane without a valid crop feeds zeros (its results are not stored)
        const float inv = ok ? 1.f / sqrtf(var + 1e-6f) : 0.f;

        f32x4 acc4[8];
#pragma unroll
        for (int i = 0; i < 8; i++) acc4[i] = f32x4{ 0.f, 0.f, 0.f, 0.f };
        for (int q3 = 0; q3 < 4; q3++) {
            f32x4 acc3[4];
#pragma unroll
            for (int i = 0; i < 4; i++) acc3[i] = f32x4{ 0.f, 0.f, 0.f, 0.f };
            // not unrolled: unrolling lets the scheduler hoist the pixel loads of
            // all 16 patches of q3 and spill
#pragma unroll 1
            for (int q2 = 0; q2 < 4; q2++) {
                f32x4 acc2[2] = { { 0.f, 0.f, 0.f, 0.f }, { 0.f, 0.f, 0.f, 0.f } };
#pragma unroll
                for (int t2 = 0; t2 < 2; t2++) {
                    bf16x8 x2;  // layer-2 B fragment: patches q1 = 2 t2, 2 t2 + 1
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int q1 = 2 * t2 + h;
                        const int py = 4 * (q3 >> 1) + 2 * (q2 >> 1) + (q1 >> 1);
                        const int px = 4 * (q3 & 1) + 2 * (q2 & 1) + (q1 & 1);
                        const float* pp = base + (size_t)(4 * py + rsub) * w + 4 * px;
                        // two rows of four pixels, in bounds for every lane
                        const f32x4u r0 = *reinterpret_cast<const f32x4u*>(pp);
                        const f32x4u r1 = *reinterpret_cast<const f32x4u*>(pp + w);
                        bf16x8 x1;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            x1[j] = (__bf16)((r0[j] - mean) * inv);
                            x1[4 + j] = (__bf16)((r1[j] - mean) * inv);
                        }
                        const f32x4 d = mfma32(w1, x1, f32x4{ 0.f, 0.f, 0.f, 0.f });
#pragma unroll
                        for (int i = 0; i < 4; i++) x2[4 * h + i] = relu_bf16(d[i] + b1[i]);
                    }
                    acc2[0] = mfma32(w2[t2], x2, acc2[0]);
                    acc2[1] = mfma32(w2[2 + t2], x2, acc2[1]);
                }
                bf16x8 x3;  // layer-3 B fragment of position q2: channels 0..15 | 16..31
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    x3[i] = relu_bf16(acc2[0][i] + b2[0][i]);
                    x3[4 + i] = relu_bf16(acc2[1][i] + b2[1][i]);
                }
#pragma unroll
                for (int mt = 0; mt < 4; mt++) acc3[mt] = mfma32(wl[(mt * 4 + q2) * 64 + lane], x3, acc3[mt]);
            }
            bf16x8 x4a, x4b;  // layer-4 B fragments of position q3: channels 0..31, 32..63
#pragma unroll
            for (int i = 0; i < 4; i++) {
                x4a[i] = relu_bf16(acc3[0][i] + p.wts.b[2][4 * g + i]);
                x4a[4 + i] = relu_bf16(acc3[1][i] + p.wts.b[2][16 + 4 * g + i]);
                x4b[i] = relu_bf16(acc3[2][i] + p.wts.b[2][32 + 4 * g + i]);
                x4b[4 + i] = relu_bf16(acc3[3][i] + p.wts.b[2][48 + 4 * g + i]);
            }
#pragma unroll
            for (int mt = 0; mt < 8; mt++) {
                const int f = 16 + mt * 8 + 2 * q3;
                acc4[mt] = mfma32(wl[f * 64 + lane], x4a, acc4[mt]);
                acc4[mt] = mfma32(wl[(f + 1) * 64 + lane], x4b, acc4[mt]);
            }
        }

        // head, FP32 MFMA 16x16x4: out[o][s] = sum_c W5[c][o] relu(acc4 + b4)[c][s];
        // lane (g, s) supplies depth g of each step = channel 16 mt + 4 g + i
        f32x4 o4 = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int mt = 0; mt < 8; mt++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float hv = fmaxf(acc4[mt][i] + p.wts.b[3][16 * mt + 4 * g + i], 0.f);
                o4 = __builtin_amdgcn_mfma_f32_16x16x4f32(p.wts.w5f[(mt * 4 + i) * 64 + lane], hv, o4, 0, 0, 0);
            }
        }
        // lane (0, s) holds outputs 0..3 of POI s, lane (1, s) outputs 4..7
        if (b < p.n) {
            if (g == 0 && p.ok) p.ok[qi] = (uint32_t)ok;
            if (ok && g < 2) {
                fi_poi& q = p.pois[qi];
                const float* b5 = p.wts.b5;
                if (g == 0) {
                    q.deformation[0] = gu + (o4[0] + b5[0]);
                    q.deformation[6] = gv + (o4[1] + b5[1]);
                    q.deformation[1] = (o4[2] + b5[2]) * (1.f / kGradScale);
                    q.deformation[2] = (o4[3] + b5[3]) * (1.f / kGradScale);
                } else {
                    q.deformation[7] = (o4[0] + b5[4]) * (1.f / kGradScale);
                    q.deformation[8] = (o4[1] + b5[5]) * (1.f / kGradScale);
                }
            }
        }
    }
}

// Packed inference weights from the FP32 parameters.  Fragment f of layer L
// (m-tile mt, depth step t, f = mt * steps + t), lane l, element j holds
// BF16(W_L[k][16 mt + l % 16]) with g = l / 16 and
//   layer 1: k = 8 g + j                                (pixel c, dy, dx)
//   layers 2-4: k = 32 t + 16 (j / 4) + 4 g + j % 4     (previous layer's D tiles)
// Head fragment mt * 4 + i, lane l: W5[16 mt + 4 g + i][l % 16] for l % 16 < 6,
// else 0.
struct PackSrc {
    const float* w[kSeedLayers];
    const float* w5;
};

__global__ void k_pack_infer(PackSrc src, uint16_t* __restrict__ frag, float* __restrict__ w5f)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < kFragTotal * 512) {
        const int f = e >> 9, r = e & 511, l = r >> 3, j = r & 7, g = l >> 4, col = l & 15;
        int layer, fl;
        if (f < kFragOff2) { layer = 0; fl = 0; }
        else if (f < kFragOff3) { layer = 1; fl = f - kFragOff2; }
        else if (f < kFragOff4) { layer = 2; fl = f - kFragOff3; }
        else { layer = 3; fl = f - kFragOff4; }
        const int steps = kSeedK[layer] / 32, mt = fl / steps, t = fl - mt * steps;
        const int k = layer == 0 ? 8 * g + j : 32 * t + 16 * (j >> 2) + 4 * g + (j & 3);
        frag[e] = to_bf16(src.w[layer][(size_t)k * kSeedN[layer] + 16 * mt + col]);
    } else if (e < kFragTotal * 512 + kW5Frags * 64) {
        const int e2 = e - kFragTotal * 512, mi = e2 >> 6, l = e2 & 63, mt = mi >> 2, i = mi & 3;
        const int o = l & 15, ch = 16 * mt + 4 * (l >> 4) + i;
        w5f[e2] = o < kSeedOut ? src.w5[ch * kSeedOut + o] : 0.f;
    }
}

size_t seed_frag_elems() { return (size_t)kFragTotal * 512; }
size_t seed_w5f_elems() { return (size_t)kW5Frags * 64; }

void launch_pack_infer(const float* const w[kSeedLayers], const float* w5, uint16_t* frag, float* w5f,
                       hipStream_t s)
{
    PackSrc src;
    for (int l = 0; l < kSeedLayers; l++) src.w[l] = w[l];
    src.w5 = w5;
    const int total = kFragTotal * 512 + kW5Frags * 64;
    hipLaunchKernelGGL(k_pack_infer, dim3((total + 255) / 256), dim3(256), 0, s, src, frag, w5f);
}

void launch_seed_infer(const SeedInferParams& p, hipStream_t s)
{
    if (p.n <= 0) return;
    const int cus = std::max(1, p.cus);
    const int groups = (p.n + kInferGroup - 1) / kInferGroup;
    int blocks = std::max(1, std::min((groups + kInferWaves - 1) / kInferWaves, kInferBlocksPerCu * cus));
    if (blocks >= 8 * kXcds) blocks -= blocks % kXcds;  // XCD-aware split (see the kernel)
    hipLaunchKernelGGL(k_seed_infer, dim3(blocks), dim3(kInferThreads), 0, s, p);
}

// ---------------------------------------------------------------------------
// Training kernels
// ---------------------------------------------------------------------------

// C (M x N, FP32, ldc = N) = A B, A(i,k) = A[i sai + k sak], B(k,j) = B[k sbk +
// j sbj] (BF16).  One wave per 16 x 16 tile; blockIdx.z splits K into
// gridDim.z chunks written to C + z M N (summed by k_sum_parts).  M, N and each
// K chunk are multiples of 16 (checked by the host).
__global__ __launch_bounds__(64) void k_gemm_bf16(const uint16_t* __restrict__ A, long sai, long sak,
                                                  const uint16_t* __restrict__ B, long sbk, long sbj,
                                                  float* __restrict__ C, int M, int N, int kchunk)
{
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    const long k_begin = (long)blockIdx.z * kchunk;
    f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
    const int i = i0 + (lane & 15), j = j0 + (lane & 15), kk = 4 * (lane >> 4);
    for (int k0 = 0; k0 < kchunk; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const long k = k_begin + k0 + kk + t;
            a[t] = (short)A[i * sai + k * sak];
            b[t] = (short)B[k * sbk + j * sbj];
        }
        acc = mfma16(a, b, acc);
    }
    float* out = C + (size_t)blockIdx.z * M * N;
#pragma unroll
    for (int t = 0; t < 4; t++) out[(size_t)(i0 + kk + t) * N + j] = acc[t];
}

__global__ void k_sum_parts(const float* __restrict__ parts, int parts_n, size_t mn, float* __restrict__ out)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (size_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < parts_n; z++) s += parts[(size_t)z * mn + e];
        out[e] = s;
    }
}

// Y = ReLU(Z + b), written as BF16 into the next layer's input (patchify on a
// g x g grid of rows per sample; g = 1 writes the plain row-major copy).
__global__ void k_bias_relu_patchify(const float* __restrict__ Z, const float* __restrict__ b, int rows, int n,
                                     int g, float* __restrict__ zb, uint16_t* __restrict__ x_next)
{
    const size_t total = (size_t)rows * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e - (size_t)r * n);
        const float z = Z[e] + b[c];
        zb[e] = z;
        const float y = fmaxf(z, 0.f);
        if (g > 1) {
            const int per = g * g, s = r / per, pp = r - s * per;
            int row, col;
            patchify(s, pp, g, n, c, row, col);
            x_next[(size_t)row * (4 * n) + col] = to_bf16(y);
        } else {
            x_next[e] = to_bf16(y);
        }
    }
}

// dZ = dY * (z > 0) with dY gathered back from the next layer's input
// gradient (un-patchify), as BF16 for the weight and input GEMMs, and FP32
// for the bias gradient.
__global__ void k_relu_bwd(const float* __restrict__ dx_next, const float* __restrict__ zb, int rows, int n, int g,
                           uint16_t* __restrict__ dz, float* __restrict__ dz32)
{
    const size_t total = (size_t)rows * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e - (size_t)r * n);
        float dy;
        if (g > 1) {
            const int per = g * g, s = r / per, pp = r - s * per;
            int row, col;
            patchify(s, pp, g, n, c, row, col);
            dy = dx_next[(size_t)row * (4 * n) + col];
        } else {
            dy = dx_next[e];
        }
        const float d = zb[e] > 0.f ? dy : 0.f;
        dz[e] = to_bf16(d);
        dz32[e] = d;
    }
}

// column sums of an FP32 rows x n matrix
__global__ void k_colsum(const float* __restrict__ x, int rows, int n, float* __restrict__ out)
{
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += x[(size_t)r * n + c];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// head: out = h W5 + b5; loss = mean over samples and outputs of (out - t)^2;
// dOut = 2 (out - t) / (B 6); dh = (dOut W5^T) * (h > 0) is left to
// k_head_bwd.  One thread per sample.
__global__ void k_head_fwd(const float* __restrict__ zb4, const float* __restrict__ w5, const float* __restrict__ b5,
                           const float* __restrict__ target, int bsz, float* __restrict__ dout,
                           float* __restrict__ loss_parts)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (s < bsz) {
        float o[kSeedOut];
#pragma unroll
        for (int k = 0; k < kSeedOut; k++) o[k] = b5[k];
        for (int c = 0; c < 128; c++) {
            const float hv = fmaxf(zb4[(size_t)s * 128 + c], 0.f);
#pragma unroll
            for (int k = 0; k < kSeedOut; k++) o[k] = __builtin_fmaf(hv, w5[c * kSeedOut + k], o[k]);
        }
#pragma unroll
        for (int k = 0; k < kSeedOut; k++) {
            const float d = o[k] - target[(size_t)s * kSeedOut + k];
            l += d * d;
            dout[(size_t)s * kSeedOut + k] = 2.f * d / (float)(bsz * kSeedOut);
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o2 = blockDim.x / 2; o2 > 0; o2 >>= 1) {
        if ((int)threadIdx.x < o2) red[threadIdx.x] += red[threadIdx.x + o2];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_parts[blockIdx.x] = red[0];
}

// dW5 (128 x 6) and db5 (blocks 0..127 and 128); dh4 = dOut W5^T (rows x 128)
__global__ void k_head_bwd_w(const float* __restrict__ zb4, const float* __restrict__ dout, int bsz,
                             float* __restrict__ gw5, float* __restrict__ gb5)
{
    const int c = blockIdx.x;  // 0..127 weights, 128 bias
    __shared__ float red[256][kSeedOut];
    float acc[kSeedOut] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
    for (int s = threadIdx.x; s < bsz; s += blockDim.x) {
        const float hv = c < 128 ? fmaxf(zb4[(size_t)s * 128 + c], 0.f) : 1.f;
#pragma unroll
        for (int k = 0; k < kSeedOut; k++) acc[k] = __builtin_fmaf(hv, dout[(size_t)s * kSeedOut + k], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < kSeedOut; k++) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < kSeedOut; k++) red[threadIdx.x][k] += red[threadIdx.x + o][k];
        __syncthreads();
    }
    if (threadIdx.x < kSeedOut) {
        if (c < 128) gw5[c * kSeedOut + threadIdx.x] = red[0][threadIdx.x];
        else gb5[threadIdx.x] = red[0][threadIdx.x];
    }
}

__global__ void k_head_bwd_x(const float* __restrict__ dout, const float* __restrict__ w5, int bsz,
                             float* __restrict__ dh4)
{
    const size_t total = (size_t)bsz * 128;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(e / 128), c = (int)(e - (size_t)s * 128);
        float d = 0.f;
#pragma unroll
        for (int k = 0; k < kSeedOut; k++) d = __builtin_fmaf(dout[(size_t)s * kSeedOut + k], w5[c * kSeedOut + k], d);
        dh4[e] = d;
    }
}

// Adam with bias correction; refreshes the BF16 N x K copies of W1..W4
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float c1, float c2, float eps)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        w[e] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    }
}

// W (K x N, FP32 row-major) -> Wt (N x K, BF16)
__global__ void k_transpose_bf16(const float* __restrict__ w, int K, int N, uint16_t* __restrict__ wt)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K * N; e += gridDim.x * blockDim.x) {
        const int k = e / N, n = e - k * N;
        wt[(size_t)n * K + k] = to_bf16(w[e]);
    }
}

// gather rows: dst[i] = src[idx[i]] for rows of `width` 16-bit elements
__global__ void k_gather_rows16(const uint16_t* __restrict__ src, const uint32_t* __restrict__ idx, int rows,
                                int width, uint16_t* __restrict__ dst)
{
    const size_t total = (size_t)rows * width;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / width), c = (int)(e - (size_t)r * width);
        dst[e] = src[(size_t)idx[r] * width + c];
    }
}

__global__ void k_gather_rows32(const float* __restrict__ src, const uint32_t* __restrict__ idx, int rows,
                                int width, float* __restrict__ dst)
{
    const size_t total = (size_t)rows * width;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / width), c = (int)(e - (size_t)r * width);
        dst[e] = src[(size_t)idx[r] * width + c];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static dim3 grid1(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)); }

void launch_gemm_bf16(const uint16_t* A, long sai, long sak, const uint16_t* B, long sbk, long sbj, float* C, int M,
                      int N, int K, int splits, float* parts, hipStream_t s)
{
    const int kchunk = K / splits;
    if (splits == 1) {
        hipLaunchKernelGGL(k_gemm_bf16, dim3(M / 16, N / 16, 1), dim3(64), 0, s, A, sai, sak, B, sbk, sbj, C, M, N,
                           kchunk);
        return;
    }
    hipLaunchKernelGGL(k_gemm_bf16, dim3(M / 16, N / 16, splits), dim3(64), 0, s, A, sai, sak, B, sbk, sbj, parts,
                       M, N, kchunk);
    hipLaunchKernelGGL(k_sum_parts, grid1((size_t)M * N), dim3(256), 0, s, parts, splits, (size_t)M * N, C);
}

void launch_bias_relu_patchify(const float* Z, const float* b, int rows, int n, int g, float* zb, uint16_t* xn,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_bias_relu_patchify, grid1((size_t)rows * n), dim3(256), 0, s, Z, b, rows, n, g, zb, xn);
}

void launch_relu_bwd(const float* dxn, const float* zb, int rows, int n, int g, uint16_t* dz, float* dz32,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_relu_bwd, grid1((size_t)rows * n), dim3(256), 0, s, dxn, zb, rows, n, g, dz, dz32);
}

void launch_colsum(const float* x, int rows, int n, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3(n), dim3(256), 0, s, x, rows, n, out);
}

void launch_head_fwd(const float* zb4, const float* w5, const float* b5, const float* t, int bsz, float* dout,
                     float* loss_parts, hipStream_t s)
{
    hipLaunchKernelGGL(k_head_fwd, dim3((bsz + 255) / 256), dim3(256), 0, s, zb4, w5, b5, t, bsz, dout, loss_parts);
}

void launch_head_bwd(const float* zb4, const float* dout, const float* w5, int bsz, float* gw5, float* gb5,
                     float* dh4, hipStream_t s)
{
    hipLaunchKernelGGL(k_head_bwd_w, dim3(129), dim3(256), 0, s, zb4, dout, bsz, gw5, gb5);
    hipLaunchKernelGGL(k_head_bwd_x, grid1((size_t)bsz * 128), dim3(256), 0, s, dout, w5, bsz, dh4);
}

void launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float c1,
                 float c2, float eps, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, grid1(n), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, c1, c2, eps);
}

void launch_transpose_bf16(const float* w, int K, int N, uint16_t* wt, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose_bf16, grid1((size_t)K * N), dim3(256), 0, s, w, K, N, wt);
}

void launch_gather_rows16(const uint16_t* src, const uint32_t* idx, int rows, int width, uint16_t* dst,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_rows16, grid1((size_t)rows * width), dim3(256), 0, s, src, idx, rows, width, dst);
}

void launch_gather_rows32(const float* src, const uint32_t* idx, int rows, int width, float* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_rows32, grid1((size_t)rows * width), dim3(256), 0, s, src, idx, rows, width, dst);
}

}  // namespace fi
