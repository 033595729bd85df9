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
constexpr int kInferPois = 16;
constexpr int kInferThreads = 256;

__global__ __launch_bounds__(kInferThreads) void k_seed_infer(SeedInferParams p)
{
    __shared__ __align__(16) uint16_t x2[kInferPois * 16 * 64];  // layer 2 input
    __shared__ __align__(16) uint16_t x3[kInferPois * 4 * 128];  // layer 3 input
    __shared__ __align__(16) uint16_t x4[kInferPois * 256];      // layer 4 input
    __shared__ float h4[kInferPois * 128];                       // layer 4 output
    __shared__ float outv[kInferPois][kSeedOut];
    __shared__ int crop[kInferPois][4];             // x0, y0 (ref), x0, y0 (target)
    __shared__ float guess[kInferPois][2];
    __shared__ int valid[kInferPois];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = blockIdx.x * kInferPois;
    const int w = p.w, h = p.h;

    if (tid < kInferPois) {
        const int b = base + tid;
        int ok = 0;
        if (b < p.n) {
            const fi_poi& q = p.pois[p.index ? p.index[b] : b];
            const float gu = rintf(q.deformation[0]), gv = rintf(q.deformation[6]);
            const float cx = floorf(q.x), cy = floorf(q.y);
            const float rx0 = cx - kSeedCropOrigin, ry0 = cy - kSeedCropOrigin;
            const float tx0 = rx0 + gu, ty0 = ry0 + gv;
            const float hi_x = (float)(w - kSeedCrop), hi_y = (float)(h - kSeedCrop);
            // finite and inside both images (NaN compares false)
            ok = (rx0 >= 0.f && rx0 <= hi_x && ry0 >= 0.f && ry0 <= hi_y && tx0 >= 0.f && tx0 <= hi_x &&
                  ty0 >= 0.f && ty0 <= hi_y) ? 1 : 0;
            crop[tid][0] = ok ? (int)rx0 : 0;
            crop[tid][1] = ok ? (int)ry0 : 0;
            crop[tid][2] = ok ? (int)tx0 : 0;
            crop[tid][3] = ok ? (int)ty0 : 0;
            guess[tid][0] = gu;
            guess[tid][1] = gv;
            if (p.ok) p.ok[b] = (uint32_t)ok;
        }
        valid[tid] = ok;
    }
    __syncthreads();

    // ---- layer 1: wave w owns POIs 4w .. 4w+3 -------------------------------
    for (int sp = 0; sp < 4; sp++) {
        const int s = wave * 4 + sp;
        float mean[2] = { 0.f, 0.f }, inv[2] = { 0.f, 0.f };
        if (valid[s]) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float* img = c ? p.tar : p.ref;
                const int x0 = crop[s][2 * c], y0 = crop[s][2 * c + 1];
                // 1024 pixels: lane reads rows (lane / 8) + 8 t, columns 4 (lane % 8) .. +3
                float sum = 0.f, sq = 0.f;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int r = (lane >> 3) + 8 * t;
                    const float* q = img + (size_t)(y0 + r) * w + x0 + 4 * (lane & 7);
                    const float v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3];
                    sum += (v0 + v1) + (v2 + v3);
                    sq += (v0 * v0 + v1 * v1) + (v2 * v2 + v3 * v3);
                }
                sum = wave_sum(sum);
                sq = wave_sum(sq);
                const float m = sum * (1.f / 1024.f);
                const float var = fmaxf(sq * (1.f / 1024.f) - m * m, 0.f);
                mean[c] = m;
                inv[c] = 1.f / sqrtf(var + 1e-6f);
            }
        }
        // 4 row tiles of 16 patches, 2 depth steps (ref, target)
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int patch = rt * 16 + (lane & 15);
                const int py = patch >> 3, px = patch & 7, dy = lane >> 4;
                bf16x4 a = { 0, 0, 0, 0 };
                if (valid[s]) {
                    const float* img = ks ? p.tar : p.ref;
                    const int x0 = crop[s][2 * ks], y0 = crop[s][2 * ks + 1];
                    const float* q = img + (size_t)(y0 + 4 * py + dy) * w + x0 + 4 * px;
#pragma unroll
                    for (int t = 0; t < 4; t++) a[t] = (short)to_bf16((q[t] - mean[ks]) * inv[ks]);
                }
                const bf16x4 bw = load4(p.wts.wt[0] + (lane & 15) * 32 + ks * 16 + 4 * (lane >> 4));
                acc = mfma16(a, bw, acc);
            }
            const int n = lane & 15;
            const float bias = p.wts.b[0][n];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int patch = rt * 16 + 4 * (lane >> 4) + j;
                int row, col;
                patchify(s, patch, 8, 16, n, row, col);
                x2[row * 64 + col] = to_bf16(fmaxf(acc[j] + bias, 0.f));
            }
        }
    }
    __syncthreads();

    // ---- layer 2: M = 256 (16 rows per POI), K = 64, N = 32 ------------------
    for (int t = 0; t < 8; t++) {
        const int rt = wave * 4 + (t >> 1), ct = t & 1;  // row tile = POI
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const bf16x4 a = load4(x2 + (rt * 16 + (lane & 15)) * 64 + ks * 16 + 4 * (lane >> 4));
            const bf16x4 bw = load4(p.wts.wt[1] + (ct * 16 + (lane & 15)) * 64 + ks * 16 + 4 * (lane >> 4));
            acc = mfma16(a, bw, acc);
        }
        const int n = ct * 16 + (lane & 15);
        const float bias = p.wts.b[1][n];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int patch = 4 * (lane >> 4) + j;
            int row, col;
            patchify(rt, patch, 4, 32, n, row, col);
            x3[row * 128 + col] = to_bf16(fmaxf(acc[j] + bias, 0.f));
        }
    }
    __syncthreads();

    // ---- layer 3: M = 64 (4 rows per POI), K = 128, N = 64 -------------------
    for (int ct = 0; ct < 4; ct++) {
        const int rt = wave;  // rows 16 w .. 16 w + 15 = POIs 4w .. 4w+3
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
            const bf16x4 a = load4(x3 + (rt * 16 + (lane & 15)) * 128 + ks * 16 + 4 * (lane >> 4));
            const bf16x4 bw = load4(p.wts.wt[2] + (ct * 16 + (lane & 15)) * 128 + ks * 16 + 4 * (lane >> 4));
            acc = mfma16(a, bw, acc);
        }
        const int n = ct * 16 + (lane & 15);
        const float bias = p.wts.b[2][n];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = rt * 16 + 4 * (lane >> 4) + j;  // POI r / 4, patch r % 4
            int row, col;
            patchify(r >> 2, r & 3, 2, 64, n, row, col);
            x4[row * 256 + col] = to_bf16(fmaxf(acc[j] + bias, 0.f));
        }
    }
    __syncthreads();

    // ---- layer 4: M = 16 POIs, K = 256, N = 128 ------------------------------
    for (int t = 0; t < 2; t++) {
        const int ct = wave * 2 + t;
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 16; ks++) {
            const bf16x4 a = load4(x4 + (lane & 15) * 256 + ks * 16 + 4 * (lane >> 4));
            const bf16x4 bw = load4(p.wts.wt[3] + (ct * 16 + (lane & 15)) * 256 + ks * 16 + 4 * (lane >> 4));
            acc = mfma16(a, bw, acc);
        }
        const int n = ct * 16 + (lane & 15);
        const float bias = p.wts.b[3][n];
#pragma unroll
        for (int j = 0; j < 4; j++) h4[(4 * (lane >> 4) + j) * 128 + n] = fmaxf(acc[j] + bias, 0.f);
    }
    __syncthreads();

    // ---- head (FP32) and seed ------------------------------------------------
    if (tid < kInferPois * kSeedOut) {
        const int s = tid / kSeedOut, o = tid - s * kSeedOut;
        float acc = p.wts.b5[o];
        for (int c = 0; c < 128; c++) acc = __builtin_fmaf(h4[s * 128 + c], p.wts.w5[c * kSeedOut + o], acc);
        outv[s][o] = acc;
    }
    __syncthreads();
    if (tid < kInferPois && valid[tid]) {
        const int b = base + tid;
        fi_poi& q = p.pois[p.index ? p.index[b] : b];
        const float* o = outv[tid];
        q.deformation[0] = guess[tid][0] + o[0];
        q.deformation[6] = guess[tid][1] + o[1];
        q.deformation[1] = o[2] * (1.f / kGradScale);
        q.deformation[2] = o[3] * (1.f / kGradScale);
        q.deformation[7] = o[4] * (1.f / kGradScale);
        q.deformation[8] = o[5] * (1.f / kGradScale);
    }
}

void launch_seed_infer(const SeedInferParams& p, hipStream_t s)
{
    if (p.n <= 0) return;
    hipLaunchKernelGGL(k_seed_infer, dim3((p.n + kInferPois - 1) / kInferPois), dim3(kInferThreads), 0, s, p);
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
