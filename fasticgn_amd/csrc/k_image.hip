// Whole-image kernels: raw -> f32 conversion, Gradient2D4 (4th-order central
// difference) and the BicubicBspline coefficient table.  Each output element
// performs the same float operations, in the same order, as the reference's
// scalar loop (this file is compiled with -ffp-contract=off), so the device
// results equal the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include "fi_device.h"
#include "fi_launch.h"

namespace fi {

// One thread per pixel over 64 x 4 pixel tiles (no per-pixel division).
template <typename T>
__global__ __launch_bounds__(256) void k_to_float(const T* __restrict__ raw, size_t pitch, int w,
                                                  int h, float scale, float* __restrict__ out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    out[(size_t)y * w + x] = load_px<T>(raw, (size_t)y * pitch + x, scale);
}

// Four pixels per thread (w and the pitch multiples of 4): one 4, 8 or 16-byte
// load and one 16-byte store, the same float(p) * scale per pixel.
template <typename T> struct Vec4;
template <> struct Vec4<uint8_t> { typedef uchar4 type; };
template <> struct Vec4<uint16_t> { typedef ushort4 type; };
template <> struct Vec4<float> { typedef float4 type; };

template <typename T>
__global__ __launch_bounds__(256) void k_to_float4(const T* __restrict__ raw, size_t pitch, int w4, int h,
                                                   float scale, float* __restrict__ out)
{
    const int x4 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x4 >= w4 || y >= h) return;
    const auto v = reinterpret_cast<const typename Vec4<T>::type*>(raw + (size_t)y * pitch)[x4];
    reinterpret_cast<float4*>(out + (size_t)y * 4 * w4)[x4] =
        make_float4((float)v.x * scale, (float)v.y * scale, (float)v.z * scale, (float)v.w * scale);
}

// Gradient2D4::getGradientX/Y, src/oc_gradient.cpp:37-79: taps applied as
// acc -= f(c+2)/12; acc += f(c+1)*2/3; acc -= f(c-1)*2/3; acc += f(c-2)/12,
// zero within 2 px of the border in the derivative's direction.
__global__ __launch_bounds__(256) void k_gradient(const float* __restrict__ img, int w, int h,
                                                  float* __restrict__ gx, float* __restrict__ gy)
{
    const float first = 1.f / 12.f;
    const float second = 2.f / 3.f;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    float ax = 0.f, ay = 0.f;
    if (x >= 2 && x < w - 2) {
        ax -= img[i + 2] * first;
        ax += img[i + 1] * second;
        ax -= img[i - 1] * second;
        ax += img[i - 2] * first;
    }
    if (y >= 2 && y < h - 2) {
        ay -= img[i + 2 * (size_t)w] * first;
        ay += img[i + (size_t)w] * second;
        ay -= img[i - (size_t)w] * second;
        ay += img[i - 2 * (size_t)w] * first;
    }
    gx[i] = ax;
    gy[i] = ay;
}

// BicubicBspline::prepare, src/oc_cubic_bspline.cpp:84-132.  One thread per
// texel; texels outside 1 <= r < h-2, 1 <= c < w-2 are zero.  Per texel the 16
// values coef[k][l] = P[3-k][3-l], P[k][l] = sum_m sum_n (BC[l][m] * BC[k][n]) *
// Q[n][m] in (m, n) order.  Row r of the table is four planes of w float4s,
// plane k holding coef[k][0..3] (bspline_table_floats), so one sample reads
// four float4s w*16 bytes apart and a wavefront reading one plane for
// consecutive texels reads one contiguous run.  Threads of the first row also
// clear the three padding planes after the last row.
template <typename T>
__global__ __launch_bounds__(256) void k_bspline(const T* __restrict__ raw, size_t pitch, int w,
                                                 int h, float scale, float* __restrict__ coef)
{
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (c >= w || r >= h) return;
    float4* tab = reinterpret_cast<float4*>(coef);
    const size_t plane = (size_t)w;
    float4* dst = tab + (size_t)4 * r * w + c;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if (r == 0) {
        float4* pad = tab + (size_t)4 * h * w + c;
        pad[0] = z; pad[plane] = z; pad[2 * plane] = z;
    }
    if (r < 1 || r >= h - 2 || c < 1 || c >= w - 2) {
        dst[0] = z; dst[plane] = z; dst[2 * plane] = z; dst[3 * plane] = z;
        return;
    }
    float q[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            q[i][j] = load_px<T>(raw, (size_t)(r - 1 + i) * pitch + (c - 1 + j), scale);
    float p[4][4];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int l = 0; l < 4; l++) {
            float acc = 0.f;
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int n = 0; n < 4; n++) acc += bc(l, m) * bc(k, n) * q[n][m];
            p[k][l] = acc;
        }
    dst[0] = make_float4(p[3][3], p[3][2], p[3][1], p[3][0]);
    dst[plane] = make_float4(p[2][3], p[2][2], p[2][1], p[2][0]);
    dst[2 * plane] = make_float4(p[1][3], p[1][2], p[1][1], p[1][0]);
    dst[3 * plane] = make_float4(p[0][3], p[0][2], p[0][1], p[0][0]);
}

static dim3 tile_grid(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }

void launch_to_float(const DevImage& img, float* out, hipStream_t s)
{
    const int esize = img.pixel == kU8 ? 1 : img.pixel == kU16 ? 2 : 4;
    if (img.width % 4 == 0 && img.pitch % 4 == 0 && reinterpret_cast<uintptr_t>(img.data) % (4 * esize) == 0) {
        const dim3 b4 = tile_grid(img.width / 4, img.height);
        switch (img.pixel) {
        case kU8:
            hipLaunchKernelGGL(k_to_float4<uint8_t>, b4, dim3(256), 0, s, (const uint8_t*)img.data, img.pitch,
                               img.width / 4, img.height, img.scale, out);
            return;
        case kU16:
            hipLaunchKernelGGL(k_to_float4<uint16_t>, b4, dim3(256), 0, s, (const uint16_t*)img.data, img.pitch,
                               img.width / 4, img.height, img.scale, out);
            return;
        default:
            hipLaunchKernelGGL(k_to_float4<float>, b4, dim3(256), 0, s, (const float*)img.data, img.pitch,
                               img.width / 4, img.height, img.scale, out);
            return;
        }
    }
    const dim3 blocks = tile_grid(img.width, img.height);
    switch (img.pixel) {
    case kU8:
        hipLaunchKernelGGL(k_to_float<uint8_t>, dim3(blocks), dim3(256), 0, s,
                           (const uint8_t*)img.data, img.pitch, img.width, img.height, img.scale, out);
        break;
    case kU16:
        hipLaunchKernelGGL(k_to_float<uint16_t>, dim3(blocks), dim3(256), 0, s,
                           (const uint16_t*)img.data, img.pitch, img.width, img.height, img.scale, out);
        break;
    default:
        hipLaunchKernelGGL(k_to_float<float>, dim3(blocks), dim3(256), 0, s,
                           (const float*)img.data, img.pitch, img.width, img.height, img.scale, out);
        break;
    }
}

void launch_gradient(const float* img, int w, int h, float* gx, float* gy, hipStream_t s)
{
    hipLaunchKernelGGL(k_gradient, tile_grid(w, h), dim3(256), 0, s, img, w, h, gx, gy);
}

void launch_bspline(const DevImage& img, float* coef, hipStream_t s)
{
    const dim3 g = tile_grid(img.width, img.height);
    switch (img.pixel) {
    case kU8:
        hipLaunchKernelGGL(k_bspline<uint8_t>, g, dim3(256), 0, s, (const uint8_t*)img.data,
                           img.pitch, img.width, img.height, img.scale, coef);
        break;
    case kU16:
        hipLaunchKernelGGL(k_bspline<uint16_t>, g, dim3(256), 0, s, (const uint16_t*)img.data,
                           img.pitch, img.width, img.height, img.scale, coef);
        break;
    default:
        hipLaunchKernelGGL(k_bspline<float>, g, dim3(256), 0, s, (const float*)img.data,
                           img.pitch, img.width, img.height, img.scale, coef);
        break;
    }
}

}  // namespace fi
