// Host runtime: last-error storage, image upload, DFT twiddles.
#include <cstring>
#include <cmath>
#include <string>

#include "fi_host.h"

namespace fi {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

const std::string& last_error() { return g_last_error; }

static size_t bytes_per_pixel(int pixel)
{
    return pixel == kU8 ? 1 : pixel == kU16 ? 2 : 4;
}

void check_image_args(const void* px, int width, int height, size_t stride_bytes, int pixel,
                      int depth12)
{
    if (px == nullptr) fail(FI_E_INVALID_ARGUMENT, "image pointer is null");
    if (width < 1 || height < 1)
        fail(FI_E_INVALID_ARGUMENT, "image size must be positive, got " + std::to_string(width) +
                                        "x" + std::to_string(height));
    if (pixel != kU8 && pixel != kU16 && pixel != kF32)
        fail(FI_E_INVALID_ARGUMENT, "unknown pixel format " + std::to_string(pixel));
    if (depth12 != FI_DEPTH_NATIVE && depth12 != FI_DEPTH12_RIGHT && depth12 != FI_DEPTH12_LEFT)
        fail(FI_E_INVALID_ARGUMENT, "unknown 12-bit alignment " + std::to_string(depth12));
    if (depth12 != FI_DEPTH_NATIVE && pixel != kU16)
        fail(FI_E_INVALID_ARGUMENT, "12-bit alignment applies to 16-bit pixels only");
    if (stride_bytes < (size_t)width * bytes_per_pixel(pixel))
        fail(FI_E_INVALID_ARGUMENT, "row stride " + std::to_string(stride_bytes) +
                                        " is shorter than a row");
}

void DeviceImage::stage(const void* px, int width, int height, size_t stride_bytes, int pixel, int depth12)
{
    check_image_args(px, width, height, stride_bytes, pixel, depth12);
    const size_t row_bytes = (size_t)width * bytes_per_pixel(pixel);
    const size_t total = row_bytes * height;
    // rows packed into pinned memory (the copy out of pageable memory is a
    // plain memcpy; the DMA then runs at the link rate)
    this->stage_buf.reserve(total);
    // a right-aligned 12-bit image is range checked on the staged rows (each
    // still in cache after its copy): one pass over the caller's pixels
    const bool check12 = depth12 == FI_DEPTH12_RIGHT && pixel == FI_PIXEL_U16;
    uint16_t mx = 0;
    for (int y = 0; y < height; y++) {
        uint8_t* dst = stage_buf.p + (size_t)y * row_bytes;
        std::memcpy(dst, static_cast<const uint8_t*>(px) + (size_t)y * stride_bytes, row_bytes);
        if (check12) {
            const uint16_t* row = reinterpret_cast<const uint16_t*>(dst);
            for (int x = 0; x < width; x++) mx = row[x] > mx ? row[x] : mx;
        }
    }
    // src/oc_image.cpp:172-178: a right-aligned 12-bit image above 4095 is rejected
    if (mx > 4095) fail(FI_E_INVALID_ARGUMENT, "12-bit right-aligned input has values above 4095");
    raw.reserve(total);
    staged_bytes = total;
    view.data = raw.p;
    view.width = width;
    view.height = height;
    view.pixel = pixel;
    view.pitch = (size_t)width;
    view.scale = (depth12 == FI_DEPTH12_LEFT) ? (1.f / 16.f) : 1.f;
}

void DeviceImage::dma(hipStream_t s)
{
    FI_HIP(hipMemcpyAsync(raw.p, stage_buf.p, staged_bytes, hipMemcpyHostToDevice, s));
    set = true;
}

void DeviceImage::upload(const void* px, int width, int height, size_t stride_bytes, int pixel,
                         int depth12, hipStream_t s)
{
    stage(px, width, height, stride_bytes, pixel, depth12);
    dma(s);
    FI_HIP(hipStreamSynchronize(s));  // the staging buffer is reused by the next upload
}

void dft_twiddles(int n, float* out)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < n; k++) {
        double a = two_pi * (double)k / (double)n;
        out[2 * k] = (float)std::cos(a);
        out[2 * k + 1] = (float)(-std::sin(a));
    }
}

}  // namespace fi
