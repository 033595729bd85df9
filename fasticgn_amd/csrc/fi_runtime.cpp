// Host runtime: last-error storage, image upload, DFT twiddles.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <string>

#include "fi_host.h"

#include <map>
#include <mutex>

namespace fi {

// ---------------------------------------------------------------------------
// Block pool
// ---------------------------------------------------------------------------
namespace {

struct PoolShelf {
    std::map<size_t, std::vector<void*>> free;  // block size -> idle blocks
    size_t kept = 0;
};

struct Pool {
    std::mutex mu;
    std::map<int, PoolShelf> shelf[2];  // [kind][device]
    std::map<int, std::vector<hipStream_t>> streams;
};

Pool& pool()
{
    static Pool* p = new Pool;  // never destroyed: blocks outlive static teardown
    return *p;
}

thread_local int recycle_depth = 0;

void raw_free(PoolKind kind, void* p)
{
    if (kind == kPoolDevice)
        (void)hipFree(p);
    else
        (void)hipHostFree(p);
}

hipError_t raw_alloc(PoolKind kind, void** q, size_t bytes)
{
    return kind == kPoolDevice ? hipMalloc(q, bytes) : hipHostMalloc(q, bytes, hipHostMallocDefault);
}

// free every kept block of (kind, dev)
void pool_trim(PoolKind kind, int dev)
{
    std::vector<void*> drop;
    {
        Pool& P = pool();
        std::lock_guard<std::mutex> lk(P.mu);
        PoolShelf& sh = P.shelf[kind][dev];
        for (auto& kv : sh.free)
            for (void* b : kv.second) drop.push_back(b);
        sh.free.clear();
        sh.kept = 0;
    }
    for (void* b : drop) raw_free(kind, b);
}

}  // namespace

int current_device()
{
    int d = 0;
    FI_HIP(hipGetDevice(&d));
    return d;
}

size_t pool_round(size_t bytes)
{
    if (bytes <= 256) return 256;
    if (bytes <= (size_t(64) << 10)) {
        size_t b = 256;
        while (b < bytes) b <<= 1;
        return b;
    }
    size_t top = size_t(1) << 16;
    while ((top << 1) <= bytes) top <<= 1;  // top <= bytes < 2 top
    const size_t step = top >> 3;
    return (bytes + step - 1) / step * step;
}

void* pool_alloc(PoolKind kind, size_t bytes, int* dev)
{
    const int d = current_device();
    *dev = d;
    {
        Pool& P = pool();
        std::lock_guard<std::mutex> lk(P.mu);
        PoolShelf& sh = P.shelf[kind][d];
        auto it = sh.free.find(bytes);
        if (it != sh.free.end() && !it->second.empty()) {
            void* b = it->second.back();
            it->second.pop_back();
            sh.kept -= bytes;
            return b;
        }
    }
    void* q = nullptr;
    hipError_t e = raw_alloc(kind, &q, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        pool_trim(kind, d);
        e = raw_alloc(kind, &q, bytes);
    }
    if (e != hipSuccess)
        fail(e == hipErrorOutOfMemory ? FI_E_NO_MEMORY : FI_E_DEVICE,
             std::string(kind == kPoolDevice ? "hipMalloc" : "hipHostMalloc") + " of " + std::to_string(bytes) +
                 " bytes failed: " + hipGetErrorString(e));
    return q;
}

void pool_release(PoolKind kind, int dev, void* p, size_t bytes)
{
    if (recycle_depth > 0) {
        Pool& P = pool();
        std::lock_guard<std::mutex> lk(P.mu);
        PoolShelf& sh = P.shelf[kind][dev];
        if (sh.kept + bytes <= (kind == kPoolDevice ? kPoolKeepDevice : kPoolKeepPinned)) {
            sh.free[bytes].push_back(p);
            sh.kept += bytes;
            return;
        }
    }
    raw_free(kind, p);
}

hipStream_t pool_stream()
{
    const int d = current_device();
    {
        Pool& P = pool();
        std::lock_guard<std::mutex> lk(P.mu);
        auto& v = P.streams[d];
        if (!v.empty()) {
            hipStream_t s = v.back();
            v.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    FI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void pool_stream_release(int dev, hipStream_t s)
{
    if (!s) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    P.streams[dev].push_back(s);
}

void RecycleScope::arm()
{
    if (!armed) {
        armed = true;
        recycle_depth++;
    }
}

RecycleScope::~RecycleScope()
{
    if (armed) recycle_depth--;
}

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

const std::string& last_error() { return g_last_error; }

static size_t bytes_per_pixel(int pixel)
{
    return pixel == kU8 ? 1 : pixel == kU16 ? 2 : 4;
}

void check_image_args(const void* px, int width, int height, size_t stride_bytes, int pixel,
                      int mode)
{
    if (px == nullptr) fail(FI_E_INVALID_ARGUMENT, "image pointer is null");
    if (width < 1 || height < 1)
        fail(FI_E_INVALID_ARGUMENT, "image size must be positive, got " + std::to_string(width) +
                                        "x" + std::to_string(height));
    if (pixel != kU8 && pixel != kU16 && pixel != kF32)
        fail(FI_E_INVALID_ARGUMENT, "unknown pixel format " + std::to_string(pixel));
    const int depth = mode & FI_DEPTH_MASK, norm = mode & FI_NORMALIZE_MASK;
    if ((mode & ~(FI_DEPTH_MASK | FI_NORMALIZE_MASK)) || depth > FI_DEPTH16 || norm > FI_NORMALIZE_UNIT)
        fail(FI_E_INVALID_ARGUMENT, "unknown image input mode " + std::to_string(mode));
    if (pixel == kF32 && mode != FI_DEPTH_NATIVE)
        fail(FI_E_INVALID_ARGUMENT, "float pixels are taken as stored (no depth or normalise mode)");
    if (stride_bytes < (size_t)width * bytes_per_pixel(pixel))
        fail(FI_E_INVALID_ARGUMENT, "row stride " + std::to_string(stride_bytes) +
                                        " is shorter than a row");
}

// loadImage2DMat (src/oc_image.cpp:125-217) on decoded pixels: the effective
// depth and alignment (inferred from the maximum where the mode says so) and
// the scale, formed in FP64 as there and applied on the device as
// float(pixel) * float(scale).  A right-aligned 12-bit image above 4095 is
// rejected (:172-178).
fi_ingest_info resolve_ingest(int pixel, int mode, unsigned max_value)
{
    fi_ingest_info in{};
    const int depth = mode & FI_DEPTH_MASK, norm = mode & FI_NORMALIZE_MASK;
    in.source_max = (double)max_value;
    in.effective_bits = pixel == kU8 ? 8 : pixel == kF32 ? 32 : 16;
    in.left_aligned = 0;
    if (pixel != kF32) {
        if (depth == FI_DEPTH12_RIGHT) in.effective_bits = 12;
        else if (depth == FI_DEPTH12_LEFT) { in.effective_bits = 12; in.left_aligned = 1; }
        else if (depth == FI_DEPTH12_AUTO) { in.effective_bits = 12; in.left_aligned = max_value > 4095; }
        else if (depth == FI_DEPTH16) in.effective_bits = 16;
        else if (depth == FI_DEPTH_AUTO && pixel == kU16) in.effective_bits = max_value <= 4095 ? 12 : 16;
        if (in.effective_bits == 12 && !in.left_aligned && max_value > 4095)
            fail(FI_E_INVALID_ARGUMENT, "12-bit right-aligned input has values above 4095");
    }
    double scale = in.left_aligned ? 1.0 / 16.0 : 1.0;
    const double logical = in.effective_bits == 8 ? 255.0 : in.effective_bits == 12 ? 4095.0 : 65535.0;
    if (norm == FI_NORMALIZE_TO8) scale *= 255.0 / logical;
    else if (norm == FI_NORMALIZE_TO16) scale *= 65535.0 / logical;
    else if (norm == FI_NORMALIZE_UNIT) scale *= 1.0 / logical;
    in.scale = scale;
    return in;
}

// The modes that need the data's maximum (a range check or an inference).
static bool needs_max(int pixel, int mode)
{
    const int depth = mode & FI_DEPTH_MASK;
    return pixel == kU16 && (depth == FI_DEPTH12_RIGHT || depth == FI_DEPTH12_AUTO || depth == FI_DEPTH_AUTO);
}
// (8-bit data never exceeds 4095: its maximum decides nothing)

void DeviceImage::stage(const void* px, int width, int height, size_t stride_bytes, int pixel, int mode)
{
    check_image_args(px, width, height, stride_bytes, pixel, mode);
    const size_t row_bytes = (size_t)width * bytes_per_pixel(pixel);
    const size_t total = row_bytes * height;
    // rows packed into pinned memory (the copy out of pageable memory is a
    // plain memcpy; the DMA then runs at the link rate)
    this->stage_buf.reserve(total);
    // the maximum a mode needs is taken on the staged rows (each still in
    // cache after its copy): one pass over the caller's pixels
    const bool scan = needs_max(pixel, mode);
    uint16_t mx = 0;
    for (int y = 0; y < height; y++) {
        uint8_t* dst = stage_buf.p + (size_t)y * row_bytes;
        std::memcpy(dst, static_cast<const uint8_t*>(px) + (size_t)y * stride_bytes, row_bytes);
        if (scan) {
            const uint16_t* row = reinterpret_cast<const uint16_t*>(dst);
            for (int x = 0; x < width; x++) mx = row[x] > mx ? row[x] : mx;
        }
    }
    info = resolve_ingest(pixel, mode, mx);
    raw.reserve(total);
    staged_bytes = total;
    view.data = raw.p;
    view.width = width;
    view.height = height;
    view.pixel = pixel;
    view.pitch = (size_t)width;
    view.scale = (float)info.scale;
}

void DeviceImage::dma(hipStream_t s)
{
    FI_HIP(hipMemcpyAsync(raw.p, stage_buf.p, staged_bytes, hipMemcpyHostToDevice, s));
    set = true;
}

void DeviceImage::upload(const void* px, int width, int height, size_t stride_bytes, int pixel,
                         int mode, hipStream_t s)
{
    stage(px, width, height, stride_bytes, pixel, mode);
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

HostSeedCells build_seed_cells(const float* xs, const float* ys, size_t n, float radius, int sample_mode)
{
    HostSeedCells hc;
    hc.cell.assign(n, -1);
    double x0 = HUGE_VAL, x1 = -HUGE_VAL, y0 = HUGE_VAL, y1 = -HUGE_VAL;
    size_t m = 0;
    for (size_t i = 0; i < n; i++) {
        const double x = xs[i], y = ys[i];
        if (!std::isfinite(x) || !std::isfinite(y)) continue;
        x0 = std::min(x0, x);
        x1 = std::max(x1, x);
        y0 = std::min(y0, y);
        y1 = std::max(y1, y);
        m++;
    }
    if (m == 0) {
        hc.start.assign(1, 0);
        return hc;
    }
    // side >= radius / 3: the neighbours within +-radius lie within
    // k = ceil(radius / side) <= 3 cells of a POI's own (7 x 7 cells, a close
    // fit to the +-radius window); >= sqrt(area / m): about one POI per cell
    // at most
    const double side = std::max((double)radius / 3.0, std::sqrt((x1 - x0 + 1.0) * (y1 - y0 + 1.0) / (double)m));
    hc.k = std::max(1, (int)std::ceil((double)radius / side));
    const double nx = std::floor((x1 - x0) / side) + 1.0, ny = std::floor((y1 - y0) / side) + 1.0;
    if (nx * ny > 4.0 * (double)m + 65536.0)
        fail(FI_E_UNSUPPORTED, "seed consensus: POI layout too sparse for its cell grid");
    hc.ncx = (int)nx;
    hc.ncy = (int)ny;
    const size_t cells = (size_t)hc.ncx * hc.ncy;
    hc.start.assign(cells + 1, 0);
    for (size_t i = 0; i < n; i++) {
        const double x = xs[i], y = ys[i];
        if (!std::isfinite(x) || !std::isfinite(y)) continue;
        const int cx = std::min((int)std::floor((x - x0) / side), hc.ncx - 1);
        const int cy = std::min((int)std::floor((y - y0) / side), hc.ncy - 1);
        hc.cell[i] = cy * hc.ncx + cx;
        hc.start[(size_t)hc.cell[i] + 1]++;
    }
    for (size_t c = 0; c < cells; c++) hc.start[c + 1] += hc.start[c];
    hc.idx.resize(m);
    std::vector<uint32_t> fill(hc.start.begin(), hc.start.end() - 1);
    for (size_t i = 0; i < n; i++)
        if (hc.cell[i] >= 0) hc.idx[fill[(size_t)hc.cell[i]]++] = (uint32_t)i;
    if (sample_mode != FI_SEED_SAMPLE_ALL) {
        hc.sample.assign(n, 0);
        for (size_t i = 0; i < n; i++) {
            const int c = hc.cell[i];
            if (c < 0) continue;
            const int cx = c % hc.ncx, cy = c / hc.ncx;
            const bool pick = sample_mode == FI_SEED_SAMPLE_CHECKER ? ((cx + cy) & 1) == 0
                                                                     : ((cx | cy) & 1) == 0;
            if (pick) {
                hc.sample[i] = 1;
                hc.sample_idx.push_back((uint32_t)i);
            }
        }
    }
    return hc;
}

void DevSeedCells::upload(const HostSeedCells& hc, float radius, float tau, hipStream_t s)
{
    start.reserve(hc.start.size());
    idx.reserve(std::max<size_t>(hc.idx.size(), 1));
    cell.reserve(std::max<size_t>(hc.cell.size(), 1));
    FI_HIP(hipMemcpyAsync(start.p, hc.start.data(), hc.start.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!hc.idx.empty())
        FI_HIP(hipMemcpyAsync(idx.p, hc.idx.data(), hc.idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!hc.cell.empty())
        FI_HIP(hipMemcpyAsync(cell.p, hc.cell.data(), hc.cell.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    sampled = hc.sample_idx.size();
    if (!hc.sample.empty()) {
        sample.reserve(hc.sample.size());
        FI_HIP(hipMemcpyAsync(sample.p, hc.sample.data(), hc.sample.size(), hipMemcpyHostToDevice, s));
        sample_idx.reserve(std::max<size_t>(sampled, 1));
        if (sampled)
            FI_HIP(hipMemcpyAsync(sample_idx.p, hc.sample_idx.data(), sampled * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, s));
    }
    FI_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope with the caller
    view.sample = hc.sample.empty() ? nullptr : sample.p;
    view.start = start.p;
    view.idx = idx.p;
    view.cell = cell.p;
    view.ncx = hc.ncx;
    view.ncy = hc.ncy;
    view.k = hc.k;
    view.m = (int)hc.idx.size();
    view.radius = radius;
    view.tau = tau;
}

}  // namespace fi
