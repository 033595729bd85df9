// Host-side runtime shared by the engines: error plumbing (no exception ever
// crosses the C ABI), device buffers, pinned staging, device-resident images.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fi_launch.h"
#include "fi_params.h"

namespace fi {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

#define FI_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            ::fi::fail(FI_E_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

void set_last_error(const std::string& msg);

// Block pool (fi_runtime.cpp).  Device and pinned host blocks, allocated in
// size classes (eight per octave above 64 KiB), are kept after release by an
// owner that has made sure no queued work still reads them, and handed to the
// next request of the same class on the same device: a driver created for
// one image pair then takes its buffers without hipMalloc / hipHostMalloc
// and returns them without hipFree.  At most kPoolKeepDevice / kPoolKeepPinned
// bytes are kept per device; an allocation that runs out of memory frees the
// kept blocks of its device and tries again.  Blocks still kept when the
// process ends are left to the process's teardown.
enum PoolKind { kPoolDevice = 0, kPoolPinned = 1 };
constexpr size_t kPoolKeepDevice = size_t(16) << 30;
constexpr size_t kPoolKeepPinned = size_t(2) << 30;
size_t pool_round(size_t bytes);
// allocate a block of pool_round(bytes) on the current device (pinned: host
// memory) from the pool or the runtime; *dev receives the current device
void* pool_alloc(PoolKind kind, size_t bytes, int* dev);
// release a block of pool_alloc: kept when a RecycleScope is active on this
// thread and the pool has room, freed otherwise
void pool_release(PoolKind kind, int dev, void* p, size_t bytes);
// streams (non-blocking) kept the same way: pool_stream_release only for an
// idle stream
hipStream_t pool_stream();
void pool_stream_release(int dev, hipStream_t s);
int current_device();

// While armed, released DevArray / PinnedArray blocks of this thread go back
// to the pool.  An owner declares one as its first member (destroyed last)
// and arms it in its destructor once its streams are idle, so that every
// member buffer destroyed after the destructor's body is recycled.
struct RecycleScope {
    bool armed = false;
    RecycleScope() = default;
    RecycleScope(const RecycleScope&) = delete;
    RecycleScope& operator=(const RecycleScope&) = delete;
    void arm();
    ~RecycleScope();
};

// Device buffer that only grows.
template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t cap = 0;
    size_t bytes = 0;  // block size
    int dev = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void reserve(size_t n)
    {
        if (n <= cap) return;
        release();
        bytes = pool_round(n * sizeof(T));
        p = static_cast<T*>(pool_alloc(kPoolDevice, bytes, &dev));
        cap = bytes / sizeof(T);
    }
    void release()
    {
        if (p) pool_release(kPoolDevice, dev, p, bytes);
        p = nullptr;
        cap = 0;
        bytes = 0;
    }
    // exchanges the two blocks (ownership included)
    void swap(DevArray& o) noexcept
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(bytes, o.bytes);
        std::swap(dev, o.dev);
    }
};

// Pinned host staging buffer that only grows.
template <typename T>
struct PinnedArray {
    T* p = nullptr;
    size_t cap = 0;
    size_t bytes = 0;
    int dev = 0;
    PinnedArray() = default;
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    ~PinnedArray() { release(); }
    void reserve(size_t n)
    {
        if (n <= cap) return;
        release();
        bytes = pool_round(n * sizeof(T));
        p = static_cast<T*>(pool_alloc(kPoolPinned, bytes, &dev));
        cap = bytes / sizeof(T);
    }
    void release()
    {
        if (p) pool_release(kPoolPinned, dev, p, bytes);
        p = nullptr;
        cap = 0;
        bytes = 0;
    }
    void swap(PinnedArray& o) noexcept
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(bytes, o.bytes);
        std::swap(dev, o.dev);
    }
};

// An image uploaded as its raw pixels (u8 / u16 / f32), converted on the fly
// by the kernels that read it.  Upload validates the 12-bit right-aligned
// range (src/oc_image.cpp:172-178).
struct DeviceImage {
    DevArray<uint8_t> raw;
    PinnedArray<uint8_t> stage_buf;  // host staging: one contiguous DMA per upload
    DevImage view;
    bool set = false;
    void upload(const void* px, int width, int height, size_t stride_bytes, int pixel, int mode,
                hipStream_t s);
    // The two halves of upload: pack the rows into the pinned staging buffer
    // (host), then one DMA on stream s that is not waited for.  stage rejects a
    // right-aligned 12-bit image with a value above 4095 (src/oc_image.cpp:172-178).
    void stage(const void* px, int width, int height, size_t stride_bytes, int pixel, int mode);
    void dma(hipStream_t s);
    size_t staged_bytes = 0;
    fi_ingest_info info{};  // how the last staged image is converted
};

// mode: FI_DEPTH_* | FI_NORMALIZE_* (include/fasticgn.h)
void check_image_args(const void* px, int width, int height, size_t stride_bytes, int pixel,
                      int mode);
fi_ingest_info resolve_ingest(int pixel, int mode, unsigned max_value);

// Selects the handle's device for the duration of a call.
struct DeviceGuard {
    explicit DeviceGuard(int device) { FI_HIP(hipSetDevice(device)); }
};

// Twiddle table of the FFTCC DFT specification: (cos, -sin)(2 pi k / n),
// each rounded from double.
void dft_twiddles(int n, float* out);

// The IC-GN loop continues while its FP64 squared convergence norm s2 is >=
// conv_threshold2(conv), which is (float)sqrt(s2) >= conv (k_icgn.hip).
double conv_threshold2(float conv);

// Scratch of fi_icgn_compute_noalloc (k_noalloc.hip): PoiPrep records for at
// most kNoAllocMaxPois POIs per call (11.5 MB per GPU), in a file-scope device
// array; its address on `device` (the current device), taken on first use.
constexpr size_t kNoAllocMaxPois = 65536;
PoiPrep* noalloc_prep_scratch(int device);

// Frame-policy kernels (k_policy.hip).
// prev2 (optional): the frame before prev, for the two-frame extrapolation
// (steps up to max_delta px); frame->temporal counts the extrapolated seeds.
void launch_seed_from_previous(const fi_poi* tmpl, const fi_poi* prev, const fi_poi* prev2, float max_delta,
                               fi_poi* out, int n, FrameCounters* frame, hipStream_t s);
// Whole-frame fallback decision (see k_policy.hip): counters[0..nclass) hold
// the classes' scans, counters[nclass] the frame entry; class_n the class
// sizes; threshold (device) the baseline bad ratio + margin.
void launch_decide(FrameCounters* counters, int nclass, const unsigned* class_n, int n, const float* threshold,
                   int collect, hipStream_t s);
// dst = src (bytes, a multiple of 4) when *gate != 0
void launch_copy_if(const void* src, void* dst, size_t bytes, const unsigned* gate, hipStream_t s);
// 36-byte raw records of src[0..n) at out + 36 * order[i] (order null: i)
void launch_pack_raw(const fi_poi* src, const uint32_t* order, int n, void* out, hipStream_t s);
void launch_scan(const fi_poi* cur, const fi_poi* prev, int n, int collect,
                 FrameCounters* counters, uint32_t* candidates, hipStream_t s);
// count (optional, device): the launch covers min(*count, n) entries; n is
// the capacity the grid is sized for
// flags (optional): flags[idx[j]] = 1 for every gathered entry
void launch_gather(const fi_poi* src, const uint32_t* idx, int n, fi_poi* dst, hipStream_t s,
                   const unsigned* count = nullptr, uint8_t* flags = nullptr);
// gate (optional, device): merge only when *gate != 0; flags (optional):
// flags[0..n) = 1 when merged
void launch_merge_frame_fallback(const fi_poi* skip, fi_poi* scratch_inout, int n, hipStream_t s,
                                 const unsigned* gate = nullptr, uint8_t* flags = nullptr);
// seeded (optional, indexed like cur): the learned seeder's flags; a POI with
// flag 0 takes its repair unconditionally, the others under
// shouldReplacePoiWithFallback
void launch_merge_repair(fi_poi* cur, const fi_poi* repaired, const uint32_t* idx, int n,
                         FrameCounters* counters, hipStream_t s, const unsigned* count = nullptr,
                         const uint32_t* seeded = nullptr);
// counters->iterations += sum of iters[0..n_iters) (the frame's per-entry
// executed iteration counts)
void launch_qa(const fi_poi* cur, int n, FrameCounters* counters, hipStream_t s, const uint32_t* iters = nullptr,
               size_t n_iters = 0);
// Neighbourhood consensus of learned seeds (see k_policy.hip): raw (2 float4
// per POI) receives the seeds as they stand, then each seeded POI (ok) takes
// the mean of its seeded neighbours' predictions transported to it, over the
// neighbours that agree within c.tau; a POI outside c.sample is seeded from its
// neighbours alone and gets its ok here.
void launch_seed_consensus(fi_poi* pois, uint32_t* ok, float4* raw, int n, const SeedCells& c,
                           hipStream_t s);
// The cells of launch_seed_consensus on the host: side max(radius / 3,
// sqrt(area / m)) over the bounding box of the m finite positions (FP64),
// searched ceil(radius / side) cells to each side; POIs of a cell in position
// order; a position with a non-finite coordinate has cell -1.  sample_mode
// picks the POIs the network seeds (FI_SEED_SAMPLE_*): every POI, those in
// the cells (cx, cy) with cx + cy even (a checkerboard), or with cx and cy
// even (a quarter); sample / sample_idx are empty for every POI.
struct HostSeedCells {
    std::vector<uint32_t> start, idx;
    std::vector<int32_t> cell;
    std::vector<uint8_t> sample;
    std::vector<uint32_t> sample_idx;
    int ncx = 0, ncy = 0, k = 1;
};
HostSeedCells build_seed_cells(const float* xs, const float* ys, size_t n, float radius, int sample_mode = 0);
// Device copy of a HostSeedCells (uploaded on s, which is synchronised).
struct DevSeedCells {
    DevArray<uint32_t> start, idx, sample_idx;
    DevArray<int32_t> cell;
    DevArray<uint8_t> sample;
    size_t sampled = 0;  // entries of sample_idx (0: every POI, no list)
    SeedCells view{};
    void upload(const HostSeedCells& hc, float radius, float tau, hipStream_t s);
};
// learned-seed fallback candidates (see k_policy.hip)
void launch_seed_scan(const fi_poi* cur, const uint32_t* ok, int n, float threshold, FrameCounters* counters,
                      uint32_t* candidates, hipStream_t s);
// p[0..n) = 0 (32-bit words), by a kernel
void launch_zero_words(uint32_t* p, size_t n, hipStream_t s);
// p[0..n) = v (bytes); with gate (device), only when *gate != 0
void launch_fill_bytes(uint8_t* p, size_t n, uint8_t v, hipStream_t s, const unsigned* gate = nullptr);
// words[0..n_words) = 0, words2[0..n_words2) = 0, flags[0..n) = v, and src
// copied to dst when given: one launch
void launch_frame_reset(uint32_t* words, size_t n_words, uint32_t* words2, size_t n_words2, uint8_t* flags,
                        size_t n, uint8_t v, hipStream_t s, const void* src = nullptr, void* dst = nullptr,
                        size_t bytes = 0);

}  // namespace fi
