// Engine implementations: host orchestration of the gfx950 kernels.
#include "fi_engines.h"
#include "fi_seeder_engine.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <utility>

using namespace fi;

namespace {

int safe_int(float v)
{
    // (int) of a non-finite or out-of-range float is undefined in C++; such a
    // coordinate can never match a prepared layout entry.
    if (!std::isfinite(v) || v >= 2147483520.f || v <= -2147483520.f) return INT_MIN;
    return (int)v;
}

std::string size_str(int w, int h) { return std::to_string(w) + "x" + std::to_string(h); }

// IcgnParams::group for a sampling mode: 1 shared coefficient tiles, 2 two
// entries per wave over shared tiles, 0 otherwise
int group_code(int sampling)
{
    return sampling == FI_SAMPLING_GROUP ? 1 : sampling == FI_SAMPLING_PAIR ? 2 : 0;
}

// The IC-GN kernels read the reference, gradient and target images through
// buffer resources with 32-bit byte offsets.
void check_image_offsets(int w, int h)
{
    if ((size_t)w * (size_t)h * sizeof(float) > (size_t)INT_MAX)
        fail(FI_E_UNSUPPORTED, "image " + size_str(w, h) + " exceeds the 2 GiB f32 image limit of IC-GN");
}

// IC-GN launch with an optional explicit lane-to-POI shape
void icgn_launch(const IcgnParams& p, int waves, int spl, hipStream_t s)
{
    check_image_offsets(p.w, p.h);
    const std::string sub = std::to_string(2 * p.rx + 1) + "x" + std::to_string(2 * p.ry + 1);
    if (waves == 0) {
        if (!launch_icgn(p, s)) fail(FI_E_UNSUPPORTED, "no IC-GN lane shape fits a " + sub + " subset");
        return;
    }
    if (!launch_icgn_shape(p, waves, spl, s))
        fail(FI_E_INVALID_ARGUMENT, "IC-GN shape " + std::to_string(waves) + "x" + std::to_string(spl) +
                                        " does not fit a " + sub + " subset");
}

// The IC-GN kernels address the B-spline table with 32-bit byte offsets.
void check_table_size(int w, int h)
{
    if ((double)bspline_table_floats(w, h) * 4.0 > 4294967295.0)
        fail(FI_E_UNSUPPORTED, "image " + size_str(w, h) + " exceeds the 4 GiB B-spline table limit");
}

size_t max_lds_per_block(int device)
{
    int v = 0;
    FI_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    return (size_t)v;
}

// An engine's stream goes back to the block pool with the engine's buffers
// once it is idle; one that fails to synchronise is destroyed instead.
void retire_stream(int device, hipStream_t s, RecycleScope& recycle)
{
    if (!s) return;
    if (hipStreamSynchronize(s) == hipSuccess) {
        pool_stream_release(device, s);
        recycle.arm();
    } else {
        (void)hipStreamDestroy(s);
    }
}

}  // namespace

void fi::check_shape(int waves, int spl)
{
    if (waves == 0 && spl == 0) return;
    if (!icgn_shape_supported(waves, spl))
        fail(FI_E_INVALID_ARGUMENT, "unsupported IC-GN shape " + std::to_string(waves) + "x" + std::to_string(spl));
}

// ===========================================================================
// FFTCC2D
// ===========================================================================
fi_fftcc::fi_fftcc(int rx, int ry, int dev) : device(dev)
{
    // src/oc_fftcc.cpp:154-158 (and src/fftcc2d_cuda.h:85-109)
    if (rx <= 0 || rx != ry || rx > 23170)
        fail(FI_E_INVALID_ARGUMENT,
             "FFTCC2D supports positive square radii only (1..23170); rectangular FFTs are unsupported");
    if (rx > 32) fail(FI_E_UNSUPPORTED, "FFTCC on the device supports radii 1..32");
    r = rx;
    DeviceGuard g(device);
    if (fftcc_lds_bytes(r, kFftccUncached) > max_lds_per_block(device))
        fail(FI_E_UNSUPPORTED, "FFTCC window of radius " + std::to_string(r) + " exceeds the LDS per block");
    stream = pool_stream();
    std::vector<float> t(2 * 2 * r);
    dft_twiddles(2 * r, t.data());
    tw.reserve(t.size());
    FI_HIP(hipMemcpy(tw.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
}

fi_fftcc::~fi_fftcc() { retire_stream(device, stream, recycle_); }

void fi_fftcc::set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    DeviceGuard g(device);
    ref.upload(px, w, h, stride, pixel, mode, stream);
    ref_f.reserve((size_t)w * h);
    launch_to_float(ref.view, ref_f.p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    ready = false;  // layout retained for a lazy rebuild (:188-195)
}

void fi_fftcc::set_tar(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    DeviceGuard g(device);
    tar.upload(px, w, h, stride, pixel, mode, stream);
    tar_f.reserve((size_t)w * h);
    launch_to_float(tar.view, tar_f.p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
}

void fi_fftcc::prepare_ref_clear()
{
    lay_x.clear();
    lay_y.clear();
    ready = false;
}

void fi_fftcc::build_cache(const fi_poi* layout_dev, size_t n)
{
    const size_t nh = (size_t)2 * r * (r + 1);
    spectra.reserve(std::max<size_t>(1, n * nh * 2));
    slots.reserve(std::max<size_t>(1, n));
    FftccParams p{};
    p.ref = ref_f.p;
    p.tar = ref_f.p;
    p.w = ref.view.width;
    p.h = ref.view.height;
    p.r = r;
    p.mode = kFftccPrepare;
    p.pois = const_cast<fi_poi*>(layout_dev);
    p.n_pois = (int)n;
    p.spectra = spectra.p;
    p.slots = slots.p;
    p.twiddles = tw.p;
    launch_fftcc(p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    ready = true;
}

void fi_fftcc::prepare_ref(const fi_poi* layout, size_t n)
{
    if (!ref.set) fail(FI_E_LOGIC, "Reference image must be set before FFTCC2D::prepareRef");
    if (n > (size_t)INT_MAX) fail(FI_E_INVALID_ARGUMENT, "queue too long");
    if (n > 0 && layout == nullptr) fail(FI_E_INVALID_ARGUMENT, "layout pointer is null");
    DeviceGuard g(device);
    lay_x.resize(n);
    lay_y.resize(n);
    for (size_t i = 0; i < n; i++) {
        lay_x[i] = safe_int(layout[i].x);
        lay_y[i] = safe_int(layout[i].y);
    }
    ready = false;
    if (n == 0) {
        ready = true;
        return;
    }
    d_pois.reserve(n);
    FI_HIP(hipMemcpyAsync(d_pois.p, layout, n * sizeof(fi_poi), hipMemcpyHostToDevice, stream));
    build_cache(d_pois.p, n);
}

// ensureReferenceCache (:205-213): rebuild the spectra of the retained layout
void fi_fftcc::ensure_cache()
{
    if (ready) return;
    const size_t n = lay_x.size();
    std::vector<fi_poi> layout(n);
    for (size_t i = 0; i < n; i++) {
        std::memset(&layout[i], 0, sizeof(fi_poi));
        layout[i].x = (float)lay_x[i];
        layout[i].y = (float)lay_y[i];
    }
    if (n == 0) {
        ready = true;
        return;
    }
    d_pois.reserve(n);
    FI_HIP(hipMemcpyAsync(d_pois.p, layout.data(), n * sizeof(fi_poi), hipMemcpyHostToDevice, stream));
    build_cache(d_pois.p, n);
}

bool fi_fftcc::layout_matches(const fi_poi* pois, size_t n) const
{
    if (lay_x.size() != n) return false;
    for (size_t i = 0; i < n; i++)
        if (lay_x[i] != safe_int(pois[i].x) || lay_y[i] != safe_int(pois[i].y)) return false;
    return true;
}

void fi_fftcc::check_sizes() const
{
    if (!ref.set || !tar.set) fail(FI_E_LOGIC, "FFTCC2D images must be set before compute");
    // src/fftcc2d_cuda.cu:694-701: a size mismatch is rejected, naming both sizes
    if (ref.view.width != tar.view.width || ref.view.height != tar.view.height)
        fail(FI_E_INVALID_ARGUMENT, "FFTCC2D reference " + size_str(ref.view.width, ref.view.height) +
                                        " and target " + size_str(tar.view.width, tar.view.height) +
                                        " sizes differ");
}

void fi_fftcc::run(fi_poi* pois, size_t n, int mode, const uint32_t* slot_host)
{
    h_pois.reserve(n);
    std::memcpy(h_pois.p, pois, n * sizeof(fi_poi));
    d_pois.reserve(n);
    FI_HIP(hipMemcpyAsync(d_pois.p, h_pois.p, n * sizeof(fi_poi), hipMemcpyHostToDevice, stream));
    if (slot_host) {
        d_slot.reserve(n);
        FI_HIP(hipMemcpyAsync(d_slot.p, slot_host, n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    }
    FftccParams p{};
    p.ref = ref_f.p;
    p.tar = tar_f.p;
    p.w = ref.view.width;
    p.h = ref.view.height;
    p.r = r;
    p.mode = mode;
    p.pois = d_pois.p;
    p.n_pois = (int)n;
    p.slot = slot_host ? d_slot.p : nullptr;
    p.spectra = spectra.p;
    p.slots = slots.p;
    p.twiddles = tw.p;
    launch_fftcc(p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipMemcpyAsync(h_pois.p, d_pois.p, n * sizeof(fi_poi), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    std::memcpy(pois, h_pois.p, n * sizeof(fi_poi));
}

// compute(vector<POI2D>&), src/oc_fftcc.cpp:529-557
void fi_fftcc::compute(fi_poi* pois, size_t n)
{
    check_sizes();
    if (n == 0) return;
    if (pois == nullptr) fail(FI_E_INVALID_ARGUMENT, "POI pointer is null");
    if (n > (size_t)INT_MAX) fail(FI_E_INVALID_ARGUMENT, "queue too long");
    DeviceGuard g(device);
    if (layout_matches(pois, n)) {
        ensure_cache();
        run(pois, n, kFftccCached, nullptr);
    } else {
        run(pois, n, kFftccUncached, nullptr);
    }
}

// compute(queue, cache_indices), src/oc_fftcc.cpp:498-527
void fi_fftcc::compute_indexed(fi_poi* pois, size_t n, const size_t* idx)
{
    if (!ref.set || !tar.set) fail(FI_E_LOGIC, "FFTCC2D images must be set before compute");
    if (n == 0) return;
    if (pois == nullptr || idx == nullptr)
        fail(FI_E_INVALID_ARGUMENT, "FFTCC2D::compute requires one cache index per POI");
    if (lay_x.empty()) fail(FI_E_LOGIC, "FFTCC2D indexed compute requires a prepared reference layout");
    std::vector<uint32_t> slot(n);
    for (size_t i = 0; i < n; i++) {
        const size_t k = idx[i];
        if (k >= lay_x.size()) fail(FI_E_OUT_OF_RANGE, "FFTCC2D reference cache index out of range");
        if (!std::isfinite(pois[i].x) || !std::isfinite(pois[i].y) || lay_x[k] != safe_int(pois[i].x) ||
            lay_y[k] != safe_int(pois[i].y))
            fail(FI_E_INVALID_ARGUMENT, "FFTCC2D reference cache index does not match POI coordinates");
        slot[i] = (uint32_t)k;
    }
    check_sizes();
    DeviceGuard g(device);
    ensure_cache();
    run(pois, n, kFftccCached, slot.data());
}

// ===========================================================================
// ICGN2D1
// ===========================================================================
fi_icgn::fi_icgn(int rx_, int ry_, float conv_, float stop_, int dev)
    : device(dev), rx(rx_), ry(ry_), conv(conv_), stop(stop_)
{
    if (rx < 1 || ry < 1) fail(FI_E_INVALID_ARGUMENT, "ICGN2D1 subset radii must be positive");
    if (rx > kIcgnMaxRadius || ry > kIcgnMaxRadius)
        fail(FI_E_UNSUPPORTED, "ICGN on the device supports radii 1.." + std::to_string(kIcgnMaxRadius));
    DeviceGuard g(device);
    stream = pool_stream();
}

fi_icgn::~fi_icgn() { retire_stream(device, stream, recycle_); }

void fi_icgn::set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    check_image_offsets(w, h);
    DeviceGuard g(device);
    ref_raw.upload(px, w, h, stride, pixel, mode, stream);
    grads_ready = false;
    ref_version++;
}

void fi_icgn::set_tar(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    DeviceGuard g(device);
    tar_raw.upload(px, w, h, stride, pixel, mode, stream);
    coef_ready = false;
}

// prepareRef (:132-142): Gradient2D4 of the reference
void fi_icgn::prepare_ref()
{
    if (!ref_raw.set) fail(FI_E_LOGIC, "ICGN2D1::prepareRef requires a reference image");
    DeviceGuard g(device);
    const size_t px = (size_t)ref_raw.view.width * ref_raw.view.height;
    ref_f.reserve(px);
    gx.reserve(px);
    gy.reserve(px);
    launch_to_float(ref_raw.view, ref_f.p, stream);
    launch_gradient(ref_f.p, ref_raw.view.width, ref_raw.view.height, gx.p, gy.p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    grads_ready = true;
    ref_version++;
}

// prepareTar (:144-153): BicubicBspline table of the target
void fi_icgn::prepare_tar()
{
    if (!tar_raw.set) fail(FI_E_LOGIC, "ICGN2D1::prepareTar requires a target image");
    DeviceGuard g(device);
    check_table_size(tar_raw.view.width, tar_raw.view.height);
    coef.reserve(bspline_table_floats(tar_raw.view.width, tar_raw.view.height));
    launch_bspline(tar_raw.view, coef.p, stream);
    tar_f.reserve((size_t)tar_raw.view.width * tar_raw.view.height);
    launch_to_float(tar_raw.view, tar_f.p, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    coef_ready = true;
}

// compute(vector<POI2D>&) (src/oc_icgn.cpp:389-397); with offsets,
// compute(queue, center_offset_queue) (:399-625), whose Hessians depend on the
// offsets and so are never taken from the prepared-layout cache.
void fi_icgn::compute(fi_poi* pois, size_t n, const float* offsets)
{
    if (!grads_ready || !coef_ready)
        fail(FI_E_LOGIC, "ICGN2D1 compute requires prepareRef and prepareTar after the images are set");
    if (ref_raw.view.width != tar_raw.view.width || ref_raw.view.height != tar_raw.view.height)
        fail(FI_E_INVALID_ARGUMENT, "ICGN2D1 reference " + size_str(ref_raw.view.width, ref_raw.view.height) +
                                        " and target " + size_str(tar_raw.view.width, tar_raw.view.height) +
                                        " sizes differ");
    if (n == 0) return;
    if (pois == nullptr) fail(FI_E_INVALID_ARGUMENT, "POI pointer is null");
    if (n > (size_t)INT_MAX) fail(FI_E_INVALID_ARGUMENT, "queue too long");
    DeviceGuard g(device);

    // radius groups (self_adaptive: per-POI subset_radius, src/oc_icgn.cpp:176-182)
    std::map<std::pair<int, int>, std::vector<uint32_t>> groups;
    std::vector<float> key(4 * n);
    for (size_t i = 0; i < n; i++) {
        int gx_r = rx, gy_r = ry;
        if (self_adaptive) {
            gx_r = safe_int(pois[i].subset_radius[0]);
            gy_r = safe_int(pois[i].subset_radius[1]);
            if (gx_r < 1 || gy_r < 1 || gx_r > kIcgnMaxRadius || gy_r > kIcgnMaxRadius)
                fail(FI_E_UNSUPPORTED, "self-adaptive subset radius outside 1.." +
                                           std::to_string(kIcgnMaxRadius) + " at POI " + std::to_string(i));
            groups[{ gx_r, gy_r }].push_back((uint32_t)i);
        }
        key[4 * i] = pois[i].x;
        key[4 * i + 1] = pois[i].y;
        key[4 * i + 2] = (float)gx_r;
        key[4 * i + 3] = (float)gy_r;
    }
    const bool prep_hit = !offsets && prep_version == ref_version && prep_key.size() == key.size() &&
                          std::memcmp(prep_key.data(), key.data(), key.size() * sizeof(float)) == 0;

    h_pois.reserve(n);
    std::memcpy(h_pois.p, pois, n * sizeof(fi_poi));
    d_pois.reserve(n);
    FI_HIP(hipMemcpyAsync(d_pois.p, h_pois.p, n * sizeof(fi_poi), hipMemcpyHostToDevice, stream));
    prep.reserve(n);

    IcgnParams p{};
    p.ref = ref_f.p;
    p.gx = gx.p;
    p.gy = gy.p;
    p.coef = coef.p;
    p.tar = sampling == FI_SAMPLING_TABLE ? nullptr : tar_f.p;
    p.group = group_code(sampling);
    p.w = ref_raw.view.width;
    p.h = ref_raw.view.height;
    p.pois = d_pois.p;
    p.prep = prep.p;
    p.conv = conv;
    p.stop = stop;
    if (offsets) {
        d_offset.reserve(2 * n);
        FI_HIP(hipMemcpyAsync(d_offset.p, offsets, 2 * n * sizeof(float), hipMemcpyHostToDevice, stream));
        p.offset = d_offset.p;
    }
    if (!self_adaptive) {
        p.index = nullptr;
        p.n = (int)n;
        p.rx = rx;
        p.ry = ry;
        if (!prep_hit) launch_icgn_prep(p, stream);
        icgn_launch(p, shape_waves, shape_spl, stream);
        FI_HIP(hipGetLastError());
    } else {
        size_t total = 0;
        for (auto& kv : groups) total += kv.second.size();
        d_index.reserve(total);
        std::vector<uint32_t> flat;
        flat.reserve(total);
        std::vector<std::pair<size_t, size_t>> spans;
        for (auto& kv : groups) {
            spans.push_back({ flat.size(), kv.second.size() });
            flat.insert(flat.end(), kv.second.begin(), kv.second.end());
        }
        FI_HIP(hipMemcpyAsync(d_index.p, flat.data(), flat.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, stream));
        p.self_adaptive = 1;
        size_t gi = 0;
        for (auto& kv : groups) {
            p.rx = kv.first.first;
            p.ry = kv.first.second;
            p.index = d_index.p + spans[gi].first;
            p.n = (int)spans[gi].second;
            if (!prep_hit) launch_icgn_prep(p, stream);
            // an explicit shape applies only where it fits the group's subset
            if (shape_waves && icgn_shape_covers(shape_waves, shape_spl, p.rx, p.ry))
                icgn_launch(p, shape_waves, shape_spl, stream);
            else
                icgn_launch(p, 0, 0, stream);
            FI_HIP(hipGetLastError());
            gi++;
        }
        // the staged index list must outlive the async copy
        FI_HIP(hipStreamSynchronize(stream));
    }
    FI_HIP(hipMemcpyAsync(h_pois.p, d_pois.p, n * sizeof(fi_poi), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    std::memcpy(pois, h_pois.p, n * sizeof(fi_poi));
    prep_n = n;
    if (offsets) {
        prep_key.clear();  // the records hold this call's offsets
    } else {
        prep_key.swap(key);
        prep_version = ref_version;
    }
}

// compute(vector<POI2D>&) on POIs the caller keeps in device memory, with
// the per-POI reference records in the library's static scratch
// (k_noalloc.hip) instead of the handle's buffers: no device or pinned
// allocation, no host staging.  Uniform subsets only (self-adaptive subsets
// need a host-side grouping), at most kNoAllocMaxPois POIs.
void fi_icgn::compute_noalloc(fi_poi* d_pois, size_t n)
{
    if (!grads_ready || !coef_ready)
        fail(FI_E_LOGIC, "ICGN2D1 compute requires prepareRef and prepareTar after the images are set");
    if (ref_raw.view.width != tar_raw.view.width || ref_raw.view.height != tar_raw.view.height)
        fail(FI_E_INVALID_ARGUMENT, "ICGN2D1 reference and target sizes differ");
    if (self_adaptive) fail(FI_E_UNSUPPORTED, "the no-allocation compute takes uniform subsets only");
    if (n > kNoAllocMaxPois)
        fail(FI_E_INVALID_ARGUMENT, "the no-allocation compute takes at most " + std::to_string(kNoAllocMaxPois) +
                                        " POIs per call, got " + std::to_string(n));
    if (n == 0) return;
    if (d_pois == nullptr) fail(FI_E_INVALID_ARGUMENT, "POI pointer is null");
    DeviceGuard g(device);
    IcgnParams p{};
    p.ref = ref_f.p;
    p.gx = gx.p;
    p.gy = gy.p;
    p.coef = coef.p;
    p.tar = sampling == FI_SAMPLING_TABLE ? nullptr : tar_f.p;
    p.group = group_code(sampling);
    p.w = ref_raw.view.width;
    p.h = ref_raw.view.height;
    p.pois = d_pois;
    p.prep = noalloc_prep_scratch(device);
    p.conv = conv;
    p.stop = stop;
    p.n = (int)n;
    p.rx = rx;
    p.ry = ry;
    launch_icgn_prep(p, stream);
    icgn_launch(p, shape_waves, shape_spl, stream);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
}

void fi_icgn::export_data(int what, void* out, size_t bytes)
{
    if (out == nullptr && bytes > 0) fail(FI_E_INVALID_ARGUMENT, "output pointer is null");
    DeviceGuard g(device);
    const int w = ref_raw.view.width, h = ref_raw.view.height;
    const size_t px = (size_t)w * h;
    auto copy_out = [&](const void* src, size_t want) {
        if (bytes != want)
            fail(FI_E_INVALID_ARGUMENT, "export needs " + std::to_string(want) + " bytes, got " + std::to_string(bytes));
        if (want) FI_HIP(hipMemcpy(out, src, want, hipMemcpyDeviceToHost));
    };
    switch (what) {
    case FI_EXPORT_REF_F32:
    case FI_EXPORT_GRAD_X:
    case FI_EXPORT_GRAD_Y:
        if (!grads_ready) fail(FI_E_LOGIC, "export of the reference side needs prepareRef");
        copy_out(what == FI_EXPORT_REF_F32 ? ref_f.p : (what == FI_EXPORT_GRAD_X ? gx.p : gy.p), px * sizeof(float));
        break;
    case FI_EXPORT_TAR_F32:
        if (!coef_ready) fail(FI_E_LOGIC, "export of the target side needs prepareTar");
        copy_out(tar_f.p, (size_t)tar_raw.view.width * tar_raw.view.height * sizeof(float));
        break;
    case FI_EXPORT_BSPLINE: {
        // table rows hold four planes of w float4 (plane k = coef[k][0..3]);
        // out is texel-major: 16 floats coef[k][l] per texel, k*4 + l
        if (!coef_ready) fail(FI_E_LOGIC, "export of the B-spline table needs prepareTar");
        const int tw = tar_raw.view.width, th = tar_raw.view.height;
        const size_t want = (size_t)16 * tw * th * sizeof(float);
        if (bytes != want)
            fail(FI_E_INVALID_ARGUMENT, "export needs " + std::to_string(want) + " bytes, got " + std::to_string(bytes));
        std::vector<float> tab((size_t)16 * tw * th);
        FI_HIP(hipMemcpy(tab.data(), coef.p, tab.size() * sizeof(float), hipMemcpyDeviceToHost));
        float* o = static_cast<float*>(out);
        for (int y = 0; y < th; y++)
            for (int k = 0; k < 4; k++)
                for (int x = 0; x < tw; x++)
                    std::memcpy(o + 16 * ((size_t)y * tw + x) + 4 * k, tab.data() + 4 * ((size_t)(4 * y + k) * tw + x),
                                4 * sizeof(float));
        break;
    }
    case FI_EXPORT_PREP:
        copy_out(prep.p, prep_n * sizeof(PoiPrep));
        break;
    default:
        fail(FI_E_INVALID_ARGUMENT, "unknown export " + std::to_string(what));
    }
}

// ===========================================================================
// Multi-frame driver
// ===========================================================================
namespace {
constexpr float kBadRatioMargin = 0.03f;  // kSkipFftccBadRatioMargin, cli:480
}

fi_frame::fi_frame(int radius, float conv_, float stop_, int dev)
    : device(dev), r(radius), conv(conv_), stop(stop_)
{
    if (radius < 1 || radius > 32)
        fail(radius < 1 ? FI_E_INVALID_ARGUMENT : FI_E_UNSUPPORTED,
             "frame driver radius must be 1..32 (FFTCC) and 1.." + std::to_string(kIcgnMaxRadius) + " (ICGN)");
    if (radius > kIcgnMaxRadius) fail(FI_E_UNSUPPORTED, "ICGN on the device supports radii 1..31");
    DeviceGuard g(device);
    if (fftcc_lds_bytes(r, kFftccCached) > max_lds_per_block(device))
        fail(FI_E_UNSUPPORTED, "FFTCC window exceeds the LDS per block");
    stream = pool_stream();
    copy_stream = pool_stream();
    aux_stream = pool_stream();
    FI_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    FI_HIP(hipEventCreateWithFlags(&tmpl_ev, hipEventDisableTiming));
    FI_HIP(hipEventCreate(&join_ev));
    for (auto& e : ev) FI_HIP(hipEventCreate(&e));
    for (auto& t : tslot) {
        FI_HIP(hipEventCreateWithFlags(&t.ready, hipEventDisableTiming));
        FI_HIP(hipEventCreateWithFlags(&t.freed, hipEventDisableTiming));
    }
    for (auto& pair : fev)
        for (auto& e : pair) FI_HIP(hipEventCreate(&e));
    rec_stream = pool_stream();
    for (auto& e : rec_ev) FI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    FI_HIP(hipEventCreateWithFlags(&pack_ev, hipEventDisableTiming));
    threshold.reserve(1);
    stager = std::thread([this] { stage_loop(); });
}

fi_frame::~fi_frame()
{
    {
        std::lock_guard<std::mutex> lk(stage_mu);
        stage_stop = true;
    }
    stage_cv.notify_all();
    if (stager.joinable()) stager.join();
    for (auto& gr : graphs) (void)hipGraphExecDestroy(gr.exec);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& t : tslot) {
        if (t.ready) (void)hipEventDestroy(t.ready);
        if (t.freed) (void)hipEventDestroy(t.freed);
    }
    for (auto& pair : fev)
        for (auto& e : pair)
            if (e) (void)hipEventDestroy(e);
    for (auto& e : rec_ev)
        if (e) (void)hipEventDestroy(e);
    if (pack_ev) (void)hipEventDestroy(pack_ev);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (tmpl_ev) (void)hipEventDestroy(tmpl_ev);
    // Every stream idle: nothing queued reads the buffers any more, so they
    // and the streams go back to the block pool.  A stream that fails to
    // synchronise is destroyed and the buffers are freed instead.
    bool idle = true;
    for (hipStream_t s : { stream, copy_stream, aux_stream, rec_stream })
        if (s) idle = hipStreamSynchronize(s) == hipSuccess && idle;
    for (hipStream_t s : { stream, copy_stream, aux_stream, rec_stream }) {
        if (!s) continue;
        if (idle)
            pool_stream_release(device, s);
        else
            (void)hipStreamDestroy(s);
    }
    if (idle) recycle_.arm();
}

// Staging thread: packs a queued target into its slot's pinned buffer and
// starts the DMA on copy_stream.  upload_target queues a slot only after the
// last frame that took it has finished (its `freed` event: frames launched by
// run_async may still be running).  The wait is the caller's, not this
// thread's: a host wait on a compute-stream event here could meet a frame
// graph being captured on that stream.
void fi_frame::stage_loop()
{
    for (;;) {
        StageJob job;
        {
            std::unique_lock<std::mutex> lk(stage_mu);
            stage_cv.wait(lk, [&] { return stage_stop || !stage_jobs.empty(); });
            if (stage_stop) return;
            job = stage_jobs.front();
            stage_jobs.pop_front();
        }
        TargetSlot& t = tslot[job.slot];
        int code = 0;
        std::string msg;
        try {
            DeviceGuard g(device);
            t.img.stage(job.px, job.w, job.h, job.stride, job.pixel, job.mode);
            t.img.dma(copy_stream);
            FI_HIP(hipEventRecord(t.ready, copy_stream));
        } catch (const Error& e) {
            code = e.code;
            msg = e.what();
        } catch (const std::exception& e) {
            code = FI_E_DEVICE;
            msg = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(stage_mu);
            t.err_code = code;
            t.err = msg;
            t.state = 2;
        }
        stage_cv.notify_all();
    }
}

// The slot holding the next frame's target once its rows are staged and its
// DMA is queued; consume: take it off the queue.
int fi_frame::take_target(bool consume)
{
    int s;
    if (!pending.empty()) {
        s = pending.front();
        if (consume) pending.pop_front();
    } else {
        s = last_slot;
    }
    if (s < 0) fail(FI_E_LOGIC, "frame driver needs an uploaded target");
    TargetSlot& t = tslot[s];
    std::unique_lock<std::mutex> lk(stage_mu);
    stage_cv.wait(lk, [&] { return t.state != 1; });
    if (t.err_code) {
        const int c = t.err_code;
        const std::string m = t.err;
        t.err_code = 0;
        t.state = 0;
        if (s == last_slot) last_slot = -1;
        lk.unlock();
        fail(c, m);
    }
    return s;
}

void fi_frame::invalidate_graphs()
{
    for (auto& gr : graphs) (void)hipGraphExecDestroy(gr.exec);
    graphs.clear();
}

void fi_frame::set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    check_image_offsets(w, h);
    DeviceGuard g(device);
    ref_raw.upload(px, w, h, stride, pixel, mode, stream);
    ref_set = true;
    prepared = false;
    reset_sequence();
}

void fi_frame::set_pois(const fi_poi* pois, size_t count)
{
    if (count > 0 && pois == nullptr) fail(FI_E_INVALID_ARGUMENT, "POI pointer is null");
    if (count > (size_t)INT_MAX) fail(FI_E_INVALID_ARGUMENT, "queue too long");
    {
        DeviceGuard g(device);
        tmpl_host.reserve(std::max<size_t>(count, 1));
    }
    if (count) std::memcpy(tmpl_host.p, pois, count * sizeof(fi_poi));
    n = count;
    pois_set = true;
    prepared = false;
    reset_sequence();
}

void fi_frame::set_self_adaptive(bool on)
{
    self_adaptive = on;
    prepared = false;
    reset_sequence();
}

// Queues one target (host pixels) for the next fi_frame_run: the rows are
// copied to pinned memory and sent to the device by the staging thread, so the
// copy overlaps a frame that is running.  The caller's pixels must stay valid
// until the next run (or upload) returns.
void fi_frame::upload_target(const void* px, int w, int h, size_t stride, int pixel, int mode)
{
    DeviceGuard g(device);
    if (ref_set && (w != ref_raw.view.width || h != ref_raw.view.height))
        fail(FI_E_INVALID_ARGUMENT, "reference " + size_str(ref_raw.view.width, ref_raw.view.height) +
                                        " and target " + size_str(w, h) + " sizes differ");
    check_image_args(px, w, h, stride, pixel, mode);
    // 12-bit range: checked by the staging thread as it copies the rows
    // (reported by the run that takes this target)
    if (pending.size() >= 2) fail(FI_E_LOGIC, "two targets are already queued; run a frame first");
    const int s = next_slot;
    // the slot's last frame (possibly still in flight) has read it
    FI_HIP(hipEventSynchronize(tslot[s].freed));
    next_slot ^= 1;
    {
        std::unique_lock<std::mutex> lk(stage_mu);
        stage_cv.wait(lk, [&] { return tslot[s].state != 1; });
        tslot[s].state = 1;
        tslot[s].err_code = 0;
        stage_jobs.push_back({ s, px, w, h, stride, pixel, mode });
    }
    stage_cv.notify_all();
    pending.push_back(s);
    if (last_slot == s) last_slot = -1;
    tar_ready = true;
}

void fi_frame::reset_sequence()
{
    have_prev = false;
    have_prev2 = false;
    rec_held[0] = rec_held[1] = -1;
    baseline_ready = false;
    baseline_bad = NAN;
    frame_index = 0;
    invalidate_graphs();
}

// prepareRef of both engines for the template (cli:4095-4104): reference f32,
// gradients, FFTCC spectra, per-POI Hessian records.  With self_adaptive the
// template is laid out on the device sorted by subset radius (stable), one
// radius class per distinct radius.
void fi_frame::prepare()
{
    if (!ref_set) fail(FI_E_LOGIC, "frame driver needs a reference image");
    if (!pois_set) fail(FI_E_LOGIC, "frame driver needs a POI template");
    cells_ready = false;  // the seed-consensus cells follow the layout
    const int w = ref_raw.view.width, h = ref_raw.view.height;
    const size_t px = (size_t)w * h;

    // radius classes
    std::vector<int> rad(n, r);
    order.clear();
    if (self_adaptive) {
        const size_t lds = max_lds_per_block(device);
        for (size_t i = 0; i < n; i++) {
            const int rx = safe_int(tmpl_host.p[i].subset_radius[0]), ry = safe_int(tmpl_host.p[i].subset_radius[1]);
            if (rx != ry)
                fail(FI_E_UNSUPPORTED, "frame driver subsets are square (FFTCC2D): POI " + std::to_string(i) +
                                           " has subset_radius " + std::to_string(rx) + "x" + std::to_string(ry));
            if (rx < 1 || rx > kIcgnMaxRadius)
                fail(FI_E_UNSUPPORTED, "self-adaptive subset radius outside 1.." + std::to_string(kIcgnMaxRadius) +
                                           " at POI " + std::to_string(i));
            if (fftcc_lds_bytes(rx, kFftccCached) > lds)
                fail(FI_E_UNSUPPORTED, "FFTCC window of radius " + std::to_string(rx) + " exceeds the LDS per block");
            rad[i] = rx;
        }
        order.resize(n);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rad[a] < rad[b]; });
    }
    classes.clear();
    fused.waves = fused.spl = 0;  // tuned for the previous classes
    fused.sampling = -1;
    fused.queue_blocks = -1;
    for (size_t i = 0; i < n; i++) {
        const int ri = order.empty() ? rad[i] : rad[order[i]];
        if (classes.empty() || classes.back()->r != ri) {
            classes.push_back(std::make_unique<FrameClass>());
            classes.back()->r = ri;
            classes.back()->off = i;
        }
        classes.back()->n++;
    }
    std::vector<fi_poi> layout;
    const fi_poi* lay = tmpl_host.p;
    if (!order.empty()) {
        layout.resize(n);
        for (size_t i = 0; i < n; i++) layout[i] = tmpl_host.p[order[i]];
        lay = layout.data();
    }

    using clock = std::chrono::steady_clock;
    const auto t_start = clock::now();
    // device buffers first (the allocation share of a cold frame)
    ref_f.reserve(px);
    gx.reserve(px);
    gy.reserve(px);
    const size_t cap = std::max<size_t>(n, 1);
    tmpl.reserve(cap);
    iters.reserve(3 * cap);
    seed_ok.reserve(cap);
    cur.reserve(cap);
    prev.reserve(cap);
    prev2.reserve(cap);
    save.reserve(cap);
    repair.reserve(cap);
    cand.reserve(cap);
    redo.reserve(cap);
    prep.reserve(cap);
    counters.reserve(nclass() + 1);
    h_counters.reserve(nclass() + 1);
    queue.reserve((size_t)kQueueWords * std::max<size_t>(nclass(), 1));
    class_n.reserve(std::max<size_t>(nclass(), 1));
    if (!order.empty()) d_order.reserve(n);
    for (auto& cp : classes) {
        FrameClass& c = *cp;
        c.tw.reserve(4 * (size_t)c.r);
        c.spectra.reserve(c.n * (size_t)2 * c.r * (c.r + 1) * 2);
        c.slots.reserve(c.n);
    }
    if (records_on)
        for (int k = 0; k < 2; k++) {
            d_rec[k].reserve(std::max<size_t>(n * (size_t)36, 36));
            h_rec[k].reserve(std::max<size_t>(n * (size_t)36, 36));
        }
    const auto t_alloc = clock::now();
    // host data: template layout (on copy_stream, beside the reference
    // conversion and gradients; pinned unless reordered), class sizes, layout
    // order, DFT twiddles (the host vectors live until the closing
    // synchronisation, which follows the template's copy)
    if (n > 0) {
        FI_HIP(hipMemcpyAsync(tmpl.p, lay, n * sizeof(fi_poi), hipMemcpyHostToDevice, copy_stream));
        FI_HIP(hipEventRecord(tmpl_ev, copy_stream));
    }
    std::vector<unsigned> sizes(nclass());
    for (size_t c = 0; c < nclass(); c++) sizes[c] = (unsigned)classes[c]->n;
    if (!sizes.empty())
        FI_HIP(hipMemcpyAsync(class_n.p, sizes.data(), sizes.size() * sizeof(unsigned), hipMemcpyHostToDevice,
                              stream));
    if (!order.empty())
        FI_HIP(hipMemcpyAsync(d_order.p, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    std::vector<std::vector<float>> tws(nclass());
    for (size_t ci = 0; ci < nclass(); ci++) {
        FrameClass& c = *classes[ci];
        tws[ci].resize(4 * c.r);
        dft_twiddles(2 * c.r, tws[ci].data());
        FI_HIP(hipMemcpyAsync(c.tw.p, tws[ci].data(), tws[ci].size() * sizeof(float), hipMemcpyHostToDevice,
                              stream));
    }
    // reference f32 and Gradient2D4, FFTCC reference spectra, per-POI records
    FI_HIP(hipEventRecord(ev[4], stream));
    launch_to_float(ref_raw.view, ref_f.p, stream);
    launch_gradient(ref_f.p, w, h, gx.p, gy.p, stream);
    FI_HIP(hipEventRecord(ev[5], stream));
    if (n > 0) FI_HIP(hipStreamWaitEvent(stream, tmpl_ev, 0));
    for (auto& cp : classes) {
        FrameClass& c = *cp;
        FftccParams fp{};
        fp.ref = ref_f.p;
        fp.tar = ref_f.p;
        fp.w = w;
        fp.h = h;
        fp.r = c.r;
        fp.mode = kFftccPrepare;
        fp.pois = tmpl.p + c.off;
        fp.n_pois = (int)c.n;
        fp.spectra = c.spectra.p;
        fp.slots = c.slots.p;
        fp.twiddles = c.tw.p;
        launch_fftcc(fp, stream);
    }
    FI_HIP(hipEventRecord(ev[6], stream));
    for (auto& cp : classes) {
        FrameClass& c = *cp;
        IcgnParams ip{};
        ip.ref = ref_f.p;
        ip.gx = gx.p;
        ip.gy = gy.p;
        ip.w = w;
        ip.h = h;
        ip.pois = tmpl.p + c.off;
        ip.n = (int)c.n;
        ip.prep = prep.p + c.off;
        ip.rx = c.r;
        ip.ry = c.r;
        launch_icgn_prep(ip, stream);
    }
    FI_HIP(hipEventRecord(ev[7], stream));
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    const auto t_end = clock::now();
    fi_prepare_stats ps{};
    ps.total_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    ps.alloc_ms = std::chrono::duration<double, std::milli>(t_alloc - t_start).count();
    float ms = 0.f;
    FI_HIP(hipEventElapsedTime(&ms, ev[4], ev[5]));
    ps.reference_ms = ms;
    FI_HIP(hipEventElapsedTime(&ms, ev[5], ev[6]));
    ps.fftcc_ms = ms;
    FI_HIP(hipEventElapsedTime(&ms, ev[6], ev[7]));
    ps.hessian_ms = ms;
    last_prepare = ps;
    invalidate_graphs();
    prepared = true;
}

fi_prepare_stats fi_frame::prepare_now()
{
    DeviceGuard g(device);
    if (!prepared) prepare();
    return last_prepare;
}

void fi_frame::fftcc_pass(const FrameClass& c, fi_poi* queue_base, int count, const uint32_t* slot_dev,
                          const unsigned* count_dev)
{
    FftccParams fp{};
    fp.ref = ref_f.p;
    fp.tar = tar_f.p;
    fp.w = ref_raw.view.width;
    fp.h = ref_raw.view.height;
    fp.r = c.r;
    fp.mode = kFftccCached;
    fp.pois = queue_base;
    fp.n_pois = count;
    fp.slot = slot_dev;
    fp.spectra = c.spectra.p;
    fp.slots = c.slots.p;
    fp.twiddles = c.tw.p;
    fp.count = count_dev;
    launch_fftcc(fp, stream);
}

void fi_frame::icgn_pass(const FrameClass& c, fi_poi* queue_base, int count, const uint32_t* prep_index,
                         const unsigned* count_dev, int region, int first)
{
    if (count <= 0) return;
    const int mode = c.sampling >= 0 ? c.sampling : sampling;
    IcgnParams ip{};
    ip.ref = ref_f.p;
    ip.gx = gx.p;
    ip.gy = gy.p;
    // the table exists only while a class samples it; the LDS-tile mode
    // samples the f32 target when a warp outgrows the tile
    ip.coef = mode == FI_SAMPLING_TABLE ? coef.p : nullptr;
    ip.tar = mode == FI_SAMPLING_TABLE ? nullptr : tar_f.p;
    ip.group = group_code(mode);
    ip.w = ref_raw.view.width;
    ip.h = ref_raw.view.height;
    ip.pois = queue_base;
    ip.n = count;
    ip.prep = prep.p + c.off;
    ip.prep_index = prep_index;
    ip.rx = c.r;
    ip.ry = c.r;
    ip.conv = conv;
    ip.stop = stop;
    ip.iters = iters.p + (size_t)region * n + c.off + first;
    int waves = c.waves, spl = c.spl;
    if (waves == 0 && shape_waves && icgn_shape_covers(shape_waves, shape_spl, c.r, c.r)) {
        waves = shape_waves;
        spl = shape_spl;
    }
    ip.count = count_dev;  // bounded on the device: one block per entry
    const int qb = c.queue_blocks >= 0 ? c.queue_blocks : persistent_blocks;
    if (qb > 0 && !count_dev) {
        int ci = 0;
        while (classes[ci].get() != &c) ci++;
        ip.queue = queue.p + kQueueWords * ci;
        ip.queue_blocks = qb;
        launch_zero_words(reinterpret_cast<uint32_t*>(ip.queue), kQueueWords, stream);
    }
    icgn_launch(ip, waves, spl, stream);
}

// The fewest-waves generic shape whose bands cover every class radius.
bool fi_frame::fused_shape(int& waves, int& spl) const
{
    for (int W : { 1, 2, 4 }) {
        bool all = true;
        for (auto& c : classes) all = all && icgn_shape_covers(W, 16, c->r, c->r);
        if (all) {
            waves = W;
            spl = 16;
            return true;
        }
    }
    return false;
}

// One launch over all classes: the kernel reads each POI's subset_radius
// (self-adaptive); the LDS tile and shape are those of the largest radius.
void fi_frame::icgn_fused(fi_poi* base, int region)
{
    const int mode = fused.sampling >= 0 ? fused.sampling : sampling;
    IcgnParams ip{};
    ip.ref = ref_f.p;
    ip.gx = gx.p;
    ip.gy = gy.p;
    // the table exists only while a class samples it; the LDS-tile mode
    // samples the f32 target when a warp outgrows the tile
    ip.coef = mode == FI_SAMPLING_TABLE ? coef.p : nullptr;
    ip.tar = mode == FI_SAMPLING_TABLE ? nullptr : tar_f.p;
    ip.group = group_code(mode);
    ip.w = ref_raw.view.width;
    ip.h = ref_raw.view.height;
    ip.pois = base;
    ip.n = (int)n;
    ip.prep = prep.p;
    ip.rx = ip.ry = classes.back()->r;
    ip.self_adaptive = 1;
    ip.conv = conv;
    ip.stop = stop;
    ip.iters = iters.p + (size_t)region * n;
    int waves = fused.waves, spl = fused.spl;
    if (waves == 0 && !fused_shape(waves, spl))
        fail(FI_E_UNSUPPORTED, "no IC-GN lane shape covers every radius class");
    const int qb = fused.queue_blocks >= 0 ? fused.queue_blocks : persistent_blocks;
    if (qb > 0) {
        ip.queue = queue.p;
        ip.queue_blocks = qb;
        launch_zero_words(reinterpret_cast<uint32_t*>(ip.queue), kQueueWords, stream);
    }
    icgn_launch(ip, waves, spl, stream);
}

void fi_frame::icgn_all(fi_poi* base, int region)
{
    if (fuse && nclass() > 1) {
        icgn_fused(base, region);
        return;
    }
    for (auto& c : classes) icgn_pass(*c, base + c->off, (int)c->n, nullptr, nullptr, region);
}

void fi_frame::fftcc_all(fi_poi* base)
{
    for (auto& c : classes) fftcc_pass(*c, base + c->off, (int)c->n, nullptr);
}

// POI-level repair of class c's m candidates (cand + off): FFTCC on the cached
// spectra, IC-GN on the resident records, merge (cli:2955-3022)
// POI-level repair of candidates [lo, m) of class c: FFTCC from the template
// on the cached spectra, IC-GN on the resident Hessians, merge (better ZNCC
// wins), redo flags.  With count_dev (lo = 0) the candidate count is read on
// the device and m is only the capacity the launches are sized for, so the
// repair can sit in a captured frame graph.
void fi_frame::repair_class(fi_poi* out, const FrameClass& c, int m, int lo, const unsigned* count_dev,
                            const uint32_t* seeded)
{
    const int k = m - lo;
    if (k <= 0) return;
    int ci = 0;
    while (classes[ci].get() != &c) ci++;
    const uint32_t* idx = cand.p + c.off + lo;
    fi_poi* rep = repair.p + c.off + lo;
    launch_gather(tmpl.p + c.off, idx, k, rep, stream, count_dev, redo.p + c.off);
    fftcc_pass(c, rep, k, idx, count_dev);
    icgn_pass(c, rep, k, idx, count_dev, 1, lo);
    launch_merge_repair(out + c.off, rep, idx, k, counters.p + ci, stream, count_dev,
                        seeded ? seeded + c.off : nullptr);
    FI_HIP(hipGetLastError());
}

// Fallback candidates a learned-seed frame repairs inside its segment (per
// class); more are repaired after it.  Fallbacks run at about 1 % of POIs.
int fi_frame::repair_cap(const FrameClass& c) const
{
    return (int)std::min<size_t>(c.n, std::max<size_t>(1024, c.n / 32));
}

void fi_frame::zero_counters()
{
    // a fill kernel, not hipMemsetAsync: this runs inside captured frame graphs
    launch_zero_words(reinterpret_cast<uint32_t*>(counters.p), (nclass() + 1) * sizeof(FrameCounters) / 4, stream);
}

FrameCounters fi_frame::read_counters(FrameCounters* per_class)
{
    FI_HIP(hipMemcpyAsync(h_counters.p, counters.p, (nclass() + 1) * sizeof(FrameCounters), hipMemcpyDeviceToHost,
                          stream));
    FI_HIP(hipStreamSynchronize(stream));
    if (per_class) std::copy(h_counters.p, h_counters.p + nclass() + 1, per_class);
    return h_counters.p[nclass()];
}

void fi_frame::prep_target(bool table)
{
    const int w = ref_raw.view.width, h = ref_raw.view.height;
    if (table) {
        check_table_size(w, h);
        coef.reserve(bspline_table_floats(w, h));
    }
    tar_f.reserve((size_t)w * h);
}

// The B-spline table is built only when some launch samples it (table mode
// set for the frame, a class or the fused launch).
bool fi_frame::table_needed() const
{
    if (sampling == FI_SAMPLING_TABLE || fused.sampling == FI_SAMPLING_TABLE) return true;
    for (auto& c : classes)
        if (c->sampling == FI_SAMPLING_TABLE) return true;
    return false;
}

// processFrame, opencorr_cli_cuda_mvp.cpp:3024-3210.  The device work of a
// frame (target prefilter, seeding, IC-GN of every radius class, fallback scan,
// the whole-frame fallback decision and, when taken, the from-scratch pass and
// per-POI merge, POI-level repair, QA, counter read-back) is one segment, run
// eagerly or replayed as a captured hipGraph.  Only repair candidates past a
// class's repair_cap run after it, eagerly.
void fi_frame::run(unsigned policy, fi_frame_stats* stats)
{
    DeviceGuard g(device);
    if (!inflight.empty() || !done.empty())
        fail(FI_E_LOGIC, "frames launched by fi_frame_run_async are waiting for fi_frame_wait");
    launch(policy);
    complete_oldest();
    if (stats) *stats = done.front();
    done.pop_front();
}

void fi_frame::run_async(unsigned policy)
{
    DeviceGuard g(device);
    launch(policy);
}

void fi_frame::wait(fi_frame_stats* stats)
{
    DeviceGuard g(device);
    if (done.empty()) {
        if (inflight.empty()) fail(FI_E_LOGIC, "no frame launched by fi_frame_run_async is waiting");
        complete_oldest();
    }
    if (stats) *stats = done.front();
    done.pop_front();
}

void fi_frame::drain()
{
    DeviceGuard g(device);
    while (!inflight.empty()) complete_oldest();
}

// The per-frame buffers and their twins trade places (set 0 <-> set 1).
void fi_frame::swap_scratch()
{
    repair.swap(repair_alt);
    cand.swap(cand_alt);
    seed_ok.swap(seed_ok_alt);
    iters.swap(iters_alt);
    redo.swap(redo_alt);
    counters.swap(counters_alt);
    h_counters.swap(h_counters_alt);
    tar_f.swap(tar_f_alt);
    scratch_set ^= 1;
}

// Launches one frame's device work (processFrame, below) and returns without
// waiting for it.  A frame is launched behind one still running only when
// neither needs the other's completion on the host: both replayed as graphs,
// not a temporal sequence (whose seeds and fallback baseline come from the
// previous frame's completion), no B-spline table (one table buffer) and no
// host-built seed cells pending.  Such a frame runs on the other set of
// per-frame buffers, so the earlier frame's late repairs (candidates past
// repair_cap, run on the host's reading of its counters) still find its own
// candidates, flags, iteration counts and converted target.
void fi_frame::launch(unsigned policy)
{
    if (!prepared) {
        drain();
        prepare();
    }
    if (!tar_ready) fail(FI_E_LOGIC, "frame driver needs an uploaded target");
    const int w = ref_raw.view.width, h = ref_raw.view.height;
    const int cnt = (int)n;
    const int C = (int)nclass();
    fi_frame_stats st{};
    st.frame_index = frame_index;
    st.baseline_bad_ratio = NAN;
    st.fallback_threshold = NAN;
    const bool table = table_needed();
    prep_target(table);
    const bool seeds_learned = (policy & FI_POLICY_LEARNED_SEED) && seeder != nullptr;
    const bool overlap = use_graph && !table && !(policy & FI_POLICY_SKIP_FFTCC) &&
                         !(seeds_learned && consensus && !cells_ready);
    if (overlap && !(repair_alt.cap >= repair.cap && cand_alt.cap >= cand.cap && seed_ok_alt.cap >= seed_ok.cap &&
                     iters_alt.cap >= iters.cap && redo_alt.cap >= redo.cap && counters_alt.cap >= counters.cap &&
                     h_counters_alt.cap >= h_counters.cap && tar_f_alt.cap >= tar_f.cap)) {
        // the twin set sized like the first: with nothing in flight, and the
        // graphs captured on the old blocks dropped
        drain();
        repair_alt.reserve(repair.cap);
        cand_alt.reserve(cand.cap);
        seed_ok_alt.reserve(seed_ok.cap);
        iters_alt.reserve(iters.cap);
        redo_alt.reserve(redo.cap);
        counters_alt.reserve(counters.cap);
        h_counters_alt.reserve(h_counters.cap);
        tar_f_alt.reserve(tar_f.cap);
        invalidate_graphs();
    }
    while (!inflight.empty() && (!overlap || inflight.size() >= 2 || !inflight.front().graph)) complete_oldest();
    // A frame that may overlap takes the other set from the frame launched
    // before it (that frame's set is the members': a completion's exchange is
    // undone), in flight or not: the sets then alternate with the record
    // slots and the staged targets, and the frame graphs' keys repeat.
    if (overlap) swap_scratch();
    const int ts = take_target(true);
    const DevImage tview = target(ts);
    FI_HIP(hipStreamWaitEvent(stream, tslot[ts].ready, 0));

    const bool skip = (policy & FI_POLICY_SKIP_FFTCC) && baseline_ready && have_prev;
    const bool learned = !skip && (policy & FI_POLICY_LEARNED_SEED) && seeder != nullptr;
    const int collect = (policy & FI_POLICY_POI_FALLBACK) ? 1 : 0;
    const int mode = learned ? 0 : (skip ? 2 : 1);
    // the frame before the previous one, for the two-frame extrapolation
    const fi_poi* hist2 = (mode == 2 && temporal && have_prev2) ? prev2.p : nullptr;
    const float max_delta = temporal_max_delta > 0.f ? temporal_max_delta : (float)r;
    if (mode == 2) {
        // the device decides the whole-frame fallback against baseline + margin
        const float thr = baseline_bad + kBadRatioMargin;
        if (!(thr == threshold_host)) {
            threshold_host = thr;
            FI_HIP(hipMemcpyAsync(threshold.p, &threshold_host, sizeof(float), hipMemcpyHostToDevice, stream));
            FI_HIP(hipStreamSynchronize(stream));
        }
    }

    // records: the slot is chosen (and its previous copy waited for) before
    // the segment, packed at the segment's end and copied after it
    const int rs = records_on ? records_slot() : -1;
    if (mode == 0 && consensus && !cells_ready) {
        // seed-consensus cells over the layout's positions, once per layout
        std::vector<float> xs(n), ys(n);
        for (size_t i = 0; i < n; i++) {
            const fi_poi& q = tmpl_host.p[order.empty() ? i : order[i]];
            xs[i] = q.x;
            ys[i] = q.y;
        }
        cells.upload(build_seed_cells(xs.data(), ys.data(), n, (float)r, seed_sample), (float)r, consensus_tau,
                     stream);
        seed_raw.reserve(2 * std::max<size_t>(n, 1));
        cells_ready = true;
    }

    auto segment = [&](bool timed) {
        // a fill kernel, not memsets: this runs inside captured frame graphs
        // FFTCC and learned seeds start from the template: copied here, in the
        // reset launch
        launch_frame_reset(reinterpret_cast<uint32_t*>(counters.p), (C + 1) * sizeof(FrameCounters) / 4, iters.p,
                           3 * n, redo.p, n, mode == 1 ? 1 : 0, stream, mode == 2 ? nullptr : tmpl.p, cur.p,
                           n * sizeof(fi_poi));
        if (timed) FI_HIP(hipEventRecord(ev[0], stream));
        // the B-spline table, when a launch samples it (HBM-write-bound), on a
        // second stream beside the seeding (FFTCC or the seeder read only the
        // f32 images); joined before IC-GN
        if (table) {
            FI_HIP(hipEventRecord(fork_ev, stream));
            FI_HIP(hipStreamWaitEvent(aux_stream, fork_ev, 0));
            launch_bspline(tview, coef.p, aux_stream);
            FI_HIP(hipEventRecord(join_ev, aux_stream));
        }
        launch_to_float(tview, tar_f.p, stream);
        if (timed) FI_HIP(hipEventRecord(ev[1], stream));
        if (mode == 0) {
            // each pass crops the target at the previous pass's rounded u, v; a
            // POI a later pass cannot crop keeps the earlier seed, and only the
            // first pass's flag decides the FFTCC fallback for an unseeded POI
            // with a sample, the network seeds the sampled POIs and the
            // consensus seeds (and flags) the rest
            const bool sampled = consensus && cells.view.sample != nullptr;
            auto network = [&](uint32_t* okp) {
                if (!sampled)
                    seeder->run_seed(ref_f.p, tar_f.p, w, h, cur.p, nullptr, cnt, okp, stream);
                else if (cells.sampled > 0)
                    seeder->run_seed(ref_f.p, tar_f.p, w, h, cur.p, cells.sample_idx.p, (int)cells.sampled, okp,
                                     stream);
            };
            network(seed_ok.p);
            if (consensus) launch_seed_consensus(cur.p, seed_ok.p, seed_raw.p, cnt, cells.view, stream);
            for (int pass = 1; pass < seed_passes; pass++) {
                network(nullptr);
                if (consensus) launch_seed_consensus(cur.p, seed_ok.p, seed_raw.p, cnt, cells.view, stream);
            }
        } else if (mode == 1) {
            fftcc_all(cur.p);
        } else {
            launch_seed_from_previous(tmpl.p, prev.p, hist2, max_delta, cur.p, cnt, counters.p + C, stream);
        }
        if (table) FI_HIP(hipStreamWaitEvent(stream, join_ev, 0));
        if (timed) FI_HIP(hipEventRecord(ev[2], stream));
        icgn_all(cur.p);
        if (timed) FI_HIP(hipEventRecord(ev[3], stream));
        for (int ci = 0; ci < C; ci++) {
            const FrameClass& c = *classes[ci];
            if (mode == 0)
                launch_seed_scan(cur.p + c.off, seed_ok.p + c.off, (int)c.n, seed_threshold, counters.p + ci,
                                 cand.p + c.off, stream);
            else if (mode == 2)
                launch_scan(cur.p + c.off, prev.p + c.off, (int)c.n, collect, counters.p + ci, cand.p + c.off,
                            stream);
        }
        if (mode == 2) {
            // bad ratio against the threshold: sets the fallback flag and the
            // per-class repair and fallback counts the launches below read
            launch_decide(counters.p, C, class_n.p, cnt, threshold.p, collect, stream);
            // the skip attempt, kept for the per-POI merge
            FI_HIP(hipMemcpyAsync(save.p, cur.p, n * sizeof(fi_poi), hipMemcpyDeviceToDevice, stream));
        }
        // POI-level repair (:2955-3022) of the scan's candidates, sized for
        // repair_cap and bounded by a count on the device (candidates past the
        // cap are repaired after the segment): always for learned seeds; for
        // temporal seeds unless the frame falls back
        if (mode == 0 || (mode == 2 && collect)) {
            for (int ci = 0; ci < C; ci++)
                if (classes[ci]->n > 0)
                    repair_class(cur.p, *classes[ci], repair_cap(*classes[ci]), 0,
                                 mode == 0 ? &counters.p[ci].candidates : &counters.p[ci].repair_n,
                                 mode == 0 ? seed_ok.p : nullptr);
        }
        if (mode == 2) {
            // whole-frame fallback with per-POI merge (:3115-3152), gated on the
            // device: the template again, FFTCC on the cached spectra, IC-GN,
            // then shouldTakeFrameFallbackResult against the skip attempt.  A
            // frame that does not fall back launches these with zero entries.
            const unsigned* flag = &counters.p[C].fallback;
            launch_copy_if(tmpl.p, cur.p, n * sizeof(fi_poi), flag, stream);
            for (int ci = 0; ci < C; ci++)
                if (classes[ci]->n > 0)
                    fftcc_pass(*classes[ci], cur.p + classes[ci]->off, (int)classes[ci]->n, nullptr,
                               &counters.p[ci].fallback_n);
            for (int ci = 0; ci < C; ci++)
                if (classes[ci]->n > 0)
                    icgn_pass(*classes[ci], cur.p + classes[ci]->off, (int)classes[ci]->n, nullptr,
                              &counters.p[ci].fallback_n, 2);
            launch_merge_frame_fallback(save.p, cur.p, cnt, stream, flag, redo.p);
        }
        launch_qa(cur.p, cnt, counters.p + C, stream, iters.p, 3 * n);
        FI_HIP(hipMemcpyAsync(h_counters.p, counters.p, (C + 1) * sizeof(FrameCounters), hipMemcpyDeviceToHost,
                              stream));
        if (rs >= 0) pack_records(cur.p, rs);
        FI_HIP(hipGetLastError());
    };

    if (use_graph) {
        const uint64_t key[12] = { (uint64_t)(uintptr_t)cand.p, (uint64_t)mode, (uint64_t)collect, (uint64_t)(uintptr_t)cur.p,
                                   (uint64_t)(uintptr_t)prev.p, (uint64_t)(uintptr_t)tview.data,
                                   (uint64_t)(uintptr_t)(table ? coef.p : nullptr), (uint64_t)(uintptr_t)tar_f.p,
                                   (uint64_t)(uintptr_t)seeder, (uint64_t)(uintptr_t)hist2,
                                   (uint64_t)(uint32_t)__builtin_bit_cast(uint32_t, max_delta),
                                   (uint64_t)(uintptr_t)(rs >= 0 ? d_rec[rs].p : nullptr) };
        hipGraphExec_t exec = nullptr;
        for (auto& gr : graphs)
            if (std::memcmp(gr.key, key, sizeof(key)) == 0) exec = gr.exec;
        if (!exec) {
            // a new capture behind an idle stream
            drain();
            FI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                segment(false);
            } catch (...) {
                hipGraph_t broken = nullptr;
                (void)hipStreamEndCapture(stream, &broken);
                if (broken) (void)hipGraphDestroy(broken);
                throw;
            }
            hipGraph_t graph = nullptr;
            FI_HIP(hipStreamEndCapture(stream, &graph));
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            FI_HIP(e);
            Graph gr;
            std::memcpy(gr.key, key, sizeof(key));
            gr.exec = exec;
            graphs.push_back(gr);
        }
        FI_HIP(hipEventRecord(fev[scratch_set][0], stream));
        FI_HIP(hipGraphLaunch(exec, stream));
        FI_HIP(hipEventRecord(fev[scratch_set][1], stream));
    } else {
        segment(true);
        FI_HIP(hipEventRecord(ev[4], stream));
    }
    if (rs >= 0) copy_records(rs, frame_index);
    // the staging thread refills this target slot only after this frame
    FI_HIP(hipEventRecord(tslot[ts].freed, stream));
    InFlight f;
    f.st = st;
    f.policy = policy;
    f.mode = mode;
    f.rs = rs;
    f.ts = ts;
    f.set = scratch_set;
    f.index = frame_index;
    f.collect = collect != 0;
    f.skip = skip;
    f.table = table;
    f.graph = use_graph;
    f.out = cur.p;
    inflight.push_back(f);
    last_slot = ts;

    // history: prev2 <- prev, prev <- this frame, cur <- the oldest buffer
    std::swap(prev2.p, prev.p);
    std::swap(prev2.cap, prev.cap);
    std::swap(cur.p, prev.p);
    std::swap(cur.cap, prev.cap);
    have_prev2 = have_prev;
    have_prev = true;
    frame_index++;
}

// The host's part of the oldest frame in flight once its device work has
// finished: statistics from the counters it copied back, repairs of the
// candidates past repair_cap (with the frame's buffer set as the members),
// QA and records again after them.
void fi_frame::complete_oldest()
{
    const InFlight f = inflight.front();
    inflight.pop_front();
    const bool other = f.set != scratch_set;
    if (other) swap_scratch();
    try {
        finish(f);
    } catch (...) {
        if (other) swap_scratch();
        throw;
    }
    if (other) swap_scratch();
}

void fi_frame::finish(const InFlight& f)
{
    FI_HIP(hipEventSynchronize(f.graph ? fev[f.set][1] : ev[4]));
    const int cnt = (int)n;
    const int C = (int)nclass();
    const int mode = f.mode, rs = f.rs, collect = f.collect ? 1 : 0;
    const bool skip = f.skip, table = f.table;
    const unsigned policy = f.policy;
    fi_frame_stats st = f.st;
    std::vector<FrameCounters> pc(h_counters.p, h_counters.p + C + 1);
    FrameCounters qa = pc[C];
    bool redo_qa = false;

    if (mode == 0) {
        unsigned long long m_all = 0;
        for (int ci = 0; ci < C; ci++) m_all += pc[ci].candidates;
        st.seed_pois = cnt;
        st.seed_fallback_count = (int64_t)m_all;
        st.fftcc_pois = (int64_t)m_all;
        for (int ci = 0; ci < C; ci++) {
            const int m = (int)pc[ci].candidates, cap = repair_cap(*classes[ci]);
            if (m > cap) {
                repair_class(f.out, *classes[ci], m, cap, nullptr, seed_ok.p);
                redo_qa = true;
            }
        }
    } else if (mode == 1) {
        st.fftcc_pois = cnt;
    } else {
        unsigned long long bad = 0, m_all = 0;
        for (int ci = 0; ci < C; ci++) {
            bad += pc[ci].bad;
            m_all += pc[ci].candidates;
        }
        const float ratio = cnt ? (float)bad / (float)cnt : 0.f;
        st.skip_fftcc_used = 1;
        st.pre_repair_bad_ratio = ratio;
        st.temporal_seed_count = (int32_t)pc[C].temporal;
        st.temporal_seed_used = pc[C].temporal > 0 ? 1 : 0;
        const bool fell_back = pc[C].fallback != 0;
        if (fell_back != (ratio > threshold_host))
            fail(FI_E_DEVICE, "whole-frame fallback decision differs between device and host");
        if (fell_back) {
            st.fallback_used = 1;
            st.fftcc_pois = cnt;
        } else if (collect && m_all > 0) {
            // the segment repaired up to repair_cap candidates per class, the
            // rest here
            std::vector<FrameCounters> rc(pc);
            for (int ci = 0; ci < C; ci++) {
                const int m = (int)pc[ci].candidates, cap = repair_cap(*classes[ci]);
                if (m > cap) {
                    repair_class(f.out, *classes[ci], m, cap);
                    redo_qa = true;
                }
            }
            if (redo_qa) read_counters(rc.data());
            unsigned long long rec = 0;
            for (int ci = 0; ci < C; ci++) rec += rc[ci].recovered;
            st.poi_fallback_count = (int32_t)m_all;
            st.poi_fallback_recovered = (int32_t)rec;
            st.fftcc_pois = (int64_t)m_all;
        }
    }
    if (redo_qa) {
        launch_zero_words(reinterpret_cast<uint32_t*>(counters.p + C), sizeof(FrameCounters) / 4, stream);
        launch_qa(f.out, cnt, counters.p + C, stream, iters.p, 3 * n);
        FI_HIP(hipGetLastError());
        qa = read_counters(nullptr);
        if (rs >= 0) {
            // repairs past the cap changed the frame: its records again, same slot
            FI_HIP(hipStreamWaitEvent(stream, rec_ev[rs], 0));
            pack_records(f.out, rs);
            copy_records(rs, f.index);
        }
    }

    const float bad_ratio = cnt ? (float)qa.bad / (float)cnt : 0.f;
    if (!skip) st.pre_repair_bad_ratio = bad_ratio;
    if ((policy & FI_POLICY_SKIP_FFTCC) && !baseline_ready && !skip) {
        baseline_bad = bad_ratio;
        baseline_ready = true;
    }
    if (baseline_ready) {
        st.baseline_bad_ratio = baseline_bad;
        st.fallback_threshold = baseline_bad + kBadRatioMargin;
    }
    st.bad_ratio = bad_ratio;
    st.valid_count = (int64_t)qa.valid;
    st.valid_ratio = cnt ? (float)qa.valid / (float)cnt : 0.f;
    st.iteration_sum = (int64_t)qa.iterations;
    st.iteration_mean = qa.iter_finite ? (float)((double)qa.iter_sum / (double)qa.iter_finite) : 0.f;
    st.iteration_max = (int32_t)qa.iter_max;
    float ms = 0.f;
    FI_HIP(hipEventElapsedTime(&ms, f.graph ? fev[f.set][0] : ev[0], f.graph ? fev[f.set][1] : ev[4]));
    st.device_ms = ms;
    if (f.graph) {
        // stages inside a graph are not timed separately
        st.bspline_ms = st.fftcc_ms = st.icgn_ms = NAN;
        st.seed_ms = mode == 0 ? NAN : 0.0;
    } else {
        ms = 0.f;
        if (table) FI_HIP(hipEventElapsedTime(&ms, ev[0], join_ev));  // on its stream, beside the seeding
        st.bspline_ms = ms;
        FI_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
        if (mode == 0)
            st.seed_ms = ms;
        else
            st.fftcc_ms = mode == 1 ? ms : 0.0;
        FI_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
        st.icgn_ms = ms;
    }
    done.push_back(st);
}

// Packs the raw records of src (the frame just run, layout order) on the
// compute stream and copies them to the next pinned slot on rec_stream, which
// overlaps the next frame.  A slot is packed again only after its previous
// copy finished (rec_ev).
void fi_frame::queue_records(const fi_poi* src)
{
    const int s = records_slot();
    pack_records(src, s);
    copy_records(s, frame_index);
}

// The next record slot, its buffers sized, the compute stream waiting for the
// slot's previous copy (outside any capture: the wait is on rec_stream work).
int fi_frame::records_slot()
{
    const int s = rec_next;
    rec_next ^= 1;
    const size_t bytes = n * (size_t)36;
    d_rec[s].reserve(std::max<size_t>(bytes, 36));
    h_rec[s].reserve(std::max<size_t>(bytes, 36));
    FI_HIP(hipStreamWaitEvent(stream, rec_ev[s], 0));
    return s;
}

// the pack kernel, on the compute stream (also inside a captured frame)
void fi_frame::pack_records(const fi_poi* src, int s)
{
    launch_pack_raw(src, order.empty() ? nullptr : d_order.p, (int)n, d_rec[s].p, stream);
}

// the copy of slot s to host memory on rec_stream, after the work queued on
// the compute stream so far
void fi_frame::copy_records(int s, int index)
{
    const size_t bytes = n * (size_t)36;
    FI_HIP(hipGetLastError());
    FI_HIP(hipEventRecord(pack_ev, stream));
    FI_HIP(hipStreamWaitEvent(rec_stream, pack_ev, 0));
    if (bytes) FI_HIP(hipMemcpyAsync(h_rec[s].p, d_rec[s].p, bytes, hipMemcpyDeviceToHost, rec_stream));
    FI_HIP(hipEventRecord(rec_ev[s], rec_stream));
    rec_held[s] = index;
}

void fi_frame::set_records(bool on)
{
    records_on = on;
}

void fi_frame::records(int age, const void** out, size_t* count)
{
    if (!have_prev) fail(FI_E_LOGIC, "no frame has been run");
    if (out == nullptr) fail(FI_E_INVALID_ARGUMENT, "output pointer is null");
    if (age < 0 || age > 1) fail(FI_E_INVALID_ARGUMENT, "records age must be 0 (last frame) or 1");
    DeviceGuard g(device);
    const int want = frame_index - 1 - age;
    int s = rec_held[0] == want ? 0 : (rec_held[1] == want ? 1 : -1);
    if (s < 0) {
        if (age != 0 || want < 0)
            fail(FI_E_LOGIC, "the records of that frame are not held (fi_frame_set_records keeps two frames)");
        // not packed by its run: pack the latest results (prev after the
        // rotation) once every frame in flight has completed
        drain();
        const int keep = frame_index;
        frame_index = want;
        try {
            queue_records(prev.p);
        } catch (...) {
            frame_index = keep;
            throw;
        }
        frame_index = keep;
        s = rec_held[0] == want ? 0 : 1;
    }
    FI_HIP(hipEventSynchronize(rec_ev[s]));
    *out = h_rec[s].p;
    if (count) *count = n;
}

// Autotuner: per radius class, every covering lane shape x sampling mode x
// launch mode (one block per POI / persistent work queue), timed on the
// FFTCC-seeded POIs of the uploaded target (median of three runs); the fastest
// is kept.  The sequence state (previous frame) is not touched.
int fi_frame::autotune(fi_tune_entry* report, size_t report_len, size_t* report_n)
{
    DeviceGuard g(device);
    if (!prepared) prepare();
    if (!tar_ready) fail(FI_E_LOGIC, "autotune needs an uploaded target");
    prep_target(true);  // table-mode candidates are timed too
    // the queued target is taken as a frame takes it (a run without a new
    // upload processes it again)
    const int ts = take_target(true);
    FI_HIP(hipStreamWaitEvent(stream, tslot[ts].ready, 0));
    launch_bspline(target(ts), coef.p, stream);
    launch_to_float(target(ts), tar_f.p, stream);
    last_slot = ts;
    // FFTCC seeds of the template in `save`
    FI_HIP(hipMemcpyAsync(save.p, tmpl.p, n * sizeof(fi_poi), hipMemcpyDeviceToDevice, stream));
    fftcc_all(save.p);
    FI_HIP(hipGetLastError());
    FI_HIP(hipStreamSynchronize(stream));
    int cus = 0;
    FI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    // Resident-block counts tried for the queue launch: 0 (one block per POI),
    // 12 waves per CU, and as many blocks as the kernel's registers and LDS
    // let every CU hold.
    auto queue_counts = [&](int r, int waves, int spl, int samp) {
        IcgnParams op{};
        op.tar = samp == FI_SAMPLING_TABLE ? nullptr : tar_f.p;
        op.w = ref_raw.view.width;
        op.h = ref_raw.view.height;
        op.rx = op.ry = r;
        std::vector<int> q = { 0, cus * 12 / waves };
        const int fill = cus * icgn_queue_per_cu(op, waves, spl);
        if (fill > 0 && fill != q[1]) q.push_back(fill);
        return q;
    };
    size_t m = 0;
    int timed = 0;
    for (auto& cp : classes) {
        FrameClass& c = *cp;
        if (c.n == 0) continue;
        int ws[64], ss[64];
        const int ns = icgn_shape_candidates(c.r, c.r, ws, ss, 64);
        double best = 1e30;
        int bw = 0, bs = 0, bsamp = -1, bq = 0;
        size_t best_entry = SIZE_MAX;
        for (int si = 0; si < ns; si++) {
            for (int samp : { FI_SAMPLING_AUTO, FI_SAMPLING_TABLE, FI_SAMPLING_GROUP, FI_SAMPLING_PAIR }) {
                // shared coefficient tiles: one-wave shapes, one block per
                // four entries (no queue); two entries per wave: once per
                // class (its layout follows from the radius), where built
                if (samp == FI_SAMPLING_GROUP && ws[si] != 1) continue;
                if (samp == FI_SAMPLING_PAIR && (si != 0 || !icgn_pair_supported(c.r, c.r))) continue;
                const std::vector<int> qbs = samp == FI_SAMPLING_GROUP || samp == FI_SAMPLING_PAIR
                                                 ? std::vector<int>{ 0 }
                                                 : queue_counts(c.r, ws[si], ss[si], samp);
                for (int qb : qbs) {
                    c.waves = ws[si];
                    c.spl = ss[si];
                    c.sampling = samp;
                    c.queue_blocks = qb;
                    float t[3];
                    for (int rep = 0; rep < 3; rep++) {
                        FI_HIP(hipMemcpyAsync(repair.p + c.off, save.p + c.off, c.n * sizeof(fi_poi),
                                              hipMemcpyDeviceToDevice, stream));
                        FI_HIP(hipEventRecord(ev[5], stream));
                        icgn_pass(c, repair.p + c.off, (int)c.n, nullptr);
                        FI_HIP(hipEventRecord(ev[6], stream));
                        FI_HIP(hipEventSynchronize(ev[6]));
                        FI_HIP(hipEventElapsedTime(&t[rep], ev[5], ev[6]));
                    }
                    std::sort(t, t + 3);
                    timed++;
                    if (report && m < report_len) {
                        fi_tune_entry& e = report[m];
                        e.radius = c.r;
                        e.waves = ws[si];
                        e.rows_per_band = ss[si];
                        e.sampling = samp;
                        e.persistent = qb;
                        e.chosen = 0;
                        e.ms = t[1];
                    }
                    if (t[1] < best) {
                        best = t[1];
                        bw = ws[si];
                        bs = ss[si];
                        bsamp = samp;
                        bq = qb;
                        best_entry = m;
                    }
                    m++;
                }
            }
        }
        c.waves = bw;
        c.spl = bs;
        c.sampling = bsamp;
        c.queue_blocks = bq;
        if (report && best_entry < report_len) report[best_entry].chosen = 1;
    }
    // one launch over every class against the per-class launches
    fuse = false;
    if (nclass() > 1) {
        double per_class = 0.0;
        {
            float t[3];
            for (int rep = 0; rep < 3; rep++) {
                FI_HIP(hipMemcpyAsync(repair.p, save.p, n * sizeof(fi_poi), hipMemcpyDeviceToDevice, stream));
                FI_HIP(hipEventRecord(ev[5], stream));
                for (auto& c : classes) icgn_pass(*c, repair.p + c->off, (int)c->n, nullptr);
                FI_HIP(hipEventRecord(ev[6], stream));
                FI_HIP(hipEventSynchronize(ev[6]));
                FI_HIP(hipEventElapsedTime(&t[rep], ev[5], ev[6]));
            }
            std::sort(t, t + 3);
            per_class = t[1];
        }
        double best = per_class;
        size_t best_entry = SIZE_MAX;
        FrameClass keep;
        for (int W : { 1, 2, 4 }) {
            bool all = true;
            for (auto& c : classes) all = all && icgn_shape_covers(W, 16, c->r, c->r);
            if (!all) continue;
            for (int samp : { FI_SAMPLING_AUTO, FI_SAMPLING_TABLE }) {
                for (int qb : queue_counts(classes.back()->r, W, 16, samp)) {
                    fused.waves = W;
                    fused.spl = 16;
                    fused.sampling = samp;
                    fused.queue_blocks = qb;
                    float t[3];
                    for (int rep = 0; rep < 3; rep++) {
                        FI_HIP(hipMemcpyAsync(repair.p, save.p, n * sizeof(fi_poi), hipMemcpyDeviceToDevice,
                                              stream));
                        FI_HIP(hipEventRecord(ev[5], stream));
                        icgn_fused(repair.p);
                        FI_HIP(hipEventRecord(ev[6], stream));
                        FI_HIP(hipEventSynchronize(ev[6]));
                        FI_HIP(hipEventElapsedTime(&t[rep], ev[5], ev[6]));
                    }
                    std::sort(t, t + 3);
                    timed++;
                    if (report && m < report_len) {
                        fi_tune_entry& e = report[m];
                        e.radius = 0;  // every class in one launch
                        e.waves = W;
                        e.rows_per_band = 16;
                        e.sampling = samp;
                        e.persistent = qb;
                        e.chosen = 0;
                        e.ms = t[1];
                    }
                    if (t[1] < best) {
                        best = t[1];
                        best_entry = m;
                        keep.waves = W;
                        keep.spl = 16;
                        keep.sampling = samp;
                        keep.queue_blocks = qb;
                    }
                    m++;
                }
            }
        }
        if (best_entry != SIZE_MAX) {
            fuse = true;
            fused.waves = keep.waves;
            fused.spl = keep.spl;
            fused.sampling = keep.sampling;
            fused.queue_blocks = keep.queue_blocks;
            // the per-class choices are no longer used
            if (report)
                for (size_t i = 0; i < std::min(m, report_len); i++)
                    if (report[i].radius != 0) report[i].chosen = 0;
            if (report && best_entry < report_len) report[best_entry].chosen = 1;
        } else {
            fused.waves = fused.spl = 0;
            fused.sampling = -1;
            fused.queue_blocks = -1;
        }
    }
    FI_HIP(hipStreamSynchronize(stream));
    invalidate_graphs();
    if (report_n) *report_n = std::min(m, report_len);
    return timed;
}

void fi_frame::download_redo(uint8_t* out, size_t count)
{
    if (!have_prev) fail(FI_E_LOGIC, "no frame has been run");
    if (count != n) fail(FI_E_INVALID_ARGUMENT, "download length must equal the template length");
    if (count == 0) return;
    if (out == nullptr) fail(FI_E_INVALID_ARGUMENT, "output pointer is null");
    DeviceGuard g(device);
    std::vector<uint8_t> h(n);
    FI_HIP(hipMemcpyAsync(h.data(), redo.p, n, hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    if (order.empty())
        std::copy(h.begin(), h.end(), out);
    else
        for (size_t i = 0; i < n; i++) out[order[i]] = h[i];
}

void fi_frame::download(fi_poi* out, size_t count)
{
    if (!have_prev) fail(FI_E_LOGIC, "no frame has been run");
    if (count != n) fail(FI_E_INVALID_ARGUMENT, "download length must equal the template length");
    if (count == 0) return;
    if (out == nullptr) fail(FI_E_INVALID_ARGUMENT, "output pointer is null");
    DeviceGuard g(device);
    if (order.empty()) {
        FI_HIP(hipMemcpyAsync(out, prev.p, n * sizeof(fi_poi), hipMemcpyDeviceToHost, stream));
        FI_HIP(hipStreamSynchronize(stream));
        return;
    }
    h_out.reserve(n);
    FI_HIP(hipMemcpyAsync(h_out.p, prev.p, n * sizeof(fi_poi), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; i++) out[order[i]] = h_out.p[i];
}
