// extern "C" boundary (include/fasticgn.h): every entry point catches
// everything, records the message for fi_last_error and returns a status.
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fi_engines.h"
#include "fi_seeder_engine.h"

namespace fi {
const std::string& last_error();
}

using namespace fi;

namespace {

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return FI_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return FI_E_NO_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return FI_E_DEVICE;
    } catch (...) {
        set_last_error("unknown failure");
        return FI_E_DEVICE;
    }
}

template <typename H>
void need(H* h)
{
    if (h == nullptr) fail(FI_E_INVALID_ARGUMENT, "null handle");
}

}  // namespace

extern "C" {

int fi_version(int* major, int* minor, int* patch)
{
    return guarded([&] {
        if (major) *major = 0;
        if (minor) *minor = 1;
        if (patch) *patch = 0;
    });
}

int fi_device_count(int* count)
{
    return guarded([&] {
        if (!count) fail(FI_E_INVALID_ARGUMENT, "null output");
        int c = 0;
        FI_HIP(hipGetDeviceCount(&c));
        *count = c;
    });
}

int fi_device_synchronize(int device)
{
    return guarded([&] {
        DeviceGuard g(device);
        FI_HIP(hipDeviceSynchronize());
    });
}

int fi_last_error(char* buf, size_t len)
{
    const std::string& m = last_error();
    if (buf && len > 0) {
        const size_t k = m.size() < len - 1 ? m.size() : len - 1;
        std::memcpy(buf, m.data(), k);
        buf[k] = '\0';
    }
    return (int)m.size();
}

int fi_image_resolve(const void* px, int width, int height, size_t stride_bytes, int pixel, int mode,
                     fi_ingest_info* out)
{
    return guarded([&] {
        check_image_args(px, width, height, stride_bytes, pixel, mode);
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        unsigned mx = 0;
        if (pixel == FI_PIXEL_U16)
            for (int y = 0; y < height; y++) {
                const uint16_t* row = reinterpret_cast<const uint16_t*>(static_cast<const uint8_t*>(px) +
                                                                       (size_t)y * stride_bytes);
                for (int x = 0; x < width; x++) mx = row[x] > mx ? row[x] : mx;
            }
        *out = resolve_ingest(pixel, mode, mx);
    });
}

int fi_dft_twiddles(int n, float* out)
{
    return guarded([&] {
        if (n < 1 || !out) fail(FI_E_INVALID_ARGUMENT, "bad twiddle request");
        dft_twiddles(n, out);
    });
}

int fi_conv_threshold2(float conv, double* out)
{
    return guarded([&] {
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = conv_threshold2(conv);
    });
}

// ---- FFTCC2D ----
int fi_fftcc_create(int rx, int ry, int device, fi_fftcc** out)
{
    return guarded([&] {
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = nullptr;
        *out = new fi_fftcc(rx, ry, device);
    });
}

int fi_fftcc_destroy(fi_fftcc* h)
{
    return guarded([&] { delete h; });
}

int fi_fftcc_set_ref(fi_fftcc* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h); h->set_ref(px, w, hh, stride, pixel, mode); });
}

int fi_fftcc_set_tar(fi_fftcc* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h); h->set_tar(px, w, hh, stride, pixel, mode); });
}

int fi_fftcc_prepare_ref_clear(fi_fftcc* h)
{
    return guarded([&] { need(h); h->prepare_ref_clear(); });
}

int fi_fftcc_prepare_ref(fi_fftcc* h, const fi_poi* layout, size_t n)
{
    return guarded([&] { need(h); h->prepare_ref(layout, n); });
}

int fi_fftcc_has_prepared_ref(fi_fftcc* h, int* out)
{
    return guarded([&] {
        need(h);
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = (h->ref.set && h->ready) ? 1 : 0;
    });
}

int fi_fftcc_compute(fi_fftcc* h, fi_poi* pois, size_t n)
{
    return guarded([&] { need(h); h->compute(pois, n); });
}

int fi_fftcc_compute_indexed(fi_fftcc* h, fi_poi* pois, size_t n, const size_t* idx)
{
    return guarded([&] { need(h); h->compute_indexed(pois, n, idx); });
}

// ---- ICGN2D1 ----
int fi_icgn_create(int rx, int ry, float conv, float stop, int device, fi_icgn** out)
{
    return guarded([&] {
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = nullptr;
        *out = new fi_icgn(rx, ry, conv, stop, device);
    });
}

int fi_icgn_destroy(fi_icgn* h)
{
    return guarded([&] { delete h; });
}

int fi_icgn_set_iteration(fi_icgn* h, float conv, float stop)
{
    return guarded([&] {
        need(h);
        h->conv = conv;
        h->stop = stop;
    });
}

int fi_icgn_set_subset(fi_icgn* h, int rx, int ry)
{
    return guarded([&] {
        need(h);
        if (rx < 1 || ry < 1) fail(FI_E_INVALID_ARGUMENT, "ICGN2D1 subset radii must be positive");
        if (rx > kIcgnMaxRadius || ry > kIcgnMaxRadius)
            fail(FI_E_UNSUPPORTED, "ICGN on the device supports radii 1.." + std::to_string(kIcgnMaxRadius));
        h->rx = rx;
        h->ry = ry;
    });
}

int fi_icgn_set_self_adaptive(fi_icgn* h, int on)
{
    return guarded([&] {
        need(h);
        h->self_adaptive = on != 0;
    });
}

int fi_icgn_set_shape(fi_icgn* h, int waves, int spl)
{
    return guarded([&] {
        need(h);
        check_shape(waves, spl);
        h->shape_waves = waves;
        h->shape_spl = spl;
    });
}

static void check_sampling(int mode)
{
    if (mode != FI_SAMPLING_AUTO && mode != FI_SAMPLING_TABLE && mode != FI_SAMPLING_GROUP && mode != FI_SAMPLING_PAIR)
        fail(FI_E_INVALID_ARGUMENT, "unknown IC-GN sampling mode " + std::to_string(mode));
}

int fi_icgn_set_sampling(fi_icgn* h, int mode)
{
    return guarded([&] {
        need(h);
        check_sampling(mode);
        h->sampling = mode;
    });
}

int fi_icgn_set_ref(fi_icgn* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h); h->set_ref(px, w, hh, stride, pixel, mode); });
}

int fi_icgn_set_tar(fi_icgn* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h); h->set_tar(px, w, hh, stride, pixel, mode); });
}

int fi_icgn_prepare_ref(fi_icgn* h)
{
    return guarded([&] { need(h); h->prepare_ref(); });
}

int fi_icgn_prepare_tar(fi_icgn* h)
{
    return guarded([&] { need(h); h->prepare_tar(); });
}

int fi_icgn_prepare(fi_icgn* h)
{
    return guarded([&] {
        need(h);
        h->prepare_ref();
        h->prepare_tar();
    });
}

int fi_icgn_compute(fi_icgn* h, fi_poi* pois, size_t n)
{
    return guarded([&] { need(h); h->compute(pois, n); });
}

int fi_icgn_compute_offset(fi_icgn* h, fi_poi* pois, size_t n, const float* center_offsets)
{
    return guarded([&] {
        need(h);
        if (n > 0 && !center_offsets) fail(FI_E_INVALID_ARGUMENT, "center offsets are null");
        h->compute(pois, n, center_offsets);
    });
}

int fi_icgn_compute_noalloc(fi_icgn* h, fi_poi* device_pois, size_t n)
{
    return guarded([&] { need(h); h->compute_noalloc(device_pois, n); });
}

int fi_icgn_export(fi_icgn* h, int what, void* out, size_t bytes)
{
    return guarded([&] { need(h); h->export_data(what, out, bytes); });
}

// ---- frame driver ----
int fi_frame_create(int radius, float conv, float stop, int device, fi_frame** out)
{
    return guarded([&] {
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = nullptr;
        *out = new fi_frame(radius, conv, stop, device);
    });
}

int fi_frame_destroy(fi_frame* h)
{
    return guarded([&] { delete h; });
}

int fi_frame_set_ref(fi_frame* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h);
        h->drain(); h->set_ref(px, w, hh, stride, pixel, mode); });
}

int fi_frame_set_pois(fi_frame* h, const fi_poi* pois, size_t n)
{
    return guarded([&] { need(h);
        h->drain(); h->set_pois(pois, n); });
}

int fi_frame_upload_target(fi_frame* h, const void* px, int w, int hh, size_t stride, int pixel, int mode)
{
    return guarded([&] { need(h); h->upload_target(px, w, hh, stride, pixel, mode); });
}

int fi_frame_run(fi_frame* h, unsigned policy, fi_frame_stats* stats)
{
    return guarded([&] { need(h); h->run(policy, stats); });
}

int fi_frame_run_async(fi_frame* h, unsigned policy)
{
    return guarded([&] { need(h); h->run_async(policy); });
}

int fi_frame_wait(fi_frame* h, fi_frame_stats* stats)
{
    return guarded([&] { need(h); h->wait(stats); });
}

int fi_frame_download(fi_frame* h, fi_poi* out, size_t n)
{
    return guarded([&] { need(h);
        h->drain(); h->download(out, n); });
}

int fi_frame_download_redo(fi_frame* h, uint8_t* out, size_t n)
{
    return guarded([&] { need(h);
        h->drain(); h->download_redo(out, n); });
}

int fi_frame_records(fi_frame* h, int age, const void** records, size_t* n_records)
{
    return guarded([&] { need(h); h->records(age, records, n_records); });
}

int fi_frame_set_records(fi_frame* h, int on)
{
    return guarded([&] { need(h);
        h->drain(); h->set_records(on != 0); });
}

int fi_frame_set_temporal_extrapolation(fi_frame* h, int on, float max_delta)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (!(max_delta >= 0.f)) fail(FI_E_INVALID_ARGUMENT, "temporal max_delta must be >= 0 (0 = the radius)");
        h->temporal = on != 0;
        h->temporal_max_delta = max_delta;
        h->invalidate_graphs();
    });
}

int fi_frame_prepare(fi_frame* h, fi_prepare_stats* stats)
{
    return guarded([&] {
        need(h);
        h->drain();
        const fi_prepare_stats ps = h->prepare_now();
        if (stats) *stats = ps;
    });
}

int fi_frame_reset_sequence(fi_frame* h)
{
    return guarded([&] { need(h);
        h->drain(); h->reset_sequence(); });
}

int fi_frame_set_shape(fi_frame* h, int waves, int spl)
{
    return guarded([&] {
        need(h);
        h->drain();
        check_shape(waves, spl);
        if (waves && !icgn_shape_covers(waves, spl, h->r, h->r))
            fail(FI_E_INVALID_ARGUMENT, "IC-GN shape " + std::to_string(waves) + "x" + std::to_string(spl) +
                                            " does not fit the driver's subset radius " + std::to_string(h->r));
        h->shape_waves = waves;
        h->shape_spl = spl;
        for (auto& c : h->classes) {  // an explicit shape replaces tuned ones
            c->waves = c->spl = 0;
            c->queue_blocks = -1;
        }
        h->invalidate_graphs();
    });
}

int fi_frame_set_sampling(fi_frame* h, int mode)
{
    return guarded([&] {
        need(h);
        h->drain();
        check_sampling(mode);
        h->sampling = mode;
        for (auto& c : h->classes) c->sampling = -1;
        h->invalidate_graphs();
    });
}

int fi_frame_set_self_adaptive(fi_frame* h, int on)
{
    return guarded([&] {
        need(h);
        h->drain();
        h->set_self_adaptive(on != 0);
    });
}

int fi_frame_set_persistent(fi_frame* h, int blocks)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (blocks < 0) fail(FI_E_INVALID_ARGUMENT, "persistent block count must be >= 0");
        h->persistent_blocks = blocks;
        for (auto& c : h->classes) c->queue_blocks = -1;
        h->invalidate_graphs();
    });
}

int fi_frame_set_fused(fi_frame* h, int on)
{
    return guarded([&] {
        need(h);
        h->drain();
        h->fuse = on != 0;
        h->fused.waves = h->fused.spl = 0;
        h->fused.sampling = -1;
        h->fused.queue_blocks = -1;
        h->invalidate_graphs();
    });
}

int fi_frame_set_seed_passes(fi_frame* h, int passes)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (passes < 1 || passes > 4) fail(FI_E_INVALID_ARGUMENT, "seeder passes must be 1..4");
        h->seed_passes = passes;
        h->invalidate_graphs();
    });
}

static void check_seed_sample(int sample)
{
    if (sample != FI_SEED_SAMPLE_ALL && sample != FI_SEED_SAMPLE_CHECKER && sample != FI_SEED_SAMPLE_QUARTER)
        fail(FI_E_INVALID_ARGUMENT, "unknown seed sample " + std::to_string(sample));
}

int fi_frame_set_seed_consensus(fi_frame* h, int on, float tau, int sample)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (!(tau > 0.f)) fail(FI_E_INVALID_ARGUMENT, "seed consensus tau must be > 0");
        check_seed_sample(sample);
        h->consensus = on != 0;
        h->consensus_tau = tau;
        h->cells.view.tau = tau;
        if (sample != h->seed_sample) h->cells_ready = false;
        h->seed_sample = sample;
        h->invalidate_graphs();
    });
}

int fi_frame_set_graph(fi_frame* h, int on)
{
    return guarded([&] {
        need(h);
        h->drain();
        h->use_graph = on != 0;
        h->invalidate_graphs();
    });
}

int fi_frame_autotune(fi_frame* h, fi_tune_entry* report, size_t report_len, size_t* report_n)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (report == nullptr && report_len > 0) fail(FI_E_INVALID_ARGUMENT, "report pointer is null");
        size_t m = 0;
        h->autotune(report, report_len, &m);
        if (report_n) *report_n = m;
    });
}

int fi_frame_set_seeder(fi_frame* h, fi_seeder* seeder, float zncc_threshold)
{
    return guarded([&] {
        need(h);
        h->drain();
        if (!(zncc_threshold <= 1.f)) fail(FI_E_INVALID_ARGUMENT, "seed ZNCC threshold must be <= 1");
        if (seeder && seeder->device != h->device)
            fail(FI_E_INVALID_ARGUMENT, "seeder and frame driver are on different devices");
        h->seeder = seeder;
        h->seed_threshold = zncc_threshold;
        h->invalidate_graphs();
    });
}

// ---- learned seeder ----
int fi_seeder_create(int device, fi_seeder** out)
{
    return guarded([&] {
        if (!out) fail(FI_E_INVALID_ARGUMENT, "null output");
        *out = nullptr;
        *out = new fi_seeder(device);
    });
}

int fi_seeder_destroy(fi_seeder* h)
{
    return guarded([&] { delete h; });
}

int fi_seeder_param_count(size_t* n)
{
    return guarded([&] {
        if (!n) fail(FI_E_INVALID_ARGUMENT, "null output");
        *n = seed_param_count();
    });
}

int fi_seeder_set_params(fi_seeder* h, const float* params, size_t n)
{
    return guarded([&] {
        need(h);
        if (!params) fail(FI_E_INVALID_ARGUMENT, "null parameters");
        h->set_params(params, n);
    });
}

int fi_seeder_get_params(fi_seeder* h, float* params, size_t n)
{
    return guarded([&] {
        need(h);
        if (!params) fail(FI_E_INVALID_ARGUMENT, "null output");
        h->get_params(params, n);
    });
}

int fi_seeder_set_training_data(fi_seeder* h, const uint16_t* x1, const float* targets, size_t n)
{
    return guarded([&] { need(h); h->set_training_data(x1, targets, n); });
}

int fi_seeder_train_epoch(fi_seeder* h, int batch, float lr, uint64_t seed, float* mean_loss)
{
    return guarded([&] { need(h); h->train_epoch(batch, lr, seed, mean_loss); });
}

int fi_seeder_gradient(fi_seeder* h, const uint16_t* x1, const float* targets, size_t n, float* grads,
                       float* loss)
{
    return guarded([&] { need(h); h->gradient(x1, targets, n, grads, loss); });
}

int fi_seeder_forward(fi_seeder* h, const uint16_t* x1, size_t n, float* out)
{
    return guarded([&] { need(h); h->forward(x1, n, out); });
}

int fi_seeder_seed(fi_seeder* h, const float* ref, const float* tar, int width, int height, fi_poi* pois, size_t n,
                   uint32_t* ok)
{
    return guarded([&] { need(h); h->seed(ref, tar, width, height, pois, n, ok); });
}

int fi_seeder_seed_consensus(fi_seeder* h, const float* ref, const float* tar, int width, int height,
                             fi_poi* pois, size_t n, uint32_t* ok, float radius, float tau, int passes,
                             int sample)
{
    return guarded([&] {
        need(h);
        if (!(radius > 0.f) || !(tau > 0.f)) fail(FI_E_INVALID_ARGUMENT, "seed consensus radius and tau must be > 0");
        if (passes < 1 || passes > 4) fail(FI_E_INVALID_ARGUMENT, "seeder passes must be 1..4");
        check_seed_sample(sample);
        h->seed(ref, tar, width, height, pois, n, ok, radius, tau, passes, sample);
    });
}

// ---- raw record, opencorr_cli_cuda_mvp.cpp:2498-2536 ----
int fi_format_raw(const fi_poi* pois, size_t n, void* out, size_t out_len, size_t* written)
{
    return guarded([&] {
        if (n > 0 && (!pois || !out)) fail(FI_E_INVALID_ARGUMENT, "null buffer");
        if (out_len < n * 36) fail(FI_E_INVALID_ARGUMENT, "output buffer shorter than n * 36 bytes");
        unsigned char* c = static_cast<unsigned char*>(out);
        // (int32_t) of a coordinate, defined for every float: truncation, and
        // INT32_MIN for NaN and values outside the int32 range (what the x86
        // conversion instruction returns; the device packer does the same)
        auto coord = [](float v) -> int32_t {
            return (v >= -2147483648.f && v < 2147483648.f) ? (int32_t)v : INT32_MIN;
        };
        for (size_t i = 0; i < n; i++) {
            const fi_poi& p = pois[i];
            const int32_t x = coord(p.x), y = coord(p.y);
            const float f[7] = { p.deformation[0], p.deformation[6], p.deformation[1], p.deformation[2],
                                 p.deformation[7], p.deformation[8], p.result[2] };
            std::memcpy(c, &x, 4);
            std::memcpy(c + 4, &y, 4);
            std::memcpy(c + 8, f, 28);
            c += 36;
        }
        if (written) *written = n * 36;
    });
}

int fi_write_raw(const fi_poi* pois, size_t n, const char* path, size_t* bytes)
{
    return guarded([&] {
        if (!path) fail(FI_E_INVALID_ARGUMENT, "null path");
        if (n > 0 && !pois) fail(FI_E_INVALID_ARGUMENT, "null POI buffer");
        std::vector<unsigned char> buf(n * 36);
        size_t w = 0;
        if (n > 0) {
            const int rc = fi_format_raw(pois, n, buf.data(), buf.size(), &w);
            if (rc != FI_OK) fail(rc, last_error());
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) fail(FI_E_INVALID_ARGUMENT, std::string("cannot open ") + path + " for writing");
        const size_t put = w ? std::fwrite(buf.data(), 1, w, f) : 0;
        const bool ok = put == w && std::fflush(f) == 0;
        if (std::fclose(f) != 0 || !ok) fail(FI_E_DEVICE, std::string("short write to ") + path);
        if (bytes) *bytes = w;
    });
}

}  // extern "C"
