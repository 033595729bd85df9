// fasticgn_opencorr.hpp -- OpenCorr-shaped C++ API over the C ABI in fasticgn.h.
//
// Drop-in shapes of the reference's classes for the 2D first-order path:
//   opencorr::POI2D    src/oc_poi.h:102-136   (same 100-byte layout as fi_poi)
//   opencorr::Image2D  src/oc_image.h         (row-major pixels instead of an Eigen matrix)
//   opencorr::FFTCC2D  src/oc_fftcc.h:56-100
//   opencorr::ICGN2D1  src/oc_icgn.h:45-77
// Errors surface as the reference's exception types: std::invalid_argument,
// std::logic_error, std::out_of_range, std::runtime_error (device failures and
// device-path limits).  Header-only; link with -lfasticgn.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fasticgn.h"

namespace opencorr {

struct Point2D {
    float x = 0.f, y = 0.f;
    Point2D() = default;
    Point2D(float x_, float y_) : x(x_), y(y_) {}
    Point2D(int x_, int y_) : x((float)x_), y((float)y_) {}
};

union DeformationVector2D {
    struct { float u, ux, uy, uxx, uxy, uyy, v, vx, vy, vxx, vxy, vyy; };
    float p[12];
};

union Result2D {
    struct { float u0, v0, zncc, iteration, convergence, feature; };
    float r[6];
};

union StrainVector2D {
    struct { float exx, eyy, exy; };
    float e[3];
};

class POI2D : public Point2D {
public:
    DeformationVector2D deformation;
    Result2D result;
    StrainVector2D strain;
    Point2D subset_radius;

    POI2D(int x_, int y_) : Point2D(x_, y_) { clear(); }
    POI2D(float x_, float y_) : Point2D(x_, y_) { clear(); }
    explicit POI2D(Point2D p) : Point2D(p) { clear(); }
    void clear()
    {
        std::memset(deformation.p, 0, sizeof(deformation.p));
        std::memset(result.r, 0, sizeof(result.r));
        std::memset(strain.e, 0, sizeof(strain.e));
        subset_radius = Point2D();
    }
};
static_assert(sizeof(POI2D) == sizeof(fi_poi), "POI2D must match fi_poi");

namespace detail {
inline void check(int status)
{
    if (status == FI_OK) return;
    char buf[1024];
    fi_last_error(buf, sizeof(buf));
    switch (status) {
    case FI_E_INVALID_ARGUMENT: throw std::invalid_argument(buf);
    case FI_E_LOGIC: throw std::logic_error(buf);
    case FI_E_OUT_OF_RANGE: throw std::out_of_range(buf);
    default: throw std::runtime_error(buf);
    }
}
inline fi_poi* raw(std::vector<POI2D>& q) { return reinterpret_cast<fi_poi*>(q.data()); }
inline const fi_poi* raw(const std::vector<POI2D>& q) { return reinterpret_cast<const fi_poi*>(q.data()); }
}  // namespace detail

// A grey image.  The caller owns the pixels and keeps them alive while a solver
// reads them (setters are change notifications, as in the reference).
class Image2D {
public:
    int width = 0, height = 0;
    int pixel = FI_PIXEL_F32;
    int depth12 = FI_DEPTH_NATIVE;
    size_t stride_bytes = 0;
    const void* data = nullptr;

    Image2D() = default;
    Image2D(const uint8_t* px, int w, int h, size_t stride = 0)
        : width(w), height(h), pixel(FI_PIXEL_U8), stride_bytes(stride ? stride : (size_t)w), data(px) {}
    Image2D(const uint16_t* px, int w, int h, int depth12_mode = FI_DEPTH_NATIVE, size_t stride = 0)
        : width(w), height(h), pixel(FI_PIXEL_U16), depth12(depth12_mode),
          stride_bytes(stride ? stride : 2 * (size_t)w), data(px) {}
    Image2D(const float* px, int w, int h, size_t stride = 0)
        : width(w), height(h), pixel(FI_PIXEL_F32), stride_bytes(stride ? stride : 4 * (size_t)w), data(px) {}
};

class FFTCC2D {
public:
    FFTCC2D(int subset_radius_x, int subset_radius_y, int device = 0)
        : subset_radius_x(subset_radius_x), subset_radius_y(subset_radius_y)
    {
        detail::check(fi_fftcc_create(subset_radius_x, subset_radius_y, device, &h_));
    }
    ~FFTCC2D() { fi_fftcc_destroy(h_); }
    FFTCC2D(const FFTCC2D&) = delete;
    FFTCC2D& operator=(const FFTCC2D&) = delete;

    void setImages(Image2D& ref, Image2D& tar) { setRefImage(ref); setTarImage(tar); }
    void setRefImage(Image2D& ref)
    {
        detail::check(fi_fftcc_set_ref(h_, ref.data, ref.width, ref.height, ref.stride_bytes, ref.pixel, ref.depth12));
        ref_img = &ref;
    }
    void setTarImage(Image2D& tar)
    {
        detail::check(fi_fftcc_set_tar(h_, tar.data, tar.width, tar.height, tar.stride_bytes, tar.pixel, tar.depth12));
        tar_img = &tar;
    }
    bool hasPreparedReference(const Image2D& ref) const
    {
        int ready = 0;
        return fi_fftcc_has_prepared_ref(h_, &ready) == FI_OK && ready && ref_img == &ref;
    }
    void prepare() { prepareRef(); prepareTar(); }
    void prepareRef() { detail::check(fi_fftcc_prepare_ref_clear(h_)); }
    void prepareRef(const std::vector<POI2D>& queue)
    {
        detail::check(fi_fftcc_prepare_ref(h_, detail::raw(queue), queue.size()));
    }
    void prepareTar() {}
    void compute(POI2D* poi) { detail::check(fi_fftcc_compute(h_, reinterpret_cast<fi_poi*>(poi), 1)); }
    void compute(std::vector<POI2D>& queue) { detail::check(fi_fftcc_compute(h_, detail::raw(queue), queue.size())); }
    void compute(std::vector<POI2D>& queue, const std::vector<size_t>& cache_indices)
    {
        if (cache_indices.size() != queue.size())
            throw std::invalid_argument("FFTCC2D::compute requires one cache index per POI");
        detail::check(fi_fftcc_compute_indexed(h_, detail::raw(queue), queue.size(), cache_indices.data()));
    }

    int subset_radius_x, subset_radius_y;
    Image2D* ref_img = nullptr;
    Image2D* tar_img = nullptr;

private:
    fi_fftcc* h_ = nullptr;
};

class ICGN2D1 {
public:
    ICGN2D1(int subset_radius_x, int subset_radius_y, float conv_criterion, float stop_condition,
            int device = 0)
        : subset_radius_x(subset_radius_x), subset_radius_y(subset_radius_y)
    {
        detail::check(fi_icgn_create(subset_radius_x, subset_radius_y, conv_criterion, stop_condition, device, &h_));
    }
    ~ICGN2D1() { fi_icgn_destroy(h_); }
    ICGN2D1(const ICGN2D1&) = delete;
    ICGN2D1& operator=(const ICGN2D1&) = delete;

    void setIteration(float conv_criterion, float stop_condition)
    {
        detail::check(fi_icgn_set_iteration(h_, conv_criterion, stop_condition));
    }
    void setIteration(POI2D* poi) { setIteration(poi->result.convergence, (float)(int)poi->result.iteration); }
    void setSelfAdaptive(bool on) { detail::check(fi_icgn_set_self_adaptive(h_, on ? 1 : 0)); }
    void setImages(Image2D& ref, Image2D& tar) { setRefImage(ref); setTarImage(tar); }
    void setRefImage(Image2D& ref)
    {
        detail::check(fi_icgn_set_ref(h_, ref.data, ref.width, ref.height, ref.stride_bytes, ref.pixel, ref.depth12));
    }
    void setTarImage(Image2D& tar)
    {
        detail::check(fi_icgn_set_tar(h_, tar.data, tar.width, tar.height, tar.stride_bytes, tar.pixel, tar.depth12));
    }
    void prepareRef() { detail::check(fi_icgn_prepare_ref(h_)); }
    void prepareTar() { detail::check(fi_icgn_prepare_tar(h_)); }
    void prepare() { detail::check(fi_icgn_prepare(h_)); }
    void compute(POI2D* poi) { detail::check(fi_icgn_compute(h_, reinterpret_cast<fi_poi*>(poi), 1)); }
    void compute(std::vector<POI2D>& queue) { detail::check(fi_icgn_compute(h_, detail::raw(queue), queue.size())); }
    // center-offset subsets (src/oc_icgn.cpp:399-625)
    void compute(POI2D* poi, Point2D& center_offset)
    {
        const float off[2] = { center_offset.x, center_offset.y };
        detail::check(fi_icgn_compute_offset(h_, reinterpret_cast<fi_poi*>(poi), 1, off));
    }
    void compute(std::vector<POI2D>& queue, std::vector<Point2D>& center_offset_queue)
    {
        if (center_offset_queue.size() != queue.size())
            throw std::invalid_argument("ICGN2D1::compute requires one center offset per POI");
        std::vector<float> off(2 * queue.size());
        for (size_t i = 0; i < queue.size(); i++) {
            off[2 * i] = center_offset_queue[i].x;
            off[2 * i + 1] = center_offset_queue[i].y;
        }
        detail::check(fi_icgn_compute_offset(h_, detail::raw(queue), queue.size(), off.data()));
    }

    int subset_radius_x, subset_radius_y;

private:
    fi_icgn* h_ = nullptr;
};

}  // namespace opencorr

// The multi-frame driver and the learned seeder: no OpenCorr class exists for
// them (the reference's sequence loop lives in opencorr_cli_cuda_mvp.cpp:
// processFrame :3024-3210, runMultiFrame :3834-4509), so they live in their
// own namespace, over the same POI2D and Image2D.
namespace fasticgn {

class LearnedSeeder {
public:
    explicit LearnedSeeder(int device = 0) { opencorr::detail::check(fi_seeder_create(device, &h_)); }
    ~LearnedSeeder() { fi_seeder_destroy(h_); }
    LearnedSeeder(const LearnedSeeder&) = delete;
    LearnedSeeder& operator=(const LearnedSeeder&) = delete;
    static size_t paramCount()
    {
        size_t n = 0;
        opencorr::detail::check(fi_seeder_param_count(&n));
        return n;
    }
    void setParams(const std::vector<float>& p) { opencorr::detail::check(fi_seeder_set_params(h_, p.data(), p.size())); }
    std::vector<float> params() const
    {
        std::vector<float> p(paramCount());
        opencorr::detail::check(fi_seeder_get_params(h_, p.data(), p.size()));
        return p;
    }
    fi_seeder* handle() const { return h_; }

private:
    fi_seeder* h_ = nullptr;
};

class FrameDriver {
public:
    FrameDriver(int subset_radius, float conv_criterion = 0.001f, float stop_condition = 10.f, int device = 0)
    {
        opencorr::detail::check(fi_frame_create(subset_radius, conv_criterion, stop_condition, device, &h_));
    }
    ~FrameDriver() { fi_frame_destroy(h_); }
    FrameDriver(const FrameDriver&) = delete;
    FrameDriver& operator=(const FrameDriver&) = delete;

    void setReference(const opencorr::Image2D& ref)
    {
        opencorr::detail::check(
            fi_frame_set_ref(h_, ref.data, ref.width, ref.height, ref.stride_bytes, ref.pixel, ref.depth12));
    }
    void setPois(const std::vector<opencorr::POI2D>& tmpl)
    {
        opencorr::detail::check(fi_frame_set_pois(h_, opencorr::detail::raw(tmpl), tmpl.size()));
        n_ = tmpl.size();
    }
    void setSelfAdaptive(bool on) { opencorr::detail::check(fi_frame_set_self_adaptive(h_, on ? 1 : 0)); }
    void setGraph(bool on) { opencorr::detail::check(fi_frame_set_graph(h_, on ? 1 : 0)); }
    void setPersistent(int blocks) { opencorr::detail::check(fi_frame_set_persistent(h_, blocks)); }
    void setSeeder(LearnedSeeder* s, float zncc_threshold = 0.8f)
    {
        opencorr::detail::check(fi_frame_set_seeder(h_, s ? s->handle() : nullptr, zncc_threshold));
    }
    // seeder passes per frame (1..4): each re-crops the target at the
    // previous pass's rounded u, v
    void setSeedPasses(int passes) { opencorr::detail::check(fi_frame_set_seed_passes(h_, passes)); }
    // Queues a target; its pixels must stay valid until the next run() returns.
    void uploadTarget(const opencorr::Image2D& tar)
    {
        opencorr::detail::check(
            fi_frame_upload_target(h_, tar.data, tar.width, tar.height, tar.stride_bytes, tar.pixel, tar.depth12));
    }
    fi_frame_stats run(unsigned policy = FI_POLICY_DEFAULT)
    {
        fi_frame_stats st{};
        opencorr::detail::check(fi_frame_run(h_, policy, &st));
        return st;
    }
    void download(std::vector<opencorr::POI2D>& out)
    {
        if (out.size() != n_) out.assign(n_, opencorr::POI2D(0.f, 0.f));
        opencorr::detail::check(fi_frame_download(h_, opencorr::detail::raw(out), out.size()));
    }
    fi_frame_stats process(const opencorr::Image2D& tar, std::vector<opencorr::POI2D>& out,
                           unsigned policy = FI_POLICY_DEFAULT)
    {
        uploadTarget(tar);
        const fi_frame_stats st = run(policy);
        download(out);
        return st;
    }
    std::vector<fi_tune_entry> autotune()
    {
        std::vector<fi_tune_entry> rep(512);
        size_t m = 0;
        opencorr::detail::check(fi_frame_autotune(h_, rep.data(), rep.size(), &m));
        rep.resize(m);
        return rep;
    }
    void resetSequence() { opencorr::detail::check(fi_frame_reset_sequence(h_)); }

private:
    fi_frame* h_ = nullptr;
    size_t n_ = 0;
};

}  // namespace fasticgn
