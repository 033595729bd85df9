// fasticgn_opencorr.hpp -- OpenCorr-shaped C++ API over the C ABI in fasticgn.h.
//
// Drop-in shapes of the reference's classes for the 2D first-order path:
//   opencorr::POI2D    src/oc_poi.h:102-136   (same 100-byte layout as fi_poi)
//   opencorr::Image2D  src/oc_image.h         (row-major pixels instead of an Eigen matrix;
//                      eg_mat(y, x) accesses them; PGM files instead of OpenCV codecs)
//   opencorr::DIC      src/oc_dic.h:44-65
//   opencorr::FFTCC2D  src/oc_fftcc.h:56-100
//   opencorr::ICGN2D1  src/oc_icgn.h:45-77
// Errors surface as the reference's exception types: std::invalid_argument,
// std::logic_error, std::out_of_range, std::runtime_error (device failures and
// device-path limits).  Header-only; link with -lfasticgn.
#pragma once

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
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

// A PGM file (P5 binary or P2 ASCII; maxval <= 255 gives 8-bit pixels, up to
// 65535 16-bit ones, returned in host byte order).  Failures throw
// std::string("Fail to load file: ..."), as the reference's loader does.
inline std::vector<uint8_t> read_pgm(const std::string& path, int& w, int& h, int& bits)
{
    const std::string fail = "Fail to load file: " + path;
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw fail;
    std::vector<uint8_t> blob;
    uint8_t buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) blob.insert(blob.end(), buf, buf + got);
    std::fclose(f);
    size_t pos = 0;
    auto token = [&]() -> std::string {
        for (;;) {
            while (pos < blob.size() && std::isspace(blob[pos])) pos++;
            if (pos < blob.size() && blob[pos] == '#') {
                while (pos < blob.size() && blob[pos] != '\n' && blob[pos] != '\r') pos++;
                continue;
            }
            break;
        }
        const size_t start = pos;
        while (pos < blob.size() && !std::isspace(blob[pos])) pos++;
        return std::string(blob.begin() + start, blob.begin() + pos);
    };
    const std::string magic = token();
    const std::string sw = token(), sh = token(), sm = token();
    if ((magic != "P5" && magic != "P2") || sw.empty() || sh.empty() || sm.empty()) throw fail;
    w = std::atoi(sw.c_str());
    h = std::atoi(sh.c_str());
    const int maxval = std::atoi(sm.c_str());
    if (w < 1 || h < 1 || maxval < 1 || maxval > 65535) throw fail;
    bits = maxval < 256 ? 8 : 16;
    const size_t n = (size_t)w * h, bpp = bits / 8;
    std::vector<uint8_t> out(n * bpp);
    if (magic == "P5") {
        pos++;  // the single whitespace after maxval
        if (blob.size() < pos + n * bpp) throw fail;
        for (size_t i = 0; i < n; i++) {
            if (bpp == 1) out[i] = blob[pos + i];
            else {
                const uint16_t v = (uint16_t)((blob[pos + 2 * i] << 8) | blob[pos + 2 * i + 1]);  // big-endian
                std::memcpy(out.data() + 2 * i, &v, 2);
            }
        }
    } else {
        for (size_t i = 0; i < n; i++) {
            const std::string t = token();
            if (t.empty()) throw fail;
            const uint16_t v = (uint16_t)std::atoi(t.c_str());
            if (bpp == 1) out[i] = (uint8_t)v;
            else std::memcpy(out.data() + 2 * i, &v, 2);
        }
    }
    return out;
}
inline const fi_poi* raw(const std::vector<POI2D>& q) { return reinterpret_cast<const fi_poi*>(q.data()); }
}  // namespace detail

// ---- images ----------------------------------------------------------------
// ImageLoadOptions / ImageLoadInfo (src/oc_image.h:26-73).
enum class ImageInputDepthMode { Compat8, Depth12, Depth16, Auto };
enum class ImageInputAlignment { Right, Left, Auto };
enum class ImageInputNormalize { None, To8, To16, Unit };

struct ImageLoadOptions {
    ImageInputDepthMode depth = ImageInputDepthMode::Compat8;
    ImageInputAlignment alignment = ImageInputAlignment::Auto;
    ImageInputNormalize normalize = ImageInputNormalize::None;
};

struct ImageLoadInfo {
    ImageInputDepthMode requested_depth = ImageInputDepthMode::Compat8;
    ImageInputDepthMode effective_depth = ImageInputDepthMode::Compat8;
    ImageInputAlignment requested_alignment = ImageInputAlignment::Auto;
    ImageInputAlignment effective_alignment = ImageInputAlignment::Auto;
    ImageInputNormalize normalize = ImageInputNormalize::None;
    int source_bits = 0;        // 8 or 16 (the decoded file)
    double source_min = 0.0, source_max = 0.0;
    double output_scale = 1.0;  // pixels become float(p) * float(output_scale)
};

class Image2D;

// eg_mat(y, x): the reference keeps the pixels in an Eigen matrix; here they
// are row-major.  Const access to a file-backed image converts the stored
// pixel; the first non-const access (it may write) turns the image into float
// pixels, as the reference's eg_mat holds them.
class PixelMatrix {
public:
    float& operator()(int r, int c);
    float operator()(int r, int c) const;
    int rows() const;
    int cols() const;

private:
    friend class Image2D;
    Image2D* img_ = nullptr;
};

// A grey image.  Either owned (Image2D(width, height), Image2D(file_path)) or
// a view of the caller's pixels, which must stay alive while a solver reads
// them.  Setters of the solvers are change notifications, as in the
// reference: pixels are uploaded when an image is set.
class Image2D {
public:
    int height = 0, width = 0;
    unsigned int size = 0;
    std::string file_path;
    PixelMatrix eg_mat;
    ImageLoadInfo load_info;
    // what the solvers upload
    int pixel = FI_PIXEL_F32;
    int depth12 = FI_DEPTH_NATIVE;  // input mode: FI_DEPTH_* | FI_NORMALIZE_*
    size_t stride_bytes = 0;
    const void* data = nullptr;

    Image2D(int width_, int height_)  // src/oc_image.cpp:219-227: zero pixels
        : height(height_), width(width_), size((unsigned)width_ * (unsigned)height_),
          stride_bytes(4 * (size_t)width_), fbuf_((size_t)width_ * height_, 0.f)
    {
        if (width_ < 1 || height_ < 1) throw std::invalid_argument("Image2D size must be positive");
        data = fbuf_.data();
        eg_mat.img_ = this;
    }
    explicit Image2D(std::string path) { eg_mat.img_ = this; load(path); }
    Image2D(const uint8_t* px, int w, int h, size_t stride = 0)
        : height(h), width(w), size((unsigned)w * (unsigned)h), pixel(FI_PIXEL_U8),
          stride_bytes(stride ? stride : (size_t)w), data(px) { eg_mat.img_ = this; }
    Image2D(const uint16_t* px, int w, int h, int mode = FI_DEPTH_NATIVE, size_t stride = 0)
        : height(h), width(w), size((unsigned)w * (unsigned)h), pixel(FI_PIXEL_U16), depth12(mode),
          stride_bytes(stride ? stride : 2 * (size_t)w), data(px) { eg_mat.img_ = this; }
    Image2D(const float* px, int w, int h, size_t stride = 0)
        : height(h), width(w), size((unsigned)w * (unsigned)h), pixel(FI_PIXEL_F32),
          stride_bytes(stride ? stride : 4 * (size_t)w), data(px) { eg_mat.img_ = this; }
    Image2D(Image2D&& o) noexcept { *this = std::move(o); }
    Image2D& operator=(Image2D&& o) noexcept
    {
        height = o.height; width = o.width; size = o.size;
        file_path = std::move(o.file_path);
        load_info = o.load_info;
        pixel = o.pixel; depth12 = o.depth12; stride_bytes = o.stride_bytes;
        const bool own_f = o.data == o.fbuf_.data() && !o.fbuf_.empty();
        const bool own_r = o.data == o.rbuf_.data() && !o.rbuf_.empty();
        fbuf_ = std::move(o.fbuf_);
        rbuf_ = std::move(o.rbuf_);
        data = own_f ? (const void*)fbuf_.data() : own_r ? (const void*)rbuf_.data() : o.data;
        eg_mat.img_ = this;
        return *this;
    }
    Image2D(const Image2D&) = delete;
    Image2D& operator=(const Image2D&) = delete;

    // Image2D::load (src/oc_image.cpp:234-258): 8-bit grey, 16-bit files
    // reduced to their high byte.  Failures throw std::string, as there.
    void load(std::string path) { loadWithOptions(path, ImageLoadOptions{}, nullptr); }

    // loadWithOptions (src/oc_image.cpp:260-285) with loadImage2DMat's
    // depth, alignment and normalise rules (:125-217), the conversion done on
    // the device as the pixels are read.  PGM files (P2 / P5, 8 or 16 bits).
    void loadWithOptions(std::string path, const ImageLoadOptions& options, ImageLoadInfo* info = nullptr)
    {
        int w = 0, h = 0, bits = 0;
        std::vector<uint8_t> px = detail::read_pgm(path, w, h, bits);
        ImageLoadInfo li;
        li.requested_depth = options.depth;
        li.requested_alignment = options.alignment;
        li.normalize = options.normalize;
        li.source_bits = bits;
        int mode = FI_DEPTH_NATIVE;
        int pix = bits == 8 ? FI_PIXEL_U8 : FI_PIXEL_U16;
        if (options.depth == ImageInputDepthMode::Compat8 && bits == 16) {
            // the 8-bit read of a 16-bit file keeps the high byte
            std::vector<uint8_t> hi((size_t)w * h);
            for (size_t i = 0; i < hi.size(); i++) hi[i] = px[2 * i + 1];
            px.swap(hi);
            pix = FI_PIXEL_U8;
        } else if (options.depth == ImageInputDepthMode::Depth12) {
            mode = options.alignment == ImageInputAlignment::Right  ? FI_DEPTH12_RIGHT
                   : options.alignment == ImageInputAlignment::Left ? FI_DEPTH12_LEFT
                                                                    : FI_DEPTH12_AUTO;
        } else if (options.depth == ImageInputDepthMode::Depth16) {
            mode = FI_DEPTH16;
        } else if (options.depth == ImageInputDepthMode::Auto && bits == 16) {
            mode = FI_DEPTH_AUTO;
        }
        mode |= options.normalize == ImageInputNormalize::To8    ? FI_NORMALIZE_TO8
                : options.normalize == ImageInputNormalize::To16 ? FI_NORMALIZE_TO16
                : options.normalize == ImageInputNormalize::Unit ? FI_NORMALIZE_UNIT
                                                                 : FI_NORMALIZE_NONE;
        const size_t bpp = pix == FI_PIXEL_U8 ? 1 : 2;
        fi_ingest_info in{};
        const int rc = fi_image_resolve(px.data(), w, h, bpp * (size_t)w, pix, mode, &in);
        if (rc != FI_OK) {
            char buf[512];
            fi_last_error(buf, sizeof(buf));
            throw std::string(std::string(buf) + ": " + path);
        }
        li.effective_depth = in.effective_bits == 8    ? ImageInputDepthMode::Compat8
                             : in.effective_bits == 12 ? ImageInputDepthMode::Depth12
                                                       : ImageInputDepthMode::Depth16;
        li.effective_alignment = in.effective_bits != 12 ? ImageInputAlignment::Auto
                                 : in.left_aligned       ? ImageInputAlignment::Left
                                                         : ImageInputAlignment::Right;
        li.output_scale = in.scale;
        double mn = 1e300, mx = 0.0;
        for (size_t i = 0; i < (size_t)w * h; i++) {
            double v = 0.0;
            if (bpp == 1) v = px[i];
            else { uint16_t s; std::memcpy(&s, px.data() + 2 * i, 2); v = s; }
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
        li.source_min = mn;
        li.source_max = mx;
        rbuf_.swap(px);
        fbuf_.clear();
        width = w;
        height = h;
        size = (unsigned)w * (unsigned)h;
        pixel = pix;
        depth12 = mode;
        stride_bytes = bpp * (size_t)w;
        data = rbuf_.data();
        file_path = path;
        load_info = li;
        if (info) *info = li;
    }

private:
    friend class PixelMatrix;
    std::vector<float> fbuf_;    // owned float pixels
    std::vector<uint8_t> rbuf_;  // owned decoded 8/16-bit pixels (file images)

    float read_px(int r, int c) const
    {
        const uint8_t* row = static_cast<const uint8_t*>(data) + (size_t)r * stride_bytes;
        if (pixel == FI_PIXEL_F32) { float v; std::memcpy(&v, row + 4 * (size_t)c, 4); return v; }
        const float s = (float)load_info.output_scale;
        if (pixel == FI_PIXEL_U8) return (float)row[c] * s;
        uint16_t v;
        std::memcpy(&v, row + 2 * (size_t)c, 2);
        return (float)v * s;
    }
    float& write_px(int r, int c)
    {
        if (data != fbuf_.data() || fbuf_.empty()) {
            // materialise float pixels (owned), converted as the device would
            std::vector<float> f((size_t)width * height);
            for (int y = 0; y < height; y++)
                for (int x = 0; x < width; x++) f[(size_t)y * width + x] = read_px(y, x);
            fbuf_.swap(f);
            rbuf_.clear();
            pixel = FI_PIXEL_F32;
            depth12 = FI_DEPTH_NATIVE;
            stride_bytes = 4 * (size_t)width;
            data = fbuf_.data();
            load_info.output_scale = 1.0;
        }
        return fbuf_[(size_t)r * width + c];
    }
};

inline float& PixelMatrix::operator()(int r, int c) { return img_->write_px(r, c); }
inline float PixelMatrix::operator()(int r, int c) const { return img_->read_px(r, c); }
inline int PixelMatrix::rows() const { return img_->height; }
inline int PixelMatrix::cols() const { return img_->width; }

// ---- the solvers ---------------------------------------------------------
// DIC (src/oc_dic.h:44-65): images, subset radii, OpenMP thread count and
// self-adaptive subsets.  The device path runs no host threads for the
// computation; thread_number is validated and kept as the reference keeps it.
class DIC {
public:
    Image2D* ref_img = nullptr;
    Image2D* tar_img = nullptr;
    int subset_radius_x = 0, subset_radius_y = 0;
    int thread_number = 1;
    bool self_adaptive = false;

    DIC() = default;
    virtual ~DIC() = default;
    virtual void setImages(Image2D& ref, Image2D& tar)
    {
        ref_img = &ref;
        tar_img = &tar;
    }
    void setSubset(int radius_x, int radius_y)
    {
        subset_radius_x = radius_x;
        subset_radius_y = radius_y;
    }
    void setSelfAdaptive(bool is_self_adaptive) { self_adaptive = is_self_adaptive; }
    virtual void prepare() = 0;
    virtual void compute(POI2D* poi) = 0;
    virtual void compute(std::vector<POI2D>& poi_queue) = 0;
};

// The device the solvers of this thread create their handles on (default 0);
// the reference's third constructor argument is the host thread count, so the
// device is chosen here or by a solver's setDevice.
inline int& default_device()
{
    static thread_local int device = 0;
    return device;
}
inline void setDefaultDevice(int device) { default_device() = device; }

class FFTCC2D : public DIC {
public:
    // src/oc_fftcc.cpp:153-170: square radii 1..23170 and a positive thread
    // count, else std::invalid_argument
    FFTCC2D(int subset_radius_x_, int subset_radius_y_, int thread_number_)
    {
        if (subset_radius_x_ <= 0 || subset_radius_x_ != subset_radius_y_ || subset_radius_x_ > 23170)
            throw std::invalid_argument(
                "FFTCC2D supports positive square radii only (1..23170); rectangular FFTs are unsupported");
        if (thread_number_ < 1) throw std::invalid_argument("FFTCC2D thread count must be positive");
        subset_radius_x = subset_radius_x_;
        subset_radius_y = subset_radius_y_;
        thread_number = thread_number_;
        allocated_radius_ = subset_radius_x_;
        detail::check(fi_fftcc_create(subset_radius_x_, subset_radius_y_, default_device(), &h_));
    }
    ~FFTCC2D() override { fi_fftcc_destroy(h_); }
    FFTCC2D(const FFTCC2D&) = delete;
    FFTCC2D& operator=(const FFTCC2D&) = delete;

    // A new handle on another device; the images set so far are set again
    // (the prepared layout is dropped).
    void setDevice(int device)
    {
        fi_fftcc* h = nullptr;
        detail::check(fi_fftcc_create(allocated_radius_, allocated_radius_, device, &h));
        fi_fftcc_destroy(h_);
        h_ = h;
        if (ref_img) setRefImage(*ref_img);
        if (tar_img) setTarImage(*tar_img);
    }

    void setImages(Image2D& ref, Image2D& tar) override
    {
        setRefImage(ref);
        setTarImage(tar);
    }
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
    bool hasPreparedReference(const Image2D& ref) const noexcept
    {
        int ready = 0;
        return fi_fftcc_has_prepared_ref(h_, &ready) == FI_OK && ready && ref_img == &ref &&
               subset_radius_x == allocated_radius_ && subset_radius_y == allocated_radius_;
    }
    void prepare() override
    {
        prepareRef();
        prepareTar();
    }
    void prepareRef()
    {
        validateGeometry();
        detail::check(fi_fftcc_prepare_ref_clear(h_));
    }
    void prepareRef(const std::vector<POI2D>& queue)
    {
        validateGeometry();
        detail::check(fi_fftcc_prepare_ref(h_, detail::raw(queue), queue.size()));
    }
    void prepareTar() {}
    void compute(POI2D* poi) override
    {
        validateGeometry();
        detail::check(fi_fftcc_compute(h_, reinterpret_cast<fi_poi*>(poi), 1));
    }
    void compute(std::vector<POI2D>& queue) override
    {
        validateGeometry();
        detail::check(fi_fftcc_compute(h_, detail::raw(queue), queue.size()));
    }
    void compute(std::vector<POI2D>& queue, const std::vector<size_t>& cache_indices)
    {
        validateGeometry();
        if (cache_indices.size() != queue.size())
            throw std::invalid_argument("FFTCC2D::compute requires one cache index per POI");
        detail::check(fi_fftcc_compute_indexed(h_, detail::raw(queue), queue.size(), cache_indices.data()));
    }

private:
    fi_fftcc* h_ = nullptr;
    int allocated_radius_ = 0;
    // src/oc_fftcc.cpp:209-213: the geometry is fixed at construction
    void validateGeometry() const
    {
        if (subset_radius_x != allocated_radius_ || subset_radius_y != allocated_radius_)
            throw std::invalid_argument("FFTCC2D subset geometry is fixed at construction; create a new solver");
    }
};

class ICGN2D1 : public DIC {
public:
    float conv_criterion = 0.001f;
    float stop_condition = 10.f;

    // src/oc_icgn.cpp:95-112.  The reference sizes a per-thread pool by
    // thread_number and does not check it; a count below 1 (which leaves it
    // no instance to compute with) is std::invalid_argument here.
    ICGN2D1(int subset_radius_x_, int subset_radius_y_, float conv_criterion_, float stop_condition_,
            int thread_number_)
        : conv_criterion(conv_criterion_), stop_condition(stop_condition_)
    {
        if (thread_number_ < 1) throw std::invalid_argument("ICGN2D1 thread count must be positive");
        subset_radius_x = subset_radius_x_;
        subset_radius_y = subset_radius_y_;
        thread_number = thread_number_;
        detail::check(fi_icgn_create(subset_radius_x_, subset_radius_y_, conv_criterion_, stop_condition_,
                                     default_device(), &h_));
        synced_rx_ = subset_radius_x_;
        synced_ry_ = subset_radius_y_;
    }
    ~ICGN2D1() override { fi_icgn_destroy(h_); }
    ICGN2D1(const ICGN2D1&) = delete;
    ICGN2D1& operator=(const ICGN2D1&) = delete;

    // A new handle on another device; the images set so far are set again and
    // need prepare() before the next compute.
    void setDevice(int device)
    {
        fi_icgn* h = nullptr;
        detail::check(fi_icgn_create(subset_radius_x, subset_radius_y, conv_criterion, stop_condition, device, &h));
        fi_icgn_destroy(h_);
        h_ = h;
        synced_rx_ = subset_radius_x;
        synced_ry_ = subset_radius_y;
        if (ref_img) setRefImage(*ref_img);
        if (tar_img) setTarImage(*tar_img);
    }

    void setIteration(float conv, float stop)  // src/oc_icgn.cpp:120-124
    {
        detail::check(fi_icgn_set_iteration(h_, conv, stop));
        conv_criterion = conv;
        stop_condition = stop;
    }
    void setIteration(POI2D* poi) { setIteration(poi->result.convergence, (float)(int)poi->result.iteration); }
    void setImages(Image2D& ref, Image2D& tar) override
    {
        setRefImage(ref);
        setTarImage(tar);
    }
    void setRefImage(Image2D& ref)
    {
        detail::check(fi_icgn_set_ref(h_, ref.data, ref.width, ref.height, ref.stride_bytes, ref.pixel, ref.depth12));
        ref_img = &ref;
    }
    void setTarImage(Image2D& tar)
    {
        detail::check(fi_icgn_set_tar(h_, tar.data, tar.width, tar.height, tar.stride_bytes, tar.pixel, tar.depth12));
        tar_img = &tar;
    }
    void prepareRef() { detail::check(fi_icgn_prepare_ref(h_)); }
    void prepareTar() { detail::check(fi_icgn_prepare_tar(h_)); }
    void prepare() override { detail::check(fi_icgn_prepare(h_)); }
    void compute(POI2D* poi) override
    {
        sync();
        detail::check(fi_icgn_compute(h_, reinterpret_cast<fi_poi*>(poi), 1));
    }
    void compute(std::vector<POI2D>& queue) override
    {
        sync();
        detail::check(fi_icgn_compute(h_, detail::raw(queue), queue.size()));
    }
    // center-offset subsets (src/oc_icgn.cpp:399-625)
    void compute(POI2D* poi, Point2D& center_offset)
    {
        sync();
        const float off[2] = { center_offset.x, center_offset.y };
        detail::check(fi_icgn_compute_offset(h_, reinterpret_cast<fi_poi*>(poi), 1, off));
    }
    void compute(std::vector<POI2D>& queue, std::vector<Point2D>& center_offset_queue)
    {
        if (center_offset_queue.size() != queue.size())
            throw std::invalid_argument("ICGN2D1::compute requires one center offset per POI");
        sync();
        std::vector<float> off(2 * queue.size());
        for (size_t i = 0; i < queue.size(); i++) {
            off[2 * i] = center_offset_queue[i].x;
            off[2 * i + 1] = center_offset_queue[i].y;
        }
        detail::check(fi_icgn_compute_offset(h_, detail::raw(queue), queue.size(), off.data()));
    }

private:
    fi_icgn* h_ = nullptr;
    int synced_rx_ = 0, synced_ry_ = 0;
    // DIC::setSubset and DIC::setSelfAdaptive only set members (they are not
    // virtual); the handle takes them when a compute starts
    void sync()
    {
        if (subset_radius_x != synced_rx_ || subset_radius_y != synced_ry_) {
            detail::check(fi_icgn_set_subset(h_, subset_radius_x, subset_radius_y));
            synced_rx_ = subset_radius_x;
            synced_ry_ = subset_radius_y;
        }
        detail::check(fi_icgn_set_self_adaptive(h_, self_adaptive ? 1 : 0));
    }
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
