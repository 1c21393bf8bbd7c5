// CudaInterface.hpp -- C++ host API of the MI355X stitch path, over the C-ABI
// in panohip.h.  It keeps the names, argument order and error conventions of
// PanoStitcher's CudaAccel layer (CudaAccel/CudaInterface.h:45-171), of the
// rig files (Warp/ZReproject.h:52-64) and of CudaPanoramaRender2
// (Task/RicohUtil.h:193-220), with project-owned image types in place of
// OpenCV's: pano::Mat (host), pano::GpuMat (device), pano::Stream, pano::Size,
// carrying the fields the reference's code reads (rows, cols, type(), data,
// step).
//
// Error conventions, as in the reference:
//   - prepare() returns false on bad input or a failed preparation
//     (CudaMultibandBlend.cpp:551-565);
//   - blend() / begin() / tile() / end() of an object that is not prepared
//     return without doing anything (CudaMultibandBlend.cpp:787-788);
//   - argument violations that the reference checks with CV_Assert throw
//     pano::Exception (the status code is kept in the exception);
//   - render() returns false on a failed call (RicohUtil.cpp:1207-1339).
// The free functions run on the given stream (asynchronously, like their
// reference counterparts); CudaPanoramaRender2::render waits for its frame.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../panohip.h"

namespace pano {

// ------------------------------------------------------------------ types --

// OpenCV type codes used at this boundary (CV_MAKETYPE values).
constexpr int kType8UC1 = PANO_8UC1, kType8UC3 = PANO_8UC3, kType8UC4 = PANO_8UC4;
constexpr int kType16SC4 = PANO_16SC4, kType32FC1 = 5, kType32FC4 = PANO_32FC4;

inline size_t elem_size(int type)
{
    switch (type) {
    case kType8UC1: return 1;
    case kType8UC3: return 3;
    case kType8UC4: return 4;
    case kType16SC4: return 8;
    case kType32FC1: return 4;
    case kType32FC4: return 16;
    default: return 0;
    }
}

struct Exception : std::runtime_error {
    int status;
    Exception(int st, const std::string& where)
        : std::runtime_error(where + ": " + pano_status_string(st)), status(st)
    {
    }
};

inline void check(int status, const char* where)
{
    if (status != PANO_OK) throw Exception(status, where);
}

struct Size {
    int width = 0, height = 0;
    Size() = default;
    Size(int w, int h) : width(w), height(h) {}
    bool operator==(const Size& o) const { return width == o.width && height == o.height; }
    bool operator!=(const Size& o) const { return !(*this == o); }
};

// Host image (cv::Mat's fields).  Owns its pixels unless made over a caller
// buffer; copies share the pixels.
struct Mat {
    int rows = 0, cols = 0;
    uint8_t* data = nullptr;
    size_t step = 0;

    Mat() = default;
    Mat(int r, int c, int t) { create(r, c, t); }
    Mat(int r, int c, int t, void* d, size_t s) : rows(r), cols(c), data((uint8_t*)d), step(s), type_(t) {}
    void create(int r, int c, int t)
    {
        if (r == rows && c == cols && t == type_ && data) return;
        const size_t es = elem_size(t);
        if (r <= 0 || c <= 0 || !es) throw Exception(PANO_ERR_ARG, "Mat::create");
        own_.reset(new uint8_t[(size_t)r * c * es](), std::default_delete<uint8_t[]>());
        rows = r;
        cols = c;
        type_ = t;
        step = (size_t)c * es;
        data = own_.get();
    }
    int type() const { return type_; }
    bool empty() const { return !data || rows <= 0 || cols <= 0; }
    Size size() const { return Size(cols, rows); }
    size_t elemSize() const { return elem_size(type_); }
    uint8_t* ptr(int y) { return data + (size_t)y * step; }
    const uint8_t* ptr(int y) const { return data + (size_t)y * step; }

private:
    int type_ = -1;
    std::shared_ptr<uint8_t> own_;
};

// A HIP stream (cv::cuda::Stream); Stream::Null() is the null stream.
class Stream {
public:
    Stream() = default;
    explicit Stream(bool create)
    {
        if (create) {
            hipStream_t s = nullptr;
            check(hipStreamCreate(&s) == hipSuccess ? PANO_OK : PANO_ERR_HIP, "Stream");
            own_.reset(s, [](hipStream_t p) { (void)hipStreamDestroy(p); });
        }
    }
    static Stream& Null()
    {
        static Stream s;
        return s;
    }
    hipStream_t handle() const { return own_ ? own_.get() : nullptr; }
    void waitForCompletion() const
    {
        check(hipStreamSynchronize(handle()) == hipSuccess ? PANO_OK : PANO_ERR_HIP, "Stream::waitForCompletion");
    }

private:
    std::shared_ptr<std::remove_pointer<hipStream_t>::type> own_;
};

// Device image (cv::cuda::GpuMat's fields): pitched device memory, shared by copies.
struct GpuMat {
    int rows = 0, cols = 0;
    uint8_t* data = nullptr;
    size_t step = 0;

    GpuMat() = default;
    GpuMat(int r, int c, int t) { create(r, c, t); }
    GpuMat(Size sz, int t) { create(sz.height, sz.width, t); }
    // over caller device memory (not owned): a cv::cuda::GpuMat's data and step
    GpuMat(int r, int c, int t, void* d, size_t s) : rows(r), cols(c), data((uint8_t*)d), step(s), type_(t) {}
    void create(int r, int c, int t)
    {
        if (r == rows && c == cols && t == type_ && data) return;
        const size_t es = elem_size(t);
        if (r <= 0 || c <= 0 || !es) throw Exception(PANO_ERR_ARG, "GpuMat::create");
        void* p = nullptr;
        size_t pitch = 0;
        if (hipMallocPitch(&p, &pitch, (size_t)c * es, (size_t)r) != hipSuccess)
            throw Exception(PANO_ERR_NOMEM, "GpuMat::create");
        own_.reset((uint8_t*)p, [](uint8_t* q) { (void)hipFree(q); });
        rows = r;
        cols = c;
        type_ = t;
        step = pitch;
        data = (uint8_t*)p;
    }
    void create(Size sz, int t) { create(sz.height, sz.width, t); }
    void release()
    {
        own_.reset();
        data = nullptr;
        rows = cols = 0;
        step = 0;
        type_ = -1;
    }
    int type() const { return type_; }
    bool empty() const { return !data || rows <= 0 || cols <= 0; }
    Size size() const { return Size(cols, rows); }
    size_t elemSize() const { return elem_size(type_); }
    long useCount() const { return own_.use_count(); }
    void upload(const Mat& m, Stream& s = Stream::Null())
    {
        create(m.rows, m.cols, m.type());
        check(hipMemcpy2DAsync(data, step, m.data, m.step, (size_t)cols * elemSize(), rows, hipMemcpyHostToDevice,
                               s.handle()) == hipSuccess
                  ? PANO_OK
                  : PANO_ERR_HIP,
              "GpuMat::upload");
    }
    void download(Mat& m, Stream& s = Stream::Null()) const
    {
        m.create(rows, cols, type_);
        check(hipMemcpy2DAsync(m.data, m.step, data, step, (size_t)cols * elemSize(), rows, hipMemcpyDeviceToHost,
                               s.handle()) == hipSuccess
                  ? PANO_OK
                  : PANO_ERR_HIP,
              "GpuMat::download");
        if (!s.handle()) s.waitForCompletion();
    }

private:
    int type_ = -1;
    std::shared_ptr<uint8_t> own_;
};

// Page-locked host image (cv::cuda::HostMem, PAGE_LOCKED): the target of
// asynchronous downloads.  Copies share the buffer; useCount() counts them.
struct HostMem {
    int rows = 0, cols = 0;
    uint8_t* data = nullptr;
    size_t step = 0;

    HostMem() = default;
    HostMem(int r, int c, int t) { create(r, c, t); }
    void create(int r, int c, int t)
    {
        if (r == rows && c == cols && t == type_ && data) return;
        const size_t es = elem_size(t);
        if (r <= 0 || c <= 0 || !es) throw Exception(PANO_ERR_ARG, "HostMem::create");
        void* p = nullptr;
        if (hipHostMalloc(&p, (size_t)r * c * es, hipHostMallocDefault) != hipSuccess)
            throw Exception(PANO_ERR_NOMEM, "HostMem::create");
        own_.reset((uint8_t*)p, [](uint8_t* q) { (void)hipHostFree(q); });
        rows = r;
        cols = c;
        type_ = t;
        step = (size_t)c * es;
        data = (uint8_t*)p;
    }
    int type() const { return type_; }
    bool empty() const { return !data; }
    Size size() const { return Size(cols, rows); }
    long useCount() const { return own_.use_count(); }
    Mat createMatHeader() const { return Mat(rows, cols, type_, data, step); }

private:
    int type_ = -1;
    std::shared_ptr<uint8_t> own_;
};

// Asynchronous download into page-locked memory (GpuMat::download(HostMem&, Stream&)).
inline void download(const GpuMat& src, HostMem& dst, Stream& stream)
{
    dst.create(src.rows, src.cols, src.type());
    check(hipMemcpy2DAsync(dst.data, dst.step, src.data, src.step, (size_t)src.cols * src.elemSize(), src.rows,
                           hipMemcpyDeviceToHost, stream.handle()) == hipSuccess
              ? PANO_OK
              : PANO_ERR_HIP,
          "download");
}

// PinnedMemoryPool / GpuMemoryPool (Task/PinnedMemoryPool.h:7-175): buffers of
// one shape; get() hands out a buffer no one else holds (the pool's copy is its
// only owner) or a new one.  Thread-safe.
template <class Buf>
class BufferPool {
public:
    bool init(int rows, int cols, int type)
    {
        clear();
        std::lock_guard<std::mutex> lock(mtx_);
        try {
            Buf test(rows, cols, type);
        } catch (const Exception&) {
            return false;
        }
        rows_ = rows;
        cols_ = cols;
        type_ = type;
        init_ = true;
        return true;
    }
    void clear()
    {
        std::lock_guard<std::mutex> lock(mtx_);
        pool_.clear();
    }
    bool get(Buf& mem)
    {
        std::lock_guard<std::mutex> lock(mtx_);
        mem = Buf();
        if (!init_) return false;
        for (Buf& b : pool_)
            if (use_count(b) == 1) {
                mem = b;
                return true;
            }
        try {
            Buf fresh(rows_, cols_, type_);
            pool_.push_back(fresh);
            mem = fresh;
        } catch (const Exception&) {
            return false;
        }
        return true;
    }
    size_t size()
    {
        std::lock_guard<std::mutex> lock(mtx_);
        return pool_.size();
    }

private:
    static long use_count(const HostMem& b) { return b.useCount(); }
    static long use_count(const GpuMat& b) { return b.useCount(); }
    std::mutex mtx_;
    std::vector<Buf> pool_;
    int rows_ = 0, cols_ = 0, type_ = 0;
    bool init_ = false;
};
using PinnedMemoryPool = BufferPool<HostMem>;
using GpuMemoryPool = BufferPool<GpuMat>;

namespace detail {
template <class M>
void need(const M& m, int type, const char* where)
{
    if (m.empty() || m.type() != type) throw Exception(PANO_ERR_ARG, where);
}
template <class M>
std::vector<const uint8_t*> ptrs(const std::vector<M>& v)
{
    std::vector<const uint8_t*> p;
    for (const M& m : v) p.push_back(m.data);
    return p;
}
template <class M>
std::vector<size_t> steps(const std::vector<M>& v)
{
    std::vector<size_t> s;
    for (const M& m : v) s.push_back(m.step);
    return s;
}
// Masks for a prepare: non-empty, 8UC1, one size (CudaMultibandBlend.cpp:551-565).
template <class M>
bool masks_ok(const std::vector<M>& masks)
{
    if (masks.empty()) return false;
    for (const M& m : masks)
        if (m.empty() || m.type() != kType8UC1 || m.size() != masks[0].size()) return false;
    return true;
}
}  // namespace detail

// ------------------------------------------------------------- rig files --

using PhotoParam = pano_photo_param;

inline void loadPhotoParamFromXML(const std::string& fileName, std::vector<PhotoParam>& params)
{
    params.clear();
    const int n = pano_load_photo_params_xml(fileName.c_str(), nullptr, 0);
    check(n < 0 ? n : PANO_OK, "loadPhotoParamFromXML");   // the reference's ticpp throws there
    params.resize(n);
    if (n) pano_load_photo_params_xml(fileName.c_str(), params.data(), n);
}

inline void loadPhotoParamFromPTS(const std::string& fileName, std::vector<PhotoParam>& params)
{
    params.clear();
    const int n = pano_load_photo_params_pts(fileName.c_str(), nullptr, 0);
    params.resize(n > 0 ? n : 0);
    if (n > 0) pano_load_photo_params_pts(fileName.c_str(), params.data(), n);
}

inline bool loadPhotoParams(const std::string& fileName, std::vector<PhotoParam>& params)
{
    params.clear();
    const int n = pano_load_photo_params(fileName.c_str(), nullptr, 0);
    if (n == PANO_ERR_ARG) return false;   // no extension: the reference prints an error
    check(n < 0 ? n : PANO_OK, "loadPhotoParams");
    params.resize(n);
    if (n) pano_load_photo_params(fileName.c_str(), params.data(), n);
    return !params.empty();
}

inline void exportPhotoParamToXML(const std::string& fileName, const std::vector<PhotoParam>& params)
{
    check(pano_export_photo_params_xml(fileName.c_str(), params.data(), (int)params.size()), "exportPhotoParamToXML");
}

inline void rotateCamera(PhotoParam& param, double yaw, double pitch, double roll)
{
    check(pano_rotate_cameras(&param, 1, yaw, pitch, roll), "rotateCamera");
}

inline void rotateCameras(std::vector<PhotoParam>& params, double yaw, double pitch, double roll)
{
    check(pano_rotate_cameras(params.data(), (int)params.size(), yaw, pitch, roll), "rotateCameras");
}

inline void rotatePhotoParamInXML(const std::string& src, const std::string& dst, double yaw, double pitch,
                                  double roll)
{
    check(pano_rotate_photo_params_xml(src.c_str(), dst.c_str(), yaw, pitch, roll), "rotatePhotoParamInXML");
}

// --------------------------------------------------------------- blenders --

class CudaTilingMultibandBlendFast {
public:
    CudaTilingMultibandBlendFast() { check(pano_blender_create(&h_), "CudaTilingMultibandBlendFast"); }
    ~CudaTilingMultibandBlendFast() { pano_blender_destroy(h_); }
    CudaTilingMultibandBlendFast(const CudaTilingMultibandBlendFast&) = delete;
    CudaTilingMultibandBlendFast& operator=(const CudaTilingMultibandBlendFast&) = delete;

    bool prepare(const std::vector<Mat>& masks, int maxLevels, int minLength)
    {
        success_ = false;
        if (!detail::masks_ok(masks) || masks.size() > 255) return false;
        const auto p = detail::ptrs(masks);
        const auto s = detail::steps(masks);
        success_ = pano_blender_prepare(h_, (int)masks.size(), masks[0].rows, masks[0].cols, p.data(), s.data(),
                                        maxLevels, minLength) == PANO_OK;
        if (success_) n_ = (int)masks.size(), size_ = masks[0].size();
        return success_;
    }
    bool prepare(const std::vector<GpuMat>& masks, int maxLevels, int minLength)
    {
        success_ = false;
        if (!detail::masks_ok(masks) || masks.size() > 255) return false;
        const auto p = detail::ptrs(masks);
        const auto s = detail::steps(masks);
        success_ = pano_blender_prepare_device(h_, (int)masks.size(), masks[0].rows, masks[0].cols, p.data(),
                                               s.data(), maxLevels, minLength) == PANO_OK;
        if (success_) n_ = (int)masks.size(), size_ = masks[0].size();
        return success_;
    }
    void begin()
    {
        if (!success_) return;
        check(pano_blender_begin(h_), "CudaTilingMultibandBlendFast::begin");
    }
    void tile(const GpuMat& image, int index, Stream& stream = Stream::Null())
    {
        if (!success_) return;
        if (image.size() != size_ || (image.type() != kType16SC4 && image.type() != kType8UC4))
            throw Exception(PANO_ERR_ARG, "CudaTilingMultibandBlendFast::tile");
        check(pano_blender_tile(h_, image.data, image.step, image.type(), index, stream.handle()),
              "CudaTilingMultibandBlendFast::tile");
    }
    void end(GpuMat& blendImage, Stream& stream = Stream::Null())
    {
        if (!success_) return;
        blendImage.create(size_, kType8UC4);
        check(pano_blender_end(h_, blendImage.data, blendImage.step, stream.handle()),
              "CudaTilingMultibandBlendFast::end");
    }
    void blend(const std::vector<GpuMat>& images, GpuMat& blendImage, Stream& stream = Stream::Null())
    {
        if (!success_) return;
        const int type = images_ok(images, "CudaTilingMultibandBlendFast::blend");
        blendImage.create(size_, kType8UC4);
        std::vector<const void*> p;
        for (const GpuMat& m : images) p.push_back(m.data);
        const auto s = detail::steps(images);
        check(pano_blender_blend(h_, n_, p.data(), s.data(), type, blendImage.data, blendImage.step,
                                 stream.handle()),
              "CudaTilingMultibandBlendFast::blend");
    }
    void blend(const std::vector<GpuMat>& images, const std::vector<GpuMat>& masks, GpuMat& blendImage,
               Stream& stream = Stream::Null())
    {
        if (!success_) return;
        const int type = images_ok(images, "CudaTilingMultibandBlendFast::blend");
        if (masks.size() != images.size() || !detail::masks_ok(masks) || masks[0].size() != size_)
            throw Exception(PANO_ERR_ARG, "CudaTilingMultibandBlendFast::blend");
        blendImage.create(size_, kType8UC4);
        std::vector<const void*> p;
        for (const GpuMat& m : images) p.push_back(m.data);
        const auto s = detail::steps(images);
        const auto mp = detail::ptrs(masks);
        const auto ms = detail::steps(masks);
        check(pano_blender_blend_custom(h_, n_, p.data(), s.data(), type, mp.data(), ms.data(), blendImage.data,
                                        blendImage.step, stream.handle()),
              "CudaTilingMultibandBlendFast::blend");
    }
    void getUniqueMasks(std::vector<GpuMat>& masks) const
    {
        masks.clear();
        if (!success_) return;
        std::vector<Mat> host(n_);
        std::vector<uint8_t*> p;
        std::vector<size_t> s;
        for (Mat& m : host) {
            m.create(size_.height, size_.width, kType8UC1);
            p.push_back(m.data);
            s.push_back(m.step);
        }
        check(pano_blender_unique_masks(h_, p.data(), s.data()), "CudaTilingMultibandBlendFast::getUniqueMasks");
        masks.resize(n_);
        for (int i = 0; i < n_; i++) masks[i].upload(host[i]);
        Stream::Null().waitForCompletion();
    }
    long long calcUsedMemorySize() const { return success_ ? (long long)pano_blender_mem_bytes(h_) : 0; }
    static long long estimateMemorySize(int width, int height, int numImages)
    {
        return (long long)pano_estimate_mem_bytes(width, height, numImages);
    }

private:
    int images_ok(const std::vector<GpuMat>& images, const char* where) const
    {
        if ((int)images.size() != n_ || images.empty()) throw Exception(PANO_ERR_ARG, where);
        const int type = images[0].type();
        if (type != kType16SC4 && type != kType8UC4) throw Exception(PANO_ERR_ARG, where);
        for (const GpuMat& m : images)
            if (m.size() != size_ || m.type() != type) throw Exception(PANO_ERR_ARG, where);
        return type;
    }
    pano_blender* h_ = nullptr;
    bool success_ = false;
    int n_ = 0;
    Size size_;
};

class CudaTilingMultibandBlendFast32F {
public:
    CudaTilingMultibandBlendFast32F() { check(pano_blender32f_create(&h_), "CudaTilingMultibandBlendFast32F"); }
    ~CudaTilingMultibandBlendFast32F() { pano_blender32f_destroy(h_); }
    CudaTilingMultibandBlendFast32F(const CudaTilingMultibandBlendFast32F&) = delete;
    CudaTilingMultibandBlendFast32F& operator=(const CudaTilingMultibandBlendFast32F&) = delete;

    bool prepare(const std::vector<Mat>& masks, int maxLevels, int minLength)
    {
        success_ = false;
        if (!detail::masks_ok(masks) || masks.size() > 255) return false;
        const auto p = detail::ptrs(masks);
        const auto s = detail::steps(masks);
        success_ = pano_blender32f_prepare(h_, (int)masks.size(), masks[0].rows, masks[0].cols, p.data(),
                                           s.data(), maxLevels, minLength) == PANO_OK;
        if (success_) n_ = (int)masks.size(), size_ = masks[0].size();
        return success_;
    }
    void blend(const std::vector<GpuMat>& images, GpuMat& blendImage, Stream& stream = Stream::Null())
    {
        if (!success_) return;
        if ((int)images.size() != n_) throw Exception(PANO_ERR_ARG, "CudaTilingMultibandBlendFast32F::blend");
        const int type = images[0].type();
        for (const GpuMat& m : images)
            if (m.size() != size_ || m.type() != type || (type != kType8UC4 && type != kType32FC4))
                throw Exception(PANO_ERR_ARG, "CudaTilingMultibandBlendFast32F::blend");
        blendImage.create(size_, kType8UC4);
        std::vector<const void*> p;
        for (const GpuMat& m : images) p.push_back(m.data);
        const auto s = detail::steps(images);
        check(pano_blender32f_blend(h_, n_, p.data(), s.data(), type, blendImage.data, blendImage.step,
                                    stream.handle()),
              "CudaTilingMultibandBlendFast32F::blend");
    }

private:
    pano_blender32f* h_ = nullptr;
    bool success_ = false;
    int n_ = 0;
    Size size_;
};

class CudaTilingLinearBlend {
public:
    CudaTilingLinearBlend() { check(pano_linear_blender_create(&h_), "CudaTilingLinearBlend"); }
    ~CudaTilingLinearBlend() { pano_linear_blender_destroy(h_); }
    CudaTilingLinearBlend(const CudaTilingLinearBlend&) = delete;
    CudaTilingLinearBlend& operator=(const CudaTilingLinearBlend&) = delete;

    bool prepare(const std::vector<Mat>& masks, int radius)
    {
        success_ = false;
        if (!detail::masks_ok(masks) || radius < 1) return false;
        const auto p = detail::ptrs(masks);
        const auto s = detail::steps(masks);
        success_ = pano_linear_blender_prepare(h_, (int)masks.size(), masks[0].rows, masks[0].cols, p.data(),
                                               s.data(), radius) == PANO_OK;
        if (success_) n_ = (int)masks.size(), size_ = masks[0].size();
        return success_;
    }
    void blend(const std::vector<GpuMat>& images, GpuMat& blendImage, Stream& stream = Stream::Null())
    {
        if (!success_) return;
        if ((int)images.size() != n_) throw Exception(PANO_ERR_ARG, "CudaTilingLinearBlend::blend");
        for (const GpuMat& m : images)
            if (m.size() != size_ || m.type() != kType8UC4) throw Exception(PANO_ERR_ARG, "CudaTilingLinearBlend::blend");
        blendImage.create(size_, kType8UC4);
        const auto p = detail::ptrs(images);
        const auto s = detail::steps(images);
        check(pano_linear_blender_blend(h_, n_, p.data(), s.data(), blendImage.data, blendImage.step, stream.handle()),
              "CudaTilingLinearBlend::blend");
    }

private:
    pano_linear_blender* h_ = nullptr;
    bool success_ = false;
    int n_ = 0;
    Size size_;
};

// --------------------------------------------------------- free functions --

// cudaTransform(src, dst, lut, stream, mask) (CudaInterface.h:117-118): one
// LUT for the three colour channels.  Masked transforms are not part of this
// path (a non-empty mask throws PANO_ERR_UNSUPPORTED).
inline void cudaTransform(const GpuMat& src, GpuMat& dst, const std::vector<std::vector<unsigned char>>& luts,
                          Stream& stream = Stream::Null(), const GpuMat& mask = GpuMat())
{
    detail::need(src, kType8UC4, "cudaTransform");
    if (!mask.empty()) throw Exception(PANO_ERR_UNSUPPORTED, "cudaTransform");
    if (luts.size() != 3) throw Exception(PANO_ERR_ARG, "cudaTransform");
    uint8_t l[768];
    for (int c = 0; c < 3; c++) {
        if (luts[c].size() != 256) throw Exception(PANO_ERR_ARG, "cudaTransform");
        std::memcpy(l + 256 * c, luts[c].data(), 256);
    }
    dst.create(src.rows, src.cols, kType8UC4);
    check(pano_transform_lut3(src.data, src.step, dst.data, dst.step, src.cols, src.rows, l, stream.handle()),
          "cudaTransform");
}

inline void cudaTransform(const GpuMat& src, GpuMat& dst, const std::vector<unsigned char>& lut,
                          Stream& stream = Stream::Null(), const GpuMat& mask = GpuMat())
{
    cudaTransform(src, dst, std::vector<std::vector<unsigned char>>(3, lut), stream, mask);
}

inline void cudaGenerateReprojectMapAndMask(const PhotoParam& param, const Size& srcSize, const Size& dstSize,
                                            GpuMat& xmap, GpuMat& ymap, GpuMat& mask)
{
    xmap.create(dstSize, kType32FC1);
    ymap.create(dstSize, kType32FC1);
    mask.create(dstSize, kType8UC1);
    if (xmap.step != ymap.step) throw Exception(PANO_ERR_ARG, "cudaGenerateReprojectMapAndMask");
    check(pano_generate_map_device(&param, srcSize.width, srcSize.height, dstSize.width, dstSize.height,
                                   (float*)xmap.data, (float*)ymap.data, xmap.step, mask.data, mask.step, nullptr),
          "cudaGenerateReprojectMapAndMask");
}

inline void cudaGenerateReprojectMap(const PhotoParam& param, const Size& srcSize, const Size& dstSize, GpuMat& xmap,
                                     GpuMat& ymap)
{
    GpuMat mask;
    cudaGenerateReprojectMapAndMask(param, srcSize, dstSize, xmap, ymap, mask);
}

inline void cudaGenerateReprojectMapsAndMasks(const std::vector<PhotoParam>& params, const Size& srcSize,
                                              const Size& dstSize, std::vector<GpuMat>& xmaps,
                                              std::vector<GpuMat>& ymaps, std::vector<GpuMat>& masks)
{
    xmaps.resize(params.size());
    ymaps.resize(params.size());
    masks.resize(params.size());
    for (size_t i = 0; i < params.size(); i++)
        cudaGenerateReprojectMapAndMask(params[i], srcSize, dstSize, xmaps[i], ymaps[i], masks[i]);
}

inline void cudaGenerateReprojectMaps(const std::vector<PhotoParam>& params, const Size& srcSize, const Size& dstSize,
                                      std::vector<GpuMat>& xmaps, std::vector<GpuMat>& ymaps)
{
    std::vector<GpuMat> masks;
    cudaGenerateReprojectMapsAndMasks(params, srcSize, dstSize, xmaps, ymaps, masks);
}

namespace detail {
inline void reproject_maps(const GpuMat& src, GpuMat& dst, const GpuMat& xmap, const GpuMat& ymap, int type,
                           Stream& stream, const char* where)
{
    need(src, kType8UC4, where);
    need(xmap, kType32FC1, where);
    need(ymap, kType32FC1, where);
    if (xmap.size() != ymap.size() || xmap.step != ymap.step) throw Exception(PANO_ERR_ARG, where);
    dst.create(xmap.size(), type);
    if (type == kType16SC4)
        check(pano_reproject_to16s(src.data, src.step, src.cols, src.rows, (const float*)xmap.data,
                                   (const float*)ymap.data, xmap.step, (int16_t*)dst.data, dst.step, dst.cols,
                                   dst.rows, stream.handle()),
              where);
    else
        check(pano_reproject(src.data, src.step, src.cols, src.rows, (const float*)xmap.data, (const float*)ymap.data,
                             xmap.step, dst.data, dst.step, dst.cols, dst.rows, stream.handle()),
              where);
}
inline void reproject_param(const GpuMat& src, GpuMat& dst, GpuMat* mask, const Size& dstSize,
                            const PhotoParam& param, int type, Stream& stream, const char* where)
{
    need(src, kType8UC4, where);
    dst.create(dstSize, type);
    if (mask) mask->create(dstSize, kType8UC1);
    check(pano_reproject_params(src.data, src.step, src.cols, src.rows, &param, dstSize.width, dstSize.height, type,
                                dst.data, dst.step, mask ? mask->data : nullptr, mask ? mask->step : 0,
                                stream.handle()),
          where);
}
}  // namespace detail

inline void cudaReproject(const GpuMat& src, GpuMat& dst, const GpuMat& xmap, const GpuMat& ymap,
                          Stream& stream = Stream::Null())
{
    detail::reproject_maps(src, dst, xmap, ymap, kType8UC4, stream, "cudaReproject");
}

inline void cudaReprojectTo16S(const GpuMat& src, GpuMat& dst, const GpuMat& xmap, const GpuMat& ymap,
                               Stream& stream = Stream::Null())
{
    detail::reproject_maps(src, dst, xmap, ymap, kType16SC4, stream, "cudaReprojectTo16S");
}

inline void cudaReproject(const GpuMat& src, GpuMat& dst, const Size& dstSize, const PhotoParam& param,
                          Stream& stream = Stream::Null())
{
    detail::reproject_param(src, dst, nullptr, dstSize, param, kType8UC4, stream, "cudaReproject");
}

inline void cudaReprojectTo16S(const GpuMat& src, GpuMat& dst, const Size& dstSize, const PhotoParam& param,
                               Stream& stream = Stream::Null())
{
    detail::reproject_param(src, dst, nullptr, dstSize, param, kType16SC4, stream, "cudaReprojectTo16S");
}

inline void cudaReproject(const GpuMat& src, GpuMat& dst, GpuMat& mask, const Size& dstSize, const PhotoParam& param,
                          Stream& stream = Stream::Null())
{
    detail::reproject_param(src, dst, &mask, dstSize, param, kType8UC4, stream, "cudaReproject");
}

inline void cudaReprojectTo16S(const GpuMat& src, GpuMat& dst, GpuMat& mask, const Size& dstSize,
                               const PhotoParam& param, Stream& stream = Stream::Null())
{
    detail::reproject_param(src, dst, &mask, dstSize, param, kType16SC4, stream, "cudaReprojectTo16S");
}

inline void cudaReprojectWeightedAccumulateTo32F(const GpuMat& src, GpuMat& dst, const GpuMat& xmap,
                                                 const GpuMat& ymap, const GpuMat& weight,
                                                 Stream& stream = Stream::Null())
{
    const char* where = "cudaReprojectWeightedAccumulateTo32F";
    detail::need(src, kType8UC4, where);
    detail::need(dst, kType32FC4, where);
    detail::need(xmap, kType32FC1, where);
    detail::need(ymap, kType32FC1, where);
    detail::need(weight, kType32FC1, where);
    if (xmap.size() != dst.size() || ymap.size() != dst.size() || weight.size() != dst.size() ||
        xmap.step != ymap.step)
        throw Exception(PANO_ERR_ARG, where);
    check(pano_reproject_weighted_accumulate(src.data, src.step, src.cols, src.rows, (const float*)xmap.data,
                                             (const float*)ymap.data, xmap.step, (const float*)weight.data,
                                             weight.step, (float*)dst.data, dst.step, dst.cols, dst.rows,
                                             stream.handle()),
          where);
}

inline void cvtBGR32ToYUV420P(const GpuMat& bgr32, GpuMat& y, GpuMat& u, GpuMat& v, Stream& stream = Stream::Null())
{
    detail::need(bgr32, kType8UC4, "cvtBGR32ToYUV420P");
    y.create(bgr32.rows, bgr32.cols, kType8UC1);
    u.create(bgr32.rows / 2, bgr32.cols / 2, kType8UC1);
    v.create(bgr32.rows / 2, bgr32.cols / 2, kType8UC1);
    check(pano_bgr32_to_yuv420p(bgr32.data, bgr32.step, bgr32.cols, bgr32.rows, y.data, y.step, u.data, u.step,
                                v.data, v.step, stream.handle()),
          "cvtBGR32ToYUV420P");
}

inline void cvtBGR32ToNV12(const GpuMat& bgr32, GpuMat& y, GpuMat& uv, Stream& stream = Stream::Null())
{
    detail::need(bgr32, kType8UC4, "cvtBGR32ToNV12");
    y.create(bgr32.rows, bgr32.cols, kType8UC1);
    uv.create(bgr32.rows / 2, bgr32.cols, kType8UC1);
    check(pano_bgr32_to_nv12(bgr32.data, bgr32.step, bgr32.cols, bgr32.rows, y.data, y.step, uv.data, uv.step,
                             stream.handle()),
          "cvtBGR32ToNV12");
}

inline void cvtYUV420PToBGR32(const GpuMat& y, const GpuMat& u, const GpuMat& v, GpuMat& bgr32,
                              Stream& stream = Stream::Null())
{
    detail::need(y, kType8UC1, "cvtYUV420PToBGR32");
    bgr32.create(y.rows, y.cols, kType8UC4);
    check(pano_yuv420p_to_bgr32(y.data, y.step, u.data, u.step, v.data, v.step, y.cols, y.rows, bgr32.data,
                                bgr32.step, stream.handle()),
          "cvtYUV420PToBGR32");
}

inline void cvtNV12ToBGR32(const GpuMat& y, const GpuMat& uv, GpuMat& bgr32, Stream& stream = Stream::Null())
{
    detail::need(y, kType8UC1, "cvtNV12ToBGR32");
    bgr32.create(y.rows, y.cols, kType8UC4);
    check(pano_nv12_to_bgr32(y.data, y.step, uv.data, uv.step, y.cols, y.rows, bgr32.data, bgr32.step,
                             stream.handle()),
          "cvtNV12ToBGR32");
}

// ------------------------------------------------------------------ render --

// CudaPanoramaRender2 (Task/RicohUtil.h:193-220, RicohUtil.cpp:1091-1339):
// prepare from a rig file; render host BGRx frames into a device canvas.
// highQualityBlend 1: multiband, blendParam = max levels; 0: linear blend,
// blendParam = radius.  Everything the frame needs stays on the device.
class CudaPanoramaRender2 {
public:
    CudaPanoramaRender2() = default;
    virtual ~CudaPanoramaRender2() { clear(); }
    CudaPanoramaRender2(const CudaPanoramaRender2&) = delete;
    CudaPanoramaRender2& operator=(const CudaPanoramaRender2&) = delete;

    virtual bool prepare(const std::string& path, int highQualityBlend, int blendParam, const Size& srcSize,
                         const Size& dstSize)
    {
        clear();
        // RicohUtil.cpp:1096-1115: W = 2 H, both even; blendParam positive; a rig
        if ((dstSize.width & 1) || (dstSize.height & 1) || dstSize.height * 2 != dstSize.width ||
            dstSize.width <= 0 || blendParam <= 0)
            return false;
        std::vector<PhotoParam> params;
        try {
            if (!loadPhotoParams(path, params)) return false;
        } catch (const Exception&) {
            return false;
        }
        const int n = (int)params.size();
        std::vector<int> sw(n, srcSize.width), sh(n, srcSize.height);
        int st;
        if (highQualityBlend) {
            if (pano_render_create(&mb_) != PANO_OK) return false;
            st = pano_render_prepare(mb_, n, params.data(), sw.data(), sh.data(), dstSize.width, dstSize.height,
                                     blendParam);
        } else {
            if (pano_linear_render_create(&lin_) != PANO_OK) return false;
            st = pano_linear_render_prepare(lin_, n, params.data(), sw.data(), sh.data(), dstSize.width,
                                            dstSize.height, blendParam);
        }
        if (st != PANO_OK) {
            clear();
            return false;
        }
        n_ = n;
        src_ = srcSize;
        dst_ = dstSize;
        frames_.resize(n);
        return true;
    }

    // src: n host CV_8UC4 frames of srcSize; luts: empty, or n x 3 x 256
    // (anything else skips the correction, RicohUtil.cpp:1240-1264).
    virtual bool render(const std::vector<Mat>& src, GpuMat& dst,
                        const std::vector<std::vector<std::vector<unsigned char>>>& luts =
                            std::vector<std::vector<std::vector<unsigned char>>>())
    {
        if (!n_ || (int)src.size() != n_) return false;
        for (const Mat& m : src)
            if (m.type() != kType8UC4 || m.size() != src_) return false;
        bool correct = (int)luts.size() == n_;
        for (size_t i = 0; correct && i < luts.size(); i++) {
            correct = luts[i].size() == 3;
            for (size_t c = 0; correct && c < 3; c++) correct = luts[i][c].size() == 256;
        }
        std::vector<uint8_t> l;
        if (correct) {
            l.resize((size_t)n_ * 768);
            for (int i = 0; i < n_; i++)
                for (int c = 0; c < 3; c++) std::memcpy(&l[(size_t)i * 768 + 256 * c], luts[i][c].data(), 256);
        }
        try {
            for (int i = 0; i < n_; i++) frames_[i].upload(src[i], stream_);
            dst.create(dst_, kType8UC4);
        } catch (const Exception&) {
            return false;
        }
        const auto p = detail::ptrs(frames_);
        const auto s = detail::steps(frames_);
        const int st = mb_ ? pano_render_render(mb_, n_, p.data(), s.data(), correct ? l.data() : nullptr, dst.data,
                                                dst.step, stream_.handle())
                           : pano_linear_render_render(lin_, n_, p.data(), s.data(), correct ? l.data() : nullptr,
                                                       dst.data, dst.step, stream_.handle());
        if (st != PANO_OK) return false;
        return hipStreamSynchronize(stream_.handle()) == hipSuccess;
    }

    virtual void clear()
    {
        if (mb_) pano_render_destroy(mb_);
        if (lin_) pano_linear_render_destroy(lin_);
        mb_ = nullptr;
        lin_ = nullptr;
        n_ = 0;
        frames_.clear();
    }
    virtual int getNumImages() const { return n_; }

private:
    pano_render* mb_ = nullptr;
    pano_linear_render* lin_ = nullptr;
    int n_ = 0;
    Size src_, dst_;
    std::vector<GpuMat> frames_;
    Stream stream_{true};
};

}  // namespace pano
