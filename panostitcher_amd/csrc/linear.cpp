// linear.cpp -- the linear-blend mode (highQualityBlend = 0) on the device:
//   pano_linear_weights      getWeightsLinearBlend32F / getWeightsLinearBlend
//                            (Blend/LinearBlend.cpp:17-73, 76-148)
//   pano_linear_blender_*    CudaTilingLinearBlend (CudaAccel/CudaInterface.h:
//                            103-115, Blend/CudaLinearBlend.cpp:15-101)
//   pano_linear_render_*     the linear branch of CudaPanoramaRender2 /
//                            CPUPanoramaRender (Task/RicohUtil.cpp:1175-1180,
//                            1269-1287, 1543-1558): weights at prepare, then per
//                            frame one launch that samples every camera listed for
//                            a canvas tile and sums the weighted samples.
// The seams (getNonIntersectingMasks), the Gaussian blur, the weights and the
// per-frame work all run on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/panohip.h"
#include "pano_host.hpp"
#include "pano_launch.hpp"

using namespace pano;

namespace pano {
// maps.cpp: getReprojectMap32FAndMask on the device (cudaGenerateReprojectMapAndMask).
int generate_map_device(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h, float* xmap,
                        float* ymap, size_t map_step, uint8_t* mask, size_t mask_step, hipStream_t s);
}  // namespace pano

#define L_TRY(expr)                                   \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return PANO_ERR_HIP;    \
    } while (0)

namespace {

// getGaussianKernel(n, sigma, CV_32F) (float taps from the double exponential,
// normalised in double), then the 8-bit fixed point of
// createSeparableLinearFilter: K = cvRound(256 k).
void gaussian_taps(int n, double sigma, std::vector<int>& K)
{
    std::vector<float> cf(n);
    const double sx = sigma > 0 ? sigma : ((n - 1) * 0.5 - 1) * 0.3 + 0.8;
    const double scale2x = -0.5 / (sx * sx);
    double sum = 0;
    for (int i = 0; i < n; i++) {
        const double x = i - (n - 1) * 0.5;
        cf[i] = (float)std::exp(scale2x * x * x);
        sum += cf[i];
    }
    sum = 1. / sum;
    K.resize(n);
    for (int i = 0; i < n; i++) {
        cf[i] = (float)(cf[i] * sum);
        K[i] = (int)std::nearbyint(cf[i] * 256.0f);
    }
}

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes)
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            p = nullptr;
            return PANO_ERR_NOMEM;
        }
        return PANO_OK;
    }
    template <class T>
    T* as() const
    {
        return (T*)p;
    }
};

int check_masks(int n, int rows, int cols, const uint8_t* const* masks, const size_t* steps)
{
    if (n <= 0 || n > kMaxSrc || rows <= 0 || cols <= 0 || !masks || !steps) return PANO_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!masks[i] || steps[i] < (size_t)cols) return PANO_ERR_ARG;
    if ((size_t)rows * cols * 4 > 0x7fffffffull * 2) return PANO_ERR_UNSUPPORTED;
    return PANO_OK;
}

// Weights of n cameras from device masks (n planes of rows x cols bytes,
// pitch cols): seams on the device, each unique mask blurred, ANDed with its
// mask, normalised.  out: n planes of rows x pitch elements (float or int32),
// rows * pitch elements apart.  Synchronous.
int device_weights_dev(int n, int rows, int cols, const uint8_t* dmasks, int radius, bool as_int, int pitch,
                       void* out)
{
    const size_t P = (size_t)rows * pitch, Pc = (size_t)rows * cols;
    std::vector<int> K;
    const int nt = 2 * radius + 1;
    gaussian_taps(nt, radius / 3.0, K);
    int k0 = 0, k1 = nt;
    while (k0 < k1 && K[k0] == 0) k0++;
    while (k1 > k0 && K[k1 - 1] == 0) k1--;
    DevBuf packed, dmask, duniq, ddist, drows, dK;
    int e = packed.alloc(Pc * n);
    if (!e) e = dmask.alloc(P * n);
    if (!e) e = duniq.alloc(P * n);
    if (!e) e = ddist.alloc(P * n);
    if (!e) e = drows.alloc(P * sizeof(int32_t));
    if (!e) e = dK.alloc(sizeof(int) * nt);
    if (e) return e;
    L_TRY(hipMemcpy(packed.p, dmasks, Pc * n, hipMemcpyDeviceToDevice));
    L_TRY(device_seams_inplace(n, rows, cols, packed.as<uint8_t>()));
    L_TRY(hipMemset(dmask.p, 0, P * n));
    L_TRY(hipMemset(duniq.p, 0, P * n));
    for (int i = 0; i < n; i++) {
        L_TRY(hipMemcpy2D(dmask.as<uint8_t>() + P * i, pitch, dmasks + Pc * i, cols, cols, rows,
                          hipMemcpyDeviceToDevice));
        L_TRY(hipMemcpy2D(duniq.as<uint8_t>() + P * i, pitch, packed.as<uint8_t>() + Pc * i, cols, cols, rows,
                          hipMemcpyDeviceToDevice));
    }
    L_TRY(hipMemcpy(dK.p, K.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
    for (int i = 0; i < n; i++)
        L_TRY(launch_blur_u8(duniq.as<uint8_t>() + P * i, pitch, cols, rows, dK.as<int>(), nt, k0, k1,
                             drows.as<int32_t>(), dmask.as<uint8_t>() + P * i, pitch,
                             ddist.as<uint8_t>() + P * i, nullptr));
    L_TRY(launch_linear_weights(as_int, ddist.as<uint8_t>(), pitch, P, n, cols, rows, out, nullptr));
    L_TRY(hipDeviceSynchronize());
    return PANO_OK;
}

// The same from host masks (n planes with the given row steps).
int device_weights(int n, int rows, int cols, const uint8_t* const* masks, const size_t* steps, int radius,
                   bool as_int, int pitch, void* out)
{
    const size_t Pc = (size_t)rows * cols;
    DevBuf dm;
    int e = dm.alloc(Pc * n);
    if (e) return e;
    for (int i = 0; i < n; i++)
        L_TRY(hipMemcpy2D(dm.as<uint8_t>() + Pc * i, cols, masks[i], steps[i], cols, rows, hipMemcpyHostToDevice));
    return device_weights_dev(n, rows, cols, dm.as<uint8_t>(), radius, as_int, pitch, out);
}

int canvas_pitch(int cols) { return (cols + 63) / 64 * 64; }

}  // namespace

struct pano_linear_blender {
    bool ok = false;
    int n = 0, rows = 0, cols = 0, pitch = 0;
    DevBuf w;   // n int32 planes
};

struct pano_linear_render {
    bool ok = false;
    int n = 0, dst_w = 0, dst_h = 0, pitch = 0;
    std::vector<int> src_w, src_h;
    DevBuf maps, wts, dcams, dtiles, dcamlist, dluts;
    std::vector<DevBuf> lsrc;           // LUT-mapped frame copies (allocated on first use)
    std::vector<size_t> lstep;
    int ntiles = 0;
    uint8_t* pinned = nullptr;          // staging of the frame's LUTs
    hipEvent_t lut_ev = nullptr;
    bool lut_pending = false;
    ~pano_linear_render()
    {
        if (lut_pending) (void)hipEventSynchronize(lut_ev);
        if (pinned) (void)hipHostFree(pinned);
        if (lut_ev) (void)hipEventDestroy(lut_ev);
    }
};

extern "C" {

int pano_linear_weights(int num, int rows, int cols, const uint8_t* const* masks, const size_t* mask_steps,
                        int radius, int as_int, void* const* out, const size_t* out_steps)
{
    int e = check_masks(num, rows, cols, masks, mask_steps);
    if (e) return e;
    if (radius < 0 || !out || !out_steps) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++)
        if (!out[i] || out_steps[i] < (size_t)cols * 4) return PANO_ERR_ARG;
    const int pitch = canvas_pitch(cols);
    DevBuf w;
    e = w.alloc((size_t)rows * pitch * 4 * num);
    if (e) return e;
    e = device_weights(num, rows, cols, masks, mask_steps, radius, as_int != 0, pitch, w.p);
    if (e) return e;
    for (int i = 0; i < num; i++)
        L_TRY(hipMemcpy2D(out[i], out_steps[i], w.as<char>() + (size_t)rows * pitch * 4 * i, (size_t)pitch * 4,
                          (size_t)cols * 4, rows, hipMemcpyDeviceToDevice));
    return PANO_OK;
}

int pano_linear_blender_create(pano_linear_blender** out)
{
    if (!out) return PANO_ERR_ARG;
    *out = new (std::nothrow) pano_linear_blender();
    return *out ? PANO_OK : PANO_ERR_NOMEM;
}

void pano_linear_blender_destroy(pano_linear_blender* b) { delete b; }

int pano_linear_blender_prepare(pano_linear_blender* b, int num, int rows, int cols, const uint8_t* const* masks,
                                const size_t* mask_steps, int radius)
{
    if (!b) return PANO_ERR_ARG;
    b->ok = false;
    // CudaLinearBlend.cpp:17-31: non-empty, radius >= 1, 8-bit masks of one size
    if (radius < 1) return PANO_ERR_ARG;
    int e = check_masks(num, rows, cols, masks, mask_steps);
    if (e) return e;
    b->n = num;
    b->rows = rows;
    b->cols = cols;
    b->pitch = canvas_pitch(cols);
    e = b->w.alloc((size_t)rows * b->pitch * 4 * num);
    if (e) return e;
    e = device_weights(num, rows, cols, masks, mask_steps, radius, true, b->pitch, b->w.p);
    if (e) return e;
    b->ok = true;
    return PANO_OK;
}

int pano_linear_blender_blend(pano_linear_blender* b, int num, const uint8_t* const* images, const size_t* steps,
                              uint8_t* out, size_t out_step, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    if (num != b->n || !images || !steps || !out || out_step < (size_t)b->cols * 4 || out_step % 4 ||
        out_step > 0xffffffffu)
        return PANO_ERR_ARG;
    SrcTable st{};
    for (int i = 0; i < num; i++) {
        if (!images[i] || steps[i] < (size_t)b->cols * 4 || steps[i] % 4 || steps[i] > 0xffffffffu)
            return PANO_ERR_ARG;
        st.p[i] = images[i];
        st.step[i] = (uint32_t)steps[i];
    }
    L_TRY(launch_linear_blend32s(st, num, b->w.as<int32_t>(), b->pitch, (size_t)b->rows * b->pitch, out, out_step,
                                 b->cols, b->rows, stream));
    return PANO_OK;
}

int pano_linear_render_create(pano_linear_render** out)
{
    if (!out) return PANO_ERR_ARG;
    *out = new (std::nothrow) pano_linear_render();
    return *out ? PANO_OK : PANO_ERR_NOMEM;
}

void pano_linear_render_destroy(pano_linear_render* r) { delete r; }

}  // extern "C"

namespace {

// The render's prepare from device maps (camera i: x plane at 2 i P, y plane
// at (2 i + 1) P, pitch dst_w, P = dst_w dst_h) and device masks (n planes of
// P bytes): weights, (x, y) map pairs, per-tile camera lists.  Arguments are
// checked by the callers.
int linear_prepare_dev(pano_linear_render* r, int num, const int* src_w, const int* src_h, int dst_w, int dst_h,
                       const float* dmaps, const uint8_t* dmasks, int radius)
{
    const int pitch = canvas_pitch(dst_w);
    const size_t P = (size_t)dst_h * pitch, Pc = (size_t)dst_h * dst_w;
    r->n = num;
    r->dst_w = dst_w;
    r->dst_h = dst_h;
    r->pitch = pitch;
    r->src_w.assign(src_w, src_w + num);
    r->src_h.assign(src_h, src_h + num);
    int e = r->wts.alloc(P * 4 * num);
    if (!e) e = r->maps.alloc(P * 8 * num);
    if (e) return e;
    e = device_weights_dev(num, dst_h, dst_w, dmasks, radius, false, pitch, r->wts.p);
    if (e) return e;
    for (int i = 0; i < num; i++)
        L_TRY(launch_interleave_map(dmaps + 2 * Pc * i, dmaps + (2 * (size_t)i + 1) * Pc, dst_w,
                                    (float2*)(r->maps.as<float>() + 2 * P * i), pitch, dst_w, dst_h, nullptr));
    std::vector<LinCam> cams(num);
    for (int i = 0; i < num; i++) {
        cams[i].map = (const float2*)(r->maps.as<float>() + 2 * P * i);
        cams[i].w = r->wts.as<float>() + P * i;
        cams[i].pitch = pitch;
        cams[i].src = i;
        cams[i].sw = src_w[i];
        cams[i].sh = src_h[i];
    }
    // tiles with the cameras whose weight is non-zero somewhere in them, in camera order
    const int tw = (dst_w + kLinTW - 1) / kLinTW, th = (dst_h + kLinTH - 1) / kLinTH;
    std::vector<uint8_t> hit((size_t)tw * th * num, 0);
    {
        DevBuf dhit;
        e = dhit.alloc(hit.size());
        if (e) return e;
        L_TRY(hipMemset(dhit.p, 0, hit.size()));
        L_TRY(launch_lin_tile_hits(r->wts.as<float>(), pitch, P, num, dst_w, dst_h, tw, dhit.as<uint8_t>(), nullptr));
        L_TRY(hipMemcpy(hit.data(), dhit.p, hit.size(), hipMemcpyDeviceToHost));
    }
    std::vector<LinTile> tiles;
    std::vector<int> camlist;
    for (int ty = 0; ty < th; ty++)
        for (int tx = 0; tx < tw; tx++) {
            LinTile t{tx * kLinTW, ty * kLinTH, (int)camlist.size(), 0};
            for (int i = 0; i < num; i++)
                if (hit[((size_t)ty * tw + tx) * num + i]) camlist.push_back(i);
            t.cnt = (int)camlist.size() - t.off;
            tiles.push_back(t);   // tiles no camera meets still write zeros
        }
    r->ntiles = (int)tiles.size();
    e = r->dcams.alloc(sizeof(LinCam) * num);
    if (!e) e = r->dtiles.alloc(sizeof(LinTile) * tiles.size());
    if (!e) e = r->dcamlist.alloc(sizeof(int) * std::max<size_t>(camlist.size(), 1));
    if (!e) e = r->dluts.alloc((size_t)num * 768);
    if (e) return e;
    L_TRY(hipMemcpy(r->dcams.p, cams.data(), sizeof(LinCam) * num, hipMemcpyHostToDevice));
    L_TRY(hipMemcpy(r->dtiles.p, tiles.data(), sizeof(LinTile) * tiles.size(), hipMemcpyHostToDevice));
    if (!camlist.empty())
        L_TRY(hipMemcpy(r->dcamlist.p, camlist.data(), sizeof(int) * camlist.size(), hipMemcpyHostToDevice));
    if (!r->pinned) {
        if (hipHostMalloc((void**)&r->pinned, (size_t)kMaxSrc * 768, hipHostMallocDefault) != hipSuccess) {
            r->pinned = nullptr;
            return PANO_ERR_NOMEM;
        }
        L_TRY(hipEventCreateWithFlags(&r->lut_ev, hipEventDisableTiming));
    }
    r->lsrc.clear();
    r->lsrc.resize(num);
    r->lstep.assign(num, 0);
    L_TRY(hipDeviceSynchronize());
    r->ok = true;
    return PANO_OK;
}

int linear_prepare_args(int num, const int* src_w, const int* src_h, int dst_w, int dst_h, int radius)
{
    if (num <= 0 || num > kMaxSrc || !src_w || !src_h || radius < 0) return PANO_ERR_ARG;
    if (dst_w <= 0 || dst_h <= 0 || (size_t)dst_w * 4 >= (1u << 24)) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++)
        if (src_w[i] <= 0 || src_h[i] <= 0) return PANO_ERR_ARG;
    return PANO_OK;
}

}  // namespace

extern "C" {

int pano_linear_render_prepare_maps(pano_linear_render* r, int num, const int* src_w, const int* src_h, int dst_w,
                                    int dst_h, const float* const* xmaps, const float* const* ymaps,
                                    const uint8_t* const* masks, int radius)
{
    if (!r) return PANO_ERR_ARG;
    r->ok = false;
    int e = linear_prepare_args(num, src_w, src_h, dst_w, dst_h, radius);
    if (e) return e;
    if (!xmaps || !ymaps || !masks) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++)
        if (!xmaps[i] || !ymaps[i] || !masks[i]) return PANO_ERR_ARG;
    const size_t Pc = (size_t)dst_h * dst_w;
    DevBuf dmaps, dmasks;
    e = dmaps.alloc(Pc * 8 * num);
    if (!e) e = dmasks.alloc(Pc * num);
    if (e) return e;
    for (int i = 0; i < num; i++) {
        L_TRY(hipMemcpy(dmaps.as<float>() + 2 * Pc * i, xmaps[i], Pc * 4, hipMemcpyHostToDevice));
        L_TRY(hipMemcpy(dmaps.as<float>() + (2 * (size_t)i + 1) * Pc, ymaps[i], Pc * 4, hipMemcpyHostToDevice));
        L_TRY(hipMemcpy(dmasks.as<uint8_t>() + Pc * i, masks[i], Pc, hipMemcpyHostToDevice));
    }
    return linear_prepare_dev(r, num, src_w, src_h, dst_w, dst_h, dmaps.as<float>(), dmasks.as<uint8_t>(), radius);
}

int pano_linear_render_prepare(pano_linear_render* r, int num, const pano_photo_param* params, const int* src_w,
                               const int* src_h, int dst_w, int dst_h, int radius)
{
    if (!r) return PANO_ERR_ARG;
    r->ok = false;
    int e = linear_prepare_args(num, src_w, src_h, dst_w, dst_h, radius);
    if (e) return e;
    if (!params) return PANO_ERR_ARG;
    const size_t Pc = (size_t)dst_h * dst_w;
    DevBuf dmaps, dmasks;
    e = dmaps.alloc(Pc * 8 * num);
    if (!e) e = dmasks.alloc(Pc * num);
    if (e) return e;
    for (int i = 0; i < num && !e; i++)
        e = generate_map_device(params[i], src_w[i], src_h[i], dst_w, dst_h, dmaps.as<float>() + 2 * Pc * i,
                                dmaps.as<float>() + (2 * (size_t)i + 1) * Pc, (size_t)dst_w * 4,
                                dmasks.as<uint8_t>() + Pc * i, (size_t)dst_w, nullptr);
    if (e) return e;
    return linear_prepare_dev(r, num, src_w, src_h, dst_w, dst_h, dmaps.as<float>(), dmasks.as<uint8_t>(), radius);
}

int pano_linear_render_render(pano_linear_render* r, int num, const uint8_t* const* srcs, const size_t* src_steps,
                              const uint8_t* luts, uint8_t* dst, size_t dst_step, hipStream_t stream)
{
    if (!r) return PANO_ERR_ARG;
    if (!r->ok) return PANO_ERR_STATE;
    if (num != r->n || !srcs || !src_steps || !dst) return PANO_ERR_ARG;
    if (dst_step < (size_t)r->dst_w * 4 || dst_step % 4 || dst_step > 0xffffffffu) return PANO_ERR_ARG;
    SrcTable st{};
    for (int i = 0; i < num; i++) {
        if (!srcs[i] || src_steps[i] < (size_t)r->src_w[i] * 4 || src_steps[i] % 4 || src_steps[i] >= (1u << 24) ||
            src_steps[i] * (uint64_t)r->src_h[i] > 0xffffffffu)
            return PANO_ERR_ARG;
        st.p[i] = srcs[i];
        st.step[i] = (uint32_t)src_steps[i];
    }
    if (luts) {
        // transform (GainAdjust.cpp:295-346) of each frame into a resident copy
        if (r->lut_pending) L_TRY(hipEventSynchronize(r->lut_ev));
        memcpy(r->pinned, luts, (size_t)num * 768);
        L_TRY(hipMemcpyAsync(r->dluts.p, r->pinned, (size_t)num * 768, hipMemcpyHostToDevice, stream));
        L_TRY(hipEventRecord(r->lut_ev, stream));
        r->lut_pending = true;
        for (int i = 0; i < num; i++) {
            const size_t step = (size_t)r->src_w[i] * 4;
            if (!r->lsrc[i].p) {
                int e = r->lsrc[i].alloc(step * r->src_h[i]);
                if (e) return e;
                r->lstep[i] = step;
            }
            L_TRY(launch_transform(srcs[i], src_steps[i], r->lsrc[i].as<uint8_t>(), step, r->src_w[i], r->src_h[i],
                                   r->dluts.as<uint8_t>() + (size_t)i * 768, stream));
            st.p[i] = r->lsrc[i].as<uint8_t>();
            st.step[i] = (uint32_t)step;
        }
    }
    L_TRY(launch_linear_render(r->dtiles.as<LinTile>(), r->ntiles, r->dcamlist.as<int>(), r->dcams.as<LinCam>(), st,
                               dst, dst_step, r->dst_w, r->dst_h, stream));
    return PANO_OK;
}

int pano_linear_render_weights(const pano_linear_render* r, int cam, float* out)
{
    if (!r || !out || cam < 0 || cam >= r->n) return PANO_ERR_ARG;
    if (!r->ok) return PANO_ERR_STATE;
    const size_t P = (size_t)r->dst_h * r->pitch;
    L_TRY(hipMemcpy2D(out, (size_t)r->dst_w * 4, r->wts.as<float>() + P * cam, (size_t)r->pitch * 4,
                      (size_t)r->dst_w * 4, r->dst_h, hipMemcpyDeviceToHost));
    return PANO_OK;
}

}  // extern "C"
