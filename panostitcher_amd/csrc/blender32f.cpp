// blender32f.cpp -- CudaTilingMultibandBlendFast32F (CudaAccel/CudaInterface.h:
// 84-101, CudaAccel/CudaMultibandBlend.cpp:1041-1179): the float-weight
// multiband blender.  prepare builds, on the device, each camera's float
// weight pyramid (from its seam-unique mask) and inverse alpha pyramid (from
// its mask), plus the inverse weight sums when the masks leave parts of the
// canvas uncovered; blend runs each camera's float Gaussian / Laplacian
// pyramid through one set of resident planes and accumulates, then collapses.
// The arithmetic is the reference kernels' (see kernels_f32.hip).
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/panohip.h"
#include "pano_host.hpp"
#include "pano_launch.hpp"

using namespace pano;

#define F_TRY(expr)                                          \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return PANO_ERR_HIP;           \
    } while (0)

namespace {

// A pitched device plane of one level.
struct FPlane {
    void* p = nullptr;
    size_t step = 0;
    int w = 0, h = 0;
};

struct Dev {
    std::vector<void*> allocs;
    ~Dev()
    {
        for (void* p : allocs) (void)hipFree(p);
    }
    int plane(FPlane& f, int w, int h, size_t px)
    {
        f.w = w;
        f.h = h;
        f.step = ((size_t)w * px + 255) / 256 * 256;
        if (hipMalloc(&f.p, f.step * h) != hipSuccess) {
            f.p = nullptr;
            return PANO_ERR_NOMEM;
        }
        allocs.push_back(f.p);
        if (hipMemset(f.p, 0, f.step * h) != hipSuccess) return PANO_ERR_HIP;
        return PANO_OK;
    }
};

}  // namespace

struct pano_blender32f {
    bool ok = false;
    int n = 0, rows = 0, cols = 0, levels = 0;
    bool full = false;
    std::unique_ptr<Dev> dev;
    std::vector<int> lw, lh;
    std::vector<FPlane> wts;     // n * (L + 1) float weights
    std::vector<FPlane> ainv;    // n * (L + 1) float inverse alphas (level 0 unused)
    std::vector<FPlane> winv;    // L + 1 inverse weight sums (partial coverage)
    std::vector<FPlane> G, R;    // L + 1 float4 Gaussian and result planes
    uint8_t* cover = nullptr;    // canvas: some mask covers the pixel (partial coverage)
};

extern "C" {

int pano_blender32f_create(pano_blender32f** out)
{
    if (!out) return PANO_ERR_ARG;
    *out = new (std::nothrow) pano_blender32f();
    return *out ? PANO_OK : PANO_ERR_NOMEM;
}

void pano_blender32f_destroy(pano_blender32f* b) { delete b; }

int pano_blender32f_prepare(pano_blender32f* b, int num, int rows, int cols,
                            const uint8_t* const* masks, const size_t* mask_steps, int max_levels,
                            int min_length)
{
    if (!b) return PANO_ERR_ARG;
    b->ok = false;
    // CudaMultibandBlend.cpp:1043-1060: non-empty, at most 255 8-bit masks of one size
    if (num <= 0 || num > 255 || rows <= 0 || cols <= 0 || !masks || !mask_steps) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++)
        if (!masks[i] || mask_steps[i] < (size_t)cols) return PANO_ERR_ARG;
    const size_t P = (size_t)rows * cols;
    if ((size_t)cols * 16 >= (1u << 31) || P * 16 > 0xffffffffull * 4) return PANO_ERR_UNSUPPORTED;
    b->dev.reset(new (std::nothrow) Dev());
    if (!b->dev) return PANO_ERR_NOMEM;
    Dev& d = *b->dev;
    const int L = true_num_levels(cols, rows, max_levels, min_length);
    if (L < 0) return PANO_ERR_ARG;  // a side shorter than min_length: no level 0 pyramid
    b->n = num;
    b->rows = rows;
    b->cols = cols;
    b->levels = L;
    b->lw.assign(L + 1, cols);
    b->lh.assign(L + 1, rows);
    for (int j = 1; j <= L; j++) {
        b->lw[j] = (b->lw[j - 1] + 1) / 2;
        b->lh[j] = (b->lh[j - 1] + 1) / 2;
    }
    std::vector<uint8_t> unique(P * num);
    if (device_nonintersecting_masks(num, rows, cols, masks, mask_steps, unique.data()) != hipSuccess)
        return PANO_ERR_HIP;
    std::vector<uint8_t> any(P, 0);
    for (int i = 0; i < num; i++)
        for (int y = 0; y < rows; y++)
            for (int x = 0; x < cols; x++)
                if (masks[i][(size_t)y * mask_steps[i] + x]) any[(size_t)y * cols + x] = 1;
    b->full = true;
    for (size_t p = 0; p < P; p++)
        if (!any[p]) {
            b->full = false;
            break;
        }
    // device scratch for the prepare: one mask, two binary alpha planes
    uint8_t* dmask = nullptr;
    if (hipMalloc((void**)&dmask, P) != hipSuccess) return PANO_ERR_NOMEM;
    d.allocs.push_back(dmask);
    FPlane bin[2];
    for (auto& f : bin) {
        int e = d.plane(f, cols, rows, 4);
        if (e) return e;
    }
    b->wts.assign((size_t)num * (L + 1), FPlane());
    b->ainv.assign((size_t)num * (L + 1), FPlane());
    hipStream_t s = nullptr;
    for (int i = 0; i < num; i++) {
        for (int j = 0; j <= L; j++) {
            int e = d.plane(b->wts[(size_t)i * (L + 1) + j], b->lw[j], b->lh[j], 4);
            if (!e && j > 0) e = d.plane(b->ainv[(size_t)i * (L + 1) + j], b->lw[j], b->lh[j], 4);
            if (e) return e;
        }
        // alpha: the mask; weights: the seam-unique mask
        F_TRY(hipMemcpy2D(dmask, cols, masks[i], mask_steps[i], cols, rows, hipMemcpyHostToDevice));
        F_TRY(launch_f32_from_mask(dmask, cols, (float*)bin[0].p, bin[0].step, cols, rows, s));
        F_TRY(hipMemcpy(dmask, unique.data() + P * i, P, hipMemcpyHostToDevice));
        FPlane& w0 = b->wts[(size_t)i * (L + 1)];
        F_TRY(launch_f32_from_mask(dmask, cols, (float*)w0.p, w0.step, cols, rows, s));
        int cur = 0;
        for (int j = 0; j < L; j++) {
            FPlane& a = b->ainv[(size_t)i * (L + 1) + j + 1];
            FPlane& nb = bin[cur ^ 1];
            F_TRY(launch_f32_down(1, false, bin[cur].p, b->lw[j], b->lh[j], bin[cur].step, a.p, a.w, a.h,
                                  a.step, nullptr, 0, s));
            F_TRY(launch_f32_alpha((float*)a.p, a.step, (float*)nb.p, nb.step, a.w, a.h, s));
            cur ^= 1;
            const FPlane& wj = b->wts[(size_t)i * (L + 1) + j];
            FPlane& wn = b->wts[(size_t)i * (L + 1) + j + 1];
            F_TRY(launch_f32_down(1, false, wj.p, wj.w, wj.h, wj.step, wn.p, wn.w, wn.h, wn.step, nullptr, 0,
                                  s));
        }
    }
    b->winv.assign(L + 1, FPlane());
    if (!b->full) {
        for (int j = 0; j <= L; j++) {
            int e = d.plane(b->winv[j], b->lw[j], b->lh[j], 4);
            if (e) return e;
            for (int i = 0; i < num; i++) {
                const FPlane& w = b->wts[(size_t)i * (L + 1) + j];
                F_TRY(launch_f32_wsum((const float*)w.p, w.step, (float*)b->winv[j].p, b->winv[j].step,
                                      w.w, w.h, false, s));
            }
            F_TRY(launch_f32_wsum(nullptr, 0, (float*)b->winv[j].p, b->winv[j].step, b->lw[j], b->lh[j],
                                  true, s));
        }
        if (hipMalloc((void**)&b->cover, P) != hipSuccess) return PANO_ERR_NOMEM;
        d.allocs.push_back(b->cover);
        F_TRY(hipMemcpy(b->cover, any.data(), P, hipMemcpyHostToDevice));
    }
    b->G.assign(L + 1, FPlane());
    b->R.assign(L + 1, FPlane());
    for (int j = 0; j <= L; j++) {
        int e = d.plane(b->R[j], b->lw[j], b->lh[j], 16);
        if (!e && j > 0) e = d.plane(b->G[j], b->lw[j], b->lh[j], 16);
        if (e) return e;
    }
    F_TRY(hipDeviceSynchronize());
    b->ok = true;
    return PANO_OK;
}

int pano_blender32f_blend(pano_blender32f* b, int num, const void* const* images, const size_t* steps,
                          int type, uint8_t* out, size_t out_step, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    if (num != b->n || !images || !steps || !out) return PANO_ERR_ARG;
    if (type != PANO_8UC4 && type != PANO_32FC4) return PANO_ERR_ARG;
    const size_t px = type == PANO_8UC4 ? 4 : 16;
    if (out_step < (size_t)b->cols * 4 || out_step % 4) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++)
        if (!images[i] || steps[i] < (size_t)b->cols * px || steps[i] % px) return PANO_ERR_ARG;
    const int L = b->levels;
    const bool u8 = type == PANO_8UC4;
    for (int j = 0; j <= L; j++)
        F_TRY(launch_zero(b->R[j].p, b->R[j].step * b->R[j].h, stream));
    for (int i = 0; i < num; i++) {
        // Gaussian pyramid: G_{j+1} = alpha_{j+1} pyrDown(G_j), G_0 the image
        for (int j = 0; j < L; j++) {
            const void* src = j == 0 ? images[i] : b->G[j].p;
            const size_t sstep = j == 0 ? steps[i] : b->G[j].step;
            const FPlane& a = b->ainv[(size_t)i * (L + 1) + j + 1];
            F_TRY(launch_f32_down(4, j == 0 && u8, src, b->lw[j], b->lh[j], sstep, b->G[j + 1].p,
                                  b->lw[j + 1], b->lh[j + 1], b->G[j + 1].step, (const float*)a.p, a.step,
                                  stream));
        }
        // R_j += W_ij (G_j - up(G_{j+1}))
        for (int j = 0; j <= L; j++) {
            const void* g = j == 0 ? images[i] : b->G[j].p;
            const size_t gstep = j == 0 ? steps[i] : b->G[j].step;
            const FPlane& w = b->wts[(size_t)i * (L + 1) + j];
            F_TRY(launch_f32_lap(j == 0 && u8, g, gstep, j < L ? b->G[j + 1].p : nullptr,
                                 j < L ? b->G[j + 1].step : 0, j < L ? b->lw[j + 1] : 0,
                                 j < L ? b->lh[j + 1] : 0, (const float*)w.p, w.step, b->R[j].p,
                                 b->R[j].step, b->lw[j], b->lh[j], stream));
        }
    }
    // collapse from the top, the inverse weight sums applied on the way
    for (int j = L; j >= 0; j--) {
        const FPlane* wi = b->full ? nullptr : &b->winv[j];
        F_TRY(launch_f32_collapse(j < L ? b->R[j + 1].p : nullptr, j < L ? b->R[j + 1].step : 0,
                                  j < L ? b->lw[j + 1] : 0, j < L ? b->lh[j + 1] : 0,
                                  wi ? (const float*)wi->p : nullptr, wi ? wi->step : 0, b->R[j].p,
                                  b->R[j].step, b->lw[j], b->lh[j], stream));
    }
    F_TRY(launch_f32_out(b->R[0].p, b->R[0].step, b->full ? nullptr : b->cover, b->cols, out, out_step,
                         b->cols, b->rows, stream));
    return PANO_OK;
}

int pano_blender32f_num_levels(const pano_blender32f* b) { return b && b->ok ? b->levels : -1; }

}  // extern "C"
