// kernels_custom.hip -- the custom-mask blend of the 16-bit blender on gfx950:
// CudaTilingMultibandBlendFast::blend(images, masks, blendImage)
// (CudaAccel/CudaInterface.h:55-56, CudaAccel/CudaMultibandBlend.cpp:825-893) with
// the integer arithmetic of the CPU TilingMultibandBlendFast::blend(images,
// masks, blendImage) (Blend/MultiBandBlend.cpp:1033-1102).
//
// Per call the masks give every camera a weight pyramid (256 where the mask is
// set, then the int16 pyrDown with rounding, createGaussPyramid :64-74) and
// the weight sums; the image pyramids are the prepared ones (alphas from the
// prepare masks).  Each level then sums W_ij L_ij over all cameras, divides by
// the weight sum (0 -> 1, C truncation, normalize :203-241) and adds up(C_{j+1});
// level 0 writes the 8-bit output, zero where no call mask is set (when the
// masks cover the canvas that zeroes nothing, as the reference's countNonZero
// test skips the setTo).  Weight planes are full level-size int16, camera
// after camera (cam_stride elements apart per level), with the level-0 row
// pitch `pitch` (elements) at every level, as are the sums and the cover bytes.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

// W_0 of every camera and the union of the masks (the masks are SrcTable
// entries 0..n-1, one byte per pixel).
__global__ __launch_bounds__(256) void k_cw_init(SrcTable masks, int n, int w, int h,
                                                 int16_t* __restrict__ w0, size_t cam_stride, int pitch,
                                                 uint8_t* __restrict__ cover)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    int any = 0;
    for (int i = 0; i < n; i++) {
        const int m = ldg(masks.p[i] + (size_t)Y * masks.step[i] + X) != 0;
        any |= m;
        w0[i * cam_stride + (size_t)Y * pitch + X] = m ? 256 : 0;
    }
    cover[(size_t)Y * pitch + X] = (uint8_t)any;
}

// pyrDown_<FixPtCast<short, 8>> of one camera's weights (blockIdx.z): the 5 x 5
// binomial sum (columns WRAP, rows REFLECT_101), (s + 128) >> 8.
__global__ __launch_bounds__(256) void k_cw_down(const int16_t* __restrict__ src, int sw, int sh, int spitch,
                                                 size_t s_stride, int16_t* __restrict__ dst, int dw, int dh,
                                                 int dpitch, size_t d_stride)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= dw || Y >= dh) return;
    const int16_t* s = src + blockIdx.z * s_stride;
    const int k[5] = {1, 4, 6, 4, 1};
    int xs[5];
    for (int b = 0; b < 5; b++) xs[b] = bwrap(2 * X - 2 + b, sw);
    int acc = 0;
    for (int a = 0; a < 5; a++) {
        const int16_t* row = s + (size_t)breflect(2 * Y - 2 + a, sh) * spitch;
        int r = 0;
        for (int b = 0; b < 5; b++) r += k[b] * ldg(row + xs[b]);
        acc += k[a] * r;
    }
    dst[blockIdx.z * d_stride + (size_t)Y * dpitch + X] = (int16_t)sat_s16((acc + 128) >> 8);
}

// Sum of the cameras' weights at one level.
__global__ __launch_bounds__(256) void k_cw_sum(const int16_t* __restrict__ wts, size_t cam_stride, int pitch,
                                                int n, int32_t* __restrict__ sum, int w, int h)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    int s = 0;
    for (int i = 0; i < n; i++) s += ldg(wts + i * cam_stride + (size_t)Y * pitch + X);
    sum[(size_t)Y * pitch + X] = s;
}

// One level of the custom blend with the collapse fused in.  planes: the
// prepared n * (L + 1) plane table whose level 0 is the caller images (kind U).
template <int U>
__global__ __launch_bounds__(256) void k_blend_custom(const Plane* __restrict__ planes,
                                                      const CanvasLevel* __restrict__ canvas, int n, int L,
                                                      int level, const int16_t* __restrict__ wts,
                                                      size_t cam_stride, int pitch,
                                                      const int32_t* __restrict__ wsum, SrcTable src,
                                                      const uint8_t* __restrict__ cover,
                                                      uint8_t* __restrict__ out, uint32_t ostep, int w, int h)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const size_t q = (size_t)Y * pitch + X;
    int a0 = 0, a1 = 0, a2 = 0;
    for (int i = 0; i < n; i++) {
        const int wv = ldg(wts + i * cam_stride + q);
        if (!wv) continue;
        const Plane P = planes[i * (L + 1) + level];
        const int4 g = level == 0 ? load_user<U>(src, P.src, X, Y) : load_plane<kPixS16>(P, X, Y);
        int l0 = g.x, l1 = g.y, l2 = g.z;
        if (level < L) {
            const Plane P1 = planes[i * (L + 1) + level + 1];
            auto f = [&](int xs, int ys) { return load_plane<kPixS16>(P1, xs, ys); };
            const int4 u = up_at(f, X, Y, P1.cw, P1.ch);
            l0 = sat_s16(g.x - (short)u.x);
            l1 = sat_s16(g.y - (short)u.y);
            l2 = sat_s16(g.z - (short)u.z);
        }
        a0 += l0 * wv;
        a1 += l1 * wv;
        a2 += l2 * wv;
    }
    int ws = ldg(wsum + q);
    if (!ws) ws = 1;
    a0 /= ws;
    a1 /= ws;
    a2 /= ws;
    if (level < L) {
        const CanvasLevel C1 = canvas[level + 1];
        auto f = [&](int xs, int ys) { return C1.c[ys * C1.pitch + xs]; };
        const int4 u = up_at(f, X, Y, C1.cw, C1.ch);
        a0 = u.x + a0;
        a1 = u.y + a1;
        a2 = u.z + a2;
    }
    if (level > 0) {
        const CanvasLevel C = canvas[level];
        C.c[Y * C.pitch + X] = make_int4(a0, a1, a2, 0);
        return;
    }
    uint32_t o = 0;
    if (ldg(cover + q))
        o = (uint32_t)sat_u8(a0) | ((uint32_t)sat_u8(a1) << 8) | ((uint32_t)sat_u8(a2) << 16);
    ((uint32_t*)(out + (size_t)Y * ostep))[X] = o;
}

namespace {
dim3 grid_of(int w, int h, int z = 1) { return dim3((w + 63) / 64, (h + 3) / 4, z); }
}  // namespace

hipError_t launch_custom_weights(const SrcTable& masks, int n, int L, const int* lw, const int* lh,
                                 int16_t* const* wts, const size_t* cam_stride, int pitch,
                                 int32_t* const* wsum, uint8_t* cover, hipStream_t s)
{
    hipLaunchKernelGGL(k_cw_init, grid_of(lw[0], lh[0]), dim3(256), 0, s, masks, n, lw[0], lh[0], wts[0],
                       cam_stride[0], pitch, cover);
    for (int j = 0; j < L; j++)
        hipLaunchKernelGGL(k_cw_down, grid_of(lw[j + 1], lh[j + 1], n), dim3(256), 0, s, wts[j], lw[j], lh[j],
                           pitch, cam_stride[j], wts[j + 1], lw[j + 1], lh[j + 1], pitch, cam_stride[j + 1]);
    for (int j = 0; j <= L; j++)
        hipLaunchKernelGGL(k_cw_sum, grid_of(lw[j], lh[j]), dim3(256), 0, s, wts[j], cam_stride[j], pitch, n,
                           wsum[j], lw[j], lh[j]);
    return hipGetLastError();
}

hipError_t launch_blend_custom(int user_kind, const Plane* planes, const CanvasLevel* canvas, int n, int L,
                               int level, const int16_t* wts, size_t cam_stride, int pitch,
                               const int32_t* wsum, const SrcTable& src, const uint8_t* cover,
                               uint8_t* out, size_t ostep, int w, int h, hipStream_t s)
{
    if (user_kind == kPixU8)
        hipLaunchKernelGGL(k_blend_custom<kPixU8>, grid_of(w, h), dim3(256), 0, s, planes, canvas, n, L, level,
                           wts, cam_stride, pitch, wsum, src, cover, out, (uint32_t)ostep,
                           w, h);
    else
        hipLaunchKernelGGL(k_blend_custom<kPixS16>, grid_of(w, h), dim3(256), 0, s, planes, canvas, n, L,
                           level, wts, cam_stride, pitch, wsum, src, cover, out,
                           (uint32_t)ostep, w, h);
    return hipGetLastError();
}

}  // namespace pano
