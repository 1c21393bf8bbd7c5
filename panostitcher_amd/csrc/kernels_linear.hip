// kernels_linear.hip -- the linear-blend mode (highQualityBlend = 0) on gfx950.
//
//   k_blur_rows / k_blur_cols   cv::GaussianBlur of an 8-bit seam mask on
//       OpenCV 3.0's 8-bit fixed-point path: integer taps K (256 x the float
//       Gaussian, rounded), rows to int32, then columns, (s + 2^15) >> 16
//       saturated; borders REFLECT_101 (Blend/LinearBlend.cpp:134-148).  The
//       result is ANDed bitwise with the camera's mask.
//   k_linear_weights            calcWeights32F (:76-132) / calcWeights (:17-57)
//       per canvas pixel over the cameras' blurred seams.
//   k_linear_render             per canvas pixel, the cameras listed for its
//       tile in order: bilinear sample (exact, as the remap path) times the
//       weight, summed in float without contraction
//       (reprojectWeightedAccumulateTo32F, Warp/ZReproject.cpp:606-691), then
//       rounded half to even and saturated (convertTo(CV_8U)).
//   k_linear_blend32s           CudaTilingLinearBlend::blend
//       (CudaAccel/CudaFeatherBlend.cu:8-29): int32 += weight x pixel over the
//       cameras, >> 16, saturated, all four channels.
// Weight and map planes are full-canvas per camera, camera after camera.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

__global__ __launch_bounds__(256) void k_blur_rows(const uint8_t* __restrict__ src, int pitch, int w, int h,
                                                   const int* __restrict__ K, int n, int k0, int k1,
                                                   int32_t* __restrict__ dst)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const uint8_t* row = src + (size_t)Y * pitch;
    const int a = n / 2;
    int s = 0;
    for (int k = k0; k < k1; k++) s += ldg(K + k) * (int)ldg(row + breflect(X + k - a, w));
    dst[(size_t)Y * pitch + X] = s;
}

__global__ __launch_bounds__(256) void k_blur_cols(const int32_t* __restrict__ rows, int pitch, int w, int h,
                                                   const int* __restrict__ K, int n, int k0, int k1,
                                                   const uint8_t* __restrict__ mask, int mpitch,
                                                   uint8_t* __restrict__ dst)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const int a = n / 2;
    int s = 0;
    for (int k = k0; k < k1; k++) s += ldg(K + k) * ldg(rows + (size_t)breflect(Y + k - a, h) * pitch + X);
    const int v = sat_u8((s + (1 << 15)) >> 16);
    dst[(size_t)Y * pitch + X] = (uint8_t)(v & ldg(mask + (size_t)Y * mpitch + X));
}

// dist: n planes (pitch bytes per row, cam_stride bytes apart); out: n planes
// of float (AS_INT: int32) with the same pitch in elements.
template <bool AS_INT>
__global__ __launch_bounds__(256) void k_linear_weights(const uint8_t* __restrict__ dist, int pitch,
                                                        size_t cam_stride, int n, int w, int h,
                                                        void* __restrict__ out)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const size_t q = (size_t)Y * pitch + X;
    float sum = 0.f;
    int nz = 0, last = 0;
    for (int k = 0; k < n; k++) {
        const int d = ldg(dist + k * cam_stride + q);
        sum = __fadd_rn(sum, (float)d);
        if (d) {
            nz++;
            last = k;
        }
    }
    const float inv = __builtin_fabsf(sum) <= __FLT_MIN__ ? 0.f : __fdiv_rn(1.0f, sum);
    for (int k = 0; k < n; k++) {
        const float d = (float)ldg(dist + k * cam_stride + q);
        if (AS_INT) {
            // (int)(d * inv * 2^16 + 1), truncated
            const float t = __fadd_rn(__fmul_rn(__fmul_rn(d, inv), 65536.0f), 1.0f);
            ((int32_t*)out)[k * cam_stride + q] = (int32_t)t;
        } else {
            float wv = 0.f;
            if (nz > 1) wv = __fmul_rn(d, inv);
            else if (nz == 1 && k == last) wv = 1.0f;
            ((float*)out)[k * cam_stride + q] = wv;
        }
    }
}

// One canvas tile of the linear render: cameras camlist[J.off .. J.off + J.cnt).
// Each thread carries kLinRows pixels (rows Y, Y + 4, ...) through every step,
// so that the loads of all of them are in flight together.
__global__ __launch_bounds__(256) void k_linear_render(const LinTile* __restrict__ tiles,
                                                       const int* __restrict__ camlist,
                                                       const LinCam* __restrict__ cams, SrcTable src,
                                                       uint8_t* __restrict__ out, uint32_t ostep, int w, int h)
{
    constexpr int R = kLinRows;
    const LinTile J = ldc(tiles + blockIdx.x);
    const int X = J.X + (threadIdx.x & 63), Y0 = J.Y + (threadIdx.x >> 6);
    if (X >= w || Y0 >= h) return;
    float a[R][3];
#pragma unroll
    for (int r = 0; r < R; r++) a[r][0] = a[r][1] = a[r][2] = 0.f;
    for (int k = 0; k < J.cnt; k++) {
        const LinCam C = ldc(cams + ldc(camlist + J.off + k));
        const uint8_t* sp = src.p[C.src];
        const uint32_t sstep = src.step[C.src];
        float wv[R];
        float2 m[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int Y = Y0 + 4 * r;
            wv[r] = Y < h ? ldg(C.w + (size_t)Y * C.pitch + X) : 0.f;
        }
        // map points in flight with the weights (tiles list a camera only where
        // its weight is non-zero somewhere, so few of these are wasted)
#pragma unroll
        for (int r = 0; r < R; r++)
            m[r] = Y0 + 4 * r < h ? ldg_f2(C.map + (size_t)(Y0 + 4 * r) * C.pitch + X) : make_float2(-1.f, -1.f);
        bool ok[R];
        uint32_t p[R][4];
#pragma unroll
        for (int r = 0; r < R; r++) {
            // a zero weight adds nothing; an invalid map point adds nothing
            ok[r] = wv[r] != 0.f && m[r].x >= 0.f && m[r].y >= 0.f && m[r].x < (float)C.sw && m[r].y < (float)C.sh;
            if (ok[r]) {
                uint32_t o[4];
                tap_offsets_fast(sstep, C.sw, C.sh, m[r].x, m[r].y, o);
#pragma unroll
                for (int t = 0; t < 4; t++) p[r][t] = ldg((const uint32_t*)(sp + o[t]));
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (!ok[r]) continue;
            bool tie = false;
            uint32_t v = bilinear_fast(__builtin_amdgcn_fractf(m[r].x), __builtin_amdgcn_fractf(m[r].y), p[r], tie);
            if (tie) v = bilinear_combine((double)m[r].x, (double)m[r].y, p[r], nullptr);
            a[r][0] = __fadd_rn(a[r][0], __fmul_rn((float)(v & 255u), wv[r]));
            a[r][1] = __fadd_rn(a[r][1], __fmul_rn((float)((v >> 8) & 255u), wv[r]));
            a[r][2] = __fadd_rn(a[r][2], __fmul_rn((float)((v >> 16) & 255u), wv[r]));
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int Y = Y0 + 4 * r;
        if (Y >= h) break;
        uint32_t o = __builtin_amdgcn_cvt_pk_u8_f32(a[r][0], 0u, 0u);
        o = __builtin_amdgcn_cvt_pk_u8_f32(a[r][1], 1u, o);
        o = __builtin_amdgcn_cvt_pk_u8_f32(a[r][2], 2u, o);
        stg((uint32_t*)(out + (size_t)Y * ostep) + X, o);
    }
}

__global__ __launch_bounds__(256) void k_linear_blend32s(SrcTable images, int n, const int32_t* __restrict__ wts,
                                                         int pitch, size_t cam_stride, uint8_t* __restrict__ out,
                                                         uint32_t ostep, int w, int h)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const size_t q = (size_t)Y * pitch + X;
    // weights are 1 .. 2^16 + 1 (calcWeights): 24-bit products suffice
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int i = 0; i < n; i++) {
        const uint32_t wv = (uint32_t)ldg(wts + i * cam_stride + q);
        const uint32_t p = ldg((const uint32_t*)(images.p[i] + (size_t)Y * images.step[i]) + X);
        s0 += __umul24(wv, p & 255u);
        s1 += __umul24(wv, (p >> 8) & 255u);
        s2 += __umul24(wv, (p >> 16) & 255u);
        s3 += __umul24(wv, p >> 24);
    }
    const uint32_t o = (uint32_t)sat_u8((int)(s0 >> 16)) | ((uint32_t)sat_u8((int)(s1 >> 16)) << 8) |
                       ((uint32_t)sat_u8((int)(s2 >> 16)) << 16) | ((uint32_t)sat_u8((int)(s3 >> 16)) << 24);
    stg((uint32_t*)(out + (size_t)Y * ostep) + X, o);
}

// reprojectWeightedAccumulateTo32F for one camera (free function): where the
// map is valid, accum (float4) += bilinear sample x weight per channel (product
// then sum, as Warp/ZReproject.cpp:606-631), channel 3 set to 0 (as
// CudaReproject.cu:604-628).
__global__ __launch_bounds__(256) void k_reproject_accum(const uint8_t* __restrict__ src, uint32_t sstep, int sw,
                                                         int sh, const float* __restrict__ xmap,
                                                         const float* __restrict__ ymap, size_t map_step,
                                                         const float* __restrict__ wt, size_t wstep,
                                                         float* __restrict__ acc, size_t astep, int w, int h)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const double x = ((const float*)((const char*)xmap + (size_t)Y * map_step))[X];
    const double y = ((const float*)((const char*)ymap + (size_t)Y * map_step))[X];
    if (!(x >= 0 && y >= 0 && x < sw && y < sh)) return;
    const uint32_t v = bilinear_bgr(src, sstep, sw, sh, x, y, nullptr);
    const float wv = ((const float*)((const char*)wt + (size_t)Y * wstep))[X];
    float* p = (float*)((char*)acc + (size_t)Y * astep) + 4 * X;
    p[0] = __fadd_rn(p[0], __fmul_rn((float)(v & 255u), wv));
    p[1] = __fadd_rn(p[1], __fmul_rn((float)((v >> 8) & 255u), wv));
    p[2] = __fadd_rn(p[2], __fmul_rn((float)((v >> 16) & 255u), wv));
    p[3] = 0.f;
}

// x and y planes (pitch sp elements) into (x, y) pairs (pitch dp pairs).
__global__ __launch_bounds__(256) void k_interleave_map(const float* __restrict__ xm, const float* __restrict__ ym,
                                                        int sp, float2* __restrict__ out, int dp, int w, int h)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const size_t q = (size_t)Y * sp + X;
    out[(size_t)Y * dp + X] = make_float2(ldg(xm + q), ldg(ym + q));
}

// hit[(tile) n + i] = 1 for the kLinTW x kLinTH canvas tiles where camera i's
// weight (n float planes of the given pitch, plane elements apart) is non-zero;
// hit zero on entry.  A camera whose weight is zero across a tile adds nothing
// there (sample x 0 = +0 leaves the non-negative sum as it is).
__global__ __launch_bounds__(256) void k_lin_tile_hits(const float* __restrict__ wts, int pitch, size_t plane, int n,
                                                       int w, int h, int tw, uint8_t* __restrict__ hit)
{
    const int X = blockIdx.x * 64 + (threadIdx.x & 63), Y = blockIdx.y * 4 + (threadIdx.x >> 6), i = blockIdx.z;
    if (X >= w || Y >= h) return;
    if (ldg(wts + plane * i + (size_t)Y * pitch + X) != 0.f)
        hit[((size_t)(Y / kLinTH) * tw + X / kLinTW) * n + i] = 1;
}

namespace {
dim3 grid_px(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }
}  // namespace

hipError_t launch_interleave_map(const float* xm, const float* ym, int sp, float2* out, int dp, int w, int h,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_interleave_map, grid_px(w, h), dim3(256), 0, s, xm, ym, sp, out, dp, w, h);
    return hipGetLastError();
}

hipError_t launch_lin_tile_hits(const float* wts, int pitch, size_t plane, int n, int w, int h, int tw, uint8_t* hit,
                                hipStream_t s)
{
    dim3 g = grid_px(w, h);
    g.z = n;
    hipLaunchKernelGGL(k_lin_tile_hits, g, dim3(256), 0, s, wts, pitch, plane, n, w, h, tw, hit);
    return hipGetLastError();
}

hipError_t launch_reproject_accum(const uint8_t* src, size_t sstep, int sw, int sh, const float* xmap,
                                  const float* ymap, size_t map_step, const float* wt, size_t wstep, float* acc,
                                  size_t astep, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(k_reproject_accum, grid_px(w, h), dim3(256), 0, s, src, (uint32_t)sstep, sw, sh, xmap, ymap,
                       map_step, wt, wstep, acc, astep, w, h);
    return hipGetLastError();
}

hipError_t launch_blur_u8(const uint8_t* src, int pitch, int w, int h, const int* K, int n, int k0, int k1,
                          int32_t* rows, const uint8_t* mask, int mpitch, uint8_t* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_blur_rows, grid_px(w, h), dim3(256), 0, s, src, pitch, w, h, K, n, k0, k1, rows);
    hipLaunchKernelGGL(k_blur_cols, grid_px(w, h), dim3(256), 0, s, rows, pitch, w, h, K, n, k0, k1, mask,
                       mpitch, dst);
    return hipGetLastError();
}

hipError_t launch_linear_weights(bool as_int, const uint8_t* dist, int pitch, size_t cam_stride, int n, int w,
                                 int h, void* out, hipStream_t s)
{
    if (as_int)
        hipLaunchKernelGGL(k_linear_weights<true>, grid_px(w, h), dim3(256), 0, s, dist, pitch, cam_stride, n, w,
                           h, out);
    else
        hipLaunchKernelGGL(k_linear_weights<false>, grid_px(w, h), dim3(256), 0, s, dist, pitch, cam_stride, n,
                           w, h, out);
    return hipGetLastError();
}

hipError_t launch_linear_render(const LinTile* tiles, int ntiles, const int* camlist, const LinCam* cams,
                                const SrcTable& src, uint8_t* out, size_t ostep, int w, int h, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_linear_render, dim3(ntiles), dim3(256), 0, s, tiles, camlist, cams, src, out,
                       (uint32_t)ostep, w, h);
    return hipGetLastError();
}

hipError_t launch_linear_blend32s(const SrcTable& images, int n, const int32_t* wts, int pitch, size_t cam_stride,
                                  uint8_t* out, size_t ostep, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(k_linear_blend32s, grid_px(w, h), dim3(256), 0, s, images, n, wts, pitch, cam_stride, out,
                       (uint32_t)ostep, w, h);
    return hipGetLastError();
}

}  // namespace pano
