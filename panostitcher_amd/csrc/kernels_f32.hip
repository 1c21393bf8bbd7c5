// kernels_f32.hip -- the float-weight multiband blender on gfx950
// (CudaTilingMultibandBlendFast32F, CudaAccel/CudaMultibandBlend.cpp:1041-1179,
// kernels CudaAccel/CudaPyramid6.cu).
//
// The reference's arithmetic is kept operation by operation in float, with the
// fused multiply-adds the CUDA compiler forms by default from `s = s + c * x`:
//   pyrDown: vertical taps 1/16 4/16 6/16 4/16 1/16 first (rows REFLECT_101),
//            then the same horizontally over the five column sums (columns WRAP)
//   pyrUp:   per source row the horizontal taps (1/16 6/16 1/16 at even output
//            columns, 4/16 4/16 at odd ones), the same over rows, times 4
//   Laplacian and accumulate: R = fma(W, G - up(G_{j+1}), R)
//   collapse: C_j = up(C_{j+1}) + invW_j R_j (invW only for partial coverage)
//   output:   round to nearest even, saturate to [0, 255] (cvt.rni.sat.u8.f32)
// Planes are full-canvas float4 (or float) images with row steps in bytes.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

namespace {

constexpr float kC1 = 0.0625f, kC4 = 0.25f, kC6 = 0.375f;

__device__ __forceinline__ int wrapc(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    p %= n;
    return p < 0 ? p + n : p;
}

__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }
__device__ __forceinline__ float4 mul4(float c, float4 v) { return make_float4(c * v.x, c * v.y, c * v.z, c * v.w); }
__device__ __forceinline__ float4 fma4(float c, float4 v, float4 s)
{
    return make_float4(__fmaf_rn(c, v.x, s.x), __fmaf_rn(c, v.y, s.y), __fmaf_rn(c, v.z, s.z),
                       __fmaf_rn(c, v.w, s.w));
}

typedef float pano_f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldf4(const void* base, uint32_t step, int y, int x)
{
    pano_f32x4 v = *(const PANO_GLOBAL pano_f32x4*)((const char*)base + (size_t)y * step + 16u * (uint32_t)x);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stf4(void* base, uint32_t step, int y, int x, float4 v)
{
    pano_f32x4 w = {v.x, v.y, v.z, v.w};
    *(PANO_GLOBAL pano_f32x4*)((char*)base + (size_t)y * step + 16u * (uint32_t)x) = w;
}
__device__ __forceinline__ float ldf1(const void* base, uint32_t step, int y, int x)
{
    return ldg((const float*)((const char*)base + (size_t)y * step) + x);
}
// A level-0 pixel of the caller's image: 8-bit BGRx converted to float, or float4.
template <bool U8>
__device__ __forceinline__ float4 ld_img(const void* base, uint32_t step, int y, int x)
{
    if (U8) {
        const uint32_t p = ldg((const uint32_t*)((const char*)base + (size_t)y * step) + x);
        return make_float4((float)(p & 255u), (float)((p >> 8) & 255u), (float)((p >> 16) & 255u),
                           (float)(p >> 24));
    }
    return ldf4(base, step, y, x);
}

// The horizontal pass of pyrUp over source row ys at output column x.
template <class Load>
__device__ __forceinline__ float4 fup_row(const Load& ld, int ys, int x, int sw)
{
    const int k = x >> 1;
    if (!(x & 1)) {
        float4 s = mul4(kC1, ld(ys, wrapc(k - 1, sw)));
        s = fma4(kC6, ld(ys, wrapc(k, sw)), s);
        return fma4(kC1, ld(ys, wrapc(k + 1, sw)), s);
    }
    const float4 s = mul4(kC4, ld(ys, wrapc(k, sw)));
    return fma4(kC4, ld(ys, wrapc(k + 1, sw)), s);
}

// pyrUp of the level-(j+1) image at level-j pixel (x, y).
template <class Load>
__device__ __forceinline__ float4 fup_at(const Load& ld, int x, int y, int sw, int sh)
{
    const int k = y >> 1;
    float4 v;
    if (!(y & 1)) {
        v = mul4(kC1, fup_row(ld, breflect(k - 1, sh), x, sw));
        v = fma4(kC6, fup_row(ld, breflect(k, sh), x, sw), v);
        v = fma4(kC1, fup_row(ld, breflect(k + 1, sh), x, sw), v);
    } else {
        v = mul4(kC4, fup_row(ld, breflect(k, sh), x, sw));
        v = fma4(kC4, fup_row(ld, breflect(k + 1, sh), x, sw), v);
    }
    return mul4(4.0f, v);
}

constexpr int kDW = 16, kDH = 16;                  // pyrDown output tile
constexpr int kSW4 = 2 * kDW + 4, kSH4 = 2 * kDH + 4;

}  // namespace

// pyrDown of a 4-channel image into dst (optionally times a per-pixel scale),
// or of a 1-channel float image (C == 1).  The source region of a 16 x 16
// output tile is staged in LDS; column sums first, then rows.
template <int C, bool U8>
__global__ __launch_bounds__(256) void k_f32_down(const void* __restrict__ src, int sw, int sh,
                                                  uint32_t sstep, void* __restrict__ dst, int dw, int dh,
                                                  uint32_t dstep, const float* __restrict__ scale,
                                                  uint32_t scstep)
{
    __shared__ float4 tile[kSH4 * kSW4];
    __shared__ float4 vs[kDH * kSW4];
    const int X0 = blockIdx.x * kDW, Y0 = blockIdx.y * kDH;
    const int sx0 = 2 * X0 - 2, sy0 = 2 * Y0 - 2;
    for (int e = threadIdx.x; e < kSH4 * kSW4; e += 256) {
        const int r = e / kSW4, c = e - r * kSW4;
        const int ys = breflect(sy0 + r, sh), xs = wrapc(sx0 + c, sw);
        float4 v;
        if (C == 1) v = f4(ldf1(src, sstep, ys, xs));
        else v = ld_img<U8>(src, sstep, ys, xs);
        tile[e] = v;
    }
    __syncthreads();
    // column sums of the rows each output row needs (source rows 2y .. 2y + 4 of the tile)
    for (int e = threadIdx.x; e < kDH * kSW4; e += 256) {
        const int y = e / kSW4, c = e - y * kSW4;
        const float4* t = tile + (2 * y) * kSW4 + c;
        float4 s = mul4(kC1, t[0]);
        s = fma4(kC4, t[kSW4], s);
        s = fma4(kC6, t[2 * kSW4], s);
        s = fma4(kC4, t[3 * kSW4], s);
        s = fma4(kC1, t[4 * kSW4], s);
        vs[e] = s;
    }
    __syncthreads();
    const int x = threadIdx.x % kDW, y = threadIdx.x / kDW;
    const int X = X0 + x, Y = Y0 + y;
    if (X >= dw || Y >= dh) return;
    const float4* v = vs + y * kSW4 + 2 * x;
    float4 h = mul4(kC1, v[0]);
    h = fma4(kC4, v[1], h);
    h = fma4(kC6, v[2], h);
    h = fma4(kC4, v[3], h);
    h = fma4(kC1, v[4], h);
    if (scale) h = mul4(ldf1(scale, scstep, Y, X), h);
    if (C == 1) stg((float*)((char*)dst + (size_t)Y * dstep) + X, h.x);
    else stf4(dst, dstep, Y, X, h);
}

// R_j = fma(W_j, G_j - up(G_{j+1}), R_j) (top level: G_j alone).
template <bool U8>
__global__ __launch_bounds__(256) void k_f32_lap(const void* __restrict__ g, uint32_t gstep,
                                                 const void* __restrict__ g1, uint32_t g1step, int w1,
                                                 int h1, const float* __restrict__ wt, uint32_t wstep,
                                                 void* __restrict__ r, uint32_t rstep, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    float4 lap = ld_img<U8>(g, gstep, y, x);
    if (g1) {
        auto ld = [&](int ys, int xs) { return ldf4(g1, g1step, ys, xs); };
        const float4 u = fup_at(ld, x, y, w1, h1);
        lap = make_float4(lap.x - u.x, lap.y - u.y, lap.z - u.z, lap.w - u.w);
    }
    const float wv = ldf1(wt, wstep, y, x);
    stf4(r, rstep, y, x, fma4(wv, lap, ldf4(r, rstep, y, x)));
}

// C_j = up(C_{j+1}) + invW_j R_j in place (top level: invW_j R_j).
__global__ __launch_bounds__(256) void k_f32_collapse(const void* __restrict__ c1, uint32_t c1step,
                                                      int w1, int h1, const float* __restrict__ winv,
                                                      uint32_t wistep, void* __restrict__ r,
                                                      uint32_t rstep, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    float4 s = ldf4(r, rstep, y, x);
    if (winv) s = mul4(ldf1(winv, wistep, y, x), s);
    if (c1) {
        auto ld = [&](int ys, int xs) { return ldf4(c1, c1step, ys, xs); };
        const float4 u = fup_at(ld, x, y, w1, h1);
        s = make_float4(u.x + s.x, u.y + s.y, u.z + s.z, u.w + s.w);
    }
    stf4(r, rstep, y, x, s);
}

// saturate_cast<uchar> of C_0, zero outside every mask (cover == 0).
__global__ __launch_bounds__(256) void k_f32_out(const void* __restrict__ c0, uint32_t cstep,
                                                 const uint8_t* __restrict__ cover, int cpitch,
                                                 uint8_t* __restrict__ out, uint32_t ostep, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    uint32_t o = 0;
    if (!cover || ldg(cover + (size_t)y * cpitch + x)) {
        const float4 v = ldf4(c0, cstep, y, x);
        o = __builtin_amdgcn_cvt_pk_u8_f32(v.x, 0u, 0u);
        o = __builtin_amdgcn_cvt_pk_u8_f32(v.y, 1u, o);
        o = __builtin_amdgcn_cvt_pk_u8_f32(v.z, 2u, o);
        o = __builtin_amdgcn_cvt_pk_u8_f32(v.w, 3u, o);
    }
    stg((uint32_t*)(out + (size_t)y * ostep) + x, o);
}

// Prepare: a 0/255 mask as 0/1 floats.
__global__ __launch_bounds__(256) void k_f32_from_mask(const uint8_t* __restrict__ m, int mpitch,
                                                       float* __restrict__ f, uint32_t fstep, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    stg((float*)((char*)f + (size_t)y * fstep) + x, ldg(m + (size_t)y * mpitch + x) ? 1.f : 0.f);
}

// Prepare: from a pyrDown of the binary alpha, the next binary alpha
// (scaledSet32FC1Mask32FC1) and the inverse alpha in place (inverse32FC1).
__global__ __launch_bounds__(256) void k_f32_alpha(float* __restrict__ a, uint32_t astep,
                                                   float* __restrict__ bin, uint32_t bstep, int w, int h)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    float* pa = (float*)((char*)a + (size_t)y * astep) + x;
    const float v = ldg(pa);
    stg((float*)((char*)bin + (size_t)y * bstep) + x, v > 0 ? 1.f : 0.f);
    stg(pa, v > 0 ? 1.0f / v : 0.f);
}

// Prepare: sum += W (accumulate32FC1), or sum = inverse (inverse32FC1).
__global__ __launch_bounds__(256) void k_f32_wsum(const float* __restrict__ wsrc, uint32_t wstep,
                                                  float* __restrict__ sum, uint32_t sstep, int w, int h,
                                                  int invert)
{
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= w || y >= h) return;
    float* ps = (float*)((char*)sum + (size_t)y * sstep) + x;
    const float s = ldg(ps);
    if (invert) stg(ps, s > 0 ? 1.0f / s : 0.f);
    else stg(ps, s + ldf1(wsrc, wstep, y, x));
}

namespace {
dim3 grid_px(int w, int h) { return dim3((w + 63) / 64, (h + 3) / 4); }
}  // namespace

hipError_t launch_f32_down(int channels, bool u8, const void* src, int sw, int sh, size_t sstep,
                           void* dst, int dw, int dh, size_t dstep, const float* scale, size_t scstep,
                           hipStream_t s)
{
    const dim3 grid((dw + kDW - 1) / kDW, (dh + kDH - 1) / kDH);
    if (channels == 1)
        hipLaunchKernelGGL((k_f32_down<1, false>), grid, dim3(256), 0, s, src, sw, sh, (uint32_t)sstep, dst,
                           dw, dh, (uint32_t)dstep, scale, (uint32_t)scstep);
    else if (u8)
        hipLaunchKernelGGL((k_f32_down<4, true>), grid, dim3(256), 0, s, src, sw, sh, (uint32_t)sstep, dst,
                           dw, dh, (uint32_t)dstep, scale, (uint32_t)scstep);
    else
        hipLaunchKernelGGL((k_f32_down<4, false>), grid, dim3(256), 0, s, src, sw, sh, (uint32_t)sstep, dst,
                           dw, dh, (uint32_t)dstep, scale, (uint32_t)scstep);
    return hipGetLastError();
}

hipError_t launch_f32_lap(bool u8, const void* g, size_t gstep, const void* g1, size_t g1step, int w1,
                          int h1, const float* wt, size_t wstep, void* r, size_t rstep, int w, int h,
                          hipStream_t s)
{
    if (u8)
        hipLaunchKernelGGL(k_f32_lap<true>, grid_px(w, h), dim3(256), 0, s, g, (uint32_t)gstep, g1,
                           (uint32_t)g1step, w1, h1, wt, (uint32_t)wstep, r, (uint32_t)rstep, w, h);
    else
        hipLaunchKernelGGL(k_f32_lap<false>, grid_px(w, h), dim3(256), 0, s, g, (uint32_t)gstep, g1,
                           (uint32_t)g1step, w1, h1, wt, (uint32_t)wstep, r, (uint32_t)rstep, w, h);
    return hipGetLastError();
}

hipError_t launch_f32_collapse(const void* c1, size_t c1step, int w1, int h1, const float* winv,
                               size_t wistep, void* r, size_t rstep, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_collapse, grid_px(w, h), dim3(256), 0, s, c1, (uint32_t)c1step, w1, h1, winv,
                       (uint32_t)wistep, r, (uint32_t)rstep, w, h);
    return hipGetLastError();
}

hipError_t launch_f32_out(const void* c0, size_t cstep, const uint8_t* cover, int cpitch, uint8_t* out,
                          size_t ostep, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_out, grid_px(w, h), dim3(256), 0, s, c0, (uint32_t)cstep, cover, cpitch, out,
                       (uint32_t)ostep, w, h);
    return hipGetLastError();
}

hipError_t launch_f32_from_mask(const uint8_t* m, int mpitch, float* f, size_t fstep, int w, int h,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_from_mask, grid_px(w, h), dim3(256), 0, s, m, mpitch, f, (uint32_t)fstep, w, h);
    return hipGetLastError();
}

hipError_t launch_f32_alpha(float* a, size_t astep, float* bin, size_t bstep, int w, int h, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_alpha, grid_px(w, h), dim3(256), 0, s, a, (uint32_t)astep, bin, (uint32_t)bstep,
                       w, h);
    return hipGetLastError();
}

hipError_t launch_f32_wsum(const float* wsrc, size_t wstep, float* sum, size_t sstep, int w, int h,
                           bool invert, hipStream_t s)
{
    hipLaunchKernelGGL(k_f32_wsum, grid_px(w, h), dim3(256), 0, s, wsrc, (uint32_t)wstep, sum,
                       (uint32_t)sstep, w, h, invert ? 1 : 0);
    return hipGetLastError();
}

}  // namespace pano
