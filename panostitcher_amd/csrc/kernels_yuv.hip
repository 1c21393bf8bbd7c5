// kernels_yuv.hip -- the encoder / decoder hand-off conversions on gfx950:
// cvtBGR32ToYUV420P, cvtBGR32ToNV12 and their inverses (CudaAccel/ImageProc.cu:
// 44-266, declared in CudaAccel/CudaInterface.h:161-171).  One thread per 2x2
// block, 64 x 4 blocks (128 x 8 pixels) per workgroup; the arithmetic is the
// BT.601 fixed point of pano_kernel_common.hpp, shared with the level-0 blend
// kernel's fused output.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

template <bool NV12>
__global__ __launch_bounds__(256) void k_bgr32_to_yuv(const uint8_t* __restrict__ src, uint32_t sstep, int bw,
                                                      int bh, uint8_t* __restrict__ y, uint32_t ystep,
                                                      uint8_t* __restrict__ u, uint32_t ustep,
                                                      uint8_t* __restrict__ v, uint32_t vstep)
{
    const int bx = blockIdx.x * 64 + (threadIdx.x & 63), by = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (bx >= bw || by >= bh) return;
    const uint8_t* t = src + (size_t)(2 * by) * sstep + 8u * (uint32_t)bx;
    const uint2 a = ldg_u2((const uint2*)t), b = ldg_u2((const uint2*)(t + sstep));
    const uint32_t q[4] = {a.x, a.y, b.x, b.y};
    uint32_t yt, yb, uu, vv;
    bt601_block(q, yt, yb, uu, vv);
    uint8_t* yr = y + (size_t)(2 * by) * ystep + 2u * (uint32_t)bx;
    stg((uint16_t*)yr, (uint16_t)yt);
    stg((uint16_t*)(yr + ystep), (uint16_t)yb);
    if (NV12) {
        stg((uint16_t*)(u + (size_t)by * ustep + 2u * (uint32_t)bx), (uint16_t)(uu | (vv << 8)));
    } else {
        stg(u + (size_t)by * ustep + bx, (uint8_t)uu);
        stg(v + (size_t)by * vstep + bx, (uint8_t)vv);
    }
}

template <bool NV12>
__global__ __launch_bounds__(256) void k_yuv_to_bgr32(const uint8_t* __restrict__ y, uint32_t ystep,
                                                      const uint8_t* __restrict__ u, uint32_t ustep,
                                                      const uint8_t* __restrict__ v, uint32_t vstep, int bw,
                                                      int bh, uint8_t* __restrict__ dst, uint32_t dstep)
{
    const int bx = blockIdx.x * 64 + (threadIdx.x & 63), by = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (bx >= bw || by >= bh) return;
    const uint8_t* yr = y + (size_t)(2 * by) * ystep + 2u * (uint32_t)bx;
    const uint32_t t = ldg((const uint16_t*)yr), b = ldg((const uint16_t*)(yr + ystep));
    const uint32_t ys[4] = {t & 255u, t >> 8, b & 255u, b >> 8};
    int uu, vv;
    if (NV12) {
        const uint32_t p = ldg((const uint16_t*)(u + (size_t)by * ustep + 2u * (uint32_t)bx));
        uu = (int)(p & 255u);
        vv = (int)(p >> 8);
    } else {
        uu = ldg(u + (size_t)by * ustep + bx);
        vv = ldg(v + (size_t)by * vstep + bx);
    }
    uint32_t q[4];
    bt601_to_bgr(ys, uu, vv, q);
    uint8_t* o = dst + (size_t)(2 * by) * dstep + 8u * (uint32_t)bx;
    stg_u2((uint2*)o, make_uint2(q[0], q[1]));
    stg_u2((uint2*)(o + dstep), make_uint2(q[2], q[3]));
}

namespace {
dim3 grid_blocks(int bw, int bh) { return dim3((bw + 63) / 64, (bh + 3) / 4); }
}  // namespace

hipError_t launch_bgr32_to_yuv(bool nv12, const uint8_t* src, size_t sstep, int w, int h, uint8_t* y,
                               size_t ystep, uint8_t* u, size_t ustep, uint8_t* v, size_t vstep, hipStream_t s)
{
    const int bw = w / 2, bh = h / 2;
    if (nv12)
        hipLaunchKernelGGL(k_bgr32_to_yuv<true>, grid_blocks(bw, bh), dim3(256), 0, s, src, (uint32_t)sstep, bw,
                           bh, y, (uint32_t)ystep, u, (uint32_t)ustep, v, (uint32_t)vstep);
    else
        hipLaunchKernelGGL(k_bgr32_to_yuv<false>, grid_blocks(bw, bh), dim3(256), 0, s, src, (uint32_t)sstep,
                           bw, bh, y, (uint32_t)ystep, u, (uint32_t)ustep, v, (uint32_t)vstep);
    return hipGetLastError();
}

hipError_t launch_yuv_to_bgr32(bool nv12, const uint8_t* y, size_t ystep, const uint8_t* u, size_t ustep,
                               const uint8_t* v, size_t vstep, int w, int h, uint8_t* dst, size_t dstep,
                               hipStream_t s)
{
    const int bw = w / 2, bh = h / 2;
    if (nv12)
        hipLaunchKernelGGL(k_yuv_to_bgr32<true>, grid_blocks(bw, bh), dim3(256), 0, s, y, (uint32_t)ystep, u,
                           (uint32_t)ustep, v, (uint32_t)vstep, bw, bh, dst, (uint32_t)dstep);
    else
        hipLaunchKernelGGL(k_yuv_to_bgr32<false>, grid_blocks(bw, bh), dim3(256), 0, s, y, (uint32_t)ystep, u,
                           (uint32_t)ustep, v, (uint32_t)vstep, bw, bh, dst, (uint32_t)dstep);
    return hipGetLastError();
}

}  // namespace pano
