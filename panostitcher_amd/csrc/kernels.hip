// kernels.hip -- gfx950 kernels of the per-frame stitch path.
//
// Arithmetic follows the reference's CPU path (the parity target):
//   remap      Warp/ZReproject.cpp:101-130   FP64 bilinear, fixed sum order
//   pyrDown    Blend/Pyramid.cpp:35-153      5x5 [1 4 6 4 1]^2, WRAP x, REFLECT_101 y
//   calcDst    Blend/MultiBandBlend.cpp:732-768  a ? ((v << 8) - v) / a : 0
//   pyrUp      Blend/Pyramid.cpp:441-536     (sum + 32) >> 6
//   blend      Blend/MultiBandBlend.cpp:987-1031  accumulate, normalize, collapse
// All steps after the remap are integer and reproduce the reference bit for bit.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

// ------------------------------------------------------------------ remap --

template <bool kLut, int K>
__global__ __launch_bounds__(256) void k_remap(const RemapJob* __restrict__ jobs,
                                               const RemapCam* __restrict__ cams,
                                               const Plane* __restrict__ planes,
                                               const uint8_t* __restrict__ luts, SrcTable src)
{
    __shared__ uint8_t lut_s[768];
    const RemapJob J = jobs[blockIdx.x];
    const RemapCam cam = cams[J.cam];
    bool use_lut = kLut && cam.lut >= 0;
    if (use_lut) {
        for (int e = threadIdx.x; e < 768; e += 256) lut_s[e] = luts[cam.lut + e];
    }
    __syncthreads();
    int tx = threadIdx.x % kRemapTW, ty = threadIdx.x / kRemapTW;
    if (tx >= J.nx || ty >= J.ny) return;
    const Plane P = planes[cam.plane];
    int idx = (J.ly + ty) * P.pitch + (J.lx + tx);
    float2 m = cam.map[idx];
    uint32_t out = (cam.mask[idx] & kFlagValid) ? (1u << 24) : 0u;
    double x = m.x, y = m.y;
    if (x >= 0 && y >= 0 && x < cam.sw && y < cam.sh)
        out |= bilinear_bgr(src.p[cam.src], src.step[cam.src], cam.sw, cam.sh, x, y,
                            use_lut ? lut_s : nullptr);
    if constexpr (K == kPixU8) {
        ((uint32_t*)P.g)[idx] = out;
    } else {
        ((short4*)P.g)[idx] = make_short4((short)(out & 255), (short)((out >> 8) & 255),
                                          (short)((out >> 16) & 255), (short)(out >> 24));
    }
}

// cudaReprojectTo16S replacement: full-canvas float32 maps -> 16SC4 (x = 0).
// One output pixel per thread: 16-bit BGRx (short4) or 8-bit BGRx (uint32),
// channel 3 zero, all zero where the map is invalid.
template <bool U8>
__global__ __launch_bounds__(256) void k_reproject(const uint8_t* __restrict__ src, uint32_t src_step,
                                                   int sw, int sh, const float* __restrict__ xmap,
                                                   const float* __restrict__ ymap, size_t map_step,
                                                   void* __restrict__ dst, size_t dst_step, int dw, int dh)
{
    int X = blockIdx.x * 64 + (threadIdx.x & 63);
    int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= dw || Y >= dh) return;
    double x = ((const float*)((const uint8_t*)xmap + (size_t)Y * map_step))[X];
    double y = ((const float*)((const uint8_t*)ymap + (size_t)Y * map_step))[X];
    uint32_t v = 0;
    if (x >= 0 && y >= 0 && x < sw && y < sh) v = bilinear_bgr(src, src_step, sw, sh, x, y, nullptr) & 0xffffffu;
    uint8_t* row = (uint8_t*)dst + (size_t)Y * dst_step;
    if (U8)
        ((uint32_t*)row)[X] = v;
    else
        ((short4*)row)[X] = make_short4((short)(v & 255), (short)((v >> 8) & 255), (short)((v >> 16) & 255), 0);
}

// cudaTransform replacement: per-channel LUT on BGRx; channel 3 copied.
__global__ __launch_bounds__(256) void k_transform(const uint8_t* __restrict__ src,
                                                   size_t src_step, uint8_t* dst, size_t dst_step,
                                                   int w, int h, const uint8_t* __restrict__ lut)
{
    __shared__ uint8_t lut_s[768];
    for (int e = threadIdx.x; e < 768; e += 256) lut_s[e] = lut[e];
    __syncthreads();
    int X = blockIdx.x * 64 + (threadIdx.x & 63);
    int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    uint32_t v = ((const uint32_t*)(src + (size_t)Y * src_step))[X];
    uint32_t o = (uint32_t)lut_s[v & 255] | ((uint32_t)lut_s[256 + ((v >> 8) & 255)] << 8) |
                 ((uint32_t)lut_s[512 + ((v >> 16) & 255)] << 16) | (v & 0xff000000u);
    ((uint32_t*)(dst + (size_t)Y * dst_step))[X] = o;
}

// ------------------------------------------------------------------- down --
//
// One output tile of kDownTW x kDownTH pixels at level j+1 per workgroup.  The
// (2 TW + 3) x (2 TH + 3) input region at level j is staged in LDS (border
// resolved: WRAP columns, REFLECT_101 rows; zero outside the camera's box),
// the horizontal 5-tap runs over it into a second LDS array, and the vertical
// 5-tap + calcDstImage writes the output.  For 8-bit planes the sums run two
// channels per 32-bit lane ((B, R) and (G, valid)): horizontal sums are at most
// 16 * 255 and the full 5x5 sum at most 256 * 255 = 65280, so no lane carries.

template <int K> struct DownLds;
template <> struct DownLds<kPixU8> {
    using In = uint32_t;
    using H = uint2;
};
template <> struct DownLds<kPixS16> {
    using In = short4;
    using H = int4;
};

// One staged input pixel.  kUser: level 0 is a caller image of kind U (its
// channel 3 is ignored) and validity comes from the prepared mask.
template <int K, bool kUser, int U>
__device__ __forceinline__ typename DownLds<K>::In fetch_in(const Plane& S, const SrcTable& src,
                                                            int X, int Y)
{
    if constexpr (kUser) {
        bool valid = S.mask[(size_t)Y * S.mask_pitch + X] != 0;
        int4 v = load_user<U>(src, S.src, X, Y);
        if constexpr (K == kPixU8) {
            return (uint32_t)v.x | ((uint32_t)v.y << 8) | ((uint32_t)v.z << 16) |
                   (valid ? (1u << 24) : 0u);
        } else {
            return make_short4((short)v.x, (short)v.y, (short)v.z, valid ? 1 : 0);
        }
    } else {
        int i = box_index(S, X, Y);
        if constexpr (K == kPixU8) {
            return i < 0 ? 0u : ((const uint32_t*)S.g)[i];
        } else {
            return i < 0 ? make_short4(0, 0, 0, 0) : ((const short4*)S.g)[i];
        }
    }
}

template <int K, bool kUser, int U>
__global__ __launch_bounds__(256) void k_down(const DownJob* __restrict__ jobs,
                                              const Plane* __restrict__ planes, SrcTable src)
{
    constexpr int IW = 2 * kDownTW + 3, IH = 2 * kDownTH + 3;
    using In = typename DownLds<K>::In;
    using H = typename DownLds<K>::H;
    __shared__ In in_s[IH * IW];
    __shared__ H hs[IH * kDownTW];
    const DownJob J = jobs[blockIdx.x];
    const Plane S = planes[J.src];
    const Plane D = planes[J.dst];
    int bx0 = 2 * J.X - 2, by0 = 2 * J.Y - 2;
    int iw = 2 * J.nx + 3, ih = 2 * J.ny + 3;
    for (int e = threadIdx.x; e < ih * iw; e += 256) {
        int r = e / iw, c = e - r * iw;
        int Xs = bwrap(bx0 + c, S.cw), Ys = breflect(by0 + r, S.ch);
        in_s[r * IW + c] = fetch_in<K, kUser, U>(S, src, Xs, Ys);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ih * kDownTW; e += 256) {
        int r = e / kDownTW, x = e - r * kDownTW;
        if (x >= J.nx) continue;
        const In* p = in_s + r * IW + 2 * x;
        if constexpr (K == kPixU8) {
            uint32_t a = p[0], b = p[1], c = p[2], d = p[3], f = p[4];
            const uint32_t M = 0x00ff00ffu;
            uint32_t lo = (a & M) + (f & M) + 4u * ((b & M) + (d & M)) + 6u * (c & M);
            uint32_t hi = ((a >> 8) & M) + ((f >> 8) & M) + 4u * (((b >> 8) & M) + ((d >> 8) & M)) +
                          6u * ((c >> 8) & M);
            hs[r * kDownTW + x] = make_uint2(lo, hi);
        } else {
            short4 a = p[0], b = p[1], c = p[2], d = p[3], f = p[4];
            hs[r * kDownTW + x] = make_int4(c.x * 6 + (b.x + d.x) * 4 + a.x + f.x,
                                            c.y * 6 + (b.y + d.y) * 4 + a.y + f.y,
                                            c.z * 6 + (b.z + d.z) * 4 + a.z + f.z,
                                            c.w * 6 + (b.w + d.w) * 4 + a.w + f.w);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < J.ny * kDownTW; e += 256) {
        int y = e / kDownTW, x = e - y * kDownTW;
        if (x >= J.nx) continue;
        const H* q = hs + (2 * y) * kDownTW + x;
        H a = q[0], b = q[kDownTW], c = q[2 * kDownTW], d = q[3 * kDownTW], f = q[4 * kDownTW];
        int di = box_index(D, J.X + x, J.Y + y);
        if (di < 0) continue;
        if constexpr (K == kPixU8) {
            uint32_t lo = a.x + f.x + 4u * (b.x + d.x) + 6u * c.x;
            uint32_t hi = a.y + f.y + 4u * (b.y + d.y) + 6u * c.y;
            int al = (int)(hi >> 16);
            uint32_t o = 0;
            if (al) {
                int vb = (int)(lo & 0xffffu), vr = (int)(lo >> 16), vg = (int)(hi & 0xffffu);
                o = (uint32_t)div_norm_u8(vb, al) | ((uint32_t)div_norm_u8(vg, al) << 8) |
                    ((uint32_t)div_norm_u8(vr, al) << 16) | (1u << 24);
            }
            ((uint32_t*)D.g)[di] = o;
        } else {
            int vx = a.x + f.x + 4 * (b.x + d.x) + 6 * c.x;
            int vy = a.y + f.y + 4 * (b.y + d.y) + 6 * c.y;
            int vz = a.z + f.z + 4 * (b.z + d.z) + 6 * c.z;
            int al = a.w + f.w + 4 * (b.w + d.w) + 6 * c.w;
            short4 o = make_short4(0, 0, 0, 0);
            if (al) {
                int alpha = al << 8;
                auto cd = [&](int v) -> short {
                    return (short)((int)(((unsigned)v << 8) - (unsigned)v) / alpha);
                };
                o = make_short4(cd(vx), cd(vy), cd(vz), 1);
            }
            ((short4*)D.g)[di] = o;
        }
    }
}

// ------------------------------------------------------------------ blend --
//
// Level j >= 1 of the blend for one canvas tile: for every camera listed for
// the tile, L = sat16(G_j - (short)up(G_{j+1})) (or G_L at the top level) is
// weighted by W_j and accumulated in int32; the sum is normalised (>> 8 when
// the masks cover the canvas, else / sum W with W = 0 -> 1) and the collapse
// C_j = N_j + up32(C_{j+1}) is fused in.

template <int K>
__global__ __launch_bounds__(256) void k_blend(const TileJob* __restrict__ jobs,
                                               const int* __restrict__ camlist,
                                               const Plane* __restrict__ planes,
                                               const CanvasLevel* __restrict__ canvas,
                                               const RenderDesc* __restrict__ rds, int level)
{
    const TileJob J = jobs[blockIdx.x];
    int tx = threadIdx.x % kTileW, ty = threadIdx.x / kTileW;
    if (tx >= J.nx || ty >= J.ny) return;
    int X = J.X + tx, Y = J.Y + ty;
    const RenderDesc R = rds[J.r];
    const bool top = J.flags & kTileTop;
    int a0 = 0, a1 = 0, a2 = 0;
    for (int k = 0; k < J.cnt; k++) {
        int pi = camlist[J.off + k];
        const Plane P = planes[pi];
        int i = box_index(P, X, Y);
        if (i < 0) continue;
        int w = P.w[i];
        if (!w) continue;
        int4 g = load_at<K>(P.g, i);
        int l0 = g.x, l1 = g.y, l2 = g.z;
        if (!top) {
            const Plane P1 = planes[pi + 1];
            auto f = [&](int xs, int ys) { return load_plane<K>(P1, xs, ys); };
            int4 u = up_at(f, X, Y, P1.cw, P1.ch);
            l0 = sat_s16(g.x - (short)u.x);
            l1 = sat_s16(g.y - (short)u.y);
            l2 = sat_s16(g.z - (short)u.z);
        }
        a0 += l0 * w;
        a1 += l1 * w;
        a2 += l2 * w;
    }
    const CanvasLevel C = canvas[R.canvas0 + level];
    if (R.full) {
        a0 >>= 8;
        a1 >>= 8;
        a2 >>= 8;
    } else {
        int ws = C.wsum[Y * C.pitch + X];
        if (!ws) ws = 1;
        a0 /= ws;
        a1 /= ws;
        a2 /= ws;
    }
    if (!top) {
        const CanvasLevel C1 = canvas[R.canvas0 + level + 1];
        auto f = [&](int xs, int ys) { return C1.c[ys * C1.pitch + xs]; };
        int4 u = up_at(f, X, Y, C1.cw, C1.ch);
        a0 = u.x + a0;
        a1 = u.y + a1;
        a2 = u.z + a2;
    }
    C.c[Y * C.pitch + X] = make_int4(a0, a1, a2, 0);
}

// Level 0: the owner camera's Laplacian (its weight is exactly 256 there and
// every other camera's is 0), plus up32(C_1), saturated to 8 bits; pixels no
// mask covers are 0 (setTo(0, maskNot)).
template <int K, bool kUser, int U>
__global__ __launch_bounds__(256) void k_final(const TileJob* __restrict__ jobs,
                                               const Plane* __restrict__ planes,
                                               const CanvasLevel* __restrict__ canvas,
                                               const RenderDesc* __restrict__ rds, SrcTable src,
                                               OutTable outs)
{
    const TileJob J = jobs[blockIdx.x];
    int tx = threadIdx.x % kTileW, ty = threadIdx.x / kTileW;
    if (tx >= J.nx || ty >= J.ny) return;
    int X = J.X + tx, Y = J.Y + ty;
    const RenderDesc R = rds[J.r];
    int o = R.label[Y * R.label_pitch + X];
    int n0 = 0, n1 = 0, n2 = 0;
    if (o != kNoOwner) {
        const Plane P0 = planes[R.plane0 + o * (R.levels + 1)];
        int4 g = kUser ? load_user<U>(src, P0.src, X, Y) : load_plane<K>(P0, X, Y);
        n0 = g.x;
        n1 = g.y;
        n2 = g.z;
        if (R.levels > 0) {
            const Plane P1 = planes[R.plane0 + o * (R.levels + 1) + 1];
            auto f = [&](int xs, int ys) { return load_plane<K>(P1, xs, ys); };
            int4 u = up_at(f, X, Y, P1.cw, P1.ch);
            n0 = sat_s16(g.x - (short)u.x);
            n1 = sat_s16(g.y - (short)u.y);
            n2 = sat_s16(g.z - (short)u.z);
        }
    }
    if (R.levels > 0) {
        const CanvasLevel C1 = canvas[R.canvas0 + 1];
        auto f = [&](int xs, int ys) { return C1.c[ys * C1.pitch + xs]; };
        int4 u = up_at(f, X, Y, C1.cw, C1.ch);
        n0 = u.x + n0;
        n1 = u.y + n1;
        n2 = u.z + n2;
    }
    uint32_t out = 0;
    if (R.full || o != kNoOwner)
        out = (uint32_t)sat_u8(n0) | ((uint32_t)sat_u8(n1) << 8) | ((uint32_t)sat_u8(n2) << 16);
    ((uint32_t*)(outs.p[R.out] + (size_t)Y * outs.step[R.out]))[X] = out;
}

// begin/tile/end: caller image (kind U) + prepared mask -> internal 16-bit
// level-0 plane with channel 3 = valid.
template <int U>
__global__ __launch_bounds__(256) void k_import(const uint8_t* __restrict__ img, size_t step,
                                                const uint8_t* __restrict__ mask, int mask_pitch,
                                                short4* __restrict__ dst, int pitch, int w, int h)
{
    int X = blockIdx.x * 64 + (threadIdx.x & 63);
    int Y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (X >= w || Y >= h) return;
    const uint8_t* row = img + (size_t)Y * step;
    int4 v;
    if (U == kPixU8) v = unpack_u8(((const uint32_t*)row)[X]);
    else { short4 s = ((const short4*)row)[X]; v = make_int4(s.x, s.y, s.z, s.w); }
    short valid = mask[(size_t)Y * mask_pitch + X] ? 1 : 0;
    dst[(size_t)Y * pitch + X] = make_short4((short)v.x, (short)v.y, (short)v.z, valid);
}

__global__ __launch_bounds__(256) void k_copy_rows(const uint8_t* __restrict__ src,
                                                   size_t src_step, uint8_t* __restrict__ dst,
                                                   size_t dst_step, int row_bytes)
{
    const uint8_t* s = src + (size_t)blockIdx.y * src_step;
    uint8_t* d = dst + (size_t)blockIdx.y * dst_step;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < row_bytes; e += gridDim.x * 256) d[e] = s[e];
}

// ------------------------------------------------------------- launchers --

hipError_t launch_remap(int kind, const RemapJob* jobs, int njobs, const RemapCam* cams,
                        const Plane* planes, const uint8_t* luts, bool use_lut,
                        const SrcTable& src, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    dim3 g(njobs), b(256);
    if (kind == kPixU8) {
        if (use_lut) hipLaunchKernelGGL((k_remap<true, kPixU8>), g, b, 0, s, jobs, cams, planes, luts, src);
        else hipLaunchKernelGGL((k_remap<false, kPixU8>), g, b, 0, s, jobs, cams, planes, luts, src);
    } else {
        if (use_lut) hipLaunchKernelGGL((k_remap<true, kPixS16>), g, b, 0, s, jobs, cams, planes, luts, src);
        else hipLaunchKernelGGL((k_remap<false, kPixS16>), g, b, 0, s, jobs, cams, planes, luts, src);
    }
    return hipGetLastError();
}

hipError_t launch_down(int kind, int user_kind, const DownJob* jobs, int njobs,
                       const Plane* planes, const SrcTable& src, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    dim3 g(njobs), b(256);
    if (user_kind < 0) {
        if (kind == kPixU8) hipLaunchKernelGGL((k_down<kPixU8, false, 0>), g, b, 0, s, jobs, planes, src);
        else hipLaunchKernelGGL((k_down<kPixS16, false, 0>), g, b, 0, s, jobs, planes, src);
    } else if (kind != kPixS16) {
        return hipErrorInvalidValue;  // caller images always go through 16-bit planes
    } else if (user_kind == kPixU8) {
        hipLaunchKernelGGL((k_down<kPixS16, true, kPixU8>), g, b, 0, s, jobs, planes, src);
    } else {
        hipLaunchKernelGGL((k_down<kPixS16, true, kPixS16>), g, b, 0, s, jobs, planes, src);
    }
    return hipGetLastError();
}

hipError_t launch_blend(int kind, const TileJob* jobs, int njobs, const int* camlist,
                        const Plane* planes, const CanvasLevel* canvas, const RenderDesc* rd,
                        int level, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    if (kind == kPixU8)
        hipLaunchKernelGGL(k_blend<kPixU8>, dim3(njobs), dim3(256), 0, s, jobs, camlist, planes,
                           canvas, rd, level);
    else
        hipLaunchKernelGGL(k_blend<kPixS16>, dim3(njobs), dim3(256), 0, s, jobs, camlist, planes,
                           canvas, rd, level);
    return hipGetLastError();
}

hipError_t launch_final(int kind, int user_kind, const TileJob* jobs, int njobs,
                        const Plane* planes, const CanvasLevel* canvas, const RenderDesc* rd,
                        const SrcTable& src, const OutTable& outs, hipStream_t s)
{
    if (njobs <= 0) return hipSuccess;
    dim3 g(njobs), b(256);
    if (user_kind < 0) {
        if (kind == kPixU8) hipLaunchKernelGGL((k_final<kPixU8, false, 0>), g, b, 0, s, jobs, planes, canvas, rd, src, outs);
        else hipLaunchKernelGGL((k_final<kPixS16, false, 0>), g, b, 0, s, jobs, planes, canvas, rd, src, outs);
    } else if (kind != kPixS16) {
        return hipErrorInvalidValue;
    } else if (user_kind == kPixU8) {
        hipLaunchKernelGGL((k_final<kPixS16, true, kPixU8>), g, b, 0, s, jobs, planes, canvas, rd, src, outs);
    } else {
        hipLaunchKernelGGL((k_final<kPixS16, true, kPixS16>), g, b, 0, s, jobs, planes, canvas, rd, src, outs);
    }
    return hipGetLastError();
}

hipError_t launch_reproject16s(const uint8_t* src, size_t src_step, int sw, int sh,
                               const float* xmap, const float* ymap, size_t map_step,
                               int16_t* dst, size_t dst_step, int dw, int dh, hipStream_t s)
{
    dim3 g((dw + 63) / 64, (dh + 3) / 4);
    hipLaunchKernelGGL(k_reproject<false>, g, dim3(256), 0, s, src, (uint32_t)src_step, sw, sh, xmap,
                       ymap, map_step, (void*)dst, dst_step, dw, dh);
    return hipGetLastError();
}

hipError_t launch_reproject8u(const uint8_t* src, size_t src_step, int sw, int sh, const float* xmap,
                              const float* ymap, size_t map_step, uint8_t* dst, size_t dst_step, int dw,
                              int dh, hipStream_t s)
{
    dim3 g((dw + 63) / 64, (dh + 3) / 4);
    hipLaunchKernelGGL(k_reproject<true>, g, dim3(256), 0, s, src, (uint32_t)src_step, sw, sh, xmap,
                       ymap, map_step, (void*)dst, dst_step, dw, dh);
    return hipGetLastError();
}

hipError_t launch_transform(const uint8_t* src, size_t src_step, uint8_t* dst, size_t dst_step,
                            int w, int h, const uint8_t* lut_dev, hipStream_t s)
{
    dim3 g((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(k_transform, g, dim3(256), 0, s, src, src_step, dst, dst_step, w, h, lut_dev);
    return hipGetLastError();
}

hipError_t launch_import(int user_kind, const void* img, size_t step, const uint8_t* mask,
                         int mask_pitch, void* dst, int pitch, int w, int h, hipStream_t s)
{
    dim3 g((w + 63) / 64, (h + 3) / 4);
    if (user_kind == kPixU8)
        hipLaunchKernelGGL(k_import<kPixU8>, g, dim3(256), 0, s, (const uint8_t*)img, step, mask,
                           mask_pitch, (short4*)dst, pitch, w, h);
    else
        hipLaunchKernelGGL(k_import<kPixS16>, g, dim3(256), 0, s, (const uint8_t*)img, step, mask,
                           mask_pitch, (short4*)dst, pitch, w, h);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_zero(uint32_t* __restrict__ p, size_t words)
{
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < words; e += (size_t)gridDim.x * 256)
        p[e] = 0;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s)
{
    size_t words = bytes / 4;
    if (!p || !words) return hipSuccess;
    size_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, words);
    return hipGetLastError();
}

hipError_t launch_copy_plane(const void* src, size_t src_step, void* dst, size_t dst_step,
                             int row_bytes, int rows, hipStream_t s)
{
    if (rows <= 0 || row_bytes <= 0) return hipSuccess;
    int gx = (row_bytes + 255) / 256;
    if (gx > 64) gx = 64;
    hipLaunchKernelGGL(k_copy_rows, dim3(gx, rows), dim3(256), 0, s, (const uint8_t*)src, src_step,
                       (uint8_t*)dst, dst_step, row_bytes);
    return hipGetLastError();
}

}  // namespace pano
