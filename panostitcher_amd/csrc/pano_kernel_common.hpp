// pano_kernel_common.hpp -- device helpers shared by the gfx950 kernels:
// OpenCV border rules, box addressing, the pyrUp tap pattern, the FP64
// bilinear sample and the exact calcDstImage division for 8-bit planes.
#pragma once
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"

namespace pano {

// ---------------------------------------------------------------- helpers --

// Loads and stores through pointers held in descriptors: the compiler cannot
// see that those point to global memory and would emit flat instructions.
#define PANO_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T ldg(const T* p)
{
    return *(const PANO_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ void stg(T* p, T v)
{
    *(PANO_GLOBAL T*)p = v;
}
// Block-uniform descriptor reads (jobs, planes, canvas levels, camera lists)
// through pointers held in other descriptors: read word by word from the
// constant address space so that they become scalar loads.
#define PANO_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ T ldc(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4");
    T v;
    uint32_t* d = (uint32_t*)&v;
    const PANO_CONST uint32_t* s = (const PANO_CONST uint32_t*)p;
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; i++) d[i] = s[i];
    return v;
}

// HIP's vector structs do not take the address-space cast as one access; go
// through the compiler's native vector types.
typedef int pano_i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 ldg4(const int4* p)
{
    pano_i32x4 v = *(const PANO_GLOBAL pano_i32x4*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg4(int4* p, int4 v)
{
    pano_i32x4 w = {v.x, v.y, v.z, v.w};
    *(PANO_GLOBAL pano_i32x4*)p = w;
}
typedef short pano_i16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ short4 ldg_s4(const short4* p)
{
    pano_i16x4 v = *(const PANO_GLOBAL pano_i16x4*)p;
    return make_short4(v.x, v.y, v.z, v.w);
}
typedef unsigned int pano_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 ldg_u2(const uint2* p)
{
    pano_u32x2 v = *(const PANO_GLOBAL pano_u32x2*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void stg_u2(uint2* p, uint2 v)
{
    pano_u32x2 w = {v.x, v.y};
    *(PANO_GLOBAL pano_u32x2*)p = w;
}
// C_j pixel of the 8-bit path: channels 0..2 as int16.
__device__ __forceinline__ int4 ldg_c16(const uint2* p)
{
    pano_u32x2 v = *(const PANO_GLOBAL pano_u32x2*)p;
    return make_int4((int)(short)(v.x & 0xffffu), (int)v.x >> 16, (int)(short)(v.y & 0xffffu), 0);
}
__device__ __forceinline__ void stg_c16(uint2* p, int a, int b, int c)
{
    pano_u32x2 w = {((uint32_t)a & 0xffffu) | ((uint32_t)b << 16), (uint32_t)c & 0xffffu};
    *(PANO_GLOBAL pano_u32x2*)p = w;
}
__device__ __forceinline__ float2 ldg_f2(const float2* p)
{
    unsigned long long v = *(const PANO_GLOBAL unsigned long long*)p;
    return make_float2(__uint_as_float((uint32_t)v), __uint_as_float((uint32_t)(v >> 32)));
}

// cv::borderInterpolate for WRAP (columns) and REFLECT_101 (rows).
__device__ __forceinline__ int bwrap(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    p %= n;
    return p < 0 ? p + n : p;
}

__device__ __forceinline__ int breflect(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    do {
        p = p < 0 ? -p : 2 * n - 2 - p;
    } while ((unsigned)p >= (unsigned)n);
    return p;
}

// Canvas pixel (X, Y), both inside the level canvas -> index into the box, or -1.
__device__ __forceinline__ int box_index(const Plane& P, int X, int Y)
{
    int ly = Y - P.y0;
    if ((unsigned)ly >= (unsigned)P.bh) return -1;
    int lx = X - P.x0;
    if (lx < 0) lx += P.cw;
    if ((unsigned)lx >= (unsigned)P.bw) return -1;
    return ly * P.pitch + lx;
}

__device__ __forceinline__ int4 unpack_u8(uint32_t v)
{
    return make_int4((int)(v & 255u), (int)((v >> 8) & 255u), (int)((v >> 16) & 255u),
                     (int)(v >> 24));
}

__device__ __forceinline__ int sat_s16(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }
__device__ __forceinline__ int sat_u8(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

template <int K>
__device__ __forceinline__ int4 load_at(const void* g, int i)
{
    if (K == kPixU8) return unpack_u8(ldg((const uint32_t*)g + i));
    short4 s = ldg_s4((const short4*)g + i);
    return make_int4(s.x, s.y, s.z, s.w);
}

// Gaussian pixel of a plane at canvas (X, Y); zero outside its box.
template <int K>
__device__ __forceinline__ int4 load_plane(const Plane& P, int X, int Y)
{
    int i = box_index(P, X, Y);
    if (i < 0) return make_int4(0, 0, 0, 0);
    return load_at<K>(P.g, i);
}

// Pixel of a caller image (level 0, full canvas).
template <int K>
__device__ __forceinline__ int4 load_user(const SrcTable& src, int s, int X, int Y)
{
    const uint8_t* row = src.p[s] + (size_t)Y * src.step[s];
    if (K == kPixU8) return unpack_u8(ldg((const uint32_t*)row + X));
    short4 v = ldg_s4((const short4*)row + X);
    return make_int4(v.x, v.y, v.z, v.w);
}

// pyrUp_<FixPtCast<T,6>> value at level-j pixel (X, Y) from the level-(j+1)
// source addressed by fetch(Xs, Ys) (border-resolved canvas coordinates):
// even output column/row taps [1 6 1] at x-1, x, x+1; odd taps [4 4] at x, x+1.
template <class Fetch>
__device__ __forceinline__ int4 up_at(const Fetch& fetch, int X, int Y, int cw1, int ch1)
{
    int x = X >> 1, y = Y >> 1;
    int xm = bwrap(x - 1, cw1), xp = bwrap(x + 1, cw1);
    int ym = breflect(y - 1, ch1), yp = breflect(y + 1, ch1);
    bool ox = X & 1, oy = Y & 1;
    auto hrow = [&](int ys) -> int4 {
        int4 c = fetch(x, ys), r = fetch(xp, ys);
        if (ox) return make_int4(4 * (c.x + r.x), 4 * (c.y + r.y), 4 * (c.z + r.z), 4 * (c.w + r.w));
        int4 l = fetch(xm, ys);
        return make_int4(l.x + 6 * c.x + r.x, l.y + 6 * c.y + r.y, l.z + 6 * c.z + r.z,
                         l.w + 6 * c.w + r.w);
    };
    int4 m = hrow(y), b = hrow(yp), v;
    if (oy) {
        v = make_int4(4 * (m.x + b.x), 4 * (m.y + b.y), 4 * (m.z + b.z), 4 * (m.w + b.w));
    } else {
        int4 t = hrow(ym);
        v = make_int4(t.x + 6 * m.x + b.x, t.y + 6 * m.y + b.y, t.z + 6 * m.z + b.z,
                      t.w + 6 * m.w + b.w);
    }
    return make_int4((v.x + 32) >> 6, (v.y + 32) >> 6, (v.z + 32) >> 6, (v.w + 32) >> 6);
}

// Byte offsets (from the image base) of the four samples of
// bilinearResampling at a valid point (x, y): (x0, y0), its right neighbour in
// row-major order, (x0, y1) and the right neighbour of that.
__device__ __forceinline__ void bilinear_offsets(uint32_t step, int sw, int sh, double x, double y,
                                                 uint32_t off[4])
{
    int x0 = (int)x, y0 = (int)y;
    int y1 = y0 + 1 > sh - 1 ? sh - 1 : y0 + 1;
    int xn0 = x0 + 1, xn1 = x0 + 1, yn0 = y0, yn1 = y1;
    if (x0 + 1 >= sw) {
        xn0 = xn1 = 0;
        yn0 = y0 + 1;
        yn1 = y1 + 1;
        if (yn0 >= sh) { yn0 = y0; xn0 = x0; }
        if (yn1 >= sh) { yn1 = y1; xn1 = x0; }
    }
    off[0] = (uint32_t)y0 * step + 4u * (uint32_t)x0;
    off[1] = (uint32_t)yn0 * step + 4u * (uint32_t)xn0;
    off[2] = (uint32_t)y1 * step + 4u * (uint32_t)x0;
    off[3] = (uint32_t)yn1 * step + 4u * (uint32_t)xn1;
}

// The four gathered BGRx samples combined with the weights of (x, y): the gain
// LUT of transform() (GainAdjust.cpp:295-346) applied to each byte, products
// and sums rounded one at a time in the reference's order (p00 w00 + p01 w01
// + p10 w10 + p11 w11), cvRound = round half to even.
__device__ __forceinline__ uint32_t bilinear_combine(double x, double y, const uint32_t p[4],
                                                     const uint8_t* lut)
{
    int x0 = (int)x, y0 = (int)y;
    double wx0 = __dadd_rn(x, -(double)x0), wx1 = __dadd_rn(1.0, -wx0);
    double wy0 = __dadd_rn(y, -(double)y0), wy1 = __dadd_rn(1.0, -wy0);
    double w00 = __dmul_rn(wx1, wy1), w01 = __dmul_rn(wx0, wy1);
    double w10 = __dmul_rn(wx1, wy0), w11 = __dmul_rn(wx0, wy0);
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int sh8 = 8 * c;
        int a = (p[0] >> sh8) & 255, b = (p[1] >> sh8) & 255;
        int d = (p[2] >> sh8) & 255, e = (p[3] >> sh8) & 255;
        if (lut) {
            a = lut[c * 256 + a];
            b = lut[c * 256 + b];
            d = lut[c * 256 + d];
            e = lut[c * 256 + e];
        }
        double r = __dmul_rn((double)a, w00);
        r = __dadd_rn(r, __dmul_rn((double)b, w01));
        r = __dadd_rn(r, __dmul_rn((double)d, w10));
        r = __dadd_rn(r, __dmul_rn((double)e, w11));
        out |= (uint32_t)sat_u8(__double2int_rn(r)) << sh8;
    }
    return out;
}

// Byte offsets of the four samples of bilinearResampling at a point inside
// the image, for the 32-bit frame addressing of the fused path (row step and
// frame size below 2^24 and 2^32 bytes): (x0, y0), its right neighbour in
// row-major order (the next row's first pixel at x0 = w - 1, the pixel itself
// at the image's last pixel), then the same for row y1 = min(y0 + 1, h - 1).
__device__ __forceinline__ void tap_offsets(uint32_t sstep, int sw, int sh, float x, float y,
                                            uint32_t o[4])
{
    const int x0 = (int)x, y0 = (int)y;
    const bool ylast = y0 + 1 >= sh;
    const uint32_t o0 = __umul24((uint32_t)y0, sstep) + 4u * (uint32_t)x0;
    const uint32_t o2 = ylast ? o0 : o0 + sstep;
    uint32_t o1 = o0 + 4u, o3 = o2 + 4u;
    if (x0 + 1 >= sw) {
        const uint32_t nxt = sstep - 4u * (uint32_t)x0;
        o1 = ylast ? o0 : o0 + nxt;
        o3 = y0 + 2 >= sh ? o2 : o2 + nxt;
    }
    o[0] = o0;
    o[1] = o1;
    o[2] = o2;
    o[3] = o3;
}

// tap_offsets with the interior case first: away from the last row and column
// the four samples are (x0, y0), (x0 + 1, y0) and the same one row down; the
// edge rules of tap_offsets run only in lanes that need them.
__device__ __forceinline__ void tap_offsets_fast(uint32_t sstep, int sw, int sh, float x, float y,
                                                 uint32_t o[4])
{
    const int x0 = (int)x, y0 = (int)y;
    if (x0 + 1 >= sw || y0 + 1 >= sh) {
        tap_offsets(sstep, sw, sh, x, y, o);
        return;
    }
    const uint32_t o0 = __umul24((uint32_t)y0, sstep) + 4u * (uint32_t)x0;
    o[0] = o0;
    o[1] = o0 + 4u;
    o[2] = o0 + sstep;
    o[3] = o0 + sstep + 4u;
}

// The four bilinear taps of (x, y), a point inside the image: away from the
// last row and column two 8-byte loads (a pixel and its right neighbour, on a
// 4-byte boundary: the hardware takes dword-aligned multi-dword loads), else
// the four loads at tap_offsets' edge rules.
typedef uint32_t pano_u32x2 __attribute__((ext_vector_type(2)));
typedef pano_u32x2 pano_u32x2_a4 __attribute__((aligned(4)));
__device__ __forceinline__ void load_taps(const uint8_t* sp, uint32_t sstep, int sw, int sh, float x, float y,
                                          uint32_t px[4])
{
    const int x0 = (int)x, y0 = (int)y;
    if (x0 + 1 < sw && y0 + 1 < sh) {
        const uint32_t o0 = __umul24((uint32_t)y0, sstep) + 4u * (uint32_t)x0;
        const pano_u32x2 a = *(const PANO_GLOBAL pano_u32x2_a4*)(sp + o0);
        const pano_u32x2 c = *(const PANO_GLOBAL pano_u32x2_a4*)(sp + o0 + sstep);
        px[0] = a.x;
        px[1] = a.y;
        px[2] = c.x;
        px[3] = c.y;
        return;
    }
    uint32_t o[4];
    tap_offsets(sstep, sw, sh, x, y, o);
#pragma unroll
    for (int q = 0; q < 4; q++) px[q] = ldg((const uint32_t*)(sp + o[q]));
}

// BT.601 studio swing in 20-bit fixed point: OpenCV's cvtColor table, which the
// encoder hand-off uses (CudaAccel/ImageProc.cu:44-60).  Forward: Y of every
// pixel of a 2x2 block, U and V from its left column (top and bottom averaged
// with one more shift, ImageProc.cu:70-104).  Inverse: one (U, V) per block.
namespace bt601 {
constexpr int kRY = 269484, kGY = 528482, kBY = 102760;
constexpr int kRU = -155188, kGU = -305135, kBU = 460324;
constexpr int kGV = -385875, kBV = -74448;   // V's red coefficient is kBU
constexpr int kY = 1220542, kUB = 2116026, kUG = -409993, kVG = -852492, kVR = 1673527;
constexpr int kShift = 20, kHalf = 1 << 19;
}  // namespace bt601

__device__ __forceinline__ uint32_t bt601_y(uint32_t p)
{
    using namespace bt601;
    const int v = kRY * (int)((p >> 16) & 255u) + kGY * (int)((p >> 8) & 255u) + kBY * (int)(p & 255u) + kHalf +
                  (16 << kShift);
    return (uint32_t)sat_u8(v >> kShift);
}

// q: the block's BGRx pixels (top-left, top-right, bottom-left, bottom-right).
// yt / yb: the two Y bytes of the top / bottom row (left in the low byte).
__device__ __forceinline__ void bt601_block(const uint32_t q[4], uint32_t& yt, uint32_t& yb, uint32_t& u,
                                            uint32_t& v)
{
    using namespace bt601;
    yt = bt601_y(q[0]) | (bt601_y(q[1]) << 8);
    yb = bt601_y(q[2]) | (bt601_y(q[3]) << 8);
    int us = 0, vs = 0;
#pragma unroll
    for (int k = 0; k < 4; k += 2) {
        const int b = (int)(q[k] & 255u), g = (int)((q[k] >> 8) & 255u), r = (int)((q[k] >> 16) & 255u);
        us += kRU * r + kGU * g + kBU * b + kHalf + (128 << kShift);
        vs += kBU * r + kGV * g + kBV * b + kHalf + (128 << kShift);
    }
    u = (uint32_t)sat_u8(us >> (kShift + 1));
    v = (uint32_t)sat_u8(vs >> (kShift + 1));
}

// The inverse for one block: ys = the four Y bytes (top-left first), channel 3 = 0.
__device__ __forceinline__ void bt601_to_bgr(const uint32_t ys[4], int uval, int vval, uint32_t q[4])
{
    using namespace bt601;
    const int u = uval - 128, v = vval - 128;
    const int ruv = kHalf + kVR * v, guv = kHalf + kVG * v + kUG * u, buv = kHalf + kUB * u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int y = max(0, (int)ys[k] - 16) * kY;
        q[k] = (uint32_t)sat_u8((y + buv) >> kShift) | ((uint32_t)sat_u8((y + guv) >> kShift) << 8) |
               ((uint32_t)sat_u8((y + ruv) >> kShift) << 16);
    }
}

// The job a block of a batched launch runs.  Workgroups are dispatched to the
// 8 XCDs round-robin (block b to XCD b mod 8), each XCD with its own L2.
// sched 0: job = block.  sched 1: each XCD gets one contiguous run of jobs, so
// neighbouring tiles, whose staged halos overlap, share an L2.  Both are
// bijections on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int block_job(int sched)
{
    const int b = blockIdx.x;
    if (sched == 0) return b;
    const int B = gridDim.x;
    const int per = (B + kXcds - 1) / kXcds;   // jobs of the first `tall` XCDs
    const int tall = B - kXcds * (per - 1);    // 1 .. kXcds
    const int x = b % kXcds, l = b / kXcds;
    return x < tall ? x * per + l : tall * per + (x - tall) * (per - 1) + l;
}

// Segment of a batched launch that owns block b (segments are few and ordered).
__device__ __forceinline__ const Seg& seg_of(const Seg* segs, int nseg, int b)
{
    int s = 0;
    while (s + 1 < nseg && b >= segs[s + 1].begin) s++;
    return segs[s];
}

// The same result from single-precision arithmetic where that is provably
// equal.  With wx = x - x0 and wy = y - y0 (exact in float), the reference's
// exact sum is V = H0 + (H1 - H0) wy with Hk = pk0 (1 - wx) + pk1 wx, because
// its weights are the exact products of (1 - wx, wx) and (1 - wy, wy).  Each
// float operation rounds once at magnitude < 256 (half an ulp <= 2^-17):
// hk = fma(wx, pk1, fma(-wx, pk0, pk0)) is within 2 * 2^-17 of Hk, and
// v = fma(h1 - h0, wy, h0) within 4 * 2^-17 = 2^-15 of V; the FP64 sum of the
// reference is within 2^-40 of V.  Unless a half-integer lies within kTieEps
// of v, both round to the integer nearest v.  tie is set when some channel is
// that close; the caller then recomputes the pixel with bilinear_combine.
constexpr float kTieEps = 0x1p-14f;
__device__ __forceinline__ uint32_t bilinear_fast(float wx, float wy, const uint32_t p[4], bool& tie)
{
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int sh8 = 8 * c;
        float a = (float)((p[0] >> sh8) & 255u), b = (float)((p[1] >> sh8) & 255u);
        float d = (float)((p[2] >> sh8) & 255u), e = (float)((p[3] >> sh8) & 255u);
        float h0 = __fmaf_rn(wx, b, __fmaf_rn(-wx, a, a));
        float h1 = __fmaf_rn(wx, e, __fmaf_rn(-wx, d, d));
        float v = __fmaf_rn(h1 - h0, wy, h0);
        float f = __builtin_amdgcn_fractf(v) - 0.5f;
        tie |= __builtin_fabsf(f) < kTieEps;
        // round to nearest and saturate into byte c (v is within 2^-15 of [0, 255])
        out = __builtin_amdgcn_cvt_pk_u8_f32(v, (uint32_t)c, out);
    }
    return out;
}

// bilinearResampling<short,3> (ZReproject.cpp:101-130) with the LUT: the
// right sample of each row is the next pixel in row-major order (the
// reference's pointer walk on a continuous image, which at x0 = w - 1 is the
// first pixel of the next row); at the image's last pixel it is the pixel itself.
__device__ __forceinline__ uint32_t bilinear_bgr(const uint8_t* src, uint32_t step, int sw, int sh,
                                                 double x, double y, const uint8_t* lut)
{
    uint32_t off[4], p[4];
    bilinear_offsets(step, sw, sh, x, y, off);
#pragma unroll
    for (int k = 0; k < 4; k++) p[k] = ldg((const uint32_t*)(src + off[k]));
    return bilinear_combine(x, y, p, lut);
}

// floor(255 v / (256 a)) for 0 <= v <= 65280, 1 <= a <= 256: t = (255 v) >> 8,
// then t / a from a reciprocal estimate corrected to the exact quotient.
__device__ __forceinline__ int div_norm_u8(int v, int a)
{
    int t = (255 * v) >> 8;
    int q = (int)((float)t * __builtin_amdgcn_rcpf((float)a));
    int r = t - q * a;
    q += (r >= a) - (r < 0);
    return q;
}


}  // namespace pano
