// kernels_u8.hip -- the fused 8-bit path of render() (gfx950).
//
// When every sampled pixel lies inside its camera's mask (as maps from
// getReprojectMapAndMask guarantee, ZReproject.cpp:206-215), every Gaussian
// level of the reference's blend is an integer in [0, 255]: level 0 is a
// saturated bilinear sample and calcDstImage (MultiBandBlend.cpp:732-768)
// returns floor(255 v / alpha) <= 254 for v <= 255 alpha / 256.  The planes
// are then 4 x u8 (B, G, R, valid) and the 5x5 / 3x3 filter sums run two
// channels per 32-bit lane (no lane exceeds 16 bits: 5x5 sums <= 65280, pyrUp
// sums <= 16320).  The arithmetic is the reference's, bit for bit:
//   k_remap_down0  LUT + FP64 bilinear remap of a level-0 region into LDS
//                  (never stored in HBM), pyrDown + calcDstImage -> level 1;
//                  the owner camera's level-0 value goes to a canvas plane
//   k_down_u8      level j -> j+1, j >= 1
//   k_blend_u8     level j >= 1: sum_i W_ij (G_ij - up(G_i,j+1)), normalize,
//                  + up32(C_j+1)
//   k_final_u8     level 0: owner's G_0 - up(G_1) + up32(C_1) -> BGRx
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

namespace {

constexpr uint32_t kLo = 0x00ff00ffu;  // (B, R) lanes of a packed pixel; >> 8 gives (G, valid)
constexpr uint32_t kHi = 0xff00ff00u;  // (G, valid) lanes in place

// Down tile geometry of shape S: output tile kTW x kTH at level j+1, input
// region kIW x kIH = (2 kTW + 3) x (2 kTH + 3) at level j, staged by kDT
// threads, kPer pixels each, with all loads of a thread issued before any is
// used (the loops are latency-bound otherwise).
template <int S>
struct DownShape {
    static constexpr int kTW = kDownTileW[S], kTH = kDownTileH[S], kDT = kDownThreads[S];
    static constexpr int kIW = 2 * kTW + 3, kIH = 2 * kTH + 3;
    static constexpr int kPer = (kIH * kIW + kDT - 1) / kDT;
};

// Horizontal 5-tap over the staged input rows, then vertical 5-tap and
// calcDstImage; writes the (nx, ny) output tile at canvas (X, Y) of plane D.
template <class SH>
__device__ __forceinline__ void down_from_lds(const uint32_t* in, uint2* hs, const DownJob& J,
                                              const Plane& D)
{
    const int ih = 2 * J.ny + 3;
    for (int e = threadIdx.x; e < SH::kIH * SH::kTW; e += SH::kDT) {
        int r = e / SH::kTW, x = e - r * SH::kTW;
        if (r >= ih || x >= J.nx) continue;
        const uint32_t* p = in + r * SH::kIW + 2 * x;
        uint32_t a = p[0], b = p[1], c = p[2], d = p[3], f = p[4];
        uint32_t lo = (a & kLo) + (f & kLo) + 4u * ((b & kLo) + (d & kLo)) + 6u * (c & kLo);
        // (G, valid) in place: G sums reach bit 19, valid (0/1) sums sit at bits 24-28
        uint32_t hi = ((a & kHi) + (f & kHi) + 4u * ((b & kHi) + (d & kHi)) + 6u * (c & kHi)) >> 8;
        hs[r * SH::kTW + x] = make_uint2(lo, hi);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < SH::kTH * SH::kTW; e += SH::kDT) {
        int y = e / SH::kTW, x = e - y * SH::kTW;
        if (y >= J.ny || x >= J.nx) continue;
        const uint2* q = hs + (2 * y) * SH::kTW + x;
        uint2 a = q[0], b = q[SH::kTW], c = q[2 * SH::kTW], d = q[3 * SH::kTW], f = q[4 * SH::kTW];
        uint32_t lo = a.x + f.x + 4u * (b.x + d.x) + 6u * c.x;
        uint32_t hi = a.y + f.y + 4u * (b.y + d.y) + 6u * c.y;
        int al = (int)(hi >> 16);
        uint32_t o = 0;
        if (al) {
            o = (uint32_t)div_norm_u8((int)(lo & 0xffffu), al) |
                ((uint32_t)div_norm_u8((int)(hi & 0xffffu), al) << 8) |
                ((uint32_t)div_norm_u8((int)(lo >> 16), al) << 16) | (1u << 24);
        }
        int di = box_index(D, J.X + x, J.Y + y);
        if (di >= 0) stg((uint32_t*)D.g + di, o);
    }
}

__device__ __forceinline__ uint32_t load_u32(const Plane& P, int X, int Y)
{
    int i = box_index(P, X, Y);
    return i < 0 ? 0u : ldg((const uint32_t*)P.g + i);
}

// Pixel (X, Y) of a plane's box as a byte offset, or -1 outside it (box
// pitches and rows stay below 2^24, the plane below 2^31 bytes).
template <class Pl>
__device__ __forceinline__ int box_off32(const Pl& P, int X, int Y)
{
    const int ly = Y - P.y0;
    int lx = X - P.x0;
    if (lx < 0) lx += P.cw;
    if ((unsigned)ly >= (unsigned)P.bh || (unsigned)lx >= (unsigned)P.bw) return -1;
    return (int)(4u * (__umul24((uint32_t)ly, (uint32_t)P.pitch) + (uint32_t)lx));
}

template <class Pl>
__device__ __forceinline__ uint32_t load_u32_off(const Pl& P, int off)
{
    return off < 0 ? 0u : ldg((const uint32_t*)((const char*)P.g + off));
}

// Level-(j+1) staging region of a quad tile: element k of this thread is
// canvas pixel (xs[k], ys[k]) (border-resolved), ok[k] whether it exists.
constexpr int kSW = kQTW + 2, kSH = kQTH + 2;
constexpr int kQT = kQTW * kQTH;   // threads of a quad-tile block, one 2x2 quad each
constexpr int kStageIt = (kSH * kSW + kQT - 1) / kQT;
struct StageCoords {
    int xs[kStageIt], ys[kStageIt];
    bool ok[kStageIt];
};
__device__ __forceinline__ StageCoords stage_coords(int sx0, int sy0, int cw, int ch)
{
    StageCoords S;
#pragma unroll
    for (int k = 0; k < kStageIt; k++) {
        const int e = threadIdx.x + k * kQT;
        const int r = e / kSW, c = e - r * kSW;
        S.ok[k] = e < kSH * kSW;
        S.xs[k] = bwrap(sx0 + c, cw);
        S.ys[k] = breflect(sy0 + r, ch);
    }
    return S;
}

// pyrUp of a 2x2 quad from packed 8-bit pixels staged in LDS: s points at the
// quad's source pixel (x, y); stride is the LDS row length.  Outputs (even
// row, even col), (even, odd), (odd, even), (odd, odd) as packed u8 lanes.
__device__ __forceinline__ void up_quad_u8(const uint32_t* s, int stride, uint32_t out[4])
{
    uint32_t he_lo[3], he_hi[3], ho_lo[3], ho_hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t* row = s + (k - 1) * stride;
        uint32_t a = row[-1], b = row[0], d = row[1];
        he_lo[k] = (a & kLo) + 6u * (b & kLo) + (d & kLo);
        he_hi[k] = ((a >> 8) & kLo) + 6u * ((b >> 8) & kLo) + ((d >> 8) & kLo);
        ho_lo[k] = 4u * ((b & kLo) + (d & kLo));
        ho_hi[k] = 4u * (((b >> 8) & kLo) + ((d >> 8) & kLo));
    }
    const uint32_t R = 0x00200020u;
    uint32_t v00l = he_lo[0] + 6u * he_lo[1] + he_lo[2], v00h = he_hi[0] + 6u * he_hi[1] + he_hi[2];
    uint32_t v01l = ho_lo[0] + 6u * ho_lo[1] + ho_lo[2], v01h = ho_hi[0] + 6u * ho_hi[1] + ho_hi[2];
    uint32_t v10l = 4u * (he_lo[1] + he_lo[2]), v10h = 4u * (he_hi[1] + he_hi[2]);
    uint32_t v11l = 4u * (ho_lo[1] + ho_lo[2]), v11h = 4u * (ho_hi[1] + ho_hi[2]);
    out[0] = (((v00l + R) >> 6) & kLo) | ((((v00h + R) >> 6) & kLo) << 8);
    out[1] = (((v01l + R) >> 6) & kLo) | ((((v01h + R) >> 6) & kLo) << 8);
    out[2] = (((v10l + R) >> 6) & kLo) | ((((v10h + R) >> 6) & kLo) << 8);
    out[3] = (((v11l + R) >> 6) & kLo) | ((((v11h + R) >> 6) & kLo) << 8);
}

// pyrUp<int> of a quad from int4 pixels staged in LDS (channels 0..2).
__device__ __forceinline__ void up_quad_i32(const int4* s, int stride, int3 out[4])
{
    int3 he[3], ho[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int4* row = s + (k - 1) * stride;
        int4 a = row[-1], b = row[0], d = row[1];
        he[k] = make_int3(a.x + 6 * b.x + d.x, a.y + 6 * b.y + d.y, a.z + 6 * b.z + d.z);
        ho[k] = make_int3(4 * (b.x + d.x), 4 * (b.y + d.y), 4 * (b.z + d.z));
    }
    auto fin = [](int v) { return (v + 32) >> 6; };
    out[0] = make_int3(fin(he[0].x + 6 * he[1].x + he[2].x), fin(he[0].y + 6 * he[1].y + he[2].y),
                       fin(he[0].z + 6 * he[1].z + he[2].z));
    out[1] = make_int3(fin(ho[0].x + 6 * ho[1].x + ho[2].x), fin(ho[0].y + 6 * ho[1].y + ho[2].y),
                       fin(ho[0].z + 6 * ho[1].z + ho[2].z));
    out[2] = make_int3(fin(4 * (he[1].x + he[2].x)), fin(4 * (he[1].y + he[2].y)),
                       fin(4 * (he[1].z + he[2].z)));
    out[3] = make_int3(fin(4 * (ho[1].x + ho[2].x)), fin(4 * (ho[1].y + ho[2].y)),
                       fin(4 * (ho[1].z + ho[2].z)));
}

__device__ __forceinline__ int lane(uint32_t v, int c) { return (int)((v >> (8 * c)) & 255u); }

}  // namespace

// ------------------------------------------------------------ level 0 -> 1 --

// kLut: the render's gain LUTs (transform(), GainAdjust.cpp:295-346) are applied
// to the four taps as they arrive, when its "LUTs active" byte is set: the same
// bytes a LUT pass over the whole frames would have produced, without that pass.
template <class SH, bool kLut>
__global__ __launch_bounds__(SH::kDT) void k_remap_down0(const Seg* __restrict__ segs, int nseg,
                                                     SrcTable src, int sched)
{
    __shared__ uint32_t in[SH::kIH * SH::kIW];
    __shared__ uint2 hs[SH::kIH * SH::kTW];
    __shared__ uint32_t lut_w[kLut ? 192 : 1];
    const int bid = block_job(sched);
    const Seg& S = seg_of(segs, nseg, bid);
    const RenderDesc R = ldc(S.rd);
    const DownJob J = ldc((const DownJob*)S.jobs + (bid - S.begin));
    const RemapCam cam = ldc(R.cams + J.src);
    const Plane P0 = ldc(R.planes + cam.plane);
    const Plane D = ldc(R.planes + J.dst);
    const uint8_t* sp = src.p[S.src_base + cam.src];
    const uint32_t sstep = src.step[S.src_base + cam.src];
    uint32_t* g0comp = R.g0comp;
    const int g0pitch = R.g0pitch;
    const int bx0 = 2 * J.X - 2, by0 = 2 * J.Y - 2;
    const int iw = 2 * J.nx + 3, ih = 2 * J.ny + 3;
    const float fsw = (float)cam.sw, fsh = (float)cam.sh;
    const uint32_t owner_byte = (uint32_t)(J.src + 1) << 24;   // byte 3 of g0comp: owner + 1
    bool use_lut = false;
    if (kLut) {
        const uint8_t* lb = R.luts;
        use_lut = ldg(lb + (size_t)R.ncams * 768) != 0;   // uniform over the block
        if (use_lut) {
            for (int e = threadIdx.x; e < 192; e += SH::kDT) lut_w[e] = ldg((const uint32_t*)(lb + cam.lut) + e);
            __syncthreads();
        }
    }
    const uint8_t* lut = (const uint8_t*)lut_w;
    // the tile's core (level-0 pixels 2X.., 2Y..) clipped to the canvas
    const uint32_t core_w = (uint32_t)min(2 * J.nx, P0.cw - 2 * J.X);
    const uint32_t core_h = (uint32_t)min(2 * J.ny, P0.ch - 2 * J.Y);
    // the staging region is one rectangle of the margined map storage
    const char* mbase = (const char*)(cam.map + ((size_t)J.my * cam.mpitch + J.mx));
    const uint32_t mrow = (uint32_t)cam.mpitch * 8u;
    // Element e = threadIdx.x + k kDT of the region is (row, col) = (r, c);
    // kDT = kStepR kIW + kStepC, so the next slot adds (kStepR, kStepC) with a carry.
    constexpr int kStepR = SH::kDT / SH::kIW, kStepC = SH::kDT % SH::kIW;
    // Software pipeline over kRounds rounds of kBatch points per thread: the
    // map points of round r+1 are in flight while round r gathers and combines.
    constexpr int kBatch = 3, kRounds = (SH::kPer + kBatch - 1) / kBatch;
    float2 m[2][kBatch];
    int rr[kRounds * kBatch], cc[kRounds * kBatch];
    {
        int r = threadIdx.x / SH::kIW, c = threadIdx.x - (threadIdx.x / SH::kIW) * SH::kIW;
#pragma unroll
        for (int k = 0; k < kRounds * kBatch; k++) {
            rr[k] = r;
            cc[k] = c;
            r += kStepR;
            c += kStepC;
            if (c >= SH::kIW) { c -= SH::kIW; r++; }
        }
    }
    auto fetch_map = [&](int round, float2* mm) {
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            const int k = round * kBatch + b;
            const bool ok = rr[k] < ih && cc[k] < iw;   // also false beyond kIH rows
            const uint32_t off = __umul24((uint32_t)rr[k], mrow) + 8u * (uint32_t)cc[k];
            mm[b] = ok ? ldg_f2((const float2*)(mbase + off)) : make_float2(__builtin_inff(), 0.f);
        }
    };
    fetch_map(0, m[0]);
#pragma unroll
    for (int round = 0; round < kRounds; round++) {
        const int cur = round & 1;
        uint32_t px[kBatch][4];
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            const float x = __builtin_fabsf(m[cur][b].x), y = __builtin_fabsf(m[cur][b].y);
            px[b][0] = px[b][1] = px[b][2] = px[b][3] = 0;
            if (x < fsw && y < fsh) load_taps(sp, sstep, cam.sw, cam.sh, x, y, px[b]);
        }
        if (round + 1 < kRounds) fetch_map(round + 1, m[cur ^ 1]);
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            const int k = round * kBatch + b;
            const int e = threadIdx.x + k * SH::kDT;
            if (e >= SH::kIH * SH::kIW) continue;
            const int r = rr[k], c = cc[k];
            const uint32_t bx = __float_as_uint(m[cur][b].x), by = __float_as_uint(m[cur][b].y);
            const float x = __builtin_fabsf(m[cur][b].x), y = __builtin_fabsf(m[cur][b].y);
            uint32_t v = 0;
            if (x < fsw && y < fsh) {
                if (kLut && use_lut)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t p = px[b][q];
                        px[b][q] = (uint32_t)lut[p & 255u] | ((uint32_t)lut[256 + ((p >> 8) & 255u)] << 8) |
                                   ((uint32_t)lut[512 + ((p >> 16) & 255u)] << 16) | (p & 0xff000000u);
                    }
                bool tie = false;
                v = bilinear_fast(__builtin_amdgcn_fractf(x), __builtin_amdgcn_fractf(y), px[b], tie);
                if (tie) v = bilinear_combine((double)x, (double)y, px[b], nullptr);
            }
            v |= (by >> 31) << 24;   // kFlagValid
            // an owned core pixel (not a border-resolved copy of one) goes to the level-0 blend
            if ((bx >> 31) && (uint32_t)(c - 2) < core_w && (uint32_t)(r - 2) < core_h) {
                const int vx = bx0 + c, vy = by0 + r;
                // canvas byte offsets fit 32 bits (checked at prepare)
                const uint32_t go = 4u * (__umul24((uint32_t)vy, (uint32_t)g0pitch) + (uint32_t)vx);
                stg((uint32_t*)((char*)g0comp + go), (v & 0xffffffu) | owner_byte);
            }
            in[e] = v;
        }
    }
    __syncthreads();
    down_from_lds<SH>(in, hs, J, D);
}

// ------------------------------------------------------- level j -> j+1 --

template <class SH>
__global__ __launch_bounds__(SH::kDT) void k_down_u8(const Seg* __restrict__ segs, int nseg, int sched)
{
    __shared__ uint32_t in[SH::kIH * SH::kIW];
    __shared__ uint2 hs[SH::kIH * SH::kTW];
    const int bid = block_job(sched);
    const Seg& Sg = seg_of(segs, nseg, bid);
    const DownJob J = ldc((const DownJob*)Sg.jobs + (bid - Sg.begin));
    const Plane* planes = ldc(&Sg.rd->planes);
    const Plane S = ldc(planes + J.src);
    const Plane D = ldc(planes + J.dst);
    const int bx0 = 2 * J.X - 2, by0 = 2 * J.Y - 2;
    const int iw = 2 * J.nx + 3, ih = 2 * J.ny + 3;
    uint32_t v[SH::kPer];
#pragma unroll
    for (int k = 0; k < SH::kPer; k++) {
        int e = threadIdx.x + k * SH::kDT;
        int r = e / SH::kIW, c = e - r * SH::kIW;
        v[k] = 0;
        if (e < SH::kIH * SH::kIW && r < ih && c < iw)
            v[k] = load_u32(S, bwrap(bx0 + c, S.cw), breflect(by0 + r, S.ch));
    }
#pragma unroll
    for (int k = 0; k < SH::kPer; k++) {
        int e = threadIdx.x + k * SH::kDT;
        if (e < SH::kIH * SH::kIW) in[e] = v[k];
    }
    __syncthreads();
    down_from_lds<SH>(in, hs, J, D);
}

// ------------------------------------------------------- blend, level >= 1 --

constexpr int kBlendStage = 4;   // cameras of a blend tile staged at once

__global__ __launch_bounds__(kQT) void k_blend_u8(const Seg* __restrict__ segs, int nseg, int level, int sched)
{
    __shared__ uint32_t gs[kBlendStage][kSH * kSW];
    __shared__ int4 cs[kSH * kSW];
    const int bid = block_job(sched);
    const Seg& Sg = seg_of(segs, nseg, bid);
    const RenderDesc R = ldc(Sg.rd);
    const TileJob J = ldc((const TileJob*)Sg.jobs + (bid - Sg.begin));
    const bool top = J.flags & kTileTop;
    const int qx = threadIdx.x % kQTW, qy = threadIdx.x / kQTW;
    const int X = J.X + 2 * qx, Y = J.Y + 2 * qy;
    const int sx0 = (J.X >> 1) - 1, sy0 = (J.Y >> 1) - 1;
    const CanvasLevel C = ldc(R.canvas + level);
    bool in[4];
#pragma unroll
    for (int d = 0; d < 4; d++) in[d] = X + (d & 1) < J.X + J.nx && Y + (d >> 1) < J.Y + J.ny;
    int ws[4] = {1, 1, 1, 1};
    if (!R.full)
#pragma unroll
        for (int d = 0; d < 4; d++)
            if (in[d]) ws[d] = ldg(C.wsum + (__umul24((uint32_t)(Y + (d >> 1)), (uint32_t)C.pitch) + X + (d & 1)));
    // stage C_{j+1} and the group's level j+1 regions with every load in flight
    const CanvasLevel C1 = ldc(R.canvas + (top ? level : level + 1));
    const StageCoords SC = stage_coords(sx0, sy0, C1.cw, C1.ch);
    if (!top) {
        int4 cv[kStageIt];
#pragma unroll
        for (int k = 0; k < kStageIt; k++)
            cv[k] = SC.ok[k] ? ldg_c16(C1.c16 + __umul24((uint32_t)SC.ys[k], (uint32_t)C1.pitch) + SC.xs[k])
                             : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < kStageIt; k++) {
            int e = threadIdx.x + k * kQT;
            if (SC.ok[k]) cs[e] = cv[k];
        }
    }
    int3 acc[4];
#pragma unroll
    for (int d = 0; d < 4; d++) acc[d] = make_int3(0, 0, 0);
    for (int g = 0; g < J.cnt; g += kBlendStage) {
        const int ng = J.cnt - g < kBlendStage ? J.cnt - g : kBlendStage;
        if (!top) {
            PlaneRef P1[kBlendStage];
#pragma unroll
            for (int s = 0; s < kBlendStage; s++) P1[s] = ldc(R.blend_refs1 + J.off + g + (s < ng ? s : 0));
            uint32_t gv[kBlendStage][kStageIt];
#pragma unroll
            for (int s = 0; s < kBlendStage; s++)
#pragma unroll
                for (int k = 0; k < kStageIt; k++)
                    gv[s][k] = (s < ng && SC.ok[k]) ? load_u32_off(P1[s], box_off32(P1[s], SC.xs[k], SC.ys[k]))
                                                    : 0u;
#pragma unroll
            for (int s = 0; s < kBlendStage; s++)
#pragma unroll
                for (int k = 0; k < kStageIt; k++) {
                    int e = threadIdx.x + k * kQT;
                    if (s < ng && SC.ok[k]) gs[s][e] = gv[s][k];
                }
        }
        __syncthreads();
        // this level's pixels and weights of the group's cameras, one camera at a time
        // (holding all of them in registers costs more occupancy than it saves)
        PlaneRef P[kBlendStage];
#pragma unroll
        for (int s = 0; s < kBlendStage; s++) P[s] = ldc(R.blend_refs + J.off + g + (s < ng ? s : 0));
#pragma unroll
        for (int s = 0; s < kBlendStage; s++) {
            if (s >= ng) continue;
            int w[4];
            uint32_t gp[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const int off = in[d] ? box_off32(P[s], X + (d & 1), Y + (d >> 1)) : -1;
                w[d] = off >= 0 ? (int)ldg((const uint16_t*)((const char*)P[s].w + (off >> 1))) : 0;
                gp[d] = load_u32_off(P[s], off);
            }
            uint32_t u[4] = {0, 0, 0, 0};
            if (!top) up_quad_u8(&gs[s][(qy + 1) * kSW + qx + 1], kSW, u);
#pragma unroll
            for (int d = 0; d < 4; d++) {
                acc[d].x += (lane(gp[d], 0) - lane(u[d], 0)) * w[d];
                acc[d].y += (lane(gp[d], 1) - lane(u[d], 1)) * w[d];
                acc[d].z += (lane(gp[d], 2) - lane(u[d], 2)) * w[d];
            }
        }
        __syncthreads();
    }
    int3 up[4];
#pragma unroll
    for (int d = 0; d < 4; d++) up[d] = make_int3(0, 0, 0);
    if (!top) {
        if (J.cnt == 0) __syncthreads();  // cs staged, no camera loop ran
        up_quad_i32(cs + (qy + 1) * kSW + qx + 1, kSW, up);
    }
#pragma unroll
    for (int d = 0; d < 4; d++) {
        if (!in[d]) continue;
        int Xp = X + (d & 1), Yp = Y + (d >> 1);
        int a0 = acc[d].x, a1 = acc[d].y, a2 = acc[d].z;
        if (R.full) {
            a0 >>= 8;
            a1 >>= 8;
            a2 >>= 8;
        } else {
            int wv = ws[d] ? ws[d] : 1;
            a0 /= wv;
            a1 /= wv;
            a2 /= wv;
        }
        stg_c16(C.c16 + (__umul24((uint32_t)Yp, (uint32_t)C.pitch) + Xp), up[d].x + a0, up[d].y + a1,
                up[d].z + a2);
    }
}

// ------------------------------------------------------------ level 0 --

constexpr int kFinStage = 4;   // owner cameras staged at once (a tile rarely has more than 2)

__global__ __launch_bounds__(kQT) void k_final_u8(const Seg* __restrict__ segs, int nseg,
                                                  OutTable outs, int sched)
{
    __shared__ uint32_t gs[kFinStage][kSH * kSW];
    __shared__ int4 cs[kSH * kSW];
    const int bid = block_job(sched);
    const Seg& Sg = seg_of(segs, nseg, bid);
    const RenderDesc R = ldc(Sg.rd);
    const TileJob J = ldc((const TileJob*)Sg.jobs + (bid - Sg.begin));
    const int L = R.levels;
    const int qx = threadIdx.x % kQTW, qy = threadIdx.x / kQTW;
    const int X = J.X + 2 * qx, Y = J.Y + 2 * qy;
    const int sx0 = (J.X >> 1) - 1, sy0 = (J.Y >> 1) - 1;
    bool in[4];
    int lab[4];
    uint32_t g0[4];
#pragma unroll
    for (int d = 0; d < 4; d++) in[d] = X + (d & 1) < J.X + J.nx && Y + (d >> 1) < J.Y + J.ny;
    // the owners' level-0 samples, a row pair at a time (X is even, the canvas
    // pitch even: 8-byte aligned); byte 3 = owner + 1 (0: no owner, never written)
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const char* row = (const char*)R.g0comp + 4u * (__umul24((uint32_t)(Y + h), (uint32_t)R.g0pitch) +
                                                        (uint32_t)X);
        if (in[2 * h + 1]) {
            const uint2 v = ldg_u2((const uint2*)row);
            g0[2 * h] = v.x;
            g0[2 * h + 1] = v.y;
        } else {
            g0[2 * h] = in[2 * h] ? ldg((const uint32_t*)row) : 0u;
            g0[2 * h + 1] = 0u;
        }
    }
#pragma unroll
    for (int d = 0; d < 4; d++) lab[d] = g0[d] >> 24 ? (int)(g0[d] >> 24) - 1 : (int)kNoOwner;
    const CanvasLevel C1 = ldc(R.canvas + (L > 0 ? 1 : 0));
    const StageCoords SC = stage_coords(sx0, sy0, C1.cw, C1.ch);
    if (L > 0) {
        int4 cv[kStageIt];
#pragma unroll
        for (int k = 0; k < kStageIt; k++)
            cv[k] = SC.ok[k] ? ldg_c16(C1.c16 + __umul24((uint32_t)SC.ys[k], (uint32_t)C1.pitch) + SC.xs[k])
                             : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < kStageIt; k++) {
            int e = threadIdx.x + k * kQT;
            if (SC.ok[k]) cs[e] = cv[k];
        }
    }
    int lap[4][3];
#pragma unroll
    for (int d = 0; d < 4; d++)
#pragma unroll
        for (int c = 0; c < 3; c++) lap[d][c] = lane(g0[d], c);
    int3 up[4];
#pragma unroll
    for (int d = 0; d < 4; d++) up[d] = make_int3(0, 0, 0);
    if (L > 0) {
        for (int g = 0; g < J.cnt; g += kFinStage) {
            const int ng = J.cnt - g < kFinStage ? J.cnt - g : kFinStage;
            // the group's owner planes: independent descriptor loads
            PlaneRef O[kFinStage];
#pragma unroll
            for (int s = 0; s < kFinStage; s++) O[s] = ldc(R.fin_refs + J.off + g + (s < ng ? s : 0));
            uint32_t gv[kFinStage][kStageIt];
#pragma unroll
            for (int s = 0; s < kFinStage; s++)
#pragma unroll
                for (int k = 0; k < kStageIt; k++)
                    gv[s][k] = (s < ng && SC.ok[k]) ? load_u32_off(O[s], box_off32(O[s], SC.xs[k], SC.ys[k]))
                                                    : 0u;
#pragma unroll
            for (int s = 0; s < kFinStage; s++)
#pragma unroll
                for (int k = 0; k < kStageIt; k++) {
                    int e = threadIdx.x + k * kQT;
                    if (s < ng && SC.ok[k]) gs[s][e] = gv[s][k];
                }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kFinStage; s++) {
                const int o = O[s].cam;
                if (s >= ng || (lab[0] != o && lab[1] != o && lab[2] != o && lab[3] != o)) continue;
                uint32_t u[4];
                up_quad_u8(&gs[s][(qy + 1) * kSW + qx + 1], kSW, u);
#pragma unroll
                for (int d = 0; d < 4; d++)
                    if (lab[d] == o)
#pragma unroll
                        for (int c = 0; c < 3; c++) lap[d][c] = lane(g0[d], c) - lane(u[d], c);
            }
            __syncthreads();
        }
        if (J.cnt == 0) __syncthreads();  // cs staged, no owner loop ran
        up_quad_i32(cs + (qy + 1) * kSW + qx + 1, kSW, up);
    }
    uint8_t* out = outs.p[Sg.out];
    const uint32_t ostep = outs.step[Sg.out];
    const bool pairs = ((ostep | (uint32_t)(uintptr_t)out) & 7u) == 0;   // 8-byte aligned rows
    uint32_t o[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        o[d] = 0;
        if (R.full || lab[d] != kNoOwner)
            o[d] = (uint32_t)sat_u8(up[d].x + lap[d][0]) | ((uint32_t)sat_u8(up[d].y + lap[d][1]) << 8) |
                   ((uint32_t)sat_u8(up[d].z + lap[d][2]) << 16);
    }
    const int fmt = outs.fmt[Sg.out];
    if (fmt != kOutBGRX) {
        // the encoder's planes straight from the quad (canvas sides are even, so
        // every quad is whole): Y pairs, and U, V (interleaved for NV12)
        if (!in[0]) return;
        uint32_t yt, yb, u, v;
        bt601_block(o, yt, yb, u, v);
        uint8_t* yr = out + (__umul24((uint32_t)Y, ostep) + (uint32_t)X);
        stg((uint16_t*)yr, (uint16_t)yt);
        stg((uint16_t*)(yr + ostep), (uint16_t)yb);
        const uint32_t cy = (uint32_t)Y >> 1;
        if (fmt == kOutNV12) {
            stg((uint16_t*)(outs.p2[Sg.out] + (__umul24(cy, outs.step2[Sg.out]) + (uint32_t)X)),
                (uint16_t)(u | (v << 8)));
        } else {
            stg(outs.p2[Sg.out] + (__umul24(cy, outs.step2[Sg.out]) + ((uint32_t)X >> 1)), (uint8_t)u);
            stg(outs.p3[Sg.out] + (__umul24(cy, outs.step3[Sg.out]) + ((uint32_t)X >> 1)), (uint8_t)v);
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t* row = (uint32_t*)(out + (__umul24((uint32_t)(Y + h), ostep) + 4u * (uint32_t)X));
        if (pairs && in[2 * h + 1]) {
            stg_u2((uint2*)row, make_uint2(o[2 * h], o[2 * h + 1]));
        } else {
            if (in[2 * h]) stg(row, o[2 * h]);
            if (in[2 * h + 1]) stg(row + 1, o[2 * h + 1]);
        }
    }
}

// ------------------------------------------------------------- launchers --

namespace {
template <int S>
void launch_down_shape(int stage, const Seg* segs, int nseg, int blocks, const SrcTable& src, int sched,
                       bool luts, hipStream_t s)
{
    using SH = DownShape<S>;
    if (stage == kStageRemapDown0 && luts)
        hipLaunchKernelGGL((k_remap_down0<SH, true>), dim3(blocks), dim3(SH::kDT), 0, s, segs, nseg, src, sched);
    else if (stage == kStageRemapDown0)
        hipLaunchKernelGGL((k_remap_down0<SH, false>), dim3(blocks), dim3(SH::kDT), 0, s, segs, nseg, src, sched);
    else
        hipLaunchKernelGGL(k_down_u8<SH>, dim3(blocks), dim3(SH::kDT), 0, s, segs, nseg, sched);
}
}  // namespace

hipError_t launch_seg(int stage, const Seg* segs, int nseg, int blocks, int level,
                      const SrcTable& src, const OutTable& outs, int sched, int shape, bool luts, hipStream_t s)
{
    if (blocks <= 0 || nseg <= 0) return hipSuccess;
    static_assert(kDownShapes == 5, "one launch_down_shape per shape");
    switch (stage) {
    case kStageRemapDown0:
    case kStageDown:
        switch (shape) {
        case 0: launch_down_shape<0>(stage, segs, nseg, blocks, src, sched, luts, s); break;
        case 1: launch_down_shape<1>(stage, segs, nseg, blocks, src, sched, luts, s); break;
        case 2: launch_down_shape<2>(stage, segs, nseg, blocks, src, sched, luts, s); break;
        case 3: launch_down_shape<3>(stage, segs, nseg, blocks, src, sched, luts, s); break;
        case 4: launch_down_shape<4>(stage, segs, nseg, blocks, src, sched, luts, s); break;
        default: return hipErrorInvalidValue;
        }
        break;
    case kStageBlend:
        hipLaunchKernelGGL(k_blend_u8, dim3(blocks), dim3(kQT), 0, s, segs, nseg, level, sched);
        break;
    case kStageFinal:
        hipLaunchKernelGGL(k_final_u8, dim3(blocks), dim3(kQT), 0, s, segs, nseg, outs, sched);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


}  // namespace pano
