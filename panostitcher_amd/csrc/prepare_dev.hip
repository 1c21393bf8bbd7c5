// prepare_dev.hip -- the tables of a render prepare built on the device.
//
// TilingMultibandBlendFast::prepare (Blend/MultiBandBlend.cpp:906-985) builds
// per camera the Gaussian pyramid of 256 * unique mask (pyrDown, FixPtCast
// <short, 8>: (s + 128) >> 8, WRAP columns, REFLECT_101 rows), the coverage
// test and the weight sums; the fused render also needs each camera's box per
// level, the alpha-validity pyramid (which tiles can hold non-zero values), the
// owner of every canvas pixel and the margined map storage.  Here the
// per-pixel work runs on the device from device masks and unique masks; the
// host keeps only the per-row / per-column / per-tile flags it needs to lay
// out the boxes and the work lists (the same layout code as the host prepare).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/panohip.h"
#include "pano_dev.hpp"
#include "pano_host.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

namespace {

// One (camera, level) plane for the kernels: its box, its offset in the
// weight / validity buffers and its down-tile flag grid (tiles of the down
// jobs that write this level: one grid over the box's two column segments).
struct DevGeom {
    int x0, y0, bw, bh, cw, ch, pitch;
    int dflag, dcols0, dcols;   // flag offset, columns of segment 0, all columns
    size_t off;                 // element offset of the plane
};

__device__ __forceinline__ int dwrap(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    p %= n;
    return p < 0 ? p + n : p;
}

__device__ __forceinline__ int dreflect(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    do {
        p = p < 0 ? -p : 2 * n - 2 - p;
    } while ((unsigned)p >= (unsigned)n);
    return p;
}

// Rows and columns of each mask that hold a set pixel.  Block (y, camera);
// colany must be zero on entry (every writer stores 1).
__global__ __launch_bounds__(256) void k_mask_proj(const uint8_t* __restrict__ masks, int rows, int cols,
                                                   size_t plane, uint8_t* __restrict__ colany,
                                                   uint8_t* __restrict__ rowany)
{
    const int i = blockIdx.y, y = blockIdx.x;
    const uint8_t* row = masks + plane * i + (size_t)y * cols;
    int any = 0;
    for (int x = threadIdx.x; x < cols; x += 256)
        if (ldg(row + x)) {
            any = 1;
            colany[(size_t)i * cols + x] = 1;
        }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) rowany[(size_t)i * rows + y] = (uint8_t)any;
}

// Owner of each canvas pixel (the camera whose unique mask holds it), the
// masks covering it as bits (cover, n <= 8, optional) and whether any pixel is
// covered by no mask (*uncovered = 1; zero on entry).
__global__ __launch_bounds__(256) void k_label_cover(const uint8_t* __restrict__ masks,
                                                     const uint8_t* __restrict__ unique, int n, size_t P,
                                                     uint8_t* __restrict__ label, uint8_t* __restrict__ cover,
                                                     int* __restrict__ uncovered)
{
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < P; p += (size_t)gridDim.x * 256) {
        uint8_t lab = kNoOwner, cov = 0;
        bool any = false;
        for (int i = 0; i < n; i++) {
            if (ldg(unique + P * i + p)) lab = (uint8_t)i;
            if (ldg(masks + P * i + p)) {
                any = true;
                if (i < 8) cov |= (uint8_t)(1u << i);
            }
        }
        label[p] = lab;
        if (cover) cover[p] = cov;
        if (!any) *uncovered = 1;
    }
}

// Level j >= 1 of every camera: weight (s + 128) >> 8 of the 5x5 binomial sum
// of level j - 1 (level 0: 256 where the unique mask is set) and validity (any
// valid pixel in the footprint; level 0: the mask), both zero outside the
// source box.  Marks the blend tile (hit) of every non-zero weight and the
// down tile (dflag) of every valid pixel.  Grid (pitch / 64, bh / 4, n).
__global__ __launch_bounds__(256) void k_level(int j, int L, const DevGeom* __restrict__ geoms,
                                               const uint8_t* __restrict__ unique, const uint8_t* __restrict__ masks,
                                               int cols, size_t plane, uint16_t* __restrict__ w,
                                               uint8_t* __restrict__ v, uint8_t* __restrict__ hit, int tw, int tiles,
                                               int ttw, int tth, int dtw, int dth, uint8_t* __restrict__ dflag)
{
    const int i = blockIdx.z;
    const DevGeom gd = geoms[i * (L + 1) + j], gs = geoms[i * (L + 1) + j - 1];
    const int lx = blockIdx.x * 64 + threadIdx.x, ly = blockIdx.y * 4 + threadIdx.y;
    if (ly >= gd.bh || lx >= gd.pitch) return;
    const size_t o = gd.off + (size_t)ly * gd.pitch + lx;
    if (lx >= gd.bw) {
        w[o] = 0;
        v[o] = 0;
        return;
    }
    int X = gd.x0 + lx;
    if (X >= gd.cw) X -= gd.cw;
    const int Y = gd.y0 + ly;
    const int k5[5] = {1, 4, 6, 4, 1};
    int s = 0, any = 0;
    for (int a = 0; a < 5; a++) {
        const int r = dreflect(2 * Y + a - 2, gs.ch);
        int hs = 0;
        for (int b = 0; b < 5; b++) {
            const int c = dwrap(2 * X + b - 2, gs.cw);
            int val, m;
            if (j == 1) {
                const size_t q = plane * i + (size_t)r * cols + c;
                val = ldg(unique + q) ? 256 : 0;
                m = ldg(masks + q) != 0;
            } else {
                const int ry = r - gs.y0;
                int rx = c - gs.x0;
                if (rx < 0) rx += gs.cw;
                const bool in = (unsigned)ry < (unsigned)gs.bh && (unsigned)rx < (unsigned)gs.bw;
                const size_t q = gs.off + (size_t)ry * gs.pitch + rx;
                val = in ? w[q] : 0;
                m = in ? v[q] : 0;
            }
            hs += k5[b] * val;
            any |= m;
        }
        s += k5[a] * hs;
    }
    const int wv = (s + 128) >> 8;
    w[o] = (uint16_t)wv;
    v[o] = (uint8_t)any;
    if (wv) hit[(size_t)i * tiles + (size_t)(Y / tth) * tw + X / ttw] = 1;
    if (any) {
        const int col = X >= gd.x0 ? lx / dtw : gd.dcols0 + X / dtw;
        dflag[gd.dflag + (ly / dth) * gd.dcols + col] = 1;
    }
}

// Level-j weight sums (MultiBandBlend.cpp:970-980) in 16 bits, as the host sums.
__global__ __launch_bounds__(256) void k_wsum(int j, int L, int n, const DevGeom* __restrict__ geoms,
                                              const uint16_t* __restrict__ w, uint16_t* __restrict__ wsum, int lw,
                                              int lh)
{
    const int X = blockIdx.x * 256 + threadIdx.x, Y = blockIdx.y;
    if (X >= lw || Y >= lh) return;
    uint16_t s = 0;
    for (int i = 0; i < n; i++) {
        const DevGeom g = geoms[i * (L + 1) + j];
        const int ly = Y - g.y0;
        int lx = X - g.x0;
        if (lx < 0) lx += g.cw;
        if ((unsigned)ly < (unsigned)g.bh && (unsigned)lx < (unsigned)g.bw)
            s = (uint16_t)(s + w[g.off + (size_t)ly * g.pitch + lx]);
    }
    wsum[(size_t)Y * lw + X] = s;
}

// Owners present in each level-0 tile as bits (n <= 32).  Block per tile.
__global__ __launch_bounds__(256) void k_fin_owners(const uint8_t* __restrict__ label, int cols, int rows, int ttw,
                                                    int tth, int tw, uint32_t* __restrict__ bits)
{
    __shared__ uint32_t acc;
    const int tile = blockIdx.x, X0 = (tile % tw) * ttw, Y0 = (tile / tw) * tth;
    const int nx = min(ttw, cols - X0), ny = min(tth, rows - Y0);
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    uint32_t b = 0;
    for (int k = threadIdx.x; k < nx * ny; k += 256) {
        const uint8_t lab = ldg(label + (size_t)(Y0 + k / nx) * cols + X0 + k % nx);
        if (lab != kNoOwner) b |= 1u << lab;
    }
    if (b) atomicOr(&acc, b);
    __syncthreads();
    if (threadIdx.x == 0) bits[tile] = acc;
}

// A point of the fused map storage: |x|, |y| with kFlagOwned / kFlagValid in
// the sign bits, +inf for a point outside the source (NaN included).
__device__ __forceinline__ float2 fused_point_dev(float x, float y, int sw, int sh, uint32_t flags)
{
    const bool inside = x >= 0.f && y >= 0.f && x < (float)sw && y < (float)sh;
    uint32_t bx = inside ? __float_as_uint(x) & 0x7fffffffu : 0x7f800000u;
    uint32_t by = inside ? __float_as_uint(y) & 0x7fffffffu : 0x7f800000u;
    if (flags & kFlagOwned) bx |= 0x80000000u;
    if (flags & kFlagValid) by |= 0x80000000u;
    return make_float2(__uint_as_float(bx), __uint_as_float(by));
}

// Margined map storage of camera cam: entry (sy, sx) holds canvas
// (x0 - m + sx, y0 - m + sy) after WRAP / REFLECT_101, or +inf when that lies
// outside the box.  Grid (mp / 256, mrows).
__global__ __launch_bounds__(256) void k_pack_map(const float* __restrict__ xm, const float* __restrict__ ym,
                                                  const uint8_t* __restrict__ mask, const uint8_t* __restrict__ label,
                                                  int cols, int cam, int x0, int y0, int bw, int bh, int cw, int ch,
                                                  int m, int mp, int sw, int sh, float2* __restrict__ out)
{
    const int sx = blockIdx.x * 256 + threadIdx.x, sy = blockIdx.y;
    if (sx >= mp) return;
    float2 r = make_float2(__uint_as_float(0x7f800000u), __uint_as_float(0x7f800000u));
    if (sx < bw + 2 * m) {
        const int X = dwrap(x0 - m + sx, cw), Y = dreflect(y0 - m + sy, ch);
        const int ly = Y - y0;
        int lx = X - x0;
        if (lx < 0) lx += cw;
        if (ly >= 0 && ly < bh && lx < bw) {
            const size_t o = (size_t)Y * cols + X;
            uint32_t f = ldg(mask + o) ? kFlagValid : 0;
            if (ldg(label + o) == cam) f |= kFlagOwned;
            r = fused_point_dev(ldg(xm + o), ldg(ym + o), sw, sh, f);
        }
    }
    out[(size_t)sy * mp + sx] = r;
}

// Valid map points outside the mask (*bad = 1; zero on entry).
__global__ __launch_bounds__(256) void k_map_consistent(const float* __restrict__ xm, const float* __restrict__ ym,
                                                        const uint8_t* __restrict__ mask, size_t P, int sw, int sh,
                                                        int* __restrict__ bad)
{
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < P; p += (size_t)gridDim.x * 256) {
        const float x = ldg(xm + p), y = ldg(ym + p);
        if (x >= 0.f && y >= 0.f && x < (float)sw && y < (float)sh && !ldg(mask + p)) *bad = 1;
    }
}

template <class T>
struct DBuf {
    T* p = nullptr;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)) == hipSuccess; }
};

#define PD_TRY(expr)                                 \
    do {                                             \
        if ((expr) != hipSuccess) return PANO_ERR_HIP; \
    } while (0)

}  // namespace

DevTables::~DevTables()
{
    if (w) (void)hipFree(w);
    for (uint16_t* p : wsum)
        if (p) (void)hipFree(p);
    if (label) (void)hipFree(label);
    if (cover) (void)hipFree(cover);
}

int prepare_tables_device(int n, int rows, int cols, const uint8_t* masks, const uint8_t* unique, int max_levels,
                          bool with_cover, int shape, Prepared& out, DevTables& dev)
{
    if (n <= 0 || n > 32 || rows <= 0 || cols <= 0 || !masks || !unique) return PANO_ERR_ARG;
    if (shape < 0 || shape >= kDownShapes) return PANO_ERR_ARG;
    const int L = true_num_levels(cols, rows, max_levels, 2);
    if (L < 0) return PANO_ERR_ARG;
    out = Prepared();
    out.kind = kPixU8;
    out.fused0 = true;
    out.shape = shape;
    out.dtw = kDownTileW[shape];
    out.dth = kDownTileH[shape];
    out.n = n;
    out.rows = rows;
    out.cols = cols;
    out.levels = L;
    out.lw.resize(L + 1);
    out.lh.resize(L + 1);
    out.lw[0] = cols;
    out.lh[0] = rows;
    for (int j = 1; j <= L; j++) {
        out.lw[j] = (out.lw[j - 1] + 1) / 2;
        out.lh[j] = (out.lh[j - 1] + 1) / 2;
    }
    const size_t P = (size_t)rows * cols;

    // boxes from the masks' row / column projections
    DBuf<uint8_t> colany, rowany;
    if (!colany.alloc((size_t)n * cols) || !rowany.alloc((size_t)n * rows)) return PANO_ERR_NOMEM;
    PD_TRY(hipMemset(colany.p, 0, (size_t)n * cols));
    hipLaunchKernelGGL(k_mask_proj, dim3(rows, n), dim3(256), 0, nullptr, masks, rows, cols, P, colany.p, rowany.p);
    PD_TRY(hipGetLastError());
    std::vector<uint8_t> hcol((size_t)n * cols), hrow((size_t)n * rows);
    PD_TRY(hipMemcpy(hcol.data(), colany.p, hcol.size(), hipMemcpyDeviceToHost));
    PD_TRY(hipMemcpy(hrow.data(), rowany.p, hrow.size(), hipMemcpyDeviceToHost));
    out.geom.assign((size_t)n * (L + 1), Geom());
    for (int i = 0; i < n; i++)
        camera_boxes(L, out.lw.data(), out.lh.data(),
                     std::vector<uint8_t>(hcol.begin() + (size_t)i * cols, hcol.begin() + (size_t)(i + 1) * cols),
                     std::vector<uint8_t>(hrow.begin() + (size_t)i * rows, hrow.begin() + (size_t)(i + 1) * rows),
                     out.geom.data() + (size_t)i * (L + 1));

    // owners, cover bits, coverage
    DBuf<int> flag;
    if (!flag.alloc(1)) return PANO_ERR_NOMEM;
    PD_TRY(hipMemset(flag.p, 0, sizeof(int)));
    if (hipMalloc((void**)&dev.label, P) != hipSuccess) return PANO_ERR_NOMEM;
    if (with_cover && n <= kGainMaxCams && hipMalloc((void**)&dev.cover, P) != hipSuccess) return PANO_ERR_NOMEM;
    const int lc_blocks = (int)std::min<size_t>((P + 255) / 256, 8192);
    hipLaunchKernelGGL(k_label_cover, dim3(lc_blocks), dim3(256), 0, nullptr, masks, unique, n, P, dev.label,
                       dev.cover, flag.p);
    PD_TRY(hipGetLastError());
    int uncovered = 0;
    PD_TRY(hipMemcpy(&uncovered, flag.p, sizeof(int), hipMemcpyDeviceToHost));
    out.full = !uncovered;

    // plane offsets, tile grids
    const int dtw = out.dtw, dth = out.dth, ttw = 2 * kQTW, tth = 2 * kQTH;
    std::vector<DevGeom> dg((size_t)n * (L + 1));
    size_t elems = 0;
    int nflags = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= L; j++) {
            const Geom& g = out.geom[(size_t)i * (L + 1) + j];
            DevGeom& d = dg[(size_t)i * (L + 1) + j];
            d.x0 = g.x0;
            d.y0 = g.y0;
            d.bw = g.bw;
            d.bh = g.bh;
            d.cw = g.cw;
            d.ch = g.ch;
            d.pitch = g.pitch;
            d.off = 0;
            d.dflag = d.dcols0 = d.dcols = 0;
            if (j == 0 || g.empty()) continue;
            d.off = elems;
            elems += g.pixels();
            const int seg0 = std::min(g.x0 + g.bw, g.cw) - g.x0, seg1 = std::max(0, g.x0 + g.bw - g.cw);
            d.dcols0 = (seg0 + dtw - 1) / dtw;
            d.dcols = d.dcols0 + (seg1 + dtw - 1) / dtw;
            d.dflag = nflags;
            nflags += d.dcols * ((g.bh + dth - 1) / dth);
        }
    dev.woff.assign((size_t)n * (L + 1), 0);
    for (size_t k = 0; k < dg.size(); k++) dev.woff[k] = dg[k].off;
    std::vector<int> tw(L + 1), tiles(L + 1), hit_base(L + 2, 0);
    for (int j = 1; j <= L; j++) {
        tw[j] = (out.lw[j] + ttw - 1) / ttw;
        tiles[j] = tw[j] * ((out.lh[j] + tth - 1) / tth);
        hit_base[j + 1] = hit_base[j] + n * tiles[j];
    }
    DBuf<DevGeom> dgeom;
    DBuf<uint8_t> valid, hit, dflag;
    if (hipMalloc((void**)&dev.w, sizeof(uint16_t) * (elems ? elems : 1)) != hipSuccess || !dgeom.alloc(dg.size()) ||
        !valid.alloc(elems) || !hit.alloc(hit_base[L + 1]) || !dflag.alloc(nflags))
        return PANO_ERR_NOMEM;
    PD_TRY(hipMemcpy(dgeom.p, dg.data(), sizeof(DevGeom) * dg.size(), hipMemcpyHostToDevice));
    PD_TRY(hipMemset(hit.p, 0, std::max(hit_base[L + 1], 1)));
    PD_TRY(hipMemset(dflag.p, 0, std::max(nflags, 1)));

    // pyramids, level by level
    for (int j = 1; j <= L; j++) {
        int mp = 0, mh = 0;
        for (int i = 0; i < n; i++) {
            const Geom& g = out.geom[(size_t)i * (L + 1) + j];
            mp = std::max(mp, g.pitch);
            mh = std::max(mh, g.bh);
        }
        if (mp == 0 || mh == 0) continue;
        hipLaunchKernelGGL(k_level, dim3((mp + 63) / 64, (mh + 3) / 4, n), dim3(64, 4), 0, nullptr, j, L, dgeom.p,
                           unique, masks, cols, P, dev.w, valid.p, hit.p + hit_base[j], tw[j], tiles[j], ttw, tth,
                           dtw, dth, dflag.p);
        PD_TRY(hipGetLastError());
    }
    dev.wsum.assign(L + 1, nullptr);
    if (!out.full)
        for (int j = 1; j <= L; j++) {
            if (hipMalloc((void**)&dev.wsum[j], (size_t)out.lw[j] * out.lh[j] * 2) != hipSuccess)
                return PANO_ERR_NOMEM;
            hipLaunchKernelGGL(k_wsum, dim3((out.lw[j] + 255) / 256, out.lh[j]), dim3(256), 0, nullptr, j, L, n,
                               dgeom.p, dev.w, dev.wsum[j], out.lw[j], out.lh[j]);
            PD_TRY(hipGetLastError());
        }
    const int ftw = (cols + ttw - 1) / ttw, fth = (rows + tth - 1) / tth;
    DBuf<uint32_t> owners;
    if (!owners.alloc((size_t)ftw * fth)) return PANO_ERR_NOMEM;
    hipLaunchKernelGGL(k_fin_owners, dim3(ftw * fth), dim3(256), 0, nullptr, dev.label, cols, rows, ttw, tth, ftw,
                       owners.p);
    PD_TRY(hipGetLastError());

    std::vector<uint8_t> hhit(hit_base[L + 1]), hflag(nflags);
    std::vector<uint32_t> hown((size_t)ftw * fth);
    PD_TRY(hipMemcpy(hhit.data(), hit.p, hhit.size(), hipMemcpyDeviceToHost));
    PD_TRY(hipMemcpy(hflag.data(), dflag.p, hflag.size(), hipMemcpyDeviceToHost));
    PD_TRY(hipMemcpy(hown.data(), owners.p, hown.size() * 4, hipMemcpyDeviceToHost));

    // the work lists from the flags
    auto tile_has_valid = [&](int i, int j, int X, int Y, int, int) {
        const Geom& g = out.geom[(size_t)i * (L + 1) + j];
        const DevGeom& d = dg[(size_t)i * (L + 1) + j];
        const int col = X >= g.x0 ? (X - g.x0) / dtw : d.dcols0 + X / dtw;
        return hflag[d.dflag + ((Y - g.y0) / dth) * d.dcols + col] != 0;
    };
    auto cam_hits = [&](int j, int i, std::vector<uint8_t>& h) {
        const uint8_t* src = hhit.data() + hit_base[j] + (size_t)i * tiles[j];
        for (int t = 0; t < tiles[j]; t++)
            if (src[t]) h[t] = 1;
    };
    auto tile_owners = [&](const TileJob& t, std::vector<int>& o) {
        const uint32_t b = hown[(size_t)(t.Y / tth) * ftw + t.X / ttw];
        for (int i = 0; i < n; i++)
            if (b & (1u << i)) o.push_back(i);
    };
    work_lists(out, tile_has_valid, cam_hits, tile_owners);
    PD_TRY(hipDeviceSynchronize());
    return PANO_OK;
}

hipError_t launch_pack_map(const float* xm, const float* ym, const uint8_t* mask, const uint8_t* label, int cols,
                           int cam, const Geom& g, int margin, int mp, int sw, int sh, float2* out, hipStream_t s)
{
    const int mrows = g.bh > 0 ? g.bh + 2 * margin : 0;
    if (mp <= 0 || mrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_map, dim3((mp + 255) / 256, mrows), dim3(256), 0, s, xm, ym, mask, label, cols, cam,
                       g.x0, g.y0, g.bw, g.bh, g.cw, g.ch, margin, mp, sw, sh, out);
    return hipGetLastError();
}

int maps_consistent_device(const float* xm, const float* ym, const uint8_t* mask, size_t P, int sw, int sh,
                           bool& consistent)
{
    DBuf<int> bad;
    if (!bad.alloc(1)) return PANO_ERR_NOMEM;
    PD_TRY(hipMemset(bad.p, 0, sizeof(int)));
    const int blocks = (int)std::min<size_t>((P + 255) / 256, 8192);
    hipLaunchKernelGGL(k_map_consistent, dim3(blocks), dim3(256), 0, nullptr, xm, ym, mask, P, sw, sh, bad.p);
    PD_TRY(hipGetLastError());
    int h = 0;
    PD_TRY(hipMemcpy(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost));
    consistent = !h;
    return PANO_OK;
}

}  // namespace pano
