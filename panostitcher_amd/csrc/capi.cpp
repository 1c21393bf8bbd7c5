// capi.cpp -- the extern "C" boundary (include/panohip.h): object lifetimes,
// device residency of prepared rigs, and the per-frame launch sequences.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <vector>

#include "../../include/panohip.h"
#include "pano_host.hpp"
#include "pano_launch.hpp"

namespace pano {
int generate_map(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h,
                 float* xmap, float* ymap, uint8_t* mask);
int generate_map_device(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h, float* xmap,
                        float* ymap, size_t map_step, uint8_t* mask, size_t mask_step, hipStream_t s);
}

using namespace pano;

#define HIP_TRY(expr)                                        \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return PANO_ERR_HIP;           \
    } while (0)

namespace {

// One device allocation carved into 256-byte aligned pieces.
struct Arena {
    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    uint8_t* base = nullptr;
    size_t size = 0, used = 0;
    ~Arena() { release(); }
    void release()
    {
        if (base) (void)hipFree(base);
        base = nullptr;
        size = used = 0;
    }
    static size_t align(size_t n) { return (n + 255) & ~(size_t)255; }
    int reserve(size_t bytes)
    {
        release();
        if (bytes == 0) bytes = 256;
        if (hipMalloc((void**)&base, bytes) != hipSuccess) {
            base = nullptr;
            return PANO_ERR_NOMEM;
        }
        size = bytes;
        used = 0;
        return PANO_OK;
    }
    void* take(size_t bytes)
    {
        size_t a = align(bytes);
        if (used + a > size) return nullptr;
        void* p = base + used;
        used += a;
        return p;
    }
};

// Sizes the arena in a first pass, hands out pointers in a second.
struct Planner {
    size_t total = 0;
    void add(size_t bytes) { total += Arena::align(bytes); }
};

// A small ring of pinned host slots for per-frame uploads (gain LUTs); a slot
// is reused only after the copy that read it has finished.
struct PinnedRing {
    PinnedRing() = default;
    PinnedRing(const PinnedRing&) = delete;
    PinnedRing& operator=(const PinnedRing&) = delete;
    static constexpr int kSlots = 4;
    uint8_t* host = nullptr;
    size_t slot_bytes = 0;
    hipEvent_t ev[kSlots] = {};
    bool used[kSlots] = {};
    int next = 0;
    ~PinnedRing() { release(); }
    void release()
    {
        if (host) (void)hipHostFree(host);
        host = nullptr;
        for (int i = 0; i < kSlots; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            ev[i] = nullptr;
            used[i] = false;
        }
    }
    int init(size_t bytes)
    {
        release();
        slot_bytes = Arena::align(bytes ? bytes : 1);
        if (hipHostMalloc((void**)&host, slot_bytes * kSlots, hipHostMallocDefault) != hipSuccess) {
            host = nullptr;
            return PANO_ERR_NOMEM;
        }
        for (int i = 0; i < kSlots; i++)
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return PANO_ERR_HIP;
        return PANO_OK;
    }
    int upload(const void* src, size_t bytes, void* dst, hipStream_t s)
    {
        int k = next;
        next = (next + 1) % kSlots;
        if (used[k]) HIP_TRY(hipEventSynchronize(ev[k]));
        uint8_t* h = host + slot_bytes * k;
        memcpy(h, src, bytes);
        HIP_TRY(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
        HIP_TRY(hipEventRecord(ev[k], s));
        used[k] = true;
        return PANO_OK;
    }
};

int host_wrap(int p, int n)
{
    p %= n;
    return p < 0 ? p + n : p;
}

int host_reflect(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    do {
        p = p < 0 ? -p : 2 * n - 2 - p;
    } while ((unsigned)p >= (unsigned)n);
    return p;
}

// Margined map storage of a level-0 box (fused path).
int map_pitch(const Geom& g) { return g.bw > 0 ? (g.bw + 2 * kMapMargin + 63) / 64 * 64 : 0; }
size_t map_cells(const Geom& g) { return g.bh > 0 ? (size_t)map_pitch(g) * (g.bh + 2 * kMapMargin) : 0; }

PlaneRef plane_ref(const Plane& p, int cam)
{
    PlaneRef f{};
    f.g = p.g;
    f.w = p.w;
    f.x0 = p.x0;
    f.y0 = p.y0;
    f.bw = p.bw;
    f.bh = p.bh;
    f.cw = p.cw;
    f.pitch = p.pitch;
    f.cam = cam;
    return f;
}


// A map point in fused storage: |x|, |y| with the flags in the sign bits, or
// +inf for a point the remap treats as outside the source (including NaN).
float2 fused_point(float x, float y, int sw, int sh, uint8_t flags)
{
    const bool inside = x >= 0.f && y >= 0.f && x < (float)sw && y < (float)sh;
    uint32_t bx, by;
    memcpy(&bx, &x, 4);
    memcpy(&by, &y, 4);
    bx = inside ? bx & 0x7fffffffu : 0x7f800000u;
    by = inside ? by & 0x7fffffffu : 0x7f800000u;
    if (flags & kFlagOwned) bx |= 0x80000000u;
    if (flags & kFlagValid) by |= 0x80000000u;
    float2 r;
    memcpy(&r.x, &bx, 4);
    memcpy(&r.y, &by, 4);
    return r;
}

template <class T>
T* upload_vec(Arena& a, const std::vector<T>& v, int& err)
{
    if (v.empty()) return nullptr;
    T* d = (T*)a.take(sizeof(T) * v.size());
    if (!d) { err = PANO_ERR_NOMEM; return nullptr; }
    if (hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
        err = PANO_ERR_HIP;
    return d;
}

// One plane of an output: rows of at least `row` bytes, step a multiple of
// `align` (as is the pointer), below 2^24, the plane below 4 GB.
bool plane_ok(const uint8_t* p, size_t step, size_t row, int rows, size_t align)
{
    return p && step >= row && step % align == 0 && (uintptr_t)p % align == 0 && step < (1u << 24) &&
           step * (uint64_t)rows <= 0xffffffffu;
}

// getNonIntersectingMasks on the device for a prepare (argument errors first,
// as prepare_rig reports them).
int device_seams(int n, int rows, int cols, const uint8_t* const* masks, const size_t* steps,
                 std::vector<uint8_t>& out)
{
    if (n <= 0 || n > 255 || rows <= 0 || cols <= 0 || !masks || !steps) return PANO_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!masks[i] || steps[i] < (size_t)cols) return PANO_ERR_ARG;
    out.resize((size_t)rows * cols * n);
    HIP_TRY(device_nonintersecting_masks(n, rows, cols, masks, steps, out.data()));
    return PANO_OK;
}

// One launch of the fused 8-bit sequence: a stage over a device segment table.
struct Launch {
    int stage, level, nseg, blocks;
    const Seg* segs;
    int sched = 0;   // block -> job order (block_job)
    int shape = 0;   // down tile shape of the renders' work lists
};

// Runs a fused sequence on a stream; ev (optional): events recorded before the
// first launch and after the remap, down, blend and final stages.  st: the
// frames the remap reads; luts: the remap applies the renders' active LUTs.
int run_sequence(const std::vector<Launch>& seq, const SrcTable& st, const OutTable& ot,
                 hipStream_t s, std::vector<hipEvent_t>* ev, bool luts = false, bool ev0_recorded = false)
{
    if (ev && !ev0_recorded) HIP_TRY(hipEventRecord((*ev)[0], s));
    int prev = kStageRemapDown0;
    for (const Launch& l : seq) {
        if (ev && l.stage != prev) {
            // stage boundaries: remap -> down -> blend -> final (down may be empty)
            for (int k = prev + 1; k <= l.stage; k++) HIP_TRY(hipEventRecord((*ev)[k], s));
            prev = l.stage;
        }
        HIP_TRY(launch_seg(l.stage, l.segs, l.nseg, l.blocks, l.level, st, ot, l.sched, l.shape, luts, s));
    }
    if (ev)
        for (int k = prev + 1; k <= 4; k++) HIP_TRY(hipEventRecord((*ev)[k], s));
    return PANO_OK;
}

// Device residency of one prepared rig (blender or render).
struct DeviceRig {
    Arena arena;
    Prepared prep;
    int kind = kPixU8;                      // pixel kind of the internal planes
    std::vector<Plane> planes;              // n * (L + 1), level 0 internal storage
    std::vector<Plane> planes_user;         // the same with level 0 = caller images
    Plane* d_planes = nullptr;
    Plane* d_planes_user = nullptr;
    CanvasLevel* d_canvas = nullptr;
    RenderDesc* d_rd = nullptr;
    std::vector<DownJob*> d_down;
    std::vector<int> n_down;
    std::vector<TileJob*> d_blend;
    std::vector<int> n_blend;
    int* d_camlist = nullptr;
    TileJob* d_fin = nullptr;
    RemapJob* d_remap = nullptr;
    RemapCam* d_cams = nullptr;
    uint8_t* d_label = nullptr;
    uint8_t* d_masks = nullptr;             // full-canvas masks (caller-image level 0)
    uint8_t* d_luts = nullptr;
    int* d_fin_owners = nullptr;
    PlaneRef* d_fin_refs = nullptr;
    PlaneRef* d_blend_refs = nullptr;
    PlaneRef* d_blend_refs1 = nullptr;
    uint32_t* d_g0comp = nullptr;           // owner's level-0 value per canvas pixel (8-bit render)
    Seg* d_segs = nullptr;                  // render() alone: one segment per fused launch
    std::vector<Launch> seq;                // render() alone: the fused launch sequence
    uint8_t* d_cover = nullptr;             // gain estimation: masks covering each canvas pixel
    unsigned long long* d_gstats = nullptr;
    double* d_gains = nullptr;
    size_t model_bytes[8] = {};
    // prepare on the device: tables and the device maps / masks the build reads
    // instead of the host ones (n planes, pitch prep.cols)
    const DevTables* dev = nullptr;
    const float* dev_map = nullptr;         // camera i: x plane at 2 i P, y plane at (2 i + 1) P
    const uint8_t* dev_masks = nullptr;

    int levels() const { return prep.levels; }

    // level0_internal: allocate level-0 planes (fused path / begin-tile-end);
    // user_level0: describe level 0 as caller images (blend()).
    // maps: host x/y planes for the fused path (nullptr otherwise); masks: host.
    int build(bool level0_internal, bool keep_masks, const float* const* xmaps,
              const float* const* ymaps, const uint8_t* const* masks, const size_t* mstep,
              const int* src_w, const int* src_h)
    {
        const int n = prep.n, L = prep.levels;
        const size_t P = (size_t)prep.rows * prep.cols;
        const size_t px = kind == kPixU8 ? 4 : 8;
        Planner pl;
        const bool fused = prep.fused0;
        if (fused) level0_internal = false;
        for (int i = 0; i < n; i++)
            for (int j = 0; j <= L; j++) {
                const Geom& g = prep.geom[(size_t)i * (L + 1) + j];
                if (j > 0 || level0_internal) pl.add(g.pixels() * px);
                if (j > 0) pl.add(g.pixels() * 2);
            }
        if (fused) pl.add(P * 4);
        const bool have_maps = xmaps || dev_map;
        const bool gain = fused && have_maps && n <= kGainMaxCams;
        if (gain) {
            pl.add(P);
            pl.add(sizeof(unsigned long long) * n * n * 5 * kGainSlots);
            pl.add(sizeof(double) * (3 * n + 1));
        }
        pl.add(sizeof(int) * prep.fin_owners.size());
        if (fused) pl.add(sizeof(PlaneRef) * prep.fin_owners.size());
        if (fused) {
            pl.add(sizeof(PlaneRef) * prep.camlist.size());
            pl.add(sizeof(PlaneRef) * prep.camlist.size());
        }
        for (int j = 1; j <= L; j++) {
            pl.add((size_t)prep.lw[j] * prep.lh[j] * (fused ? 8 : 16));
            if (!prep.full) pl.add((size_t)prep.lw[j] * prep.lh[j] * 2);
        }
        pl.add(sizeof(Plane) * n * (L + 1));
        pl.add(sizeof(Plane) * n * (L + 1));
        pl.add(sizeof(CanvasLevel) * (L + 1));
        pl.add(sizeof(RenderDesc));
        for (int j = 0; j < L; j++) pl.add(sizeof(DownJob) * prep.down[j].size());
        for (int j = 1; j <= L; j++) pl.add(sizeof(TileJob) * prep.blend[j].size());
        pl.add(sizeof(int) * prep.camlist.size());
        pl.add(sizeof(TileJob) * prep.fin.size());
        pl.add(P);  // label
        if (keep_masks) pl.add(P * n);
        if (have_maps) {
            pl.add(sizeof(RemapJob) * prep.remap.size());
            pl.add(sizeof(RemapCam) * n);
            for (int i = 0; i < n; i++) {
                const Geom& g = prep.geom[(size_t)i * (L + 1)];
                size_t cells = fused ? map_cells(g) : g.pixels();
                pl.add(cells * 8);
                if (!fused) pl.add(cells);
            }
            pl.add((size_t)n * 768 + 4);
        }
        pl.add(sizeof(Seg) * (2 * L + 2));
        pl.add(64 << 10);  // slack for per-piece alignment
        int err = arena.reserve(pl.total);
        if (err) return err;

        planes.assign((size_t)n * (L + 1), Plane());
        for (int i = 0; i < n; i++)
            for (int j = 0; j <= L; j++) {
                const Geom& g = prep.geom[(size_t)i * (L + 1) + j];
                Plane& p = planes[(size_t)i * (L + 1) + j];
                p.x0 = g.x0; p.y0 = g.y0; p.bw = g.bw; p.bh = g.bh;
                p.cw = g.cw; p.ch = g.ch; p.pitch = g.pitch; p.src = -1;
                if (j > 0 || level0_internal) {
                    p.g = arena.take(g.pixels() * px);
                    if (!p.g && g.pixels()) return PANO_ERR_NOMEM;
                    if (p.g) HIP_TRY(hipMemset(p.g, 0, g.pixels() * px));
                }
                if (j > 0) {
                    p.w = (uint16_t*)arena.take(g.pixels() * 2);
                    if (dev) {
                        if (g.pixels())
                            HIP_TRY(hipMemcpy(p.w, dev->w + dev->woff[(size_t)i * (L + 1) + j], g.pixels() * 2,
                                              hipMemcpyDeviceToDevice));
                    } else {
                        const std::vector<uint16_t>& w = prep.weight[(size_t)i * (L + 1) + j];
                        if (!w.empty()) HIP_TRY(hipMemcpy(p.w, w.data(), w.size() * 2, hipMemcpyHostToDevice));
                    }
                }
            }
        std::vector<CanvasLevel> canvas(L + 1);
        for (int j = 1; j <= L; j++) {
            CanvasLevel& c = canvas[j];
            c.cw = prep.lw[j];
            c.ch = prep.lh[j];
            c.pitch = prep.lw[j];
            if (fused) c.c16 = (uint2*)arena.take((size_t)c.cw * c.ch * 8);
            else c.c = (int4*)arena.take((size_t)c.cw * c.ch * 16);
            if (!prep.full) {
                c.wsum = (uint16_t*)arena.take((size_t)c.cw * c.ch * 2);
                if (dev)
                    HIP_TRY(hipMemcpy(c.wsum, dev->wsum[j], (size_t)c.cw * c.ch * 2, hipMemcpyDeviceToDevice));
                else
                    HIP_TRY(hipMemcpy(c.wsum, prep.wsum[j].data(), prep.wsum[j].size() * 2,
                                      hipMemcpyHostToDevice));
            }
        }
        canvas[0].cw = prep.cols;
        canvas[0].ch = prep.rows;
        d_label = (uint8_t*)arena.take(P);
        HIP_TRY(hipMemcpy(d_label, dev ? dev->label : prep.label.data(), P,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
        if (keep_masks) {
            d_masks = (uint8_t*)arena.take(P * n);
            for (int i = 0; i < n; i++)
                HIP_TRY(hipMemcpy2D(d_masks + P * i, prep.cols, masks[i], mstep[i], prep.cols,
                                    prep.rows, hipMemcpyHostToDevice));
        }
        planes_user = planes;
        for (int i = 0; i < n; i++) {
            Plane& p = planes_user[(size_t)i * (L + 1)];
            p.g = nullptr;
            p.src = i;
            if (d_masks) {
                p.mask = d_masks + P * i;
                p.mask_pitch = prep.cols;
            }
            if (d_masks) {
                planes[(size_t)i * (L + 1)].mask = d_masks + P * i;
                planes[(size_t)i * (L + 1)].mask_pitch = prep.cols;
            }
        }
        d_planes = upload_vec(arena, planes, err);
        d_planes_user = upload_vec(arena, planes_user, err);
        d_canvas = upload_vec(arena, canvas, err);
        if (fused && L > 0) {
            for (DownJob& d : prep.down[0]) {
                const Geom& g = prep.geom[(size_t)d.src * (L + 1)];
                // staging region: level-0 columns 2X - 2 .., rows 2Y - 2 ..
                int mx = host_wrap(2 * d.X - 2 - (g.x0 - kMapMargin), g.cw);
                int my = 2 * d.Y - 2 - (g.y0 - kMapMargin);
                if (my < 0 || my + 2 * d.ny + 3 > g.bh + 2 * kMapMargin ||
                    mx + 2 * d.nx + 3 > g.bw + 2 * kMapMargin)
                    return PANO_ERR_ARG;  // a staging region outside the margined box
                d.mx = mx;
                d.my = my;
            }
        }
        d_down.assign(L, nullptr);
        n_down.assign(L, 0);
        for (int j = 0; j < L; j++) {
            d_down[j] = upload_vec(arena, prep.down[j], err);
            n_down[j] = (int)prep.down[j].size();
        }
        d_blend.assign(L + 1, nullptr);
        n_blend.assign(L + 1, 0);
        for (int j = 1; j <= L; j++) {
            d_blend[j] = upload_vec(arena, prep.blend[j], err);
            n_blend[j] = (int)prep.blend[j].size();
        }
        d_camlist = upload_vec(arena, prep.camlist, err);
        d_fin = upload_vec(arena, prep.fin, err);
        d_fin_owners = upload_vec(arena, prep.fin_owners, err);
        if (fused && L > 0) {
            std::vector<PlaneRef> refs(prep.fin_owners.size());
            for (size_t k = 0; k < refs.size(); k++) {
                const int o = prep.fin_owners[k];
                refs[k] = plane_ref(planes[(size_t)o * (L + 1) + 1], o);
            }
            d_fin_refs = upload_vec(arena, refs, err);
            // blend tiles: camlist entries are plane indices cam (L + 1) + j
            std::vector<PlaneRef> r0(prep.camlist.size()), r1(prep.camlist.size());
            for (size_t k = 0; k < r0.size(); k++) {
                const int p = prep.camlist[k], cam = p / (L + 1), j = p % (L + 1);
                r0[k] = plane_ref(planes[p], cam);
                r1[k] = j < L ? plane_ref(planes[p + 1], cam) : r0[k];
            }
            d_blend_refs = upload_vec(arena, r0, err);
            d_blend_refs1 = upload_vec(arena, r1, err);
        }
        if (fused) {
            d_g0comp = (uint32_t*)arena.take(P * 4);
            if (!d_g0comp) return PANO_ERR_NOMEM;
            HIP_TRY(hipMemset(d_g0comp, 0, P * 4));
        }
        if (err) return err;
        if (have_maps) {
            d_remap = upload_vec(arena, prep.remap, err);
            std::vector<RemapCam> cams(n);
            for (int i = 0; i < n; i++) {
                const Geom& g = prep.geom[(size_t)i * (L + 1)];
                RemapCam& c = cams[i];
                c.sw = src_w[i];
                c.sh = src_h[i];
                c.src = i;
                c.lut = i * 768;
                c.plane = i * (L + 1);
                const int m = fused ? kMapMargin : 0;
                const int mp = fused ? map_pitch(g) : g.pitch;
                const int mrows = g.bh > 0 ? g.bh + 2 * m : 0;
                const size_t cells = (size_t)mp * mrows;
                float2* dm = (float2*)arena.take(cells * 8);
                uint8_t* dk = fused ? nullptr : (uint8_t*)arena.take(cells);
                if (!dm || (!fused && !dk && cells)) return PANO_ERR_NOMEM;
                c.map = dm;
                c.mask = dk;
                c.mpitch = mp;
                if (dev_map) {
                    // fused storage packed on the device from the device maps
                    if (!fused) return PANO_ERR_ARG;
                    HIP_TRY(launch_pack_map(dev_map + 2 * P * i, dev_map + (2 * (size_t)i + 1) * P,
                                            dev_masks + P * i, d_label, prep.cols, i, g, m, mp, src_w[i], src_h[i],
                                            dm, nullptr));
                    continue;
                }
                std::vector<float2> hm(cells, fused ? fused_point(-1.f, -1.f, 1, 1, 0) : make_float2(-1.f, -1.f));
                std::vector<uint8_t> hk(fused ? 0 : cells, 0);
                // entry (sy, sx) holds canvas (x0 - m + sx, y0 - m + sy) after border
                // resolution, or nothing when that lies outside the box
                for (int sy = 0; sy < mrows; sy++)
                    for (int sx = 0; sx < g.bw + 2 * m && g.bw > 0; sx++) {
                        int X = host_wrap(g.x0 - m + sx, g.cw), Y = host_reflect(g.y0 - m + sy, g.ch);
                        int ly = Y - g.y0, lx = X - g.x0;
                        if (lx < 0) lx += g.cw;
                        if (ly < 0 || ly >= g.bh || lx >= g.bw) continue;
                        size_t o = (size_t)Y * prep.cols + X;
                        uint8_t f = masks[i][(size_t)Y * mstep[i] + X] ? kFlagValid : 0;
                        if (prep.label[o] == i) f |= kFlagOwned;
                        if (fused) {
                            hm[(size_t)sy * mp + sx] =
                                fused_point(xmaps[i][o], ymaps[i][o], src_w[i], src_h[i], f);
                        } else {
                            hm[(size_t)sy * mp + sx] = make_float2(xmaps[i][o], ymaps[i][o]);
                            hk[(size_t)sy * mp + sx] = f;
                        }
                    }
                if (cells) {
                    HIP_TRY(hipMemcpy(dm, hm.data(), cells * 8, hipMemcpyHostToDevice));
                    if (dk) HIP_TRY(hipMemcpy(dk, hk.data(), cells, hipMemcpyHostToDevice));
                }
                c.map = dm;
                c.mask = dk;
                c.mpitch = mp;
            }
            d_cams = upload_vec(arena, cams, err);
            // n LUTs, then one byte: whether this frame applies them
            d_luts = (uint8_t*)arena.take((size_t)n * 768 + 4);
            if (!d_luts) return PANO_ERR_NOMEM;
            HIP_TRY(hipMemset(d_luts, 0, (size_t)n * 768 + 4));
            if (err) return err;
        }
        if (gain) {
            if (dev) {
                d_cover = (uint8_t*)arena.take(P);
                if (!d_cover || !dev->cover) return PANO_ERR_NOMEM;
                HIP_TRY(hipMemcpy(d_cover, dev->cover, P, hipMemcpyDeviceToDevice));
            } else {
                std::vector<uint8_t> cover(P, 0);
                for (int i = 0; i < n; i++)
                    for (int y = 0; y < prep.rows; y++) {
                        const uint8_t* m = masks[i] + (size_t)y * mstep[i];
                        uint8_t* c = cover.data() + (size_t)y * prep.cols;
                        for (int x = 0; x < prep.cols; x++)
                            if (m[x]) c[x] |= (uint8_t)(1u << i);
                    }
                d_cover = upload_vec(arena, cover, err);
            }
            d_gstats = (unsigned long long*)arena.take(sizeof(unsigned long long) * n * n * 5 * kGainSlots);
            d_gains = (double*)arena.take(sizeof(double) * (3 * n + 1));
            if (!d_gstats || !d_gains) return PANO_ERR_NOMEM;
            HIP_TRY(hipMemset(d_gstats, 0, sizeof(unsigned long long) * n * n * 5 * kGainSlots));
            HIP_TRY(hipMemset(d_gains, 0, sizeof(double) * (3 * n + 1)));
            if (err) return err;
        }
        RenderDesc rd{};
        rd.label = d_label;
        rd.cover = d_cover;
        rd.gstats = d_gstats;
        rd.gains = d_gains;
        rd.planes = d_planes;
        rd.canvas = d_canvas;
        rd.camlist = d_camlist;
        rd.owners = d_fin_owners;
        rd.fin_refs = d_fin_refs;
        rd.blend_refs = d_blend_refs;
        rd.blend_refs1 = d_blend_refs1;
        rd.cams = d_cams;
        rd.luts = d_luts;
        rd.g0comp = d_g0comp;
        rd.label_pitch = prep.cols;
        rd.g0pitch = prep.cols;
        rd.full = prep.full;
        rd.levels = L;
        rd.plane0 = 0;
        rd.canvas0 = 0;
        rd.out = 0;
        rd.ncams = n;
        d_rd = upload_vec(arena, std::vector<RenderDesc>{rd}, err);
        if (err) return err;
        if (fused) {
            // one segment per launch for render() on its own
            std::vector<Seg> segs;
            auto add = [&](const void* jobs, int count) {
                Seg g{};
                g.rd = d_rd;
                g.jobs = jobs;
                g.begin = 0;
                g.count = count;
                segs.push_back(g);
            };
            std::vector<Launch> ls;
            auto launch = [&](int stage, int level, int count) {
                ls.push_back(Launch{stage, level, 1, count, nullptr});
            };
            add(d_down[0], n_down[0]);
            launch(kStageRemapDown0, 0, n_down[0]);
            for (int j = 1; j < L; j++) {
                add(d_down[j], n_down[j]);
                launch(kStageDown, j, n_down[j]);
            }
            for (int j = L; j >= 1; j--) {
                add(d_blend[j], n_blend[j]);
                launch(kStageBlend, j, n_blend[j]);
            }
            add(d_fin, (int)prep.fin.size());
            launch(kStageFinal, 0, (int)prep.fin.size());
            d_segs = upload_vec(arena, segs, err);
            if (err) return err;
            for (size_t k = 0; k < ls.size(); k++) {
                ls[k].segs = d_segs + k;
                ls[k].shape = prep.shape;
            }
            seq = ls;
        }
        compute_model_bytes(have_maps, src_w, src_h);
        return err;
    }

    // Algorithmic bytes per stage: every input read once, every output written
    // once (staging halos and cache re-reads are not counted).  Stage 0 is the
    // level-0 work (remap, and for the 8-bit path the fused level 0 -> 1 down),
    // 1 the remaining down levels, 2 the blend of levels >= 1, 3 the level-0 blend.
    void compute_model_bytes(bool fused_maps, const int* src_w, const int* src_h)
    {
        const int n = prep.n, L = prep.levels;
        const size_t px = kind == kPixU8 ? 4 : 8;
        const size_t P = (size_t)prep.rows * prep.cols;
        const size_t cb = prep.fused0 ? 8 : 16;   // bytes per C_j pixel
        size_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        if (fused_maps)
            for (int i = 0; i < n; i++) s0 += (size_t)src_w[i] * src_h[i] * 4;
        for (int j = 0; j < L; j++)
            for (const DownJob& d : prep.down[j]) {
                size_t outp = (size_t)d.nx * d.ny, inp = 4 * outp;
                if (j == 0 && prep.fused0) s0 += inp * 8 + outp * px;
                else if (j == 0 && fused_maps) s0 += inp * (8 + 1 + px) + inp * px + outp * px;
                else s1 += inp * px + outp * px;
            }
        if (prep.fused0) s0 += P * 4;  // owners' level-0 values
        for (int j = 1; j <= L; j++) {
            for (const TileJob& t : prep.blend[j]) {
                size_t tp = (size_t)t.nx * t.ny;
                s2 += tp * cb + (prep.full ? 0 : tp * 2);                 // C_j, sum W
                s2 += (size_t)t.cnt * tp * (px + 2);                       // G_j, W_j
                if (j < L) s2 += (size_t)t.cnt * tp * px / 4 + tp * cb / 4;  // G_j+1, C_j+1
            }
        }
        // level 0: owner label (a byte of the owners' samples on the 8-bit path),
        // the owner's G_0, G_1 and C_1, the output
        s3 = P * ((prep.fused0 ? 0 : 1) + px + 4) + (L > 0 ? P * px / 4 + P * cb / 4 : 0);
        model_bytes[0] = s0;
        model_bytes[1] = s1;
        model_bytes[2] = s2;
        model_bytes[3] = s3;
        model_bytes[4] = s0 + s1 + s2 + s3;
    }

    // ev: optional events 2..4 recorded after the down, blend and final stages.
    int run_pyramid(int user_kind, const SrcTable& st, const OutTable& ot, hipStream_t s,
                    std::vector<hipEvent_t>* ev = nullptr)
    {
        const int L = prep.levels;
        const Plane* pl = user_kind >= 0 ? d_planes_user : d_planes;
        for (int j = 0; j < L; j++)
            HIP_TRY(launch_down(kind, j == 0 ? user_kind : -1, d_down[j], n_down[j], pl, st, s));
        if (ev) HIP_TRY(hipEventRecord((*ev)[2], s));
        for (int j = L; j >= 1; j--)
            HIP_TRY(launch_blend(kind, d_blend[j], n_blend[j], d_camlist, pl, d_canvas, d_rd, j, s));
        if (ev) HIP_TRY(hipEventRecord((*ev)[3], s));
        HIP_TRY(launch_final(kind, user_kind, d_fin, (int)prep.fin.size(), pl, d_canvas, d_rd, st,
                             ot, s));
        if (ev) HIP_TRY(hipEventRecord((*ev)[4], s));
        return PANO_OK;
    }
};

}  // namespace

// ------------------------------------------------------------------ objects --

namespace {
// Device buffers of blend(images, masks): per level the cameras' weight planes
// (one after another, stride[j] elements apart) and the weight sums, all with
// the level-0 row pitch, and the masks' union.  Allocated on first use.
struct CustomBufs {
    int pitch = 0;
    std::vector<int16_t*> w;
    std::vector<int32_t*> wsum;
    std::vector<size_t> stride;
    uint8_t* cover = nullptr;
    ~CustomBufs()
    {
        for (int16_t* p : w) (void)hipFree(p);
        for (int32_t* p : wsum) (void)hipFree(p);
        if (cover) (void)hipFree(cover);
    }
    int alloc(const Prepared& pr)
    {
        const int L = pr.levels;
        pitch = (pr.cols + 63) / 64 * 64;
        w.assign(L + 1, nullptr);
        wsum.assign(L + 1, nullptr);
        stride.assign(L + 1, 0);
        for (int j = 0; j <= L; j++) {
            stride[j] = (size_t)pitch * pr.lh[j];
            if (hipMalloc((void**)&w[j], stride[j] * pr.n * sizeof(int16_t)) != hipSuccess ||
                hipMalloc((void**)&wsum[j], stride[j] * sizeof(int32_t)) != hipSuccess)
                return PANO_ERR_NOMEM;
        }
        if (hipMalloc((void**)&cover, stride[0]) != hipSuccess) return PANO_ERR_NOMEM;
        return PANO_OK;
    }
};
}  // namespace

struct pano_blender {
    std::unique_ptr<DeviceRig> rig;
    bool ok = false;
    bool tiling = false;
    std::vector<uint8_t> tiled;
    std::unique_ptr<CustomBufs> custom;
};

// Events between the stages of each timed frame.
struct StageTimer {
    static constexpr int kMarks = 5;
    bool on = false;
    std::vector<std::vector<hipEvent_t>> pending, spare;
    ~StageTimer()
    {
        for (auto* v : {&pending, &spare})
            for (auto& set : *v)
                for (hipEvent_t e : set) (void)hipEventDestroy(e);
    }
    std::vector<hipEvent_t>* begin_frame()
    {
        if (!on) return nullptr;
        std::vector<hipEvent_t> set;
        if (!spare.empty()) {
            set = spare.back();
            spare.pop_back();
        } else {
            set.resize(kMarks);
            for (auto& e : set)
                if (hipEventCreate(&e) != hipSuccess) return nullptr;
        }
        pending.push_back(set);
        return &pending.back();
    }
    int collect(double* ms, int* frames)
    {
        for (int k = 0; k < kMarks - 1; k++) ms[k] = 0;
        int nf = 0;
        for (auto& set : pending) {
            HIP_TRY(hipEventSynchronize(set[kMarks - 1]));
            for (int k = 0; k < kMarks - 1; k++) {
                float t = 0;
                HIP_TRY(hipEventElapsedTime(&t, set[k], set[k + 1]));
                ms[k] += t;
            }
            nf++;
            spare.push_back(set);
        }
        pending.clear();
        *frames = nf;
        return PANO_OK;
    }
};

struct pano_render {
    std::unique_ptr<DeviceRig> rig;
    bool ok = false;
    std::vector<int> src_w, src_h;
    int dst_w = 0, dst_h = 0;
    PinnedRing ring;
    StageTimer timer;
    int lut_state = 0;                // the device's "apply LUTs" byte
    std::vector<uint8_t> lut_block;   // n x 768 LUTs + the flag byte
    uint8_t* bgrx_scratch = nullptr;  // 16-bit path with encoder-plane output: the BGRx canvas
    double prep_ms[4] = {};           // last prepare: map support, seams, tables, device build
    bool prepare_on_device = true;    // fused-path tables, map storage and seams built on the device
    int tile_shape = 0;               // down tile shape of the next prepare (kDownShapes)
    ~pano_render()
    {
        if (bgrx_scratch) (void)hipFree(bgrx_scratch);
    }

    // Stream-ordered update of the LUT buffer for this frame (fused path).
    int upload_luts(const uint8_t* luts, hipStream_t s)
    {
        const size_t n = rig->prep.n;
        if (luts) {
            lut_block.resize(n * 768 + 1);
            memcpy(lut_block.data(), luts, n * 768);
            lut_block[n * 768] = 1;
            lut_state = 1;
            return ring.upload(lut_block.data(), n * 768 + 1, rig->d_luts, s);
        }
        if (lut_state) {
            static const uint8_t zero = 0;
            lut_state = 0;
            return ring.upload(&zero, 1, rig->d_luts + n * 768, s);
        }
        return PANO_OK;
    }
};

// A batch's launch tables live in one of kSlots device buffers.  A new
// composition is uploaded (stream-ordered, from pinned memory) into the next
// slot at its first render; a slot is reused only after the renders that read
// it have finished, so changing the composition never waits for the device
// unless kSlots compositions are still in flight.
struct pano_batch {
    static constexpr int kSlots = 4;
    static constexpr size_t kSlotBytes = 64 << 10;   // >= 32 launches x 32 renders of Seg
    struct Slot {
        uint8_t* dev = nullptr;
        hipEvent_t ev = nullptr;
        bool used = false;
    };
    std::vector<pano_render*> renders;
    int total_cams = 0;
    // host tables of the current composition: segs of every launch
    std::vector<Seg> segs;
    std::vector<size_t> seg_off;         // per launch: first Seg
    std::vector<Launch> seq;             // segs pointers into the current slot
    bool dirty = false;                  // tables not yet uploaded
    Slot slots[kSlots];
    int cur = -1;
    PinnedRing ring;
    StageTimer timer;
    int sched[4] = {0, 0, 0, 0};        // per stage: block -> job order (autotuned)
    void apply_sched()
    {
        for (Launch& l : seq) l.sched = sched[l.stage];
    }
    // captured launch sequence (pano_batch_capture) and the tables it reads
    hipGraphExec_t exec = nullptr;
    int exec_slot = -1;
    std::vector<char> exec_luts;          // per render: captured with its LUTs applied
    void drop_graph()
    {
        if (!exec) return;
        if (exec_slot >= 0 && slots[exec_slot].used) (void)hipEventSynchronize(slots[exec_slot].ev);
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        exec_slot = -1;
    }
    ~pano_batch()
    {
        drop_graph();
        for (Slot& sl : slots) {
            if (sl.used) (void)hipEventSynchronize(sl.ev);
            if (sl.dev) (void)hipFree(sl.dev);
            if (sl.ev) (void)hipEventDestroy(sl.ev);
        }
    }
    int init()
    {
        for (Slot& sl : slots) {
            if (hipMalloc((void**)&sl.dev, kSlotBytes) != hipSuccess) return PANO_ERR_NOMEM;
            if (hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) return PANO_ERR_HIP;
        }
        return ring.init(kSlotBytes);
    }
    // Upload the current tables into the next slot on stream s.
    int upload(hipStream_t s)
    {
        const size_t sb = sizeof(Seg) * segs.size();
        if (sb > kSlotBytes) return PANO_ERR_UNSUPPORTED;
        const int k = (cur + 1) % kSlots;
        Slot& sl = slots[k];
        if (sl.used) HIP_TRY(hipEventSynchronize(sl.ev));
        int e = ring.upload(segs.data(), sb, sl.dev, s);
        if (e) return e;
        for (size_t i = 0; i < seq.size(); i++) seq[i].segs = (const Seg*)sl.dev + seg_off[i];
        cur = k;
        dirty = false;
        return PANO_OK;
    }
};

extern "C" {

const char* pano_status_string(int status)
{
    switch (status) {
    case PANO_OK: return "ok";
    case PANO_ERR_ARG: return "invalid argument";
    case PANO_ERR_STATE: return "not prepared";
    case PANO_ERR_HIP: return "HIP runtime error";
    case PANO_ERR_NOMEM: return "out of memory";
    case PANO_ERR_UNSUPPORTED: return "unsupported";
    case PANO_ERR_BUDGET: return "device memory budget exceeded";
    case PANO_ERR_IO: return "file could not be opened or written";
    case PANO_ERR_FORMAT: return "rig file field is not a number";
    default: return "unknown status";
    }
}

int pano_version(void) { return 100; }

// ------------------------------------------------------------------ blender --

int pano_blender_create(pano_blender** out)
{
    if (!out) return PANO_ERR_ARG;
    *out = new (std::nothrow) pano_blender();
    return *out ? PANO_OK : PANO_ERR_NOMEM;
}

void pano_blender_destroy(pano_blender* b) { delete b; }

int pano_blender_prepare(pano_blender* b, int num, int rows, int cols, const uint8_t* const* masks,
                         const size_t* mask_steps, int max_levels, int min_length)
{
    if (!b) return PANO_ERR_ARG;
    b->ok = false;
    b->rig.reset(new (std::nothrow) DeviceRig());
    if (!b->rig) return PANO_ERR_NOMEM;
    b->rig->kind = kPixS16;
    std::vector<uint8_t> unique;
    int err = device_seams(num, rows, cols, masks, mask_steps, unique);
    if (err) return err;
    err = prepare_rig(num, rows, cols, masks, mask_steps, nullptr, max_levels, min_length, true, kPixS16,
                      false, b->rig->prep, unique.data());
    if (err) return err;
    err = b->rig->build(true, true, nullptr, nullptr, masks, mask_steps, nullptr, nullptr);
    if (err) return err;
    b->ok = true;
    return PANO_OK;
}

int pano_blender_prepare_device(pano_blender* b, int num, int rows, int cols,
                                const uint8_t* const* masks, const size_t* mask_steps,
                                int max_levels, int min_length)
{
    if (!b || num <= 0 || rows <= 0 || cols <= 0 || !masks || !mask_steps) return PANO_ERR_ARG;
    std::vector<std::vector<uint8_t>> host(num, std::vector<uint8_t>((size_t)rows * cols));
    std::vector<const uint8_t*> ptrs(num);
    std::vector<size_t> steps(num, (size_t)cols);
    for (int i = 0; i < num; i++) {
        if (!masks[i]) return PANO_ERR_ARG;
        HIP_TRY(hipMemcpy2D(host[i].data(), cols, masks[i], mask_steps[i], cols, rows,
                            hipMemcpyDeviceToHost));
        ptrs[i] = host[i].data();
    }
    return pano_blender_prepare(b, num, rows, cols, ptrs.data(), steps.data(), max_levels,
                                min_length);
}

int pano_blender_blend(pano_blender* b, int num, const void* const* images, const size_t* steps,
                       int type, uint8_t* out, size_t out_step, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    DeviceRig& r = *b->rig;
    if (num != r.prep.n || !images || !steps || !out) return PANO_ERR_ARG;
    if (type != PANO_16SC4 && type != PANO_8UC4) return PANO_ERR_ARG;
    if (num > kMaxSrc) return PANO_ERR_UNSUPPORTED;
    size_t px = type == PANO_16SC4 ? 8 : 4;
    if (out_step < (size_t)r.prep.cols * 4 || out_step % 4) return PANO_ERR_ARG;
    SrcTable st{};
    for (int i = 0; i < num; i++) {
        if (!images[i] || steps[i] < (size_t)r.prep.cols * px || steps[i] % px) return PANO_ERR_ARG;
        st.p[i] = (const uint8_t*)images[i];
        st.step[i] = (uint32_t)steps[i];
    }
    OutTable ot{};
    ot.p[0] = out;
    ot.step[0] = (uint32_t)out_step;
    return r.run_pyramid(type == PANO_16SC4 ? kPixS16 : kPixU8, st, ot, stream);
}

int pano_blender_blend_custom(pano_blender* b, int num, const void* const* images, const size_t* steps,
                              int type, const uint8_t* const* masks, const size_t* mask_steps, uint8_t* out,
                              size_t out_step, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    DeviceRig& r = *b->rig;
    const Prepared& pr = r.prep;
    if (num != pr.n || !images || !steps || !masks || !mask_steps || !out) return PANO_ERR_ARG;
    if (type != PANO_16SC4 && type != PANO_8UC4) return PANO_ERR_ARG;
    if (num > kMaxSrc) return PANO_ERR_UNSUPPORTED;
    const size_t px = type == PANO_16SC4 ? 8 : 4;
    if (out_step < (size_t)pr.cols * 4 || out_step % 4 || out_step > 0xffffffffu) return PANO_ERR_ARG;
    SrcTable st{}, mt{};
    for (int i = 0; i < num; i++) {
        if (!images[i] || steps[i] < (size_t)pr.cols * px || steps[i] % px || steps[i] > 0xffffffffu)
            return PANO_ERR_ARG;
        if (!masks[i] || mask_steps[i] < (size_t)pr.cols || mask_steps[i] > 0xffffffffu) return PANO_ERR_ARG;
        st.p[i] = (const uint8_t*)images[i];
        st.step[i] = (uint32_t)steps[i];
        mt.p[i] = masks[i];
        mt.step[i] = (uint32_t)mask_steps[i];
    }
    if (!b->custom) {
        std::unique_ptr<CustomBufs> c(new (std::nothrow) CustomBufs());
        if (!c) return PANO_ERR_NOMEM;
        int e = c->alloc(pr);
        if (e) return e;
        b->custom = std::move(c);
    }
    CustomBufs& cb = *b->custom;
    const int L = pr.levels, user = type == PANO_16SC4 ? kPixS16 : kPixU8;
    // the image pyramids with the prepared alphas
    for (int j = 0; j < L; j++)
        HIP_TRY(launch_down(r.kind, j == 0 ? user : -1, r.d_down[j], r.n_down[j], r.d_planes_user, st, stream));
    HIP_TRY(launch_custom_weights(mt, num, L, pr.lw.data(), pr.lh.data(), cb.w.data(), cb.stride.data(),
                                  cb.pitch, cb.wsum.data(), cb.cover, stream));
    for (int j = L; j >= 0; j--)
        HIP_TRY(launch_blend_custom(user, r.d_planes_user, r.d_canvas, num, L, j, cb.w[j], cb.stride[j],
                                    cb.pitch, cb.wsum[j], st, cb.cover, out, out_step, pr.lw[j], pr.lh[j],
                                    stream));
    return PANO_OK;
}

int pano_blender_begin(pano_blender* b)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    b->tiling = true;
    b->tiled.assign(b->rig->prep.n, 0);
    return PANO_OK;
}

int pano_blender_tile(pano_blender* b, const void* image, size_t step, int type, int index,
                      hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok || !b->tiling) return PANO_ERR_STATE;
    DeviceRig& r = *b->rig;
    if (index < 0 || index >= r.prep.n || !image) return PANO_ERR_ARG;
    if (type != PANO_16SC4 && type != PANO_8UC4) return PANO_ERR_ARG;
    size_t px = type == PANO_16SC4 ? 8 : 4;
    if (step < (size_t)r.prep.cols * px || step % px) return PANO_ERR_ARG;
    const Plane& p = r.planes[(size_t)index * (r.prep.levels + 1)];
    HIP_TRY(launch_import(type == PANO_16SC4 ? kPixS16 : kPixU8, image, step, p.mask,
                          p.mask_pitch, p.g, p.pitch, r.prep.cols, r.prep.rows, stream));
    b->tiled[index] = 1;
    return PANO_OK;
}

int pano_blender_end(pano_blender* b, uint8_t* out, size_t out_step, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->ok || !b->tiling) return PANO_ERR_STATE;
    DeviceRig& r = *b->rig;
    if (!out || out_step < (size_t)r.prep.cols * 4 || out_step % 4) return PANO_ERR_ARG;
    for (int i = 0; i < r.prep.n; i++)
        if (!b->tiled[i]) {
            // a camera that was never tiled contributes a black image
            const Plane& p = r.planes[(size_t)i * (r.prep.levels + 1)];
            HIP_TRY(launch_zero(p.g, (size_t)p.pitch * p.bh * 8, stream));
        }
    SrcTable st{};
    OutTable ot{};
    ot.p[0] = out;
    ot.step[0] = (uint32_t)out_step;
    int err = r.run_pyramid(-1, st, ot, stream);
    b->tiling = false;
    return err;
}

int pano_blender_unique_masks(const pano_blender* b, uint8_t* const* out, const size_t* steps)
{
    if (!b || !out || !steps) return PANO_ERR_ARG;
    if (!b->ok) return PANO_ERR_STATE;
    const Prepared& p = b->rig->prep;
    const size_t P = (size_t)p.rows * p.cols;
    for (int i = 0; i < p.n; i++) {
        if (!out[i] || steps[i] < (size_t)p.cols) return PANO_ERR_ARG;
        for (int y = 0; y < p.rows; y++)
            memcpy(out[i] + (size_t)y * steps[i], p.unique.data() + P * i + (size_t)y * p.cols, p.cols);
    }
    return PANO_OK;
}

int pano_blender_num_levels(const pano_blender* b) { return b && b->ok ? b->rig->prep.levels : -1; }

int64_t pano_blender_mem_bytes(const pano_blender* b)
{
    return b && b->ok ? (int64_t)b->rig->arena.used : 0;
}

int64_t pano_estimate_mem_bytes(int width, int height, int num)
{
    if (width <= 0 || height <= 0 || num <= 0) return 0;
    // full-canvas 16-bit pyramids (x 4/3) + weights (levels >= 1) + level-0 copies,
    // collapsed result pyramid (32-bit x 4, levels >= 1), masks and label
    double P = (double)width * height;
    double per_cam = P * 8 * (1.0 / 3) + P * 2 * (1.0 / 3) + P * 8 + P;
    double canvas = P * 16 * (1.0 / 3) + P * 2 * (1.0 / 3) + P;
    return (int64_t)(num * per_cam + canvas);
}

// ------------------------------------------------------------ free functions --

int pano_reproject_to16s(const uint8_t* src, size_t src_step, int src_w, int src_h,
                         const float* xmap, const float* ymap, size_t map_step, int16_t* dst,
                         size_t dst_step, int dst_w, int dst_h, hipStream_t stream)
{
    if (!src || !xmap || !ymap || !dst || src_w <= 0 || src_h <= 0 || dst_w <= 0 || dst_h <= 0)
        return PANO_ERR_ARG;
    if (src_step < (size_t)src_w * 4 || src_step % 4 || map_step < (size_t)dst_w * 4 ||
        map_step % 4 || dst_step < (size_t)dst_w * 8 || dst_step % 8 || src_step > 0xffffffffu)
        return PANO_ERR_ARG;
    HIP_TRY(launch_reproject16s(src, src_step, src_w, src_h, xmap, ymap, map_step, dst, dst_step,
                                dst_w, dst_h, stream));
    return PANO_OK;
}

int pano_nonintersecting_masks(int num, int rows, int cols, const uint8_t* const* masks, const size_t* mask_steps,
                               uint8_t* const* out, const size_t* out_steps)
{
    if (!out || !out_steps) return PANO_ERR_ARG;
    std::vector<uint8_t> unique;
    int e = device_seams(num, rows, cols, masks, mask_steps, unique);
    if (e) return e;
    for (int i = 0; i < num; i++) {
        if (!out[i] || out_steps[i] < (size_t)cols) return PANO_ERR_ARG;
        for (int y = 0; y < rows; y++)
            memcpy(out[i] + (size_t)y * out_steps[i], unique.data() + ((size_t)i * rows + y) * cols, cols);
    }
    return PANO_OK;
}

int pano_reproject(const uint8_t* src, size_t src_step, int src_w, int src_h, const float* xmap,
                   const float* ymap, size_t map_step, uint8_t* dst, size_t dst_step, int dst_w, int dst_h,
                   hipStream_t stream)
{
    if (!src || !xmap || !ymap || !dst || src_w <= 0 || src_h <= 0 || dst_w <= 0 || dst_h <= 0)
        return PANO_ERR_ARG;
    if (src_step < (size_t)src_w * 4 || src_step % 4 || map_step < (size_t)dst_w * 4 ||
        map_step % 4 || dst_step < (size_t)dst_w * 4 || dst_step % 4 || src_step > 0xffffffffu)
        return PANO_ERR_ARG;
    HIP_TRY(launch_reproject8u(src, src_step, src_w, src_h, xmap, ymap, map_step, dst, dst_step, dst_w,
                               dst_h, stream));
    return PANO_OK;
}

int pano_bgr32_to_yuv420p(const uint8_t* src, size_t src_step, int w, int h, uint8_t* y, size_t y_step,
                          uint8_t* u, size_t u_step, uint8_t* v, size_t v_step, hipStream_t stream)
{
    if (w <= 0 || h <= 0 || (w | h) & 1 || !plane_ok(src, src_step, (size_t)w * 4, h, 8) ||
        !plane_ok(y, y_step, w, h, 2) || !plane_ok(u, u_step, w / 2, h / 2, 1) ||
        !plane_ok(v, v_step, w / 2, h / 2, 1))
        return PANO_ERR_ARG;
    HIP_TRY(launch_bgr32_to_yuv(false, src, src_step, w, h, y, y_step, u, u_step, v, v_step, stream));
    return PANO_OK;
}

int pano_bgr32_to_nv12(const uint8_t* src, size_t src_step, int w, int h, uint8_t* y, size_t y_step,
                       uint8_t* uv, size_t uv_step, hipStream_t stream)
{
    if (w <= 0 || h <= 0 || (w | h) & 1 || !plane_ok(src, src_step, (size_t)w * 4, h, 8) ||
        !plane_ok(y, y_step, w, h, 2) || !plane_ok(uv, uv_step, w, h / 2, 2))
        return PANO_ERR_ARG;
    HIP_TRY(launch_bgr32_to_yuv(true, src, src_step, w, h, y, y_step, uv, uv_step, nullptr, 0, stream));
    return PANO_OK;
}

int pano_yuv420p_to_bgr32(const uint8_t* y, size_t y_step, const uint8_t* u, size_t u_step, const uint8_t* v,
                          size_t v_step, int w, int h, uint8_t* dst, size_t dst_step, hipStream_t stream)
{
    if (w <= 0 || h <= 0 || (w | h) & 1 || !plane_ok(dst, dst_step, (size_t)w * 4, h, 8) ||
        !plane_ok(y, y_step, w, h, 2) || !plane_ok(u, u_step, w / 2, h / 2, 1) ||
        !plane_ok(v, v_step, w / 2, h / 2, 1))
        return PANO_ERR_ARG;
    HIP_TRY(launch_yuv_to_bgr32(false, y, y_step, u, u_step, v, v_step, w, h, dst, dst_step, stream));
    return PANO_OK;
}

int pano_nv12_to_bgr32(const uint8_t* y, size_t y_step, const uint8_t* uv, size_t uv_step, int w, int h,
                       uint8_t* dst, size_t dst_step, hipStream_t stream)
{
    if (w <= 0 || h <= 0 || (w | h) & 1 || !plane_ok(dst, dst_step, (size_t)w * 4, h, 8) ||
        !plane_ok(y, y_step, w, h, 2) || !plane_ok(uv, uv_step, w, h / 2, 2))
        return PANO_ERR_ARG;
    HIP_TRY(launch_yuv_to_bgr32(true, y, y_step, uv, uv_step, nullptr, 0, w, h, dst, dst_step, stream));
    return PANO_OK;
}

int pano_reproject_weighted_accumulate(const uint8_t* src, size_t src_step, int src_w, int src_h, const float* xmap,
                                       const float* ymap, size_t map_step, const float* weight, size_t weight_step,
                                       float* accum, size_t accum_step, int dst_w, int dst_h, hipStream_t stream)
{
    if (!src || !xmap || !ymap || !weight || !accum || src_w <= 0 || src_h <= 0 || dst_w <= 0 || dst_h <= 0)
        return PANO_ERR_ARG;
    if (src_step < (size_t)src_w * 4 || src_step % 4 || src_step > 0xffffffffu || map_step < (size_t)dst_w * 4 ||
        map_step % 4 || weight_step < (size_t)dst_w * 4 || weight_step % 4 || accum_step < (size_t)dst_w * 16 ||
        accum_step % 16)
        return PANO_ERR_ARG;
    HIP_TRY(launch_reproject_accum(src, src_step, src_w, src_h, xmap, ymap, map_step, weight, weight_step, accum,
                                   accum_step, dst_w, dst_h, stream));
    return PANO_OK;
}

int pano_reproject_params(const uint8_t* src, size_t src_step, int src_w, int src_h, const pano_photo_param* param,
                          int dst_w, int dst_h, int type, void* dst, size_t dst_step, uint8_t* mask,
                          size_t mask_step, hipStream_t stream)
{
    if (!param || !dst || dst_w <= 0 || dst_h <= 0) return PANO_ERR_ARG;
    if (type != PANO_8UC4 && type != PANO_16SC4) return PANO_ERR_ARG;
    if (mask && mask_step < (size_t)dst_w) return PANO_ERR_ARG;
    const size_t P = (size_t)dst_w * dst_h;
    float* maps = nullptr;
    uint8_t* m = mask;
    HIP_TRY(hipMallocAsync((void**)&maps, P * 8 + (mask ? 0 : P), stream));
    if (!m) m = (uint8_t*)(maps + 2 * P);
    int e = generate_map_device(*param, src_w, src_h, dst_w, dst_h, maps, maps + P, (size_t)dst_w * 4, m,
                                mask ? mask_step : (size_t)dst_w, stream);
    if (!e)
        e = type == PANO_8UC4 ? pano_reproject(src, src_step, src_w, src_h, maps, maps + P, (size_t)dst_w * 4,
                                               (uint8_t*)dst, dst_step, dst_w, dst_h, stream)
                              : pano_reproject_to16s(src, src_step, src_w, src_h, maps, maps + P,
                                                     (size_t)dst_w * 4, (int16_t*)dst, dst_step, dst_w, dst_h,
                                                     stream);
    const hipError_t f = hipFreeAsync(maps, stream);
    if (e) return e;
    return f == hipSuccess ? PANO_OK : PANO_ERR_HIP;
}

int pano_transform_lut3(const uint8_t* src, size_t src_step, uint8_t* dst, size_t dst_step, int w,
                        int h, const uint8_t* luts, hipStream_t stream)
{
    if (!src || !dst || !luts || w <= 0 || h <= 0) return PANO_ERR_ARG;
    if (src_step < (size_t)w * 4 || dst_step < (size_t)w * 4 || src_step % 4 || dst_step % 4)
        return PANO_ERR_ARG;
    // the LUT travels with the work on the stream: a device copy owned by the
    // stream order (freed asynchronously after the kernel)
    uint8_t* d = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d, 768, stream));
    HIP_TRY(hipMemcpyAsync(d, luts, 768, hipMemcpyHostToDevice, stream));
    hipError_t e = launch_transform(src, src_step, dst, dst_step, w, h, d, stream);
    hipError_t f = hipFreeAsync(d, stream);
    return (e != hipSuccess || f != hipSuccess) ? PANO_ERR_HIP : PANO_OK;
}

int pano_get_lut(uint8_t* lut, double k)
{
    if (!lut || !(k > 0)) return PANO_ERR_ARG;
    for (int i = 0; i < 256; i++) {
        double v = k * i;
        long r = (v > -2147483648.5 && v < 2147483647.5) ? lrint(v) : -2147483648L;
        lut[i] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
    }
    return PANO_OK;
}

int pano_exposure_luts(int num, const double* exposures, const double* red_ratios,
                       const double* blue_ratios, uint8_t* luts)
{
    if (num <= 0 || !exposures || !red_ratios || !blue_ratios || !luts) return PANO_ERR_ARG;
    for (int i = 0; i < num; i++) {
        int e = pano_get_lut(luts + (size_t)i * 768, exposures[i] * blue_ratios[i]);
        if (!e) e = pano_get_lut(luts + (size_t)i * 768 + 256, exposures[i]);
        if (!e) e = pano_get_lut(luts + (size_t)i * 768 + 512, exposures[i] * red_ratios[i]);
        if (e) return e;
    }
    return PANO_OK;
}

int pano_generate_map(const pano_photo_param* param, int src_w, int src_h, int dst_w, int dst_h,
                      float* xmap, float* ymap, uint8_t* mask)
{
    if (!param) return PANO_ERR_ARG;
    return generate_map(*param, src_w, src_h, dst_w, dst_h, xmap, ymap, mask);
}

// ------------------------------------------------------------------- render --

int pano_render_create(pano_render** out)
{
    if (!out) return PANO_ERR_ARG;
    *out = new (std::nothrow) pano_render();
    return *out ? PANO_OK : PANO_ERR_NOMEM;
}

void pano_render_destroy(pano_render* r) { delete r; }

}  // extern "C"

namespace {

// Device buffers of one device prepare: maps (camera i: x plane at 2 i P, y
// plane at (2 i + 1) P), masks and unique masks (n planes of P bytes).
struct PrepBufs {
    float* maps = nullptr;
    uint8_t* masks = nullptr;
    uint8_t* unique = nullptr;
    ~PrepBufs()
    {
        if (maps) (void)hipFree(maps);
        if (masks) (void)hipFree(masks);
        if (unique) (void)hipFree(unique);
    }
    int alloc(int n, size_t P)
    {
        if (hipMalloc((void**)&maps, P * 8 * n) != hipSuccess || hipMalloc((void**)&masks, P * n) != hipSuccess ||
            hipMalloc((void**)&unique, P * n) != hipSuccess)
            return PANO_ERR_NOMEM;
        return PANO_OK;
    }
};

using prep_clock = std::chrono::steady_clock;

double ms_between(prep_clock::time_point a, prep_clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// The fused render's prepare with seams, tables and map storage built on the
// device from consistent device maps and masks.  t0: start of the prepare
// (map generation or checks included in the first timing).
int render_prepare_device(pano_render* r, int num, const int* src_w, const int* src_h, int dst_w, int dst_h,
                          PrepBufs& b, int max_levels, prep_clock::time_point t0)
{
    const size_t P = (size_t)dst_w * dst_h;
    const auto t1 = prep_clock::now();
    HIP_TRY(hipMemcpy(b.unique, b.masks, P * num, hipMemcpyDeviceToDevice));
    HIP_TRY(device_seams_inplace(num, dst_h, dst_w, b.unique));
    const auto t2 = prep_clock::now();
    r->rig.reset(new (std::nothrow) DeviceRig());
    if (!r->rig) return PANO_ERR_NOMEM;
    r->rig->kind = kPixU8;
    DevTables dt;
    int err = prepare_tables_device(num, dst_h, dst_w, b.masks, b.unique, max_levels, true, r->tile_shape,
                                    r->rig->prep, dt);
    if (err) return err;
    const auto t3 = prep_clock::now();
    std::vector<size_t> steps(num, (size_t)dst_w);
    r->rig->dev = &dt;
    r->rig->dev_map = b.maps;
    r->rig->dev_masks = b.masks;
    err = r->rig->build(true, false, nullptr, nullptr, nullptr, steps.data(), src_w, src_h);
    r->rig->dev = nullptr;
    r->rig->dev_map = nullptr;
    r->rig->dev_masks = nullptr;
    if (err) return err;
    HIP_TRY(hipDeviceSynchronize());
    const auto t4 = prep_clock::now();
    r->prep_ms[0] = ms_between(t0, t1);
    r->prep_ms[1] = ms_between(t1, t2);
    r->prep_ms[2] = ms_between(t2, t3);
    r->prep_ms[3] = ms_between(t3, t4);
    err = r->ring.init((size_t)num * 768 + 1);
    if (err) return err;
    r->src_w.assign(src_w, src_w + num);
    r->src_h.assign(src_h, src_h + num);
    r->dst_w = dst_w;
    r->dst_h = dst_h;
    r->ok = true;
    return PANO_OK;
}

// Size limits of the render path (shared by both prepares).
int render_prepare_limits(int num, const int* src_w, const int* src_h, int dst_w, int dst_h)
{
    if ((size_t)dst_w * dst_h * 4 > 0xffffffffu) return PANO_ERR_UNSUPPORTED;   // 32-bit canvas offsets
    for (int i = 0; i < num; i++) {
        if (src_w[i] <= 0 || src_h[i] <= 0) return PANO_ERR_ARG;
        // 24-bit row offsets in the remap, 32-bit frame offsets
        if ((size_t)src_w[i] * 4 >= (1u << 24) || (size_t)src_w[i] * 4 * (uint64_t)src_h[i] > 0xffffffffu)
            return PANO_ERR_UNSUPPORTED;
    }
    return PANO_OK;
}

}  // namespace

extern "C" {

int pano_render_set_prepare_on_device(pano_render* r, int on)
{
    if (!r) return PANO_ERR_ARG;
    r->prepare_on_device = on != 0;
    return PANO_OK;
}

int pano_render_set_tile_shape(pano_render* r, int shape)
{
    if (!r || shape < 0 || shape >= kDownShapes) return PANO_ERR_ARG;
    r->tile_shape = shape;
    return PANO_OK;
}

int pano_render_tile_shape(const pano_render* r)
{
    if (!r || !r->ok) return PANO_ERR_STATE;
    return r->rig->prep.shape;
}

int pano_render_autotune_tile_shape(pano_render* r, int num, const pano_photo_param* params, const int* src_w,
                                    const int* src_h, int dst_w, int dst_h, int max_levels,
                                    const uint8_t* const* srcs, const size_t* src_steps, uint8_t* dst,
                                    size_t dst_step, int iters, int* best, float* ms_out, hipStream_t stream)
{
    if (!r || !best || iters <= 0) return PANO_ERR_ARG;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return PANO_ERR_HIP;
    }
    int err = PANO_OK, pick = -1;
    float best_ms = 0, shape0_ms = 0;
    for (int sh = 0; sh < kDownShapes && !err; sh++) {
        r->tile_shape = sh;
        err = pano_render_prepare(r, num, params, src_w, src_h, dst_w, dst_h, max_levels);
        if (!err) err = pano_render_render(r, num, srcs, src_steps, nullptr, dst, dst_step, stream);   // warm-up
        if (!err && hipEventRecord(e0, stream) != hipSuccess) err = PANO_ERR_HIP;
        for (int k = 0; k < iters && !err; k++)
            err = pano_render_render(r, num, srcs, src_steps, nullptr, dst, dst_step, stream);
        float ms = 0;
        if (!err && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            err = PANO_ERR_HIP;
        if (err) break;
        ms /= iters;
        if (ms_out) ms_out[sh] = ms;
        if (sh == 0) shape0_ms = ms;
        if (pick < 0 || ms < best_ms) {
            pick = sh;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err) return err;
    // the default shape unless another is clearly faster (timing noise is ~1%)
    if (pick != 0 && shape0_ms <= best_ms * 1.02f) pick = 0;
    *best = pick;
    r->tile_shape = pick;
    // leave the render prepared with the chosen shape
    return pick == kDownShapes - 1 ? PANO_OK
                                   : pano_render_prepare(r, num, params, src_w, src_h, dst_w, dst_h, max_levels);
}

int pano_render_prepare_maps(pano_render* r, int num, const int* src_w, const int* src_h,
                             int dst_w, int dst_h, const float* const* xmaps,
                             const float* const* ymaps, const uint8_t* const* masks,
                             int max_levels)
{
    if (!r) return PANO_ERR_ARG;
    r->ok = false;
    if (num <= 0 || num > kMaxSrc || !src_w || !src_h || !xmaps || !ymaps || !masks)
        return PANO_ERR_ARG;
    if ((dst_w & 1) || (dst_h & 1) || dst_h * 2 != dst_w || dst_w <= 0) return PANO_ERR_ARG;
    if (max_levels <= 0) return PANO_ERR_ARG;
    const size_t P = (size_t)dst_w * dst_h;
    if (P * 4 > 0xffffffffu) return PANO_ERR_UNSUPPORTED;   // 32-bit canvas offsets
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    std::vector<size_t> steps(num, (size_t)dst_w);
    // pixels whose level-0 value can be non-zero: mask or a valid map point
    std::vector<std::vector<uint8_t>> support(num, std::vector<uint8_t>(P));
    std::vector<const uint8_t*> sp(num);
    for (int i = 0; i < num; i++) {
        if (!xmaps[i] || !ymaps[i] || !masks[i] || src_w[i] <= 0 || src_h[i] <= 0) return PANO_ERR_ARG;
        // 24-bit row offsets in the remap, 32-bit frame offsets
        if ((size_t)src_w[i] * 4 >= (1u << 24) || (size_t)src_w[i] * 4 * (uint64_t)src_h[i] > 0xffffffffu)
            return PANO_ERR_UNSUPPORTED;
        sp[i] = support[i].data();
    }
    std::atomic<bool> consistent{true};
    parallel_for(num, [&](int i) {
        bool ok = true;
        for (size_t p = 0; p < P; p++) {
            double x = xmaps[i][p], y = ymaps[i][p];
            bool v = x >= 0 && y >= 0 && x < src_w[i] && y < src_h[i];
            support[i][p] = v;
            if (v && !masks[i][p]) ok = false;
        }
        if (!ok) consistent = false;
    });
    // 8-bit planes hold every Gaussian level exactly when sampled pixels lie
    // inside the masks (as getReprojectMapAndMask produces them); otherwise the
    // reference's pyramid can leave [0, 255] and 16-bit planes are used.
    const bool fused = consistent && true_num_levels(dst_w, dst_h, max_levels, 2) >= 1;
    if (fused && r->prepare_on_device && num <= 32) {
        PrepBufs b;
        int e = b.alloc(num, P);
        if (e) return e;
        for (int i = 0; i < num; i++) {
            HIP_TRY(hipMemcpy(b.maps + 2 * P * i, xmaps[i], P * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(b.maps + (2 * (size_t)i + 1) * P, ymaps[i], P * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(b.masks + P * i, masks[i], P, hipMemcpyHostToDevice));
        }
        return render_prepare_device(r, num, src_w, src_h, dst_w, dst_h, b, max_levels, t0);
    }
    r->rig.reset(new (std::nothrow) DeviceRig());
    if (!r->rig) return PANO_ERR_NOMEM;
    r->rig->kind = fused ? kPixU8 : kPixS16;
    const auto t1 = clock::now();
    std::vector<uint8_t> unique;
    int err = device_seams(num, dst_h, dst_w, masks, steps.data(), unique);
    if (err) return err;
    const auto t2 = clock::now();
    err = prepare_rig(num, dst_h, dst_w, masks, steps.data(), sp.data(), max_levels, 2, false, r->rig->kind,
                      fused, r->rig->prep, unique.data(), fused ? r->tile_shape : 0);
    if (err) return err;
    const auto t3 = clock::now();
    err = r->rig->build(true, false, xmaps, ymaps, masks, steps.data(), src_w, src_h);
    if (err) return err;
    const auto t4 = clock::now();
    auto ms = [](clock::time_point a, clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    r->prep_ms[0] = ms(t0, t1);
    r->prep_ms[1] = ms(t1, t2);
    r->prep_ms[2] = ms(t2, t3);
    r->prep_ms[3] = ms(t3, t4);
    err = r->ring.init((size_t)num * 768 + 1);
    if (err) return err;
    r->src_w.assign(src_w, src_w + num);
    r->src_h.assign(src_h, src_h + num);
    r->dst_w = dst_w;
    r->dst_h = dst_h;
    r->ok = true;
    return PANO_OK;
}

int pano_generate_map_device(const pano_photo_param* param, int src_w, int src_h, int dst_w, int dst_h,
                             float* xmap, float* ymap, size_t map_step, uint8_t* mask, size_t mask_step,
                             hipStream_t stream)
{
    if (!param) return PANO_ERR_ARG;
    return generate_map_device(*param, src_w, src_h, dst_w, dst_h, xmap, ymap, map_step, mask, mask_step, stream);
}

int pano_render_prepare(pano_render* r, int num, const pano_photo_param* params, const int* src_w,
                        const int* src_h, int dst_w, int dst_h, int max_levels)
{
    if (!r || num <= 0 || !params || !src_w || !src_h) return PANO_ERR_ARG;
    if ((dst_w & 1) || (dst_h & 1) || dst_h * 2 != dst_w || dst_w <= 0) return PANO_ERR_ARG;
    const size_t P = (size_t)dst_w * dst_h;
    if (r->prepare_on_device && num <= 32) {
        if (max_levels <= 0) return PANO_ERR_ARG;
        int e = render_prepare_limits(num, src_w, src_h, dst_w, dst_h);
        if (e) return e;
        r->ok = false;
        const auto t0 = prep_clock::now();
        PrepBufs b;
        e = b.alloc(num, P);
        if (e) return e;
        for (int i = 0; i < num && !e; i++)
            e = generate_map_device(params[i], src_w[i], src_h[i], dst_w, dst_h, b.maps + 2 * P * i,
                                    b.maps + (2 * (size_t)i + 1) * P, (size_t)dst_w * 4, b.masks + P * i,
                                    (size_t)dst_w, nullptr);
        if (e) return e;
        // masks made by the generator hold every valid map point; checked all the same
        bool all = true;
        for (int i = 0; i < num && !e; i++) {
            bool c = true;
            e = maps_consistent_device(b.maps + 2 * P * i, b.maps + (2 * (size_t)i + 1) * P, b.masks + P * i, P,
                                       src_w[i], src_h[i], c);
            all = all && c;
        }
        if (e) return e;
        if (all && true_num_levels(dst_w, dst_h, max_levels, 2) >= 1)
            return render_prepare_device(r, num, src_w, src_h, dst_w, dst_h, b, max_levels, t0);
        // otherwise the host prepare below, on maps read back
    }
    std::vector<std::vector<float>> xm(num, std::vector<float>(P)), ym(num, std::vector<float>(P));
    std::vector<std::vector<uint8_t>> mk(num, std::vector<uint8_t>(P));
    std::vector<const float*> xp(num), yp(num);
    std::vector<const uint8_t*> kp(num);
    // maps and masks on the device (cudaGenerateReprojectMapsAndMasks), read back
    // for the host tables of the prepare
    float* dmap = nullptr;
    uint8_t* dmask = nullptr;
    if (hipMalloc((void**)&dmap, P * 8) != hipSuccess) return PANO_ERR_NOMEM;
    if (hipMalloc((void**)&dmask, P) != hipSuccess) {
        (void)hipFree(dmap);
        return PANO_ERR_NOMEM;
    }
    int e = PANO_OK;
    for (int i = 0; i < num && !e; i++) {
        e = generate_map_device(params[i], src_w[i], src_h[i], dst_w, dst_h, dmap, dmap + P, (size_t)dst_w * 4,
                                dmask, (size_t)dst_w, nullptr);
        if (!e && (hipMemcpy(xm[i].data(), dmap, P * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(ym[i].data(), dmap + P, P * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                   hipMemcpy(mk[i].data(), dmask, P, hipMemcpyDeviceToHost) != hipSuccess))
            e = PANO_ERR_HIP;
        xp[i] = xm[i].data();
        yp[i] = ym[i].data();
        kp[i] = mk[i].data();
    }
    (void)hipFree(dmap);
    (void)hipFree(dmask);
    if (e) return e;
    return pano_render_prepare_maps(r, num, src_w, src_h, dst_w, dst_h, xp.data(), yp.data(),
                                    kp.data(), max_levels);
}

}  // extern "C"

namespace {

// Checks an output of a w x h canvas and enters it as OutTable entry k.
int output_entry(const pano_output& o, int w, int h, OutTable& ot, int k)
{
    switch (o.format) {
    case PANO_OUT_BGRX:
        if (!o.plane[0] || o.step[0] < (size_t)w * 4 || o.step[0] % 4 || o.step[0] >= (1u << 24) ||
            o.step[0] * (uint64_t)h > 0xffffffffu)
            return PANO_ERR_ARG;
        ot.fmt[k] = kOutBGRX;
        break;
    case PANO_OUT_NV12:
        if (!plane_ok(o.plane[0], o.step[0], w, h, 2) || !plane_ok(o.plane[1], o.step[1], w, h / 2, 2))
            return PANO_ERR_ARG;
        ot.fmt[k] = kOutNV12;
        break;
    case PANO_OUT_YUV420P:
        if (!plane_ok(o.plane[0], o.step[0], w, h, 2) || !plane_ok(o.plane[1], o.step[1], w / 2, h / 2, 1) ||
            !plane_ok(o.plane[2], o.step[2], w / 2, h / 2, 1))
            return PANO_ERR_ARG;
        ot.fmt[k] = kOutYUV420P;
        break;
    default:
        return PANO_ERR_ARG;
    }
    if ((w | h) & 1) return PANO_ERR_ARG;   // canvases are even (W = 2H, H even)
    ot.p[k] = o.plane[0];
    ot.step[k] = (uint32_t)o.step[0];
    ot.p2[k] = o.plane[1];
    ot.step2[k] = (uint32_t)o.step[1];
    ot.p3[k] = o.plane[2];
    ot.step3[k] = (uint32_t)o.step[2];
    return PANO_OK;
}

pano_output bgrx_output(uint8_t* dst, size_t step)
{
    pano_output o{};
    o.format = PANO_OUT_BGRX;
    o.plane[0] = dst;
    o.step[0] = step;
    return o;
}

// Argument checks of a render call; fills the frame and output tables.
int render_args(const pano_render* r, int num, const uint8_t* const* srcs, const size_t* src_steps,
                const pano_output& out, SrcTable& st, OutTable& ot)
{
    if (!r) return PANO_ERR_ARG;
    if (!r->ok) return PANO_ERR_STATE;
    const DeviceRig& d = *r->rig;
    if (num != d.prep.n || !srcs || !src_steps) return PANO_ERR_ARG;
    ot = OutTable{};
    int e = output_entry(out, r->dst_w, r->dst_h, ot, 0);
    if (e) return e;
    st = SrcTable{};
    for (int i = 0; i < num; i++) {
        if (!srcs[i] || src_steps[i] < (size_t)r->src_w[i] * 4 || src_steps[i] % 4 ||
            src_steps[i] * (uint64_t)r->src_h[i] > 0xffffffffu || src_steps[i] >= (1u << 24))
            return PANO_ERR_ARG;
        st.p[i] = srcs[i];
        st.step[i] = (uint32_t)src_steps[i];
    }
    return PANO_OK;
}

// The fused sequence with the remap applying the render's LUTs to its taps.
int run_fused_with_luts(DeviceRig& d, const SrcTable& st, const OutTable& ot, hipStream_t stream,
                        std::vector<hipEvent_t>* ev, bool ev0_recorded)
{
    return run_sequence(d.seq, st, ot, stream, ev, true, ev0_recorded);
}

}  // namespace

extern "C" {

int pano_render_render_to(pano_render* r, int num, const uint8_t* const* srcs, const size_t* src_steps,
                          const uint8_t* luts, const pano_output* out, hipStream_t stream)
{
    if (!out) return PANO_ERR_ARG;
    SrcTable st;
    OutTable ot;
    int err = render_args(r, num, srcs, src_steps, *out, st, ot);
    if (err) return err;
    DeviceRig& d = *r->rig;
    if (!d.prep.fused0 && out->format != PANO_OUT_BGRX) {
        // the 16-bit kernels write BGRx: render into a resident canvas, then convert
        const size_t step = (size_t)r->dst_w * 4;
        if (!r->bgrx_scratch && hipMalloc((void**)&r->bgrx_scratch, step * r->dst_h) != hipSuccess) {
            r->bgrx_scratch = nullptr;
            return PANO_ERR_NOMEM;
        }
        const pano_output canvas = bgrx_output(r->bgrx_scratch, step);
        err = pano_render_render_to(r, num, srcs, src_steps, luts, &canvas, stream);
        if (err) return err;
        const bool nv12 = out->format == PANO_OUT_NV12;
        HIP_TRY(launch_bgr32_to_yuv(nv12, r->bgrx_scratch, step, r->dst_w, r->dst_h, out->plane[0], out->step[0],
                                    out->plane[1], out->step[1], out->plane[2], out->step[2], stream));
        return PANO_OK;
    }
    std::vector<hipEvent_t>* ev = r->timer.begin_frame();
    if (d.prep.fused0) {
        int e = r->upload_luts(luts, stream);
        if (e) return e;
        if (!luts) return run_sequence(d.seq, st, ot, stream, ev);
        return run_fused_with_luts(d, st, ot, stream, ev, false);
    }
    if (luts) {
        int e = r->ring.upload(luts, (size_t)num * 768, d.d_luts, stream);
        if (e) return e;
    }
    if (ev) HIP_TRY(hipEventRecord((*ev)[0], stream));
    HIP_TRY(launch_remap(d.kind, d.d_remap, (int)d.prep.remap.size(), d.d_cams, d.d_planes,
                         d.d_luts, luts != nullptr, st, stream));
    if (ev) HIP_TRY(hipEventRecord((*ev)[1], stream));
    return d.run_pyramid(-1, st, ot, stream, ev);
}

int pano_render_render(pano_render* r, int num, const uint8_t* const* srcs,
                       const size_t* src_steps, const uint8_t* luts, uint8_t* dst,
                       size_t dst_step, hipStream_t stream)
{
    const pano_output o = bgrx_output(dst, dst_step);
    return pano_render_render_to(r, num, srcs, src_steps, luts, &o, stream);
}

int pano_render_render_autogain(pano_render* r, int num, const uint8_t* const* srcs,
                                const size_t* src_steps, int flags, uint8_t* dst,
                                size_t dst_step, hipStream_t stream)
{
    SrcTable st;
    OutTable ot;
    int err = render_args(r, num, srcs, src_steps, bgrx_output(dst, dst_step), st, ot);
    if (err) return err;
    DeviceRig& d = *r->rig;
    if (!d.prep.fused0 || !d.d_cover || d.seq.empty()) return PANO_ERR_UNSUPPORTED;
    if (flags & ~PANO_GAIN_TINT) return PANO_ERR_ARG;
    std::vector<hipEvent_t>* ev = r->timer.begin_frame();
    if (ev) HIP_TRY(hipEventRecord((*ev)[0], stream));
    const int n = d.prep.n;
    HIP_TRY(launch_zero(d.d_gstats, sizeof(unsigned long long) * n * n * 5 * kGainSlots, stream));
    HIP_TRY(launch_gain(d.seq[0].segs, 1, d.seq[0].blocks, d.d_rd,
                        (flags & PANO_GAIN_TINT) ? kGainTint : 0, st, stream));
    r->lut_state = 1;   // the solve set the device's "apply LUTs" byte
    return run_fused_with_luts(d, st, ot, stream, ev, true);
}

int pano_render_gains(const pano_render* r, double* gains, double* red, double* blue, int* anchor)
{
    if (!r) return PANO_ERR_ARG;
    if (!r->ok) return PANO_ERR_STATE;
    const DeviceRig& d = *r->rig;
    if (!d.d_gains) return PANO_ERR_UNSUPPORTED;
    const int n = d.prep.n;
    std::vector<double> h(3 * n + 1);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h.data(), d.d_gains, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    if (gains) memcpy(gains, h.data(), sizeof(double) * n);
    if (red) memcpy(red, h.data() + n, sizeof(double) * n);
    if (blue) memcpy(blue, h.data() + 2 * n, sizeof(double) * n);
    if (anchor) *anchor = (int)h[3 * n];
    return PANO_OK;
}

int pano_render_gain_sums(const pano_render* r, int64_t* sums)
{
    if (!r || !sums) return PANO_ERR_ARG;
    if (!r->ok) return PANO_ERR_STATE;
    const DeviceRig& d = *r->rig;
    if (!d.d_gstats) return PANO_ERR_UNSUPPORTED;
    const int n = d.prep.n;
    HIP_TRY(hipDeviceSynchronize());
    std::vector<int64_t> slots((size_t)kGainSlots * n * n * 5);
    HIP_TRY(hipMemcpy(slots.data(), d.d_gstats, sizeof(int64_t) * slots.size(), hipMemcpyDeviceToHost));
    for (int t = 0; t < n * n * 5; t++) {
        int64_t v = 0;
        for (int k = 0; k < kGainSlots; k++) v += slots[(size_t)k * n * n * 5 + t];
        sums[t] = v;
    }
    return PANO_OK;
}

int pano_render_set_timing(pano_render* r, int enable)
{
    if (!r) return PANO_ERR_ARG;
    r->timer.on = enable != 0;
    return PANO_OK;
}

int pano_render_stage_ms(pano_render* r, double* ms4, int* frames)
{
    if (!r || !ms4 || !frames) return PANO_ERR_ARG;
    return r->timer.collect(ms4, frames);
}

int pano_render_fused(const pano_render* r) { return r && r->ok && r->rig->prep.fused0 ? 1 : 0; }

int pano_render_prepare_ms(const pano_render* r, double* ms4)
{
    if (!r || !ms4) return PANO_ERR_ARG;
    for (int k = 0; k < 4; k++) ms4[k] = r->prep_ms[k];
    return PANO_OK;
}

int pano_render_num_images(const pano_render* r) { return r && r->ok ? r->rig->prep.n : 0; }
int pano_render_num_levels(const pano_render* r) { return r && r->ok ? r->rig->prep.levels : -1; }
int64_t pano_render_mem_bytes(const pano_render* r)
{
    return r && r->ok ? (int64_t)r->rig->arena.used : 0;
}
int64_t pano_render_model_bytes(const pano_render* r, int stage)
{
    if (!r || !r->ok || stage < 0 || stage > 4) return -1;
    return (int64_t)r->rig->model_bytes[stage];
}

// -------------------------------------------------------------------- batch --

int pano_batch_create(pano_batch** out)
{
    if (!out) return PANO_ERR_ARG;
    pano_batch* b = new (std::nothrow) pano_batch();
    if (!b) return PANO_ERR_NOMEM;
    int e = b->init();
    if (e) {
        delete b;
        return e;
    }
    *out = b;
    return PANO_OK;
}

void pano_batch_destroy(pano_batch* b) { delete b; }

int pano_batch_set(pano_batch* b, int num, pano_render* const* renders)
{
    if (!b || num <= 0 || num > kMaxBatch || num > kMaxOut || !renders) return PANO_ERR_ARG;
    int cams = 0, maxL = 0;
    for (int i = 0; i < num; i++) {
        pano_render* r = renders[i];
        if (!r || !r->ok) return PANO_ERR_STATE;
        if (!r->rig->prep.fused0) return PANO_ERR_UNSUPPORTED;
        // one launch per stage for all renders: one down tile shape
        if (r->rig->prep.shape != renders[0]->rig->prep.shape) return PANO_ERR_ARG;
        cams += r->rig->prep.n;
        maxL = std::max(maxL, r->rig->prep.levels);
    }
    if (cams > kMaxSrc) return PANO_ERR_UNSUPPORTED;
    // segment tables: one per launch, each render's share of the launch in order
    std::vector<std::vector<Seg>> tables;
    std::vector<Launch> ls;
    auto build = [&](int stage, int level, auto pick) {
        std::vector<Seg> t;
        int begin = 0, src_base = 0;
        for (int i = 0; i < num; i++) {
            DeviceRig& d = *renders[i]->rig;
            const void* jobs = nullptr;
            int count = 0;
            pick(d, jobs, count);
            if (count > 0) {
                Seg g{};
                g.rd = d.d_rd;
                g.jobs = jobs;
                g.begin = begin;
                g.count = count;
                g.src_base = src_base;
                g.out = i;
                t.push_back(g);
                begin += count;
            }
            src_base += d.prep.n;
        }
        if (!t.empty()) {
            ls.push_back(Launch{stage, level, (int)t.size(), begin, nullptr});
            tables.push_back(t);
        }
    };
    build(kStageRemapDown0, 0, [&](DeviceRig& d, const void*& j, int& c) {
        if (d.prep.levels >= 1) { j = d.d_down[0]; c = d.n_down[0]; }
    });
    for (int lv = 1; lv < maxL; lv++)
        build(kStageDown, lv, [&](DeviceRig& d, const void*& j, int& c) {
            if (lv < d.prep.levels) { j = d.d_down[lv]; c = d.n_down[lv]; }
        });
    for (int lv = maxL; lv >= 1; lv--)
        build(kStageBlend, lv, [&](DeviceRig& d, const void*& j, int& c) {
            if (lv <= d.prep.levels) { j = d.d_blend[lv]; c = d.n_blend[lv]; }
        });
    build(kStageFinal, 0, [&](DeviceRig& d, const void*& j, int& c) {
        j = d.d_fin;
        c = (int)d.prep.fin.size();
    });
    std::vector<Seg> flat;
    std::vector<size_t> offs;
    for (auto& t : tables) {
        offs.push_back(flat.size());
        flat.insert(flat.end(), t.begin(), t.end());
    }
    if (sizeof(Seg) * flat.size() > pano_batch::kSlotBytes)
        return PANO_ERR_UNSUPPORTED;
    b->drop_graph();
    for (Launch& l : ls) {
        l.sched = b->sched[l.stage];
        l.shape = renders[0]->rig->prep.shape;
    }
    b->segs.swap(flat);
    b->seg_off.swap(offs);
    b->seq = ls;
    b->renders.assign(renders, renders + num);
    b->total_cams = cams;
    b->dirty = true;
    return PANO_OK;
}

}  // extern "C"

namespace {

// The frame and canvas tables of one batched render, after checking the
// arguments; uploads the renders' LUTs and the batch's tables on the stream.
// any: whether the remap applies LUTs for some render this call.
int batch_prepare(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                  const uint8_t* const* luts, const pano_output* outs, hipStream_t stream, SrcTable& st,
                  OutTable& ot, bool& any)
{
    if (!b) return PANO_ERR_ARG;
    if (b->renders.empty()) return PANO_ERR_STATE;
    if (!srcs || !src_steps || !outs) return PANO_ERR_ARG;
    st = SrcTable{};
    ot = OutTable{};
    int k = 0;
    for (size_t i = 0; i < b->renders.size(); i++) {
        pano_render* r = b->renders[i];
        int e = output_entry(outs[i], r->dst_w, r->dst_h, ot, (int)i);
        if (e) return e;
        for (int c = 0; c < r->rig->prep.n; c++, k++) {
            if (!srcs[k] || src_steps[k] < (size_t)r->src_w[c] * 4 || src_steps[k] % 4 ||
                src_steps[k] * (uint64_t)r->src_h[c] > 0xffffffffu || src_steps[k] >= (1u << 24))
                return PANO_ERR_ARG;
            st.p[k] = srcs[k];
            st.step[k] = (uint32_t)src_steps[k];
        }
    }
    // each render's LUTs and its "LUTs active" byte; any: the remap applies LUTs
    any = false;
    for (size_t i = 0; i < b->renders.size(); i++) {
        const uint8_t* lt = luts ? luts[i] : nullptr;
        int e = b->renders[i]->upload_luts(lt, stream);
        if (e) return e;
        any = any || lt;
    }
    if (b->dirty) {
        int e = b->upload(stream);
        if (e) return e;
    }
    return PANO_OK;
}


}  // namespace

extern "C" {

int pano_batch_render_to(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                         const uint8_t* const* luts, const pano_output* outs, hipStream_t stream)
{
    SrcTable st;
    OutTable ot;
    bool any = false;
    int e = batch_prepare(b, srcs, src_steps, luts, outs, stream, st, ot, any);
    if (e) return e;
    e = run_sequence(b->seq, st, ot, stream, b->timer.begin_frame(), any);
    if (e) return e;
    pano_batch::Slot& sl = b->slots[b->cur];
    HIP_TRY(hipEventRecord(sl.ev, stream));
    sl.used = true;
    return PANO_OK;
}

int pano_batch_capture_to(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                          const uint8_t* const* luts, const pano_output* outs, hipStream_t stream)
{
    if (!b || !stream) return PANO_ERR_ARG;   // the null stream cannot be captured
    b->drop_graph();
    SrcTable st;
    OutTable ot;
    bool any = false;
    // uploads happen here, outside the capture: the graph holds kernel launches only
    int e = batch_prepare(b, srcs, src_steps, luts, outs, stream, st, ot, any);
    if (e) return e;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    e = run_sequence(b->seq, st, ot, stream, nullptr, any);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(stream, &graph);
    if (!e && ce != hipSuccess) e = PANO_ERR_HIP;
    if (!e && hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        b->exec = nullptr;
        e = PANO_ERR_HIP;
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (e) return e;
    b->exec_slot = b->cur;
    b->exec_luts.assign(b->renders.size(), 0);
    for (size_t i = 0; i < b->renders.size(); i++) b->exec_luts[i] = luts && luts[i];
    return PANO_OK;
}

namespace {
// BGRx outputs of a batch as pano_output records (null when an argument is missing).
bool bgrx_outputs(const pano_batch* b, uint8_t* const* dsts, const size_t* dst_steps,
                  std::vector<pano_output>& outs)
{
    if (!b || !dsts || !dst_steps) return false;
    outs.resize(b->renders.size());
    for (size_t i = 0; i < outs.size(); i++) outs[i] = bgrx_output(dsts[i], dst_steps[i]);
    return true;
}
}  // namespace

int pano_batch_render(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                      const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                      hipStream_t stream)
{
    std::vector<pano_output> outs;
    if (!bgrx_outputs(b, dsts, dst_steps, outs)) return PANO_ERR_ARG;
    return pano_batch_render_to(b, srcs, src_steps, luts, outs.data(), stream);
}

int pano_batch_capture(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                       const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                       hipStream_t stream)
{
    std::vector<pano_output> outs;
    if (!bgrx_outputs(b, dsts, dst_steps, outs)) return PANO_ERR_ARG;
    return pano_batch_capture_to(b, srcs, src_steps, luts, outs.data(), stream);
}

int pano_batch_replay(pano_batch* b, const uint8_t* const* luts, hipStream_t stream)
{
    if (!b) return PANO_ERR_ARG;
    if (!b->exec || b->dirty || b->cur != b->exec_slot) return PANO_ERR_STATE;
    for (size_t i = 0; i < b->renders.size(); i++) {
        const uint8_t* lt = luts ? luts[i] : nullptr;
        if (lt && !b->exec_luts[i]) return PANO_ERR_ARG;   // captured with the LUT-free remap
        if (lt) {
            int e = b->renders[i]->upload_luts(lt, stream);
            if (e) return e;
        }
    }
    std::vector<hipEvent_t>* ev = b->timer.begin_frame();
    if (ev) HIP_TRY(hipEventRecord((*ev)[0], stream));
    HIP_TRY(hipGraphLaunch(b->exec, stream));
    if (ev)
        for (int k = 1; k <= 4; k++) HIP_TRY(hipEventRecord((*ev)[k], stream));
    pano_batch::Slot& sl = b->slots[b->cur];
    HIP_TRY(hipEventRecord(sl.ev, stream));
    sl.used = true;
    return PANO_OK;
}

int pano_batch_set_sched(pano_batch* b, const int* sched4)
{
    if (!b || !sched4) return PANO_ERR_ARG;
    for (int k = 0; k < 4; k++)
        if (sched4[k] < 0 || sched4[k] >= kSchedModes) return PANO_ERR_ARG;
    b->drop_graph();
    for (int k = 0; k < 4; k++) b->sched[k] = sched4[k];
    b->apply_sched();
    return PANO_OK;
}

int pano_batch_get_sched(const pano_batch* b, int* sched4)
{
    if (!b || !sched4) return PANO_ERR_ARG;
    for (int k = 0; k < 4; k++) sched4[k] = b->sched[k];
    return PANO_OK;
}

int pano_batch_autotune(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                        const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                        hipStream_t stream, int iters, double* ms_out)
{
    if (!b || iters <= 0 || iters > 1000) return PANO_ERR_ARG;
    if (b->renders.empty()) return PANO_ERR_STATE;
    b->drop_graph();
    const bool was_on = b->timer.on;
    double scratch[4];
    int frames = 0;
    if (was_on) {
        int e = b->timer.collect(scratch, &frames);   // the caller's pending frames are dropped
        if (e) return e;
    }
    double ms[kSchedModes][4] = {};
    int e = PANO_OK;
    for (int m = 0; m < kSchedModes && !e; m++) {
        for (int k = 0; k < 4; k++) b->sched[k] = m;
        b->apply_sched();
        b->timer.on = false;
        e = pano_batch_render(b, srcs, src_steps, luts, dsts, dst_steps, stream);   // warm-up
        b->timer.on = true;
        for (int it = 0; it < iters && !e; it++)
            e = pano_batch_render(b, srcs, src_steps, luts, dsts, dst_steps, stream);
        if (!e) e = b->timer.collect(ms[m], &frames);
    }
    b->timer.on = was_on;
    if (e) return e;
    for (int k = 0; k < 4; k++) {
        int best = 0;
        for (int m = 1; m < kSchedModes; m++)
            if (ms[m][k] < ms[best][k]) best = m;
        b->sched[k] = best;
    }
    b->apply_sched();
    if (ms_out)
        for (int m = 0; m < kSchedModes; m++)
            for (int k = 0; k < 4; k++) ms_out[m * 4 + k] = ms[m][k] / iters;
    return PANO_OK;
}

int pano_batch_set_timing(pano_batch* b, int enable)
{
    if (!b) return PANO_ERR_ARG;
    b->timer.on = enable != 0;
    return PANO_OK;
}

int pano_batch_stage_ms(pano_batch* b, double* ms4, int* frames)
{
    if (!b || !ms4 || !frames) return PANO_ERR_ARG;
    return b->timer.collect(ms4, frames);
}

}  // extern "C"
