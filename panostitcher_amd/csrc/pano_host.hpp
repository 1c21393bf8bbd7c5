// pano_host.hpp -- host-side prepare of a rig: seams, pyramid geometry,
// blend weights and work lists (everything the per-frame kernels read).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <vector>

#include "pano_dev.hpp"

namespace pano {

// f(0) .. f(n - 1) on up to 16 host threads.
template <class F>
void parallel_for(int n, F f)
{
    int hw = (int)std::thread::hardware_concurrency();
    int nt = std::max(1, std::min(n, std::min(hw > 0 ? hw : 1, 16)));
    if (nt <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

// Geometry of one camera plane at one level (a box of the level canvas).
struct Geom {
    int x0 = 0, y0 = 0, bw = 0, bh = 0;  // box origin / size on the level canvas
    int cw = 0, ch = 0;                  // level canvas size
    int pitch = 0;                       // pixels per row
    size_t pixels() const { return (size_t)pitch * bh; }
    bool empty() const { return bw <= 0 || bh <= 0; }
};

struct Prepared {
    int n = 0, rows = 0, cols = 0, levels = 0, full = 0;
    std::vector<int> lw, lh;                     // level canvas sizes
    std::vector<uint8_t> unique;                 // n * rows * cols, 0 / 255
    std::vector<uint8_t> label;                  // rows * cols, owner or kNoOwner
    std::vector<Geom> geom;                      // n * (levels + 1)
    std::vector<std::vector<uint16_t>> weight;   // n * (levels + 1), box-local; level 0 empty
    std::vector<std::vector<uint16_t>> wsum;     // levels + 1 canvas planes (pitch lw); level 0 empty
    std::vector<std::vector<DownJob>> down;      // [j]: level j -> j+1
    std::vector<std::vector<TileJob>> blend;     // [j], j >= 1
    std::vector<int> camlist;                    // plane indices referenced by blend jobs
    std::vector<TileJob> fin;                    // level-0 canvas tiles
    std::vector<int> fin_owners;                 // owner cameras referenced by fin jobs (8-bit path)
    std::vector<RemapJob> remap;                 // level-0 box tiles (16-bit render path)
    int max_tile_cams = 0;
    int kind = kPixU8;                           // tile geometry of the work lists
    bool fused0 = false;                         // down[0] jobs carry camera indices (8-bit render)
    int shape = 0;                               // down tile shape (8-bit path), kDownShapes
    int dtw = kFTW, dth = kFTH;                  // its output tile
};

// cv::distanceTransform(mask, dst, CV_DIST_L1, 3) semantics (OpenCV 3.0
// distanceTransform_3x3 with metrics {1, 2}); dst as float distances.
void distance_transform_l1(const uint8_t* mask, size_t step, int w, int h, float* dst);

// getNonIntersectingMasks (Blend/Mask.cpp:376-397).  masks: n host planes.
void nonintersecting_masks(int n, int rows, int cols, const uint8_t* const* masks,
                           const size_t* steps, uint8_t* out);

int true_num_levels(int width, int height, int max_levels, int min_length);

// Boxes of one camera per level 0..L: colf / rowf flag the level-0 columns /
// rows its support meets; each level dilates them by the 5-tap footprint
// (WRAP columns, REFLECT_101 rows) and takes the smallest cyclic column
// interval and the row span.  lw / lh: level sizes.
void camera_boxes(int L, const int* lw, const int* lh, std::vector<uint8_t> colf, std::vector<uint8_t> rowf,
                  Geom* geom);

// Work lists of a prepared rig from its geometry (n, rows, cols, levels, lw,
// lh, geom, kind, fused0 set):
//   tile_has_valid(i, j, X, Y, nx, ny): camera i's level-j alpha is non-zero
//     somewhere in the canvas tile (down jobs of other tiles are dropped);
//   cam_hits(j, i, hit): sets hit[tile] for the level-j blend tiles where camera
//     i's weight is non-zero (row-major tile grid of the level canvas);
//   tile_owners(t, owners): appends the owners of level-0 tile t (8-bit path).
using TileValidFn = std::function<bool(int, int, int, int, int, int)>;
using CamHitsFn = std::function<void(int, int, std::vector<uint8_t>&)>;
using TileOwnersFn = std::function<void(const TileJob&, std::vector<int>&)>;
void work_lists(Prepared& out, const TileValidFn& tile_has_valid, const CamHitsFn& cam_hits,
                const TileOwnersFn& tile_owners);

// Full prepare.  support0 (optional, n planes rows x cols): pixels whose level-0
// value may be non-zero besides the mask (e.g. where a map is valid); with
// full_boxes every plane spans the whole canvas (caller images of any content).
// Returns PANO_OK or a negative status.
// kind selects the work-list geometry (kPixU8: fused 8-bit kernels, quad
// tiles; kPixS16: generic kernels); fused0: the 8-bit render path, whose
// level 0 -> 1 jobs address cameras instead of planes.
// unique (optional): the seams already computed (n planes rows x cols).
// shape: the down tile shape of the 8-bit path (kDownShapes).
int prepare_rig(int n, int rows, int cols, const uint8_t* const* masks, const size_t* steps,
                const uint8_t* const* support0, int max_levels, int min_length, bool full_boxes,
                int kind, bool fused0, Prepared& out, const uint8_t* unique = nullptr, int shape = 0);

}  // namespace pano
