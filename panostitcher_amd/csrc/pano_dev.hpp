// pano_dev.hpp -- descriptors shared by the host code and the gfx950 kernels.
//
// Layout in HBM (per prepared rig):
//   * per camera i and pyramid level j a Plane: the bounding box of the
//     camera's support on the level-j canvas (wrapping horizontally), holding
//     Gaussian pixels G_ij as 4 x u8 (B, G, R, valid) or 4 x s16, and, for
//     j >= 1, the blend weight W_ij (u16, 0..256) with the same geometry;
//   * per level a CanvasLevel: the collapsed result C_j (4 x s32) and, when
//     the masks do not cover the canvas, sum_i W_ij (u16);
//   * a level-0 owner label per canvas pixel (u8, 255 = uncovered): at level 0
//     the reference's weights are exactly 256 on each camera's seam region, so
//     the blended Laplacian there is the owner camera's own;
//   * work lists of tiles for each stage.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace pano {

constexpr int kMaxSrc = 128;    // source images addressable by one launch
constexpr int kMaxLevels = 16;  // getTrueNumLevels caps at 16 (MultiBandBlend.cpp:325)
constexpr uint8_t kNoOwner = 255;

// Tile shapes (pixels).  Down: output tile at level j+1; blend/final: canvas tile.
constexpr int kDownTW = 64, kDownTH = 8;
constexpr int kTileW = 64, kTileH = 4;
constexpr int kRemapTW = 64, kRemapTH = 4;

// 8-bit fused path: output tile (level j+1) of the down kernels, and quad
// tiles of the blend kernels (kQTW x kQTH quads = 2 kQTW x 2 kQTH pixels,
// one 2x2 quad per thread).  The down tile is one of kDownShapes shapes,
// chosen per rig at prepare (shape 0 by default, or by the tile autotuner):
// kDownTileW[s] x kDownTileH[s] outputs, kDownThreads[s] threads per block.
constexpr int kDownShapes = 5;
constexpr int kDownTileW[kDownShapes] = {64, 64, 128, 32, 64};
constexpr int kDownTileH[kDownShapes] = {16, 32, 16, 32, 8};
constexpr int kDownThreads[kDownShapes] = {512, 1024, 1024, 512, 256};
constexpr int kFTW = kDownTileW[0], kFTH = kDownTileH[0];
constexpr int kQTW = 64, kQTH = 4;
constexpr int kMaxStage = 8;    // owner cameras staged in LDS at once (level 0)

enum PixKind : int { kPixU8 = 0, kPixS16 = 1 };

// Per-pixel flags of a camera's level-0 box (fused path).
enum RemapFlags : uint8_t { kFlagValid = 1, kFlagOwned = 2 };

struct Plane {
    void* g;          // Gaussian pixels (uchar4 or short4), box-local, row pitch in pixels
    uint16_t* w;      // blend weights (levels >= 1) or nullptr; same geometry
    const uint8_t* mask;  // level 0 only, when the pixels come from a caller image
    int pitch;        // pixels per row of g and w
    int mask_pitch;   // bytes per row of mask
    int x0, y0;       // box origin on the level canvas
    int bw, bh;       // box size; bw == cw means the full width (wraps)
    int cw, ch;       // level canvas size
    int src;          // >= 0: level-0 pixels are caller image src (SrcTable index)
    int pad;
};

struct CanvasLevel {
    int4* c;          // collapsed result C_j, row pitch in pixels (16-bit planes)
    uint2* c16;       // C_j as 4 x int16 (8-bit planes: |C_j| <= 256 (L - j + 1))
    uint16_t* wsum;   // sum of the cameras' weights, or nullptr when the masks cover the canvas
    int pitch;
    int cw, ch;
    int pad;
};

struct RemapCam;

// A camera's plane as one tile of a stage reads it (the owner cameras of a
// level-0 tile, the cameras of a blend tile), built at prepare so that each
// is one independent descriptor load.
struct PlaneRef {
    const void* g;
    const uint16_t* w;   // weights (blend tiles' level-j planes)
    int x0, y0, bw, bh, cw, pitch;
    int cam;             // camera index (the owner id of level-0 pixels)
    int pad;
};

// Per render (rig), in device memory.  The generic 16-bit kernels index the
// launch's arrays with plane0 / canvas0 / out; the 8-bit kernels reach every
// array of their render through the pointers.
struct RenderDesc {
    const uint8_t* label;        // level-0 owner per canvas pixel
    const Plane* planes;         // n * (levels + 1)
    const CanvasLevel* canvas;   // levels + 1
    const int* camlist;          // blend cameras (plane indices)
    const int* owners;           // level-0 owner cameras
    const PlaneRef* fin_refs;    // level-1 planes of the owners of each level-0 tile
    const PlaneRef* blend_refs;  // per camlist entry: the camera's level-j plane
    const PlaneRef* blend_refs1; // per camlist entry: its level-(j+1) plane (j < L)
    const RemapCam* cams;        // remap inputs
    const uint8_t* luts;         // n x 768 LUT bytes, then 1 byte: LUTs active this frame
    uint32_t* g0comp;            // owners' level-0 samples; byte 3 = owner + 1, 0 = no owner
    const uint8_t* cover;        // gain estimation: bit j = mask j covers the canvas pixel
    unsigned long long* gstats;  // gain estimation sums [kGainSlots][n][n][5] (count, gray, B, G, R)
    double* gains;               // estimated gains: n exposure, n red, n blue, anchor
    int label_pitch;
    int g0pitch;
    int full;                    // masks cover the whole canvas
    int levels;
    int plane0;                  // (generic kernels) index of (camera 0, level 0)
    int canvas0;                 // (generic kernels) index of level 0
    int out;                     // (generic kernels) OutTable index
    int ncams;
};

// One render's share of a batched launch: blocks [begin, begin + count) run its
// jobs; its cameras' frames start at SrcTable index src_base; its canvas is
// OutTable entry out.
struct Seg {
    const RenderDesc* rd;
    const void* jobs;
    int begin, count, src_base, out;
};

constexpr int kMaxBatch = 32;   // renders per batched launch
constexpr int kSchedModes = 2;  // block -> job orders of the batched kernels (block_job)

// Per camera of the render path (resident).  map/mask cover the camera's
// level-0 box.  For the 8-bit fused path they carry a kMapMargin frame of
// border-resolved copies (WRAP columns, REFLECT_101 rows, invalid outside the
// box) so that a tile's staging region is one contiguous rectangle; for the
// 16-bit path they are box-local with the level-0 plane pitch.
//
// Fused storage packs the flags into the map: a point outside the source
// image is stored as +inf, and the sign bit of x is kFlagOwned, the sign bit
// of y kFlagValid (the map's own values are then |x|, |y|, never negative).
constexpr int kMapMargin = 5;
struct RemapCam {
    const float2* map;     // (x, y) map points
    const uint8_t* mask;   // flags (kFlagValid | kFlagOwned); null for fused storage
    int sw, sh;            // source image size
    int src;               // SrcTable index of its frame
    int lut;               // byte offset of its 3 x 256 LUT in the LUT buffer, -1 = none
    int plane;             // its level-0 plane
    int mpitch;            // elements per row of map and mask
};

struct RemapJob { int cam, lx, ly, nx, ny, pad0, pad1, pad2; };   // box-local tile, level 0
// Canvas tile at level j+1.  Fused level 0: src = camera, and (mx, my) is the
// position of the tile's staging region in the camera's margined map storage.
struct DownJob { int src, dst, X, Y, nx, ny, mx, my; };
struct TileJob { int X, Y, nx, ny, r, off, cnt, flags; };         // canvas tile at level j

enum TileFlags : int { kTileTop = 1 };

constexpr int kGainMaxCams = 8;   // cameras of a rig with per-frame gain estimation
constexpr int kGainSlots = 16;    // copies of the gain sums the statistics blocks spread their atomics over
enum GainFlags : int { kGainTint = 1 };

struct SrcTable {
    const uint8_t* p[kMaxSrc];
    uint32_t step[kMaxSrc];
};

// Linear-blend render: a camera's full-canvas map and weight planes (pitch in
// elements) and its frame (SrcTable index, size); a canvas tile of 64 x kLinTH
// pixels with its cameras camlist[off .. off + cnt).
struct LinCam {
    const float2* map;
    const float* w;
    int pitch, src, sw, sh;
};
struct LinTile { int X, Y, off, cnt; };
constexpr int kLinRows = 4;                       // canvas rows per thread (4 apart)
constexpr int kLinTW = 64, kLinTH = 4 * kLinRows;  // a 256-thread block

constexpr int kMaxOut = 32;     // output canvases addressable by one launch
// Output formats of the level-0 kernel: the BGRx canvas, or the encoder's
// planes (cvtBGR32ToNV12 / cvtBGR32ToYUV420P of that canvas) written directly.
enum OutFormat : int { kOutBGRX = 0, kOutNV12 = 1, kOutYUV420P = 2 };
struct OutTable {
    uint8_t* p[kMaxOut];    // BGRx canvas, or the Y plane
    uint32_t step[kMaxOut];
    uint8_t* p2[kMaxOut];   // NV12: interleaved UV; YUV420P: U
    uint8_t* p3[kMaxOut];   // YUV420P: V
    uint32_t step2[kMaxOut], step3[kMaxOut];
    int fmt[kMaxOut];       // OutFormat (zero-initialised tables are BGRx)
};

}  // namespace pano
