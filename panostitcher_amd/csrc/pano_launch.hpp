// pano_launch.hpp -- host-side launchers of the gfx950 kernels (kernels.hip).
// Each returns hipSuccess or the launch error; no launch allocates memory.
#pragma once
#include <vector>

#include "pano_dev.hpp"
#include "pano_host.hpp"

namespace pano {

hipError_t launch_remap(int kind, const RemapJob* jobs, int njobs, const RemapCam* cams,
                        const Plane* planes, const uint8_t* luts, bool use_lut,
                        const SrcTable& src, hipStream_t s);

// user_kind < 0: level j planes are internal; else level 0 is caller images of
// that kind (then kind must be kPixS16).
hipError_t launch_down(int kind, int user_kind, const DownJob* jobs, int njobs,
                       const Plane* planes, const SrcTable& src, hipStream_t s);

hipError_t launch_blend(int kind, const TileJob* jobs, int njobs, const int* camlist,
                        const Plane* planes, const CanvasLevel* canvas, const RenderDesc* rd,
                        int level, hipStream_t s);

hipError_t launch_final(int kind, int user_kind, const TileJob* jobs, int njobs,
                        const Plane* planes, const CanvasLevel* canvas, const RenderDesc* rd,
                        const SrcTable& src, const OutTable& outs, hipStream_t s);

hipError_t launch_import(int user_kind, const void* img, size_t step, const uint8_t* mask,
                         int mask_pitch, void* dst, int pitch, int w, int h, hipStream_t s);

hipError_t launch_reproject16s(const uint8_t* src, size_t src_step, int sw, int sh,
                               const float* xmap, const float* ymap, size_t map_step,
                               int16_t* dst, size_t dst_step, int dw, int dh, hipStream_t s);

hipError_t launch_reproject8u(const uint8_t* src, size_t src_step, int sw, int sh, const float* xmap,
                              const float* ymap, size_t map_step, uint8_t* dst, size_t dst_step, int dw,
                              int dh, hipStream_t s);

hipError_t launch_transform(const uint8_t* src, size_t src_step, uint8_t* dst, size_t dst_step,
                            int w, int h, const uint8_t* lut_dev, hipStream_t s);

// Fused 8-bit path (kernels_u8.hip): one launch of a stage over the segments
// of a batch (blocks = sum of the segments' job counts).
enum Stage : int { kStageRemapDown0 = 0, kStageDown = 1, kStageBlend = 2, kStageFinal = 3 };
// sched: block -> job order (block_job): 0 in order, 1 contiguous per XCD;
// shape: the down tile shape the jobs of the down stages were laid out for.
// luts: the remap stage applies each render's gain LUTs where its "LUTs active"
// byte is set (the frames in src are the unmapped ones).
hipError_t launch_seg(int stage, const Seg* segs, int nseg, int blocks, int level,
                      const SrcTable& src, const OutTable& outs, int sched, int shape, bool luts,
                      hipStream_t s);

// Per-frame gain estimation (kernels_gain.hip): statistics over the fused
// level-0 jobs of segs (the remap launch's table), then the solve writing the
// render's gains and LUTs (kGainTint: also the tint ratios).  The statistics
// buffer must be zero.
hipError_t launch_gain(const Seg* segs, int nseg, int blocks, const RenderDesc* rd, int flags,
                       const SrcTable& src, hipStream_t s);

// Zero-fill with a kernel (stays correct inside captured graphs).
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);

// Float-weight blender (kernels_f32.hip).  Planes are pitched float4 images
// (float for weights, alphas and sums); the 8-bit level-0 image is read as
// BGRx bytes when u8 is set.
hipError_t launch_f32_down(int channels, bool u8, const void* src, int sw, int sh, size_t sstep,
                           void* dst, int dw, int dh, size_t dstep, const float* scale, size_t scstep,
                           hipStream_t s);
hipError_t launch_f32_lap(bool u8, const void* g, size_t gstep, const void* g1, size_t g1step, int w1,
                          int h1, const float* wt, size_t wstep, void* r, size_t rstep, int w, int h,
                          hipStream_t s);
hipError_t launch_f32_collapse(const void* c1, size_t c1step, int w1, int h1, const float* winv,
                               size_t wistep, void* r, size_t rstep, int w, int h, hipStream_t s);
hipError_t launch_f32_out(const void* c0, size_t cstep, const uint8_t* cover, int cpitch, uint8_t* out,
                          size_t ostep, int w, int h, hipStream_t s);
hipError_t launch_f32_from_mask(const uint8_t* m, int mpitch, float* f, size_t fstep, int w, int h,
                                hipStream_t s);
hipError_t launch_f32_alpha(float* a, size_t astep, float* bin, size_t bstep, int w, int h, hipStream_t s);
hipError_t launch_f32_wsum(const float* wsrc, size_t wstep, float* sum, size_t sstep, int w, int h,
                           bool invert, hipStream_t s);

// Custom-mask blend of the 16-bit blender (kernels_custom.hip).  Weights,
// sums and cover share the row pitch `pitch` (elements) at every level.
hipError_t launch_custom_weights(const SrcTable& masks, int n, int L, const int* lw, const int* lh,
                                 int16_t* const* wts, const size_t* cam_stride, int pitch,
                                 int32_t* const* wsum, uint8_t* cover, hipStream_t s);
hipError_t launch_blend_custom(int user_kind, const Plane* planes, const CanvasLevel* canvas, int n, int L,
                               int level, const int16_t* wts, size_t cam_stride, int pitch,
                               const int32_t* wsum, const SrcTable& src, const uint8_t* cover,
                               uint8_t* out, size_t ostep, int w, int h, hipStream_t s);

// Encoder / decoder hand-off (kernels_yuv.hip); w, h even; nv12: u is the
// interleaved UV plane and v unused.
hipError_t launch_bgr32_to_yuv(bool nv12, const uint8_t* src, size_t sstep, int w, int h, uint8_t* y,
                               size_t ystep, uint8_t* u, size_t ustep, uint8_t* v, size_t vstep, hipStream_t s);
hipError_t launch_yuv_to_bgr32(bool nv12, const uint8_t* y, size_t ystep, const uint8_t* u, size_t ustep,
                               const uint8_t* v, size_t vstep, int w, int h, uint8_t* dst, size_t dstep,
                               hipStream_t s);

// Linear-blend mode (kernels_linear.hip).  Blur: src / dst / rows share the
// row pitch (elements); taps K[k0 .. k1) are the non-zero ones of n.
hipError_t launch_blur_u8(const uint8_t* src, int pitch, int w, int h, const int* K, int n, int k0, int k1,
                          int32_t* rows, const uint8_t* mask, int mpitch, uint8_t* dst, hipStream_t s);
hipError_t launch_linear_weights(bool as_int, const uint8_t* dist, int pitch, size_t cam_stride, int n, int w,
                                 int h, void* out, hipStream_t s);
hipError_t launch_interleave_map(const float* xm, const float* ym, int sp, float2* out, int dp, int w, int h,
                                 hipStream_t s);
hipError_t launch_lin_tile_hits(const float* wts, int pitch, size_t plane, int n, int w, int h, int tw, uint8_t* hit,
                                hipStream_t s);
hipError_t launch_linear_render(const LinTile* tiles, int ntiles, const int* camlist, const LinCam* cams,
                                const SrcTable& src, uint8_t* out, size_t ostep, int w, int h, hipStream_t s);
hipError_t launch_reproject_accum(const uint8_t* src, size_t sstep, int sw, int sh, const float* xmap,
                                  const float* ymap, size_t map_step, const float* wt, size_t wstep, float* acc,
                                  size_t astep, int w, int h, hipStream_t s);
hipError_t launch_linear_blend32s(const SrcTable& images, int n, const int32_t* wts, int pitch, size_t cam_stride,
                                  uint8_t* out, size_t ostep, int w, int h, hipStream_t s);

// getNonIntersectingMasks on device planes (pitch cols), masks replaced by the
// unique masks; synchronous.
hipError_t device_seams_inplace(int n, int rows, int cols, uint8_t* M);
// The same with host masks in and host unique masks out (n planes rows x cols).
hipError_t device_nonintersecting_masks(int n, int rows, int cols, const uint8_t* const* masks,
                                        const size_t* steps, uint8_t* out);

// ---- render prepare on the device (prepare_dev.hip) ----------------------
// Device results of prepare_tables_device, read by the rig build: weights of
// every (camera, level >= 1) plane, box-local with the plane pitch, at element
// woff[i (L + 1) + j] of w; per level the weight sums (pitch lw[j], only when
// the canvas is not fully covered); owner per canvas pixel; mask bits per
// canvas pixel (n <= kGainMaxCams, when asked for).
struct DevTables {
    uint16_t* w = nullptr;
    std::vector<size_t> woff;
    std::vector<uint16_t*> wsum;
    uint8_t* label = nullptr;
    uint8_t* cover = nullptr;
    DevTables() = default;
    DevTables(const DevTables&) = delete;
    DevTables& operator=(const DevTables&) = delete;
    ~DevTables();
};
// The prepare of the fused 8-bit render (boxes, work lists, weight pyramids,
// weight sums, owners) from device masks and unique masks (n planes, pitch
// cols), n <= 32; synchronous.  Fills out as prepare_rig(..., kPixU8, fused0)
// does, except out.unique / label / weight / wsum, which stay on the device.
int prepare_tables_device(int n, int rows, int cols, const uint8_t* masks, const uint8_t* unique, int max_levels,
                          bool with_cover, int shape, Prepared& out, DevTables& dev);
// Margined fused map storage of camera cam (box g at level 0) from device maps
// and masks (pitch cols) and the owner plane.
hipError_t launch_pack_map(const float* xm, const float* ym, const uint8_t* mask, const uint8_t* label, int cols,
                           int cam, const Geom& g, int margin, int mp, int sw, int sh, float2* out, hipStream_t s);
// Whether every valid map point (inside the source) lies inside the mask.
int maps_consistent_device(const float* xm, const float* ym, const uint8_t* mask, size_t P, int sw, int sh,
                           bool& consistent);

hipError_t launch_copy_plane(const void* src, size_t src_step, void* dst, size_t dst_step,
                             int row_bytes, int rows, hipStream_t s);

}  // namespace pano
