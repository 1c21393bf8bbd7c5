/*
 * panohip.h -- C-ABI of the MI355X (gfx950) per-frame panorama stitch path.
 *
 * A drop-in for PanoStitcher's CudaAccel layer on the path
 *     per-camera gain LUTs -> bilinear remap through float32 maps ->
 *     integer Gaussian/Laplacian multiband blend -> 8-bit BGRx canvas.
 * Results follow the reference's CPU path (CPUPanoramaRender), which is the
 * parity target: every step is integer and bit-exact except the FP64
 * bilinear sample, which is evaluated in the same order without contraction.
 *
 * Conventions
 *   - every entry point returns an int status: PANO_OK (0) or a negative code;
 *     pano_status_string() names it.  Nothing throws across this boundary.
 *   - device images are pitched: a device pointer plus a row step in bytes.
 *     Colour images are 4-channel BGRx (CV_8UC4 / CV_16SC4), as in the
 *     reference's GPU API; masks are single-channel bytes (0 / non-zero).
 *   - "host" arguments are read during the call only; "device" arguments must
 *     stay valid until the work queued on the given stream has finished.
 *   - stream: any hipStream_t (0 = the null stream).  Work of one object is
 *     queued in order on the stream passed to each call; one object must not
 *     be used from two threads at once (the reference classes are not
 *     thread-safe either, CudaMultibandBlend.cpp:785-823 keeps per-call
 *     scratch in members).
 */
#ifndef PANOHIP_H
#define PANOHIP_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PANO_OK 0
#define PANO_ERR_ARG (-1)         /* bad argument: sizes, types, null pointers */
#define PANO_ERR_STATE (-2)       /* object not prepared / prepare failed       */
#define PANO_ERR_HIP (-3)         /* a HIP runtime call failed                  */
#define PANO_ERR_NOMEM (-4)       /* device or host allocation failed           */
#define PANO_ERR_UNSUPPORTED (-5) /* valid request this build does not handle   */
#define PANO_ERR_BUDGET (-6)      /* admission refused: device memory budget    */
#define PANO_ERR_IO (-7)          /* a file could not be opened or written      */
#define PANO_ERR_FORMAT (-8)      /* a rig file field is not a number           */

/* ---- image type codes (OpenCV CV_MAKETYPE values) ---------------------- */
#define PANO_8UC1 0
#define PANO_8UC3 16
#define PANO_8UC4 24
#define PANO_16SC4 27
#define PANO_32FC4 29

/* ---- output of a render: the BGRx canvas or the encoder's planes --------
 * PANO_OUT_NV12 / PANO_OUT_YUV420P write what cvtBGR32ToNV12 /
 * cvtBGR32ToYUV420P (CudaAccel/CudaInterface.h:161-166) would make of the
 * BGRx canvas, straight from the level-0 kernel.  plane / step: BGRX {canvas};
 * NV12 {Y, interleaved UV}; YUV420P {Y, U, V}.  Y and UV planes need even
 * steps and 2-byte aligned pointers. */
#define PANO_OUT_BGRX 0
#define PANO_OUT_NV12 1
#define PANO_OUT_YUV420P 2
typedef struct pano_output {
    int format;
    uint8_t* plane[3];
    size_t step[3];
} pano_output;

const char* pano_status_string(int status);
int pano_version(void);

/* ---- camera model: PhotoParam (Warp/ZReproject.h:17-50), same fields ---- */
typedef struct pano_photo_param {
    int imageType; /* 0 rectilinear, 1 full-frame fisheye, 2 drum, 3 circular */
    int cropMode, cropX, cropY, cropWidth, cropHeight;
    int circleX, circleY, circleR;
    double hfov, vfov, alpha, beta, gamma, shiftX, shiftY, shearX, shearY;
    double yaw, pitch, roll;
} pano_photo_param;

/* ---- rig files: PhotoParam XML / PTGui .pts (Warp/PhotoParam.cpp) --------
 * Loaders write up to cap params to out and return how many the file holds
 * (0 when the file is missing, is not one of the layouts or holds no camera:
 * the reference's empty result / false), or a negative status.
 *
 * loadPhotoParams (PhotoParam.cpp:756-770): extension "pts" -> PTS, any other
 * -> XML; no extension -> PANO_ERR_ARG. */
int pano_load_photo_params(const char* path, pano_photo_param* out, int cap);
/* loadPhotoParamFromXML (PhotoParam.cpp:506-511): the per-camera layout
 * (:273-504), else the shared-lens layout (:17-271).  An element whose text is
 * not a number -> PANO_ERR_FORMAT (the reference throws there). */
int pano_load_photo_params_xml(const char* path, pano_photo_param* out, int cap);
/* loadPhotoParamFromPTS (PhotoParam.cpp:681-754). */
int pano_load_photo_params_pts(const char* path, pano_photo_param* out, int cap);
/* exportPhotoParamToXML (PhotoParam.cpp:772-925): the per-camera layout. */
int pano_export_photo_params_xml(const char* path, const pano_photo_param* params, int num);
/* rotateCameras (PhotoParam.cpp:931-967): each camera's pose pre-multiplied by
 * the rotation (yaw, pitch, roll), given in radians. */
int pano_rotate_cameras(pano_photo_param* params, int num, double yaw, double pitch, double roll);
/* rotatePhotoParamInXML (PhotoParam.cpp:969-975). */
int pano_rotate_photo_params_xml(const char* src, const char* dst, double yaw, double pitch, double roll);

/* ---- multiband blender: CudaTilingMultibandBlendFast -------------------
 * Replaces the class in CudaAccel/CudaInterface.h:45-82 (implementation
 * CudaAccel/CudaMultibandBlend.cpp:549-823), with the arithmetic of the CPU
 * TilingMultibandBlendFast (Blend/MultiBandBlend.cpp:906-1031). */
typedef struct pano_blender pano_blender;

int pano_blender_create(pano_blender** out);
void pano_blender_destroy(pano_blender* b);

/* prepare(const vector<Mat>& masks, int maxLevels, int minLength)
 * (CudaInterface.h:49).  masks: host, num planes of rows x cols bytes with
 * the given row steps.  Fails (PANO_ERR_ARG) on empty or mismatched masks or
 * more than 255 masks, like CudaMultibandBlend.cpp:551-565. */
int pano_blender_prepare(pano_blender* b, int num, int rows, int cols,
                         const uint8_t* const* masks, const size_t* mask_steps, int max_levels,
                         int min_length);

/* prepare(const vector<GpuMat>& masks, ...) (CudaInterface.h:50): the same with
 * device masks; the masks are read back and prepared on the host. */
int pano_blender_prepare_device(pano_blender* b, int num, int rows, int cols,
                                const uint8_t* const* masks, const size_t* mask_steps,
                                int max_levels, int min_length);

/* blend(const vector<GpuMat>& images, GpuMat& blendImage) (CudaInterface.h:54).
 * images: device, num images of rows x cols, type PANO_16SC4 or PANO_8UC4.
 * out: device rows x cols PANO_8UC4 (channel 3 written as 0). */
int pano_blender_blend(pano_blender* b, int num, const void* const* images, const size_t* steps,
                       int type, uint8_t* out, size_t out_step, hipStream_t stream);

/* blend(const vector<GpuMat>& images, const vector<GpuMat>& masks, GpuMat&
 * blendImage) (CudaInterface.h:55-56, CudaMultibandBlend.cpp:825-893): the
 * weights come from this call's masks (device, num planes of rows x cols
 * bytes), always divided by their sums, with the integer arithmetic of
 * Blend/MultiBandBlend.cpp:1033-1102; the image pyramids keep the prepared
 * alphas.  The output is zero where none of these masks is set.  Device
 * buffers for the weights are allocated on the first call and kept. */
int pano_blender_blend_custom(pano_blender* b, int num, const void* const* images, const size_t* steps,
                              int type, const uint8_t* const* masks, const size_t* mask_steps, uint8_t* out,
                              size_t out_step, hipStream_t stream);

/* begin() / tile(image, index) / end(blendImage) (CudaInterface.h:51-53).
 * tile() consumes the image before it returns on the stream: the caller may
 * reuse the buffer for the next camera once the stream reaches that point. */
int pano_blender_begin(pano_blender* b);
int pano_blender_tile(pano_blender* b, const void* image, size_t step, int type, int index,
                      hipStream_t stream);
int pano_blender_end(pano_blender* b, uint8_t* out, size_t out_step, hipStream_t stream);

/* getUniqueMasks (CudaInterface.h:57): host output, num planes of rows x cols. */
int pano_blender_unique_masks(const pano_blender* b, uint8_t* const* out, const size_t* steps);
int pano_blender_num_levels(const pano_blender* b);
/* calcUsedMemorySize (CudaInterface.h:58) and estimateMemorySize (:59). */
int64_t pano_blender_mem_bytes(const pano_blender* b);
int64_t pano_estimate_mem_bytes(int width, int height, int num);

/* ---- float-weight multiband blender: CudaTilingMultibandBlendFast32F -----
 * Replaces the class in CudaAccel/CudaInterface.h:84-101 (implementation
 * CudaAccel/CudaMultibandBlend.cpp:1041-1179, kernels CudaAccel/CudaPyramid6.cu),
 * with its float arithmetic: float weight and inverse-alpha pyramids, float
 * Laplacian pyramids, round-to-nearest-even 8-bit output. */
typedef struct pano_blender32f pano_blender32f;

int pano_blender32f_create(pano_blender32f** out);
void pano_blender32f_destroy(pano_blender32f* b);

/* prepare(const vector<Mat>& masks, int maxLevels, int minLength)
 * (CudaInterface.h:88): host masks, num planes of rows x cols bytes.  The
 * pyramids are built on the device; the call returns when they are done. */
int pano_blender32f_prepare(pano_blender32f* b, int num, int rows, int cols,
                            const uint8_t* const* masks, const size_t* mask_steps, int max_levels,
                            int min_length);

/* blend(const vector<GpuMat>& images, GpuMat& blendImage) (CudaInterface.h:89).
 * images: device, num images of rows x cols, type PANO_8UC4 or PANO_32FC4.
 * out: device rows x cols PANO_8UC4 (all four channels blended). */
int pano_blender32f_blend(pano_blender32f* b, int num, const void* const* images, const size_t* steps,
                          int type, uint8_t* out, size_t out_step, hipStream_t stream);
int pano_blender32f_num_levels(const pano_blender32f* b);

/* ---- linear-blend mode (highQualityBlend = 0) ----------------------------
 * Weights: getWeightsLinearBlend32F (as_int 0: float, calcWeights32F) or
 * getWeightsLinearBlend (as_int 1: int32 in 2^16 units, calcWeights),
 * Blend/LinearBlend.cpp:17-148: seams, a (2 radius + 1)^2 Gaussian blur of
 * each seam mask (sigma radius / 3, OpenCV 3.0's 8-bit fixed-point path),
 * ANDed with the mask, normalised per pixel.  masks: host; out: device planes
 * of rows x cols elements (4 bytes each).  Synchronous. */
int pano_linear_weights(int num, int rows, int cols, const uint8_t* const* masks, const size_t* mask_steps,
                        int radius, int as_int, void* const* out, const size_t* out_steps);

/* CudaTilingLinearBlend (CudaAccel/CudaInterface.h:103-115,
 * Blend/CudaLinearBlend.cpp:15-101): prepare(masks, radius) (false on empty or
 * mismatched masks or radius < 1), blend(images, blendImage): images device
 * PANO_8UC4, out device PANO_8UC4 (all four channels: sum of weight x pixel,
 * >> 16, saturated). */
typedef struct pano_linear_blender pano_linear_blender;
int pano_linear_blender_create(pano_linear_blender** out);
void pano_linear_blender_destroy(pano_linear_blender* b);
int pano_linear_blender_prepare(pano_linear_blender* b, int num, int rows, int cols, const uint8_t* const* masks,
                                const size_t* mask_steps, int radius);
int pano_linear_blender_blend(pano_linear_blender* b, int num, const uint8_t* const* images, const size_t* steps,
                              uint8_t* out, size_t out_step, hipStream_t stream);

/* The render's linear branch (prepare with highQualityBlend = 0, blendParam =
 * radius; render: Task/RicohUtil.cpp:1269-1287, CPU twin :1543-1558): per
 * frame, for each camera in order, the bilinear sample times its float weight
 * is added in float, then the sum is rounded to 8 bits.  Maps and masks as
 * for pano_render_prepare_maps (host, dst_w x dst_h); luts as for
 * pano_render_render.  dst: device BGRx canvas (channel 3 written as 0). */
typedef struct pano_linear_render pano_linear_render;
int pano_linear_render_create(pano_linear_render** out);
void pano_linear_render_destroy(pano_linear_render* r);
int pano_linear_render_prepare_maps(pano_linear_render* r, int num, const int* src_w, const int* src_h, int dst_w,
                                    int dst_h, const float* const* xmaps, const float* const* ymaps,
                                    const uint8_t* const* masks, int radius);
/* The same from camera parameters: maps and masks generated on the device
 * (pano_generate_map_device), then prepared there (CudaPanoramaRender2::prepare
 * with highQualityBlend = 0, RicohUtil.cpp:1175-1180). */
int pano_linear_render_prepare(pano_linear_render* r, int num, const pano_photo_param* params, const int* src_w,
                               const int* src_h, int dst_w, int dst_h, int radius);
int pano_linear_render_render(pano_linear_render* r, int num, const uint8_t* const* srcs, const size_t* src_steps,
                              const uint8_t* luts, uint8_t* dst, size_t dst_step, hipStream_t stream);
/* The prepared float weights of camera cam (host dst_w x dst_h floats). */
int pano_linear_render_weights(const pano_linear_render* r, int cam, float* out);

/* ---- free functions ----------------------------------------------------- */

/* cudaReprojectTo16S(src, dst, xmap, ymap, stream) (CudaInterface.h:150-151),
 * with the CPU path's bilinear arithmetic (Warp/ZReproject.cpp:101-130) rather
 * than the reference GPU's bicubic texture fetch.  src: device PANO_8UC4;
 * xmap/ymap: device float32 planes of dst_w x dst_h (invalid = -1); dst:
 * device PANO_16SC4 dst_w x dst_h, zero where the map is invalid. */
int pano_reproject_to16s(const uint8_t* src, size_t src_step, int src_w, int src_h,
                         const float* xmap, const float* ymap, size_t map_step, int16_t* dst,
                         size_t dst_step, int dst_w, int dst_h, hipStream_t stream);

/* getNonIntersectingMasks (Blend/Mask.cpp:376-397, over cv::distanceTransform
 * L1 3x3) computed on the device: host masks in, host unique masks (0 / 255,
 * camera 0 keeps its mask values) out.  Synchronous.  The blender and render
 * prepares use the same device path. */
int pano_nonintersecting_masks(int num, int rows, int cols, const uint8_t* const* masks, const size_t* mask_steps,
                               uint8_t* const* out, const size_t* out_steps);

/* cudaReproject(src, dst, xmap, ymap, stream) (CudaInterface.h:147-148), with
 * the same bilinear arithmetic as pano_reproject_to16s (the reference GPU
 * samples bicubic, CudaReproject.cu:714-741): dst is device
 * PANO_8UC4 dst_w x dst_h (channel 3 zero), zero where the map is invalid. */
int pano_reproject(const uint8_t* src, size_t src_step, int src_w, int src_h, const float* xmap,
                   const float* ymap, size_t map_step, uint8_t* dst, size_t dst_step, int dst_w, int dst_h,
                   hipStream_t stream);

/* cudaReprojectWeightedAccumulateTo32F(src, dst, xmap, ymap, weight, stream)
 * (CudaInterface.h:153-155) with the CPU path's bilinear samples and float
 * arithmetic (Warp/ZReproject.cpp:606-631): where the map is valid, accum
 * (device float BGRx, dst_w x dst_h) += sample x weight per channel and
 * channel 3 is set to 0.  weight: device float plane. */
int pano_reproject_weighted_accumulate(const uint8_t* src, size_t src_step, int src_w, int src_h, const float* xmap,
                                       const float* ymap, size_t map_step, const float* weight, size_t weight_step,
                                       float* accum, size_t accum_step, int dst_w, int dst_h, hipStream_t stream);

/* cudaReproject / cudaReprojectTo16S(src, dst, [mask,] dstSize, param, stream)
 * (CudaInterface.h:135-145): the map generated on the device for this call,
 * then the bilinear reprojection (type PANO_8UC4 or PANO_16SC4); mask (device,
 * optional) receives the validity mask. */
int pano_reproject_params(const uint8_t* src, size_t src_step, int src_w, int src_h, const pano_photo_param* param,
                          int dst_w, int dst_h, int type, void* dst, size_t dst_step, uint8_t* mask,
                          size_t mask_step, hipStream_t stream);

/* cudaTransform(src, dst, luts, stream) (CudaInterface.h:120-121):
 * per-channel LUTs on BGRx (channel 3 copied).  luts: host, 3 x 256 bytes
 * (B, G, R).  src and dst may alias. */
int pano_transform_lut3(const uint8_t* src, size_t src_step, uint8_t* dst, size_t dst_step, int w,
                        int h, const uint8_t* luts, hipStream_t stream);

/* Encoder / decoder hand-off: cvtBGR32ToYUV420P, cvtBGR32ToNV12,
 * cvtYUV420PToBGR32, cvtNV12ToBGR32 (CudaAccel/CudaInterface.h:161-171,
 * CudaAccel/ImageProc.cu:44-266), BT.601 studio swing in 20-bit fixed point.
 * w and h even; BGRx rows 8-byte aligned; Y (and NV12 UV) rows 2-byte aligned.
 * The BGRx outputs of the inverses have channel 3 = 0. */
int pano_bgr32_to_yuv420p(const uint8_t* src, size_t src_step, int w, int h, uint8_t* y, size_t y_step,
                          uint8_t* u, size_t u_step, uint8_t* v, size_t v_step, hipStream_t stream);
int pano_bgr32_to_nv12(const uint8_t* src, size_t src_step, int w, int h, uint8_t* y, size_t y_step,
                       uint8_t* uv, size_t uv_step, hipStream_t stream);
int pano_yuv420p_to_bgr32(const uint8_t* y, size_t y_step, const uint8_t* u, size_t u_step, const uint8_t* v,
                          size_t v_step, int w, int h, uint8_t* dst, size_t dst_step, hipStream_t stream);
int pano_nv12_to_bgr32(const uint8_t* y, size_t y_step, const uint8_t* uv, size_t uv_step, int w, int h,
                       uint8_t* dst, size_t dst_step, hipStream_t stream);

/* getLUT (Blend/ExposureColorOptimize.cpp:1217-1223): lut[v] = sat_u8(rint(k v)). */
int pano_get_lut(uint8_t* lut, double k);
/* getExposureColorOptimizeLUTs (ExposureColorOptimize.cpp:1425-1443) without
 * the max-scale clamp the reference computes but does not apply:
 * luts[i] = {getLUT(e b), getLUT(e), getLUT(e r)}, n x 3 x 256 bytes. */
int pano_exposure_luts(int num, const double* exposures, const double* red_ratios,
                       const double* blue_ratios, uint8_t* luts);

/* getReprojectMap32FAndMask (Warp/ZReproject.cpp:231-240) on the host, rows
 * in parallel: xmap/ymap float32 dst_w x dst_h (-1 where invalid), mask 0/255. */
int pano_generate_map(const pano_photo_param* param, int src_w, int src_h, int dst_w, int dst_h,
                      float* xmap, float* ymap, uint8_t* mask);

/* cudaGenerateReprojectMapAndMask(param, srcSize, dstSize, xmap, ymap, mask)
 * (CudaInterface.h:129-130, CudaReproject.cu:266-385): the same FP64 chain on the
 * device, one canvas pixel per thread.  xmap / ymap: device float planes of
 * dst_w x dst_h (row step map_step bytes), mask: device bytes (mask_step).
 * The device's FP64 sin / cos / atan2 / tan may differ from the host library's
 * in the last bits; pano_render_prepare generates its maps this way. */
int pano_generate_map_device(const pano_photo_param* param, int src_w, int src_h, int dst_w, int dst_h,
                             float* xmap, float* ymap, size_t map_step, uint8_t* mask, size_t mask_step,
                             hipStream_t stream);

/* ---- render: CudaPanoramaRender2 (Task/RicohUtil.h:134-161) -------------
 * prepare + per-frame render of one rig: maps, masks, seams and weight
 * pyramids stay resident on the device; render() runs the fused path
 * (LUT + remap -> pyramid -> blend) with one launch per stage for all cameras.
 * Cameras may have different source sizes. */
typedef struct pano_render pano_render;

int pano_render_create(pano_render** out);
void pano_render_destroy(pano_render* r);

/* prepare(path, highQualityBlend = 1, blendParam = maxLevels, srcSize, dstSize)
 * (RicohUtil.cpp:1091-1205) with the parameters given directly instead of a
 * rig file.  dst_w must equal 2 * dst_h, both even (RicohUtil.cpp:1098-1103). */
int pano_render_prepare(pano_render* r, int num, const pano_photo_param* params, const int* src_w,
                        const int* src_h, int dst_w, int dst_h, int max_levels);

/* The same with caller-made maps and masks (host float32 planes dst_w x dst_h
 * and host masks), the inputs getReprojectMaps32FAndMasks would produce. */
int pano_render_prepare_maps(pano_render* r, int num, const int* src_w, const int* src_h,
                             int dst_w, int dst_h, const float* const* xmaps,
                             const float* const* ymaps, const uint8_t* const* masks,
                             int max_levels);

/* render(src, dst, luts) (RicohUtil.cpp:1207-1339), multiband branch.
 * srcs: device PANO_8UC4 images (sizes as prepared).  luts: host, num x 3 x
 * 256 bytes, or NULL for no correction.  dst: device PANO_8UC4 canvas. */
int pano_render_render(pano_render* r, int num, const uint8_t* const* srcs,
                       const size_t* src_steps, const uint8_t* luts, uint8_t* dst,
                       size_t dst_step, hipStream_t stream);
/* The same frame written as `out` (the canvas, or the encoder's NV12 / YUV420P
 * planes that the tasks produce next, PanoramaLocalDiskTask.cpp:2053-2071). */
int pano_render_render_to(pano_render* r, int num, const uint8_t* const* srcs, const size_t* src_steps,
                          const uint8_t* luts, const pano_output* out, hipStream_t stream);

/* Wall time of the last prepare, in ms: maps (generation or upload) and their
 * check, seams (on the device), tables (boxes, weight pyramids, weight sums,
 * owners, work lists), build (device residency, fused map storage). */
int pano_render_prepare_ms(const pano_render* r, double* ms4);
/* Where the fused 8-bit render's prepare builds its tables, weight pyramids
 * and map storage: on the device (on != 0, the default, up to 32 cameras) or
 * on the host (the reference prepares on the host; both give the same rig). */
int pano_render_set_prepare_on_device(pano_render* r, int on);
/* Down tile shape of the fused path (the level j -> j+1 output tile of the
 * remap and pyrDown kernels): 0 = 64 x 16 (512 threads), 1 = 64 x 32 (1024),
 * 2 = 128 x 16 (1024), 3 = 32 x 32 (512), 4 = 64 x 8 (256).  Takes effect at
 * the next prepare; renders batched together need one shape.
 * pano_render_tile_shape: the prepared one. */
#define PANO_TILE_SHAPES 5
int pano_render_set_tile_shape(pano_render* r, int shape);
int pano_render_tile_shape(const pano_render* r);
/* Tile autotuner: prepares the rig (pano_render_prepare) with each shape,
 * times iters frames of pano_render_render (no LUTs) on the stream after one
 * warm-up frame, and leaves r prepared with the fastest shape (shape 0 unless
 * another is more than 2% faster).  best: the shape; ms_out (optional):
 * PANO_TILE_SHAPES device ms per frame. */
int pano_render_autotune_tile_shape(pano_render* r, int num, const pano_photo_param* params, const int* src_w,
                                    const int* src_h, int dst_w, int dst_h, int max_levels,
                                    const uint8_t* const* srcs, const size_t* src_steps, uint8_t* dst,
                                    size_t dst_step, int iters, int* best, float* ms_out, hipStream_t stream);
int pano_render_num_images(const pano_render* r);
int pano_render_num_levels(const pano_render* r);
int64_t pano_render_mem_bytes(const pano_render* r);
/* Algorithmic HBM bytes of one render(): every stage input read once and every
 * stage output written once.  stage 0 remap, 1 down, 2 blend levels >= 1,
 * 3 level-0 blend, 4 total. */
int64_t pano_render_model_bytes(const pano_render* r, int stage);

/* 1 when the render runs the fused 8-bit path (its maps sample only inside
 * their masks and it has at least one pyramid level), 0 for the 16-bit path. */
int pano_render_fused(const pano_render* r);

/* ---- per-frame gain estimation ---------------------------------------------
 * render() with gains estimated from the same frame: the overlap-means
 * exposure estimate getTransformsMeanApproxSiftPanoPaper (Blend/Compensate.cpp:
 * 501-563) with its anchor rescale (:565-575, as GainCompensate::prepare uses
 * it, :614-646), optionally the tint estimate
 * getTintTransformsMeanApproxMimicSiftPanoPaper (Blend/TintAdjust.cpp:123-215),
 * both on the frame's 8-bit reprojection (before any LUT) and the rig's
 * original masks, then the LUTs of getExposureColorOptimizeLUTs
 * (Blend/ExposureColorOptimize.cpp:1425-1443) applied to the same frame.
 * Statistics, solve and LUTs stay on the device (no host round trip).  Fused
 * renders of up to 8 cameras (PANO_ERR_UNSUPPORTED otherwise). */
#define PANO_GAIN_TINT 1
int pano_render_render_autogain(pano_render* r, int num, const uint8_t* const* srcs,
                                const size_t* src_steps, int flags, uint8_t* dst,
                                size_t dst_step, hipStream_t stream);
/* The last estimate (waits for the device): gains, red and blue ratio gains
 * (num each, any may be NULL) and the anchor camera. */
int pano_render_gains(const pano_render* r, double* gains, double* red, double* blue, int* anchor);
/* The last estimate's sums (waits for the device): num x num x 5 int64 --
 * count, gray, B, G, R of camera i's reprojection over mask_i & mask_j. */
int pano_render_gain_sums(const pano_render* r, int64_t* sums);

/* ---- batched render: several rigs per launch ------------------------------
 * One launch per pyramid stage covers every camera of every render in the
 * batch, each render's tiles taken from its own resident work lists (no
 * counterpart in the reference, which stitches one rig per call,
 * RicohUtil.cpp:1207-1339; this is for servers that stitch many streams).
 * Up to 32 renders and 128 cameras in total, all on the fused 8-bit path.
 * The renders must outlive the batch.  pano_batch_set only builds the new
 * launch tables on the host; the next pano_batch_render uploads them
 * (stream-ordered) into one of 4 device table slots, waiting only when all 4
 * still serve renders in flight. */
typedef struct pano_batch pano_batch;

int pano_batch_create(pano_batch** out);
void pano_batch_destroy(pano_batch* b);
int pano_batch_set(pano_batch* b, int num, pano_render* const* renders);
/* srcs/src_steps: the cameras of render 0, then render 1, ... (device BGRx);
 * luts: per render a host n x 3 x 256 block or NULL (no correction), or NULL
 * for none at all; dsts/dst_steps: per render a device BGRx canvas. */
int pano_batch_render(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                      const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                      hipStream_t stream);
/* The same with one pano_output per render. */
int pano_batch_render_to(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                         const uint8_t* const* luts, const pano_output* outs, hipStream_t stream);
/* Block -> tile order of each stage's launch (remap, down, blend, final):
 * 0 in job order, 1 one contiguous run of jobs per XCD (each XCD has its own
 * L2, and neighbouring tiles share halos).  No reference counterpart: the
 * reference launches one camera at a time.  The order changes no result. */
#define PANO_SCHED_MODES 2
int pano_batch_set_sched(pano_batch* b, const int* sched4);
int pano_batch_get_sched(const pano_batch* b, int* sched4);
/* Autotuner: renders the given frames iters times under each order, keeps the
 * fastest order per stage (device time from events), and returns the mean ms
 * per stage and order in ms_out (PANO_SCHED_MODES x 4, may be NULL).  The
 * outputs are written as by pano_batch_render; a captured graph is dropped. */
int pano_batch_autotune(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                        const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                        hipStream_t stream, int iters, double* ms_out);
int pano_batch_set_timing(pano_batch* b, int enable);
int pano_batch_stage_ms(pano_batch* b, double* ms4, int* frames);
/* The per-frame sequence of one pano_batch_render as a HIP graph (no
 * counterpart in the reference): capture records the launches for these
 * frame and canvas buffers (and uploads LUTs and tables outside the graph);
 * each replay re-runs them on the stream, for frames refreshed in place.
 * Renders captured with LUTs take new LUT contents at replay (luts[i], or NULL
 * to keep the previous ones); the others must pass NULL.  pano_batch_set drops
 * the graph.  With timing enabled, a replay's time is reported as stage 0.
 * The stream of a capture must not be the null stream. */
int pano_batch_capture(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                       const uint8_t* const* luts, uint8_t* const* dsts, const size_t* dst_steps,
                       hipStream_t stream);
int pano_batch_capture_to(pano_batch* b, const uint8_t* const* srcs, const size_t* src_steps,
                          const uint8_t* const* luts, const pano_output* outs, hipStream_t stream);
int pano_batch_replay(pano_batch* b, const uint8_t* const* luts, hipStream_t stream);

/* ---- serving loop ----------------------------------------------------------
 * Replaces the per-stream worker threads of the live task
 * (Task/PanoramaLiveStreamTask2.cpp:340-350, 979-983, each calling
 * CudaPanoramaRender2::render for its own stream) with one server: prepared
 * renders are admitted as streams while their resident device memory
 * (pano_render_mem_bytes) fits the HBM budget; submitted frames are batched
 * continuously across streams -- whenever fewer than max_inflight batches are
 * queued on the server's stream, the oldest pending frame of every stream with
 * nothing in flight is launched in one pano_batch_render -- and each frame's
 * submit-to-ready latency is reported.  Inputs and outputs of a submitted frame
 * must stay valid until it is reported done.  Calls are thread-safe. */
typedef struct pano_server pano_server;
typedef struct pano_server_done {
    uint64_t tag;         /* as submitted */
    double latency_ms;    /* submit -> output ready (host clock) */
    int stream_id;
    int status;           /* PANO_OK or the launch error */
    int batch_frames;     /* frames in the batch it ran in */
    int pad;
} pano_server_done;

int pano_server_create(pano_server** out, int64_t hbm_budget_bytes, int max_inflight);
/* Waits for the batches in flight; frames not yet launched are dropped. */
void pano_server_destroy(pano_server* s);
/* PANO_ERR_BUDGET when the render does not fit the remaining budget. */
int pano_server_admit(pano_server* s, pano_render* r, int* stream_id);
/* Waits for the stream's submitted frames, then releases its budget. */
int pano_server_retire(pano_server* s, int stream_id);
/* srcs/src_steps: the stream's cameras (device BGRx); luts: host num x 3 x 256
 * (copied) or NULL; dst: device BGRx canvas. */
int pano_server_submit(pano_server* s, int stream_id, const uint8_t* const* srcs,
                       const size_t* src_steps, const uint8_t* luts, uint8_t* dst, size_t dst_step,
                       uint64_t tag);
/* Moves up to max_out completed frames into out (waiting up to wait_ms for
 * the first); returns their number or a negative status. */
int pano_server_poll(pano_server* s, pano_server_done* out, int max_out, int wait_ms);
/* Waits until every submitted frame is done. */
int pano_server_drain(pano_server* s);
int pano_server_stats(pano_server* s, int64_t* used_bytes, int64_t* frames_done, int64_t* batches);

/* Per-stage device timing of render() (no counterpart in the reference, whose
 * timers are host-side, Tool/Timer.h:8-143): when enabled, HIP events are
 * recorded on the render's stream between its stages.  pano_render_stage_ms
 * waits for the recorded frames and returns the summed milliseconds of each
 * stage -- 0 remap, 1 pyramid down (all levels), 2 blend levels >= 1, 3 level-0
 * blend -- and the number of frames, then starts a new period. */
int pano_render_set_timing(pano_render* r, int enable);
int pano_render_stage_ms(pano_render* r, double* ms4, int* frames);

#ifdef __cplusplus
}
#endif
#endif
