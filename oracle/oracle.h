/*
 * oracle.h -- CPU restatement of PanoStitcher's per-frame stitch path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry and bench.py's cpu_baseline leg load it to produce the
 * expected results (or to time the reference's CPU algorithm).  The product
 * library (panostitcher_amd/csrc) never includes, links or calls it.
 *
 * Every function names the reference file:line it restates.  Pixel layouts
 * follow the reference CPU path: interleaved channels (BGR, cn = 3) with the
 * row step given in elements unless noted.  Arithmetic is integer except the
 * bilinear reprojection (FP64, no contraction: built with -ffp-contract=off)
 * and the map generator (FP64 libm).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   orc_border                 known values + mod/reflection identities
 *   orc_pyrdown_*              numpy 5x5 convolution on padded images
 *   orc_pyrup_*                numpy zero-insertion + convolution
 *   orc_distance_transform_l1  brute-force L1 distance (BFS) incl. border cap
 *   orc_nonintersecting_masks  hand-made seam cases + partition invariants
 *   orc_blender_*              independent numpy multiband restatement,
 *                              single-camera identity, permutation invariance
 *   orc_reproject_*            integer / half-pixel closed forms, edge reads
 *   orc_get_lut / transform    rint(k*v) closed form incl. ties and clamps
 *   orc_true_num_levels        closed-form values
 *   orc_map_*                  geometric invariants of the equirect chain
 *   orc_gain_* / orc_cv_solve  see gain.c
 *   orc_*32f                   see blend32f.c
 */
#ifndef PANO_ORACLE_H
#define PANO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BORDER_REFLECT_101 = 4, ORC_BORDER_WRAP = 3 }; /* OpenCV codes */

/* cv::borderInterpolate for the two border types the path uses. */
int orc_border(int p, int len, int type);

/* Pyramid primitives (Blend/Pyramid.cpp).  Element steps, cn channels. */
void orc_pyrdown_to32s_s16(const int16_t* src, int sw, int sh, int sstep, int cn,
                           int32_t* dst, int dw, int dh, int dstep, int hb, int vb);
void orc_pyrdown_s16(const int16_t* src, int sw, int sh, int sstep, int cn,
                     int16_t* dst, int dw, int dh, int dstep, int hb, int vb);
void orc_pyrup_s16(const int16_t* src, int sw, int sh, int sstep, int cn,
                   int16_t* dst, int dw, int dh, int dstep, int hb, int vb);
void orc_pyrup_s32(const int32_t* src, int sw, int sh, int sstep, int cn,
                   int32_t* dst, int dw, int dh, int dstep, int hb, int vb);

int orc_true_num_levels(int width, int height, int max_levels, int min_length);

/* cv::distanceTransform(src, dst, CV_DIST_L1, 3) -> float distances. */
void orc_distance_transform_l1(const uint8_t* src, int w, int h, float* dst);

/* getNonIntersectingMasks: n masks of w*h bytes stored back to back. */
void orc_nonintersecting_masks(int n, int w, int h, const uint8_t* masks, uint8_t* out);

/* TilingMultibandBlendFast (CPU, 3-channel 16S in, 8U out). */
typedef struct orc_blender orc_blender;
orc_blender* orc_blender_create(void);
void orc_blender_destroy(orc_blender* b);
int orc_blender_prepare(orc_blender* b, int n, int rows, int cols, const uint8_t* masks,
                        int max_levels, int min_length);
/* images: n planes of rows*cols*3 int16 back to back; out rows*cols*3 u8.
 * nthreads > 1 builds the per-camera pyramids concurrently, the way
 * TilingMultibandBlendFastParallel does; the arithmetic is identical. */
int orc_blender_blend(orc_blender* b, const int16_t* images, uint8_t* out, int nthreads);
int orc_blender_blend_custom(orc_blender* b, const int16_t* images, const uint8_t* masks,
                             uint8_t* out, int nthreads);
int orc_blender_num_levels(const orc_blender* b);
int orc_blender_full_mask(const orc_blender* b);
void orc_blender_unique_masks(const orc_blender* b, uint8_t* out);
/* Read back prepared pyramids: weight (int16) and alpha (int32, level>=1). */
int orc_blender_level_size(const orc_blender* b, int level, int* w, int* h);
int orc_blender_weight(const orc_blender* b, int cam, int level, int16_t* out);
int orc_blender_alpha(const orc_blender* b, int cam, int level, int32_t* out);

/* Warp (Warp/ZReproject.cpp). src is u8 with cn channels and a row step in
 * bytes; the map is two float32 planes (x, y) of dw*dh promoted to double. */
void orc_reproject_to16s(const uint8_t* src, int sw, int sh, size_t sstep, int cn,
                         const float* xmap, const float* ymap, int dw, int dh,
                         int16_t* dst, int nthreads);
void orc_reproject_to16s_map64(const uint8_t* src, int sw, int sh, size_t sstep, int cn,
                               const double* map_xy, int dw, int dh, int16_t* dst,
                               int nthreads);

/* Gains (Blend/ExposureColorOptimize.cpp, Blend/GainAdjust.cpp). */
void orc_get_lut(uint8_t* lut, double k);
void orc_transform(const uint8_t* src, uint8_t* dst, int w, int h, size_t step, int cn,
                   const uint8_t* lut3);

/* Camera model (Warp/ZReproject.h:17-50) and map generation. */
typedef struct orc_photo_param {
    int imageType, cropMode, cropX, cropY, cropWidth, cropHeight;
    int circleX, circleY, circleR;
    double hfov, vfov, alpha, beta, gamma, shiftX, shiftY, shearX, shearY;
    double yaw, pitch, roll;
} orc_photo_param;

/* Remap::init + remapImage for one canvas pixel (x, y) -> camera (sx, sy). */
void orc_remap_point(const orc_photo_param* p, int src_w, int src_h, int dst_w, int dst_h,
                     double x, double y, double* sx, double* sy);
/* getReprojectMapAndMask: map_xy is dst_w*dst_h*2 doubles, mask dst_w*dst_h. */
void orc_reproject_map_and_mask(const orc_photo_param* p, int src_w, int src_h, int dst_w,
                                int dst_h, double* map_xy, uint8_t* mask, int nthreads);

/* CPUPanoramaRender::render (Task/RicohUtil.cpp:1481-1587), multiband branch.
 * srcs[i]: src_w[i]*src_h[i]*3 BGR bytes (contiguous); xmaps/ymaps float32
 * canvas planes; luts: n*3*256 bytes or NULL (no correction); dst W*H*3. */
int orc_render(orc_blender* b, int n, const uint8_t* const* srcs, const int* src_w,
               const int* src_h, const float* const* xmaps, const float* const* ymaps,
               const uint8_t* luts, uint8_t* dst, int nthreads);

/* Per-frame gain estimation (gain.c).  bgr[i]: rows*cols*3 reprojected BGR
 * bytes of camera i, masks[i]: its original mask (rows*cols bytes). */
int orc_gray_bgr(int b, int g, int r);
int orc_cv_solve(int n, double* A, const double* b, double* x);
/* sums[(i*n + j)*5 + f], f = count, gray, B, G, R of image i over
 * mask_i & mask_j (mask_i alone for j = i). */
void orc_gain_sums(int n, int rows, int cols, const uint8_t* const* bgr,
                   const uint8_t* const* masks, int64_t* sums);
int orc_gain_from_sums(int n, const int64_t* sums, int tint, double* gains, double* rg,
                       double* bg, int* max_index, uint8_t* luts);
int orc_gain_estimate(int n, int rows, int cols, const uint8_t* const* bgr,
                      const uint8_t* const* masks, int tint, double* gains, double* rg, double* bg,
                      int* max_index, uint8_t* luts);

/* Float-weight blender (blend32f.c): CudaTilingMultibandBlendFast32F. */
void orc_pyrdown32f(const float* src, int sw, int sh, int cn, float* dst);
void orc_pyrup32f(const float* src, int sw, int sh, int cn, float* dst, int dw, int dh);
int orc_blend32f(int n, int rows, int cols, const uint8_t* masks, const float* images,
                 int max_levels, int min_length, uint8_t* out);

/* Linear-blend mode (linear.c).  masks / weights: n planes rows x cols;
 * as_int: int32 weights (calcWeights, 2^16 units) else float (calcWeights32F).
 * srcs: BGR (3-byte) frames; out: dw x dh x 3.  images (blend32s): n planes of
 * rows x cols x cn bytes. */
void orc_gaussian_kernel_q8(int n, double sigma, int* K);
void orc_gaussian_blur_u8(const uint8_t* src, int w, int h, int n, double sigma, uint8_t* dst);
int orc_linear_weights(int n, int rows, int cols, const uint8_t* masks, int radius, int as_int, void* out);
int orc_linear_render(int n, const uint8_t* const* srcs, const int* sw, const int* sh,
                      const float* const* xmaps, const float* const* ymaps, const float* weights,
                      int dw, int dh, const uint8_t* luts, uint8_t* out, int nthreads);
int orc_linear_blend32s(int n, int rows, int cols, int cn, const uint8_t* images, const int32_t* weights,
                        uint8_t* out);

/* Encoder / decoder hand-off (yuv.c): CudaAccel/ImageProc.cu.  nv12: u is the
 * interleaved UV plane (w bytes per row, v unused); else u and v are w/2 x h/2.
 * y is w x h; w and h even.  Return 0 on bad sizes. */
int orc_bgr32_to_yuv(const uint8_t* bgr, size_t step, int w, int h, int nv12, uint8_t* y,
                     uint8_t* u, uint8_t* v);
int orc_yuv_to_bgr32(const uint8_t* y, const uint8_t* u, const uint8_t* v, int w, int h, int nv12,
                     uint8_t* bgr, size_t step);

#ifdef __cplusplus
}
#endif
#endif
