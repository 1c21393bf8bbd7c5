/*
 * linear.c -- CPU restatement of the linear-blend mode (highQualityBlend = 0).
 * TEST INFRASTRUCTURE ONLY (part of the oracle library; see oracle.h).
 *
 *   weights   getWeightsLinearBlend32F (Blend/LinearBlend.cpp:134-148 with
 *             calcWeights32F :76-132) and getWeightsLinearBlend (:59-73 with
 *             calcWeights :17-57): seams (getNonIntersectingMasks), each unique
 *             mask Gaussian-blurred with a (2r+1)^2 kernel, sigma r / 3, ANDed
 *             bitwise with its mask, then per pixel normalised.
 *   blur      cv::GaussianBlur of an 8-bit plane as OpenCV 3.0 runs it:
 *             getGaussianKernel(n, sigma, CV_32F) (float taps from the double
 *             exponential, normalised in double); the kernel is smooth and
 *             symmetric, so createSeparableLinearFilter takes its 8-bit fixed
 *             point path: taps K = cvRound(256 k), a row pass to int32 and a
 *             column pass, (s + 2^15) >> 16 saturated to 8 bits; borders
 *             REFLECT_101 both ways (BORDER_DEFAULT).
 *   render    CPUPanoramaRender::render, linear branch (Task/RicohUtil.cpp:
 *             1543-1558): per camera in order, transform (LUTs) then
 *             reprojectWeightedAccumulateParallelTo32F (Warp/ZReproject.cpp:
 *             606-691): accum += bilinear sample * weight in float, without
 *             contraction; then convertTo(CV_8U) (round half to even,
 *             saturate).
 *   blend32s  TilingLinearBlend::blend (LinearBlend.cpp:553-566) and the GPU
 *             CudaTilingLinearBlend::blend (CudaLinearBlend.cpp:87-101,
 *             CudaFeatherBlend.cu:8-29): accum (int32) += pixel * weight, >> 16,
 *             saturate to 8 bits; cn channels per pixel.
 *
 * Pinned by tests/test_oracle_linear.py: the kernel against the normalised
 * Gaussian (|K - 256 g| <= 1/2), the blur against a brute-force 2-D sum over a
 * numpy reflect-padded image with those taps, the weights' partition of unity
 * and their single-camera and zero cases, and the render against float64
 * sums of the same samples.  The OpenCV 3.0 boundary itself (the fixed-point
 * path choice, its rounding) is restated from its published source: parity
 * unpinned there, as for the other OpenCV routines.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void orc_gaussian_kernel_q8(int n, double sigma, int* K)
{
    float* cf = (float*)malloc(sizeof(float) * n);
    const double sx = sigma > 0 ? sigma : ((n - 1) * 0.5 - 1) * 0.3 + 0.8;
    const double scale2x = -0.5 / (sx * sx);
    double sum = 0;
    for (int i = 0; i < n; i++) {
        const double x = i - (n - 1) * 0.5;
        cf[i] = (float)exp(scale2x * x * x);
        sum += cf[i];
    }
    sum = 1. / sum;
    for (int i = 0; i < n; i++) {
        cf[i] = (float)(cf[i] * sum);
        K[i] = (int)nearbyintf(cf[i] * 256.0f);
    }
    free(cf);
}

void orc_gaussian_blur_u8(const uint8_t* src, int w, int h, int n, double sigma, uint8_t* dst)
{
    int* K = (int*)malloc(sizeof(int) * n);
    orc_gaussian_kernel_q8(n, sigma, K);
    const int a = n / 2;
    int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (size_t)w * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t s = 0;
            for (int k = 0; k < n; k++)
                s += K[k] * src[(size_t)y * w + orc_border(x + k - a, w, ORC_BORDER_REFLECT_101)];
            rows[(size_t)y * w + x] = s;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int32_t s = 0;
            for (int k = 0; k < n; k++)
                s += K[k] * rows[(size_t)orc_border(y + k - a, h, ORC_BORDER_REFLECT_101) * w + x];
            const int v = (s + (1 << 15)) >> 16;
            dst[(size_t)y * w + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    free(rows);
    free(K);
}

int orc_linear_weights(int n, int rows, int cols, const uint8_t* masks, int radius, int as_int, void* out)
{
    if (n <= 0 || rows <= 0 || cols <= 0 || radius < 0) return 0;
    const size_t P = (size_t)rows * cols;
    uint8_t* unique = (uint8_t*)malloc(P * n);
    uint8_t* dist = (uint8_t*)malloc(P * n);
    orc_nonintersecting_masks(n, cols, rows, masks, unique);
    for (int i = 0; i < n; i++) {
        orc_gaussian_blur_u8(unique + P * i, cols, rows, 2 * radius + 1, radius / 3.0, dist + P * i);
        for (size_t p = 0; p < P; p++) dist[P * i + p] &= masks[P * i + p];
    }
    for (size_t p = 0; p < P; p++) {
        float sum = 0;
        int nz = 0, last = 0;
        for (int k = 0; k < n; k++) {
            sum += dist[P * k + p];
            if (dist[P * k + p]) {
                nz++;
                last = k;
            }
        }
        if (as_int) {
            /* calcWeights: d * (1 / sum) * 2^16 + 1, truncated */
            int32_t* w = (int32_t*)out;
            const float inv = fabsf(sum) <= FLT_MIN ? 0.f : 1.0f / sum;
            for (int k = 0; k < n; k++) w[P * k + p] = (int32_t)(dist[P * k + p] * inv * 65536 + 1);
        } else {
            /* calcWeights32F: several cameras share by d / sum, a lone one gets 1 */
            float* w = (float*)out;
            for (int k = 0; k < n; k++) w[P * k + p] = 0.f;
            if (nz > 1) {
                const float inv = fabsf(sum) <= FLT_MIN ? 0.f : 1.0f / sum;
                for (int k = 0; k < n; k++) w[P * k + p] = dist[P * k + p] * inv;
            } else if (nz == 1) {
                w[P * last + p] = 1.0f;
            }
        }
    }
    free(unique);
    free(dist);
    return 1;
}

int orc_linear_render(int n, const uint8_t* const* srcs, const int* sw, const int* sh,
                      const float* const* xmaps, const float* const* ymaps, const float* weights,
                      int dw, int dh, const uint8_t* luts, uint8_t* out, int nthreads)
{
    const size_t P = (size_t)dw * dh;
    float* acc = (float*)calloc(P * 3, sizeof(float));
    int16_t* sample = (int16_t*)malloc(sizeof(int16_t) * P * 3);
    for (int i = 0; i < n; i++) {
        const uint8_t* src = srcs[i];
        uint8_t* adj = NULL;
        if (luts) {
            adj = (uint8_t*)malloc((size_t)sw[i] * sh[i] * 3);
            orc_transform(src, adj, sw[i], sh[i], (size_t)sw[i] * 3, 3, luts + (size_t)i * 768);
            src = adj;
        }
        /* saturate_cast<uchar> of the FP64 bilinear sum, 0 where the map is invalid
           (an invalid point adds nothing; adding 0 * w leaves the sum as it is) */
        orc_reproject_to16s(src, sw[i], sh[i], (size_t)sw[i] * 3, 3, xmaps[i], ymaps[i], dw, dh, sample,
                            nthreads);
        const float* w = weights + P * i;
        for (size_t p = 0; p < P; p++)
            for (int c = 0; c < 3; c++) {
                const float prod = (float)sample[p * 3 + c] * w[p];
                acc[p * 3 + c] = acc[p * 3 + c] + prod;
            }
        free(adj);
    }
    for (size_t e = 0; e < P * 3; e++) {
        const float r = nearbyintf(acc[e]);
        out[e] = (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
    }
    free(acc);
    free(sample);
    return 1;
}

int orc_linear_blend32s(int n, int rows, int cols, int cn, const uint8_t* images, const int32_t* weights,
                        uint8_t* out)
{
    const size_t P = (size_t)rows * cols;
    for (size_t p = 0; p < P; p++)
        for (int c = 0; c < cn; c++) {
            int32_t s = 0;
            for (int i = 0; i < n; i++) s += images[(P * i + p) * cn + c] * weights[P * i + p];
            const int v = s >> 16;
            out[p * cn + c] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    return 1;
}
