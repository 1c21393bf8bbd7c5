/*
 * blend32f.c -- CPU restatement of CudaTilingMultibandBlendFast32F, the
 * reference's float-weight multiband blender (CudaAccel/CudaMultibandBlend.cpp:
 * 1041-1179 with the kernels of CudaAccel/CudaPyramid6.cu).  TEST
 * INFRASTRUCTURE ONLY (part of the oracle library; see oracle.h).
 *
 * The reference exists only as CUDA; its arithmetic is restated operation by
 * operation in float, with the fused multiply-adds its compiler forms from
 * `sum = sum + c * x` (CUDA contracts them by default) written as fmaf:
 *   pyrDown32F (CudaPyramid6.cu:21-221): vertical 5-tap sums first, then the
 *     horizontal 5-tap over them; taps 1/16 4/16 6/16; rows REFLECT_101,
 *     columns WRAP; size (w + 1) / 2.
 *   pyrUp32FC4 (:229-310): horizontal taps of the source rows (1/16 6/16 1/16
 *     at even columns, 4/16 4/16 at odd ones), the same vertically, times 4;
 *     taps of weight zero add nothing and are left out.
 *   prepare (:1041-1133): alpha_{j+1} = 1 / pyrDown(binary alpha_j) where
 *     positive, 0 elsewhere; binary alpha_{j+1} = pyrDown(...) > 0; weights
 *     W_0 = unique mask (0/1), W_{j+1} = pyrDown(W_j); when the masks do not
 *     cover the canvas, the inverse of sum_i W_ij per level.
 *   blend (:1135-1179): G_0 = the image as float; G_{j+1} = alpha_{j+1}
 *     pyrDown(G_j); L_j = G_j - pyrUp(G_{j+1}); R_j = fma(W_ij, L_ij, R_j);
 *     R_j *= inverse weight sum (partial coverage); C_j = pyrUp(C_{j+1}) + R_j
 *     from the top; out = saturate_cast<uchar>(C_0) (round to nearest even),
 *     0 outside every mask.
 * Images are 4-channel float (channel 3 is carried like the others).
 *
 * Pinned by tests/test_oracle_blend32f.py: pyrDown / pyrUp against numpy
 * float64 convolutions (closed form up to float rounding), the single-camera
 * identity, agreement with the integer blender within a few levels, and
 * the partial-coverage zeroing.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static const float C1 = 0.0625f, C4 = 0.25f, C6 = 0.375f;

static int wrapi(int p, int n)
{
    p %= n;
    return p < 0 ? p + n : p;
}

static int refl(int p, int n)
{
    return orc_border(p, n, ORC_BORDER_REFLECT_101);
}

/* pyrDown of an interleaved float image with cn channels. */
void orc_pyrdown32f(const float* src, int sw, int sh, int cn, float* dst)
{
    const int dw = (sw + 1) / 2, dh = (sh + 1) / 2;
    static const float taps[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    for (int y = 0; y < dh; y++)
        for (int x = 0; x < dw; x++)
            for (int c = 0; c < cn; c++) {
                float h = 0.f;
                for (int k = 0; k < 5; k++) {
                    const int xs = wrapi(2 * x - 2 + k, sw);
                    const float* col = src + (size_t)xs * cn + c;
                    float s = C1 * col[(size_t)refl(2 * y - 2, sh) * sw * cn];
                    s = fmaf(C4, col[(size_t)refl(2 * y - 1, sh) * sw * cn], s);
                    s = fmaf(C6, col[(size_t)(2 * y) * sw * cn], s);
                    s = fmaf(C4, col[(size_t)refl(2 * y + 1, sh) * sw * cn], s);
                    s = fmaf(C1, col[(size_t)refl(2 * y + 2, sh) * sw * cn], s);
                    h = k == 0 ? taps[0] * s : fmaf(taps[k], s, h);
                }
                dst[((size_t)y * dw + x) * cn + c] = h;
            }
}

/* The horizontal pass of pyrUp for one source row at output column x. */
static float up_row(const float* row, int sw, int cn, int c, int x)
{
    const int k = x >> 1;
    if (!(x & 1)) {
        float s = C1 * row[(size_t)wrapi(k - 1, sw) * cn + c];
        s = fmaf(C6, row[(size_t)wrapi(k, sw) * cn + c], s);
        return fmaf(C1, row[(size_t)wrapi(k + 1, sw) * cn + c], s);
    }
    const float s = C4 * row[(size_t)wrapi(k, sw) * cn + c];
    return fmaf(C4, row[(size_t)wrapi(k + 1, sw) * cn + c], s);
}

/* pyrUp32FC4 of a float image (sw x sh) to dw x dh. */
void orc_pyrup32f(const float* src, int sw, int sh, int cn, float* dst, int dw, int dh)
{
    for (int y = 0; y < dh; y++) {
        const int k = y >> 1;
        for (int x = 0; x < dw; x++)
            for (int c = 0; c < cn; c++) {
                float v;
                if (!(y & 1)) {
                    v = C1 * up_row(src + (size_t)refl(k - 1, sh) * sw * cn, sw, cn, c, x);
                    v = fmaf(C6, up_row(src + (size_t)refl(k, sh) * sw * cn, sw, cn, c, x), v);
                    v = fmaf(C1, up_row(src + (size_t)refl(k + 1, sh) * sw * cn, sw, cn, c, x), v);
                } else {
                    v = C4 * up_row(src + (size_t)refl(k, sh) * sw * cn, sw, cn, c, x);
                    v = fmaf(C4, up_row(src + (size_t)refl(k + 1, sh) * sw * cn, sw, cn, c, x), v);
                }
                dst[((size_t)y * dw + x) * cn + c] = 4.0f * v;
            }
    }
}

static void level_sizes(int w, int h, int L, int* lw, int* lh)
{
    lw[0] = w;
    lh[0] = h;
    for (int j = 1; j <= L; j++) {
        lw[j] = (lw[j - 1] + 1) / 2;
        lh[j] = (lh[j - 1] + 1) / 2;
    }
}

/* CudaTilingMultibandBlendFast32F prepare + blend in one call.  masks: n planes
 * rows x cols; images: n planes rows x cols x 4 floats (an 8-bit BGRx image
 * converted exactly, or a CV_32FC4 one); out rows x cols x 4 bytes.  Returns
 * the number of levels, or -1 on bad arguments. */
int orc_blend32f(int n, int rows, int cols, const uint8_t* masks, const float* images,
                 int max_levels, int min_length, uint8_t* out)
{
    if (n <= 0 || rows <= 0 || cols <= 0) return -1;
    const int L = orc_true_num_levels(cols, rows, max_levels, min_length);
    if (L < 0) return -1;
    int lw[64], lh[64];
    level_sizes(cols, rows, L, lw, lh);
    const size_t P = (size_t)rows * cols;
    uint8_t* unique = (uint8_t*)malloc(P * n);
    orc_nonintersecting_masks(n, cols, rows, masks, unique);
    /* per camera: inverse alphas (levels 1..L) and weights (levels 0..L) */
    float** ainv = (float**)calloc((size_t)n * (L + 1), sizeof(float*));
    float** wts = (float**)calloc((size_t)n * (L + 1), sizeof(float*));
    for (int i = 0; i < n; i++) {
        float* bin = (float*)malloc(sizeof(float) * P);
        float* w0 = (float*)malloc(sizeof(float) * P);
        for (size_t p = 0; p < P; p++) {
            bin[p] = masks[(size_t)i * P + p] ? 1.f : 0.f;
            w0[p] = unique[(size_t)i * P + p] ? 1.f : 0.f;
        }
        wts[(size_t)i * (L + 1)] = w0;
        for (int j = 0; j < L; j++) {
            const size_t q = (size_t)lw[j + 1] * lh[j + 1];
            float* a = (float*)malloc(sizeof(float) * q);
            orc_pyrdown32f(bin, lw[j], lh[j], 1, a);
            float* nb = (float*)malloc(sizeof(float) * q);
            for (size_t p = 0; p < q; p++) {
                nb[p] = a[p] > 0 ? 1.f : 0.f;
                a[p] = a[p] > 0 ? 1.0f / a[p] : 0.f;
            }
            free(bin);
            bin = nb;
            ainv[(size_t)i * (L + 1) + j + 1] = a;
            float* w = (float*)malloc(sizeof(float) * q);
            orc_pyrdown32f(wts[(size_t)i * (L + 1) + j], lw[j], lh[j], 1, w);
            wts[(size_t)i * (L + 1) + j + 1] = w;
        }
        free(bin);
    }
    int full = 1;
    for (size_t p = 0; p < P && full; p++) {
        int any = 0;
        for (int i = 0; i < n; i++) any |= masks[(size_t)i * P + p] != 0;
        full = any;
    }
    float** winv = (float**)calloc(L + 1, sizeof(float*));
    if (!full)
        for (int j = 0; j <= L; j++) {
            const size_t q = (size_t)lw[j] * lh[j];
            winv[j] = (float*)calloc(q, sizeof(float));
            for (int i = 0; i < n; i++)
                for (size_t p = 0; p < q; p++) winv[j][p] = winv[j][p] + wts[(size_t)i * (L + 1) + j][p];
            for (size_t p = 0; p < q; p++) winv[j][p] = winv[j][p] > 0 ? 1.0f / winv[j][p] : 0.f;
        }
    float** R = (float**)calloc(L + 1, sizeof(float*));
    float** G = (float**)calloc(L + 1, sizeof(float*));
    for (int j = 0; j <= L; j++) {
        R[j] = (float*)calloc((size_t)lw[j] * lh[j] * 4, sizeof(float));
        G[j] = (float*)malloc(sizeof(float) * (size_t)lw[j] * lh[j] * 4);
    }
    float* up = (float*)malloc(sizeof(float) * P * 4);
    for (int i = 0; i < n; i++) {
        memcpy(G[0], images + (size_t)i * P * 4, sizeof(float) * P * 4);
        for (int j = 0; j < L; j++) {
            const size_t q = (size_t)lw[j + 1] * lh[j + 1];
            orc_pyrdown32f(G[j], lw[j], lh[j], 4, G[j + 1]);
            const float* a = ainv[(size_t)i * (L + 1) + j + 1];
            for (size_t p = 0; p < q; p++)
                for (int c = 0; c < 4; c++) G[j + 1][p * 4 + c] = a[p] * G[j + 1][p * 4 + c];
        }
        for (int j = 0; j <= L; j++) {
            const size_t q = (size_t)lw[j] * lh[j];
            const float* w = wts[(size_t)i * (L + 1) + j];
            if (j < L) orc_pyrup32f(G[j + 1], lw[j + 1], lh[j + 1], 4, up, lw[j], lh[j]);
            for (size_t p = 0; p < q; p++)
                for (int c = 0; c < 4; c++) {
                    const float lap = j < L ? G[j][p * 4 + c] - up[p * 4 + c] : G[j][p * 4 + c];
                    R[j][p * 4 + c] = fmaf(w[p], lap, R[j][p * 4 + c]);
                }
        }
    }
    if (!full)
        for (int j = 0; j <= L; j++) {
            const size_t q = (size_t)lw[j] * lh[j];
            for (size_t p = 0; p < q; p++)
                for (int c = 0; c < 4; c++) R[j][p * 4 + c] = winv[j][p] * R[j][p * 4 + c];
        }
    for (int j = L - 1; j >= 0; j--) {
        const size_t q = (size_t)lw[j] * lh[j];
        orc_pyrup32f(R[j + 1], lw[j + 1], lh[j + 1], 4, up, lw[j], lh[j]);
        for (size_t p = 0; p < q * 4; p++) R[j][p] = up[p] + R[j][p];
    }
    for (size_t p = 0; p < P; p++) {
        int covered = full;
        for (int i = 0; i < n && !covered; i++) covered = masks[(size_t)i * P + p] != 0;
        for (int c = 0; c < 4; c++) {
            float v = R[0][p * 4 + c];
            float r = nearbyintf(v);
            int o = r < 0 ? 0 : (r > 255 ? 255 : (int)r);
            if (v != v) o = 0;
            out[p * 4 + c] = covered ? (uint8_t)o : 0;
        }
    }
    for (int j = 0; j <= L; j++) {
        free(R[j]);
        free(G[j]);
        free(winv[j]);
    }
    for (size_t k = 0; k < (size_t)n * (L + 1); k++) {
        free(ainv[k]);
        free(wts[k]);
    }
    free(R);
    free(G);
    free(winv);
    free(ainv);
    free(wts);
    free(up);
    free(unique);
    return L;
}
