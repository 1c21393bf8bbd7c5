/*
 * oracle.c -- CPU restatement of PanoStitcher's per-frame stitch path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by smoke() and
 * by bench.py's cpu_baseline leg; the product never links it.
 *
 * Citations are relative to the reference source tree (xpngzhng/PanoStitcher).
 * Built with -fwrapv (int overflow wraps as on the reference's MSVC build)
 * and -ffp-contract=off (no FMA contraction in the FP64 bilinear sum).
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* OpenCV primitives the reference calls (third-party boundary, OpenCV 3.0).  */

/* cv::borderInterpolate, used by Pyramid.cpp:56-57, 83, 463-464, 488. */
int orc_border(int p, int len, int type)
{
    if ((unsigned)p < (unsigned)len)
        return p;
    if (type == ORC_BORDER_REFLECT_101) {
        if (len == 1)
            return 0;
        do {
            if (p < 0)
                p = -p;
            else
                p = 2 * len - 2 - p;
        } while ((unsigned)p >= (unsigned)len);
        return p;
    }
    /* ORC_BORDER_WRAP */
    if (p < 0)
        p -= ((p - len + 1) / len) * len;
    if (p >= len)
        p %= len;
    return p;
}

/* cvRound on SSE2 (cvtsd2si, round half to even); out-of-range -> INT_MIN. */
static int orc_round(double v)
{
    if (!(v > -2147483648.5 && v < 2147483647.5))
        return INT_MIN;
    return (int)lrint(v);
}

static int16_t sat_s16(int v) { return (int16_t)(v < -32768 ? -32768 : v > 32767 ? 32767 : v); }
static uint8_t sat_u8(int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }

/* ------------------------------------------------------------------------ */
/* Pyramids: Blend/Pyramid.cpp.                                              */

/*
 * pyrDown_ (Pyramid.cpp:35-153 for the 32S sum, :319-439 for the casted
 * variants).  Horizontal 5-tap [1 4 6 4 1] with decimation into a ring of 5
 * row buffers, then the vertical 5-tap, exactly as the reference's loops; the
 * tab of source columns is borderInterpolate(2x + k - 2) for every dst column
 * (the reference's tabL / direct / tabR split yields the same indexes).
 * CASTK/DELTA: 0/0 = no cast (int32 sum); 8/128 = FixPtCast<T,8>: (s + 128) >> 8.
 */
#define ORC_DEFINE_PYRDOWN(NAME, ST, DT, CASTK, DELTA)                                   \
    void NAME(const ST* src, int sw, int sh, int sstep, int cn, DT* dst, int dw, int dh,  \
              int dstep, int hb, int vb)                                                  \
    {                                                                                     \
        int W = dw * cn;                                                                  \
        int* tab = (int*)malloc(sizeof(int) * 5 * (size_t)W);                             \
        int* buf = (int*)malloc(sizeof(int) * 5 * (size_t)W);                             \
        for (int x = 0; x < dw; x++)                                                      \
            for (int k = 0; k < 5; k++) {                                                 \
                int sx = orc_border(2 * x + k - 2, sw, hb);                               \
                for (int c = 0; c < cn; c++)                                              \
                    tab[k * W + x * cn + c] = sx * cn + c;                                \
            }                                                                             \
        int sy = -2;                                                                      \
        for (int y = 0; y < dh; y++) {                                                    \
            for (; sy <= 2 * y + 2; sy++) {                                               \
                int* row = buf + ((sy + 2) % 5) * W;                                      \
                const ST* s = src + (size_t)orc_border(sy, sh, vb) * sstep;               \
                for (int x = 0; x < W; x++)                                               \
                    row[x] = s[tab[2 * W + x]] * 6 + (s[tab[W + x]] + s[tab[3 * W + x]]) * 4 + \
                             s[tab[x]] + s[tab[4 * W + x]];                               \
            }                                                                             \
            const int* r0 = buf + ((2 * y + 0) % 5) * W;                                  \
            const int* r1 = buf + ((2 * y + 1) % 5) * W;                                  \
            const int* r2 = buf + ((2 * y + 2) % 5) * W;                                  \
            const int* r3 = buf + ((2 * y + 3) % 5) * W;                                  \
            const int* r4 = buf + ((2 * y + 4) % 5) * W;                                  \
            DT* d = dst + (size_t)y * dstep;                                              \
            for (int x = 0; x < W; x++) {                                                 \
                int s = r2[x] * 6 + (r1[x] + r3[x]) * 4 + r0[x] + r4[x];                  \
                d[x] = (DT)((s + DELTA) >> CASTK);                                        \
            }                                                                             \
        }                                                                                 \
        free(tab);                                                                        \
        free(buf);                                                                        \
    }

ORC_DEFINE_PYRDOWN(orc_pyrdown_to32s_s16, int16_t, int32_t, 0, 0)
ORC_DEFINE_PYRDOWN(orc_pyrdown_s16, int16_t, int16_t, 8, 128)

/*
 * pyrUp_<FixPtCast<T,6>> (Pyramid.cpp:441-536).  Each source row is expanded
 * horizontally into a ring of 3 row buffers (even dst column: s[x-1] + 6 s[x] +
 * s[x+1], odd: 4 (s[x] + s[x+1]), with lx = B(-1), rx = B(w) at the ends and
 * the width-1 special case 8 s); then dst row 2y = row(y-1) + 6 row(y) + row(y+1),
 * dst row 2y+1 = 4 (row(y) + row(y+1)), both (v + 32) >> 6; a dst row past the
 * end collapses onto row 2y (dst1 = dst0, written second).
 */
#define ORC_DEFINE_PYRUP(NAME, T)                                                        \
    void NAME(const T* src, int sw, int sh, int sstep, int cn, T* dst, int dw, int dh,    \
              int dstep, int hb, int vb)                                                  \
    {                                                                                     \
        int SW = sw * cn, DW = dw * cn;                                                   \
        int bstep = (dw + 1) * cn;                                                        \
        if (bstep < 2 * SW) bstep = 2 * SW;                                               \
        int* buf = (int*)calloc((size_t)3 * bstep, sizeof(int));                          \
        int lx = orc_border(-1, sw, hb) * cn;                                             \
        int rx = orc_border(sw, sw, hb) * cn;                                             \
        int sy = -1;                                                                      \
        for (int y = 0; y < sh; y++) {                                                    \
            T* d0 = dst + (size_t)(2 * y) * dstep;                                        \
            T* d1 = dst + (size_t)(2 * y + 1) * dstep;                                    \
            if (2 * y + 1 >= dh)                                                          \
                d1 = d0;                                                                  \
            for (; sy <= y + 1; sy++) {                                                   \
                int* row = buf + ((sy + 1) % 3) * bstep;                                  \
                const T* s = src + (size_t)orc_border(sy, sh, vb) * sstep;                \
                if (SW == cn) {                                                           \
                    for (int x = 0; x < cn; x++)                                          \
                        row[x] = row[x + cn] = s[x] * 8;                                  \
                    continue;                                                             \
                }                                                                         \
                for (int x = 0; x < cn; x++) {                                            \
                    int dx = x;                                                           \
                    int t0 = s[lx + x] + s[x] * 6 + s[x + cn];                            \
                    int t1 = (s[x] + s[x + cn]) * 4;                                      \
                    row[dx] = t0;                                                         \
                    row[dx + cn] = t1;                                                    \
                    int sx = SW - cn + x;                                                 \
                    dx = (sw - 1) * 2 * cn + x;                                           \
                    t0 = s[sx - cn] + s[sx] * 6 + s[rx + x];                              \
                    t1 = (s[sx] + s[rx + x]) * 4;                                         \
                    row[dx] = t0;                                                         \
                    row[dx + cn] = t1;                                                    \
                }                                                                         \
                for (int x = cn; x < SW - cn; x++) {                                      \
                    int dx = (x / cn) * 2 * cn + x % cn;                                  \
                    row[dx] = s[x - cn] + s[x] * 6 + s[x + cn];                           \
                    row[dx + cn] = (s[x] + s[x + cn]) * 4;                                \
                }                                                                         \
            }                                                                             \
            const int* r0 = buf + ((y + 0) % 3) * bstep;                                  \
            const int* r1 = buf + ((y + 1) % 3) * bstep;                                  \
            const int* r2 = buf + ((y + 2) % 3) * bstep;                                  \
            for (int x = 0; x < DW; x++) {                                                \
                T t1 = (T)(((r1[x] + r2[x]) * 4 + 32) >> 6);                              \
                T t0 = (T)((r0[x] + r1[x] * 6 + r2[x] + 32) >> 6);                        \
                d1[x] = t1;                                                               \
                d0[x] = t0;                                                               \
            }                                                                             \
        }                                                                                 \
        free(buf);                                                                        \
    }

ORC_DEFINE_PYRUP(orc_pyrup_s16, int16_t)
ORC_DEFINE_PYRUP(orc_pyrup_s32, int32_t)

/* getTrueNumLevels (MultiBandBlend.cpp:325-337). */
int orc_true_num_levels(int width, int height, int max_levels, int min_length)
{
    int num = max_levels < 0 ? 16 : (max_levels > 16 ? 16 : max_levels);
    int side = min_length < 0 ? 2 : (min_length > 64 ? 64 : min_length);
    int lw = (int)(log((double)width / side) / log(2.0));
    int lh = (int)(log((double)height / side) / log(2.0));
    int m = lw < lh ? lw : lh;
    return num < m ? num : m;
}

/* ------------------------------------------------------------------------ */
/* Seams: Blend/Mask.cpp:376-397 over cv::distanceTransform(L1, 3x3).        */

/*
 * OpenCV 3.0 distanceTransform_3x3 with the L1 metrics {1, 2}: fixed point
 * (16 fractional bits), a one-pixel frame initialised to INT_MAX >> 2 so that
 * outside the image counts as far from any zero, a forward pass over the
 * upper/left neighbours and a backward pass over the lower/right ones, then
 * float(t) * 2^-16.
 */
void orc_distance_transform_l1(const uint8_t* src, int w, int h, float* dst)
{
    const int HV = 1 << 16, DIAG = 2 << 16, INIT = INT_MAX >> 2;
    const float scale = 1.f / (1 << 16);
    int tw = w + 2;
    int* t = (int*)malloc(sizeof(int) * (size_t)tw * (h + 2));
    for (int j = 0; j < tw; j++) {
        t[j] = INIT;
        t[(size_t)(h + 1) * tw + j] = INIT;
    }
    for (int i = 0; i < h; i++) {
        const uint8_t* s = src + (size_t)i * w;
        int* r = t + (size_t)(i + 1) * tw + 1;
        r[-1] = r[w] = INIT;
        for (int j = 0; j < w; j++) {
            if (!s[j]) {
                r[j] = 0;
            } else {
                int v = r[j - tw - 1] + DIAG, u;
                u = r[j - tw] + HV; if (v > u) v = u;
                u = r[j - tw + 1] + DIAG; if (v > u) v = u;
                u = r[j - 1] + HV; if (v > u) v = u;
                r[j] = v;
            }
        }
    }
    for (int i = h - 1; i >= 0; i--) {
        int* r = t + (size_t)(i + 1) * tw + 1;
        float* d = dst + (size_t)i * w;
        for (int j = w - 1; j >= 0; j--) {
            int v = r[j], u;
            if (v > HV) {
                u = r[j + tw + 1] + DIAG; if (v > u) v = u;
                u = r[j + tw] + HV; if (v > u) v = u;
                u = r[j + tw - 1] + DIAG; if (v > u) v = u;
                u = r[j + 1] + HV; if (v > u) v = u;
                r[j] = v;
            }
            d[j] = (float)v * scale;
        }
    }
    free(t);
}

/* getNonIntersectingMasks (Mask.cpp:376-397). */
void orc_nonintersecting_masks(int n, int w, int h, const uint8_t* masks, uint8_t* out)
{
    size_t P = (size_t)w * h;
    memcpy(out, masks, P * n);
    if (n <= 0)
        return;
    uint8_t* acc = (uint8_t*)malloc(P);
    float* dist = (float*)malloc(sizeof(float) * P);
    float* cur = (float*)malloc(sizeof(float) * P);
    memcpy(acc, out, P);
    for (int i = 1; i < n; i++) {
        uint8_t* ui = out + P * i;
        orc_distance_transform_l1(acc, w, h, dist);
        orc_distance_transform_l1(ui, w, h, cur);
        for (size_t p = 0; p < P; p++)
            ui[p] = cur[p] > dist[p] ? 255 : 0;
        for (int j = 0; j < i; j++) {
            uint8_t* uj = out + P * j;
            for (size_t p = 0; p < P; p++)
                if (ui[p])
                    uj[p] = 0;
        }
        for (size_t p = 0; p < P; p++)
            acc[p] |= ui[p];
    }
    free(acc);
    free(dist);
    free(cur);
}

/* ------------------------------------------------------------------------ */
/* TilingMultibandBlendFast (MultiBandBlend.cpp:906-1031).                   */

struct orc_blender {
    int n, rows, cols, levels, full;
    int lw[17], lh[17];
    uint8_t* unique;     /* n * rows * cols */
    int16_t** weight;    /* [n * (levels+1)] */
    int32_t** alpha;     /* [n * (levels+1)], level 0 unused */
    int32_t** rweight;   /* [levels+1] when !full */
    uint8_t* covered;    /* OR of masks */
};

orc_blender* orc_blender_create(void) { return (orc_blender*)calloc(1, sizeof(orc_blender)); }

static void orc_blender_clear(orc_blender* b)
{
    int L1 = b->levels + 1;
    if (b->weight)
        for (int i = 0; i < b->n * L1; i++) free(b->weight[i]);
    if (b->alpha)
        for (int i = 0; i < b->n * L1; i++) free(b->alpha[i]);
    if (b->rweight)
        for (int i = 0; i < L1; i++) free(b->rweight[i]);
    free(b->weight);
    free(b->alpha);
    free(b->rweight);
    free(b->unique);
    free(b->covered);
    memset(b, 0, sizeof(*b));
}

void orc_blender_destroy(orc_blender* b)
{
    if (!b) return;
    orc_blender_clear(b);
    free(b);
}

/* prepare (MultiBandBlend.cpp:906-985). */
int orc_blender_prepare(orc_blender* b, int n, int rows, int cols, const uint8_t* masks,
                        int max_levels, int min_length)
{
    orc_blender_clear(b);
    if (n <= 0 || rows <= 0 || cols <= 0 || !masks)
        return 0;
    size_t P = (size_t)rows * cols;
    b->n = n;
    b->rows = rows;
    b->cols = cols;
    b->unique = (uint8_t*)malloc(P * n);
    orc_nonintersecting_masks(n, cols, rows, masks, b->unique);
    b->levels = orc_true_num_levels(cols, rows, max_levels, min_length);
    int L = b->levels;
    b->lw[0] = cols;
    b->lh[0] = rows;
    for (int j = 1; j <= L; j++) {
        b->lw[j] = (b->lw[j - 1] + 1) / 2;
        b->lh[j] = (b->lh[j - 1] + 1) / 2;
    }
    b->weight = (int16_t**)calloc((size_t)n * (L + 1), sizeof(int16_t*));
    b->alpha = (int32_t**)calloc((size_t)n * (L + 1), sizeof(int32_t*));
    int16_t* tmp16 = (int16_t*)malloc(sizeof(int16_t) * P);
    for (int i = 0; i < n; i++) {
        /* weight Gaussian pyramid of 256 * unique (:936-943, createGaussPyramid :64-74) */
        int16_t** wp = b->weight + (size_t)i * (L + 1);
        wp[0] = (int16_t*)malloc(sizeof(int16_t) * P);
        const uint8_t* u = b->unique + P * i;
        for (size_t p = 0; p < P; p++) wp[0][p] = u[p] ? 256 : 0;
        for (int j = 0; j < L; j++) {
            wp[j + 1] = (int16_t*)malloc(sizeof(int16_t) * (size_t)b->lw[j + 1] * b->lh[j + 1]);
            orc_pyrdown_s16(wp[j], b->lw[j], b->lh[j], b->lw[j], 1, wp[j + 1], b->lw[j + 1],
                            b->lh[j + 1], b->lw[j + 1], ORC_BORDER_WRAP, ORC_BORDER_REFLECT_101);
        }
        /* alpha pyramid, re-binarised to 256 at each level (:945-959, :714-730) */
        int32_t** ap = b->alpha + (size_t)i * (L + 1);
        const uint8_t* m = masks + P * i;
        for (size_t p = 0; p < P; p++) tmp16[p] = m[p] ? 256 : 0;
        for (int j = 0; j < L; j++) {
            size_t np = (size_t)b->lw[j + 1] * b->lh[j + 1];
            ap[j + 1] = (int32_t*)malloc(sizeof(int32_t) * np);
            orc_pyrdown_to32s_s16(tmp16, b->lw[j], b->lh[j], b->lw[j], 1, ap[j + 1],
                                  b->lw[j + 1], b->lh[j + 1], b->lw[j + 1], ORC_BORDER_WRAP,
                                  ORC_BORDER_REFLECT_101);
            for (size_t p = 0; p < np; p++) tmp16[p] = ap[j + 1][p] ? 256 : 0;
        }
    }
    free(tmp16);
    /* coverage (:961-981) */
    b->covered = (uint8_t*)calloc(P, 1);
    size_t cnt = 0;
    for (size_t p = 0; p < P; p++) {
        uint8_t v = 0;
        for (int i = 0; i < n; i++) v |= masks[P * i + p];
        b->covered[p] = v;
        cnt += v != 0;
    }
    b->full = cnt == P;
    if (!b->full) {
        b->rweight = (int32_t**)calloc((size_t)L + 1, sizeof(int32_t*));
        for (int j = 0; j <= L; j++) {
            size_t np = (size_t)b->lw[j] * b->lh[j];
            b->rweight[j] = (int32_t*)calloc(np, sizeof(int32_t));
            for (int i = 0; i < n; i++) {
                const int16_t* w = b->weight[(size_t)i * (L + 1) + j];
                for (size_t p = 0; p < np; p++) b->rweight[j][p] += w[p];
            }
        }
    }
    return 1;
}

/* calcDstImage (MultiBandBlend.cpp:732-768): a ? ((v << 8) - v) / a : 0. */
static void orc_calc_dst(const int32_t* s32, const int32_t* alpha, int16_t* dst, size_t np)
{
    for (size_t p = 0; p < np; p++) {
        int a = alpha[p];
        for (int c = 0; c < 3; c++) {
            int v = s32[p * 3 + c];
            dst[p * 3 + c] = a ? (int16_t)((int)(((unsigned)v << 8) - (unsigned)v) / a) : 0;
        }
    }
}

/* Per-camera Laplacian pyramid (:1004-1019, TilingMultibandBlendFastParallel::buildPyramid
 * :1462-1476): down with calcDstImage, then L_j = sat16(G_j - up(G_{j+1})). */
static void orc_build_laplace(const orc_blender* b, int i, const int16_t* image, int16_t** pyr)
{
    int L = b->levels;
    size_t P = (size_t)b->rows * b->cols;
    memcpy(pyr[0], image, sizeof(int16_t) * P * 3);
    int32_t* s32 = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)b->lw[1] * b->lh[1]);
    for (int j = 0; j < L; j++) {
        orc_pyrdown_to32s_s16(pyr[j], b->lw[j], b->lh[j], b->lw[j] * 3, 3, s32, b->lw[j + 1],
                              b->lh[j + 1], b->lw[j + 1] * 3, ORC_BORDER_WRAP,
                              ORC_BORDER_REFLECT_101);
        orc_calc_dst(s32, b->alpha[(size_t)i * (L + 1) + j + 1], pyr[j + 1],
                     (size_t)b->lw[j + 1] * b->lh[j + 1]);
    }
    free(s32);
    int16_t* up = (int16_t*)malloc(sizeof(int16_t) * P * 3);
    for (int j = 0; j < L; j++) {
        orc_pyrup_s16(pyr[j + 1], b->lw[j + 1], b->lh[j + 1], b->lw[j + 1] * 3, 3, up, b->lw[j],
                      b->lh[j], b->lw[j] * 3, ORC_BORDER_WRAP, ORC_BORDER_REFLECT_101);
        size_t ne = (size_t)b->lw[j] * b->lh[j] * 3;
        for (size_t e = 0; e < ne; e++) pyr[j][e] = sat_s16(pyr[j][e] - up[e]);
    }
    free(up);
}

/* blend (MultiBandBlend.cpp:987-1031) with the weight pyramids `weight`
 * (n * (L + 1) planes); rweight NULL normalises by >> 8 (fullMask), else by the
 * weight sums; pixels with covered[p] == 0 are zeroed unless full. */
static int blend_with(orc_blender* b, const int16_t* images, int16_t* const* weight,
                      int32_t* const* rweight, const uint8_t* covered, int full, uint8_t* out,
                      int nthreads)
{
    int L = b->levels, n = b->n;
    size_t P = (size_t)b->rows * b->cols;
    int32_t** R = (int32_t**)calloc((size_t)L + 1, sizeof(int32_t*));
    for (int j = 0; j <= L; j++)
        R[j] = (int32_t*)calloc((size_t)b->lw[j] * b->lh[j] * 3, sizeof(int32_t));
    int par = nthreads > 1 ? (nthreads < n ? nthreads : n) : 1;
    int16_t*** pyrs = (int16_t***)calloc((size_t)par, sizeof(int16_t**));
    for (int t = 0; t < par; t++) {
        pyrs[t] = (int16_t**)calloc((size_t)L + 1, sizeof(int16_t*));
        for (int j = 0; j <= L; j++)
            pyrs[t][j] = (int16_t*)malloc(sizeof(int16_t) * 3 * (size_t)b->lw[j] * b->lh[j]);
    }
    for (int i0 = 0; i0 < n; i0 += par) {
        int cnt = n - i0 < par ? n - i0 : par;
#ifdef _OPENMP
#pragma omp parallel for num_threads(cnt) if (cnt > 1)
#endif
        for (int t = 0; t < cnt; t++)
            orc_build_laplace(b, i0 + t, images + P * 3 * (size_t)(i0 + t), pyrs[t]);
        /* accumulate (:106-139): R_j += L_j * W_j, int32 += int16 * int16 */
        for (int t = 0; t < cnt; t++) {
            int i = i0 + t;
            for (int j = 0; j <= L; j++) {
                size_t np = (size_t)b->lw[j] * b->lh[j];
                const int16_t* w = weight[(size_t)i * (L + 1) + j];
                const int16_t* l = pyrs[t][j];
                int32_t* r = R[j];
                for (size_t p = 0; p < np; p++) {
                    int wv = w[p];
                    r[p * 3 + 0] += l[p * 3 + 0] * wv;
                    r[p * 3 + 1] += l[p * 3 + 1] * wv;
                    r[p * 3 + 2] += l[p * 3 + 2] * wv;
                }
            }
        }
    }
    for (int t = 0; t < par; t++) {
        for (int j = 0; j <= L; j++) free(pyrs[t][j]);
        free(pyrs[t]);
    }
    free(pyrs);
    /* normalize (:141-161 for fullMask, :203-241 otherwise) */
    for (int j = 0; j <= L; j++) {
        size_t np = (size_t)b->lw[j] * b->lh[j];
        int32_t* r = R[j];
        if (!rweight) {
            for (size_t e = 0; e < np * 3; e++) r[e] = r[e] >> 8;
        } else {
            const int32_t* w = rweight[j];
            for (size_t p = 0; p < np; p++) {
                int wv = w[p];
                if (!wv) wv++;
                r[p * 3 + 0] /= wv;
                r[p * 3 + 1] /= wv;
                r[p * 3 + 2] /= wv;
            }
        }
    }
    /* restoreImageFromLaplacePyramid (:770-780): C_{j-1} = R_{j-1} + up32(C_j) */
    int32_t* up = (int32_t*)malloc(sizeof(int32_t) * P * 3);
    for (int j = L; j > 0; j--) {
        orc_pyrup_s32(R[j], b->lw[j], b->lh[j], b->lw[j] * 3, 3, up, b->lw[j - 1], b->lh[j - 1],
                      b->lw[j - 1] * 3, ORC_BORDER_WRAP, ORC_BORDER_REFLECT_101);
        size_t ne = (size_t)b->lw[j - 1] * b->lh[j - 1] * 3;
        for (size_t e = 0; e < ne; e++) R[j - 1][e] = up[e] + R[j - 1][e];
    }
    free(up);
    /* convertTo(CV_8U) and setTo(0, maskNot) (:1028-1030) */
    for (size_t p = 0; p < P; p++)
        for (int c = 0; c < 3; c++)
            out[p * 3 + c] = (!full && !covered[p]) ? 0 : sat_u8(R[0][p * 3 + c]);
    for (int j = 0; j <= L; j++) free(R[j]);
    free(R);
    return 1;
}

int orc_blender_blend(orc_blender* b, const int16_t* images, uint8_t* out, int nthreads)
{
    if (!b || !b->n)
        return 0;
    return blend_with(b, images, b->weight, b->full ? NULL : b->rweight, b->covered, b->full, out,
                      nthreads);
}

/* blend(images, masks, blendImage) (MultiBandBlend.cpp:1033-1102): weights
 * from the call's masks (Gaussian pyramids of 256 * mask, createGaussPyramid
 * :64-74), always normalised by their sums (normalize(pyr, weight) :203-241),
 * alphas from prepare; zero where no call mask covers the pixel unless they
 * cover the canvas. */
int orc_blender_blend_custom(orc_blender* b, const int16_t* images, const uint8_t* masks,
                             uint8_t* out, int nthreads)
{
    if (!b || !b->n)
        return 0;
    const int L = b->levels, n = b->n;
    const size_t P = (size_t)b->rows * b->cols;
    int16_t** w = (int16_t**)calloc((size_t)n * (L + 1), sizeof(int16_t*));
    int32_t** rw = (int32_t**)calloc((size_t)L + 1, sizeof(int32_t*));
    uint8_t* cov = (uint8_t*)calloc(P, 1);
    size_t cnt = 0;
    for (size_t p = 0; p < P; p++) {
        uint8_t v = 0;
        for (int i = 0; i < n; i++) v |= masks[P * i + p];
        cov[p] = v;
        cnt += v != 0;
    }
    for (int i = 0; i < n; i++) {
        int16_t** wp = w + (size_t)i * (L + 1);
        wp[0] = (int16_t*)malloc(sizeof(int16_t) * P);
        for (size_t p = 0; p < P; p++) wp[0][p] = masks[P * i + p] ? 256 : 0;
        for (int j = 0; j < L; j++) {
            wp[j + 1] = (int16_t*)malloc(sizeof(int16_t) * (size_t)b->lw[j + 1] * b->lh[j + 1]);
            orc_pyrdown_s16(wp[j], b->lw[j], b->lh[j], b->lw[j], 1, wp[j + 1], b->lw[j + 1],
                            b->lh[j + 1], b->lw[j + 1], ORC_BORDER_WRAP, ORC_BORDER_REFLECT_101);
        }
    }
    for (int j = 0; j <= L; j++) {
        size_t np = (size_t)b->lw[j] * b->lh[j];
        rw[j] = (int32_t*)calloc(np, sizeof(int32_t));
        for (int i = 0; i < n; i++)
            for (size_t p = 0; p < np; p++) rw[j][p] += w[(size_t)i * (L + 1) + j][p];
    }
    int ok = blend_with(b, images, w, rw, cov, cnt == P, out, nthreads);
    for (size_t k = 0; k < (size_t)n * (L + 1); k++) free(w[k]);
    for (int j = 0; j <= L; j++) free(rw[j]);
    free(w);
    free(rw);
    free(cov);
    return ok;
}

int orc_blender_num_levels(const orc_blender* b) { return b->levels; }
int orc_blender_full_mask(const orc_blender* b) { return b->full; }
void orc_blender_unique_masks(const orc_blender* b, uint8_t* out)
{
    memcpy(out, b->unique, (size_t)b->rows * b->cols * b->n);
}
int orc_blender_level_size(const orc_blender* b, int level, int* w, int* h)
{
    if (level < 0 || level > b->levels) return 0;
    *w = b->lw[level];
    *h = b->lh[level];
    return 1;
}
int orc_blender_weight(const orc_blender* b, int cam, int level, int16_t* out)
{
    if (cam < 0 || cam >= b->n || level < 0 || level > b->levels) return 0;
    memcpy(out, b->weight[(size_t)cam * (b->levels + 1) + level],
           sizeof(int16_t) * (size_t)b->lw[level] * b->lh[level]);
    return 1;
}
int orc_blender_alpha(const orc_blender* b, int cam, int level, int32_t* out)
{
    if (cam < 0 || cam >= b->n || level < 1 || level > b->levels) return 0;
    memcpy(out, b->alpha[(size_t)cam * (b->levels + 1) + level],
           sizeof(int32_t) * (size_t)b->lw[level] * b->lh[level]);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Warp: ZReproject.cpp.                                                     */

/*
 * bilinearResampling<short, cn> (ZReproject.cpp:101-130) for one valid map
 * point.  x0 = (int)x, y1 = min(y0 + 1, h - 1); the second sample of each row
 * is the next pixel in row-major order (the reference advances its pointer
 * past x0 and never uses its clamped x1, ZReproject.cpp:117-126, on images
 * that are continuous -- transform() allocates a continuous image): at
 * x0 = w - 1 that is pixel (0, y + 1); at the last pixel of the image, where
 * the reference would read past its buffer, the pixel itself is used.
 * Sum order p00 w00, p01 w01, p10 w10, p11 w11; saturate_cast<short> = cvRound.
 * Callers here always pass continuous images (sstep == sw * cn).
 */
static void orc_bilinear_s16(const uint8_t* src, int sw, int sh, size_t sstep, int cn, double x,
                             double y, int16_t* d)
{
    int x0 = (int)x, y0 = (int)y, y1 = y0 + 1;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (y1 > sh - 1) y1 = sh - 1;
    double wx0 = x - x0, wx1 = 1 - wx0;
    double wy0 = y - y0, wy1 = 1 - wy0;
    double w00 = wx1 * wy1, w01 = wx0 * wy1;
    double w10 = wx1 * wy0, w11 = wx0 * wy0;
    size_t end = sstep * (size_t)(sh - 1) + (size_t)sw * cn;
    size_t o0 = sstep * (size_t)y0 + (size_t)x0 * cn, o1 = sstep * (size_t)y1 + (size_t)x0 * cn;
    size_t n0 = o0 + cn < end ? o0 + cn : o0, n1 = o1 + cn < end ? o1 + cn : o1;
    for (int c = 0; c < cn; c++) {
        double r = 0;
        r += src[o0 + c] * w00;
        r += src[n0 + c] * w01;
        r += src[o1 + c] * w10;
        r += src[n1 + c] * w11;
        d[c] = sat_s16(orc_round(r));
    }
}

/* reprojectParallelTo16S (ZReproject.cpp:535-567, loop :448-491) with a map
 * built from two float32 planes. */
void orc_reproject_to16s(const uint8_t* src, int sw, int sh, size_t sstep, int cn,
                         const float* xmap, const float* ymap, int dw, int dh, int16_t* dst,
                         int nthreads)
{
    memset(dst, 0, sizeof(int16_t) * (size_t)dw * dh * cn);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int h = 0; h < dh; h++) {
        for (int w = 0; w < dw; w++) {
            double x = xmap[(size_t)h * dw + w], y = ymap[(size_t)h * dw + w];
            if (x >= 0 && y >= 0 && x < sw && y < sh)
                orc_bilinear_s16(src, sw, sh, sstep, cn, x, y, dst + ((size_t)h * dw + w) * cn);
        }
    }
}

void orc_reproject_to16s_map64(const uint8_t* src, int sw, int sh, size_t sstep, int cn,
                               const double* map_xy, int dw, int dh, int16_t* dst, int nthreads)
{
    memset(dst, 0, sizeof(int16_t) * (size_t)dw * dh * cn);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int h = 0; h < dh; h++) {
        for (int w = 0; w < dw; w++) {
            double x = map_xy[((size_t)h * dw + w) * 2], y = map_xy[((size_t)h * dw + w) * 2 + 1];
            if (x >= 0 && y >= 0 && x < sw && y < sh)
                orc_bilinear_s16(src, sw, sh, sstep, cn, x, y, dst + ((size_t)h * dw + w) * cn);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Gains.                                                                    */

/* getLUT (ExposureColorOptimize.cpp:1217-1223): saturate_cast<uchar>(k * i). */
void orc_get_lut(uint8_t* lut, double k)
{
    for (int i = 0; i < 256; i++)
        lut[i] = sat_u8(orc_round(k * i));
}

/* transform (GainAdjust.cpp:295-346, no mask): per-channel LUT on BGR;
 * channels beyond the third are copied. */
void orc_transform(const uint8_t* src, uint8_t* dst, int w, int h, size_t step, int cn,
                   const uint8_t* lut3)
{
    for (int y = 0; y < h; y++) {
        const uint8_t* s = src + step * y;
        uint8_t* d = dst + step * y;
        for (int x = 0; x < w; x++)
            for (int c = 0; c < cn; c++)
                d[x * cn + c] = c < 3 ? lut3[c * 256 + s[x * cn + c]] : s[x * cn + c];
    }
}

/* ------------------------------------------------------------------------ */
/* Camera model: Warp/Remap.cpp.                                             */

static const double ORC_PI = 3.14159265358979323846264338327950288;
#define ORC_DEG_TO_RAD(x) ((x) * 2.0 * ORC_PI / 360.0)

typedef struct {
    int type; /* PanoTools image type: 0 rectilinear, 2 circular fisheye, 3 full frame */
    double srcTX, srcTY, destTX, destTY;
    double scale[2], shear[2], rot[2], rad[6], mt[3][3];
    double distance, horizontal, vertical;
} orc_remap;

static void orc_mat_mult(double a[3][3], double b[3][3], double r[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++)
            r[i][k] = a[i][0] * b[0][k] + a[i][1] * b[1][k] + a[i][2] * b[2][k];
}

/* SetMatrix (Remap.cpp:22-43). */
static void orc_set_matrix(double a, double b, double c, double m[3][3], int cl)
{
    double mx[3][3], my[3][3], mz[3][3], t[3][3];
    mx[0][0] = 1.0; mx[0][1] = 0.0; mx[0][2] = 0.0;
    mx[1][0] = 0.0; mx[1][1] = cos(a); mx[1][2] = sin(a);
    mx[2][0] = 0.0; mx[2][1] = -mx[1][2]; mx[2][2] = mx[1][1];
    my[0][0] = cos(b); my[0][1] = 0.0; my[0][2] = -sin(b);
    my[1][0] = 0.0; my[1][1] = 1.0; my[1][2] = 0.0;
    my[2][0] = -my[0][2]; my[2][1] = 0.0; my[2][2] = my[0][0];
    mz[0][0] = cos(c); mz[0][1] = sin(c); mz[0][2] = 0.0;
    mz[1][0] = -mz[0][1]; mz[1][1] = mz[0][0]; mz[1][2] = 0.0;
    mz[2][0] = 0.0; mz[2][1] = 0.0; mz[2][2] = 1.0;
    if (cl)
        orc_mat_mult(mz, mx, t);
    else
        orc_mat_mult(mx, mz, t);
    orc_mat_mult(t, my, m);
}

/* squareZero / cubeZero / smallestRoot (Remap.cpp:45-135). */
static void orc_square_zero(double* a, int* n, double* root)
{
    if (a[2] == 0.0) {
        if (a[1] == 0.0) {
            if (a[0] == 0.0) { *n = 1; root[0] = 0.0; }
            else *n = 0;
        } else {
            *n = 1;
            root[0] = -a[0] / a[1];
        }
    } else {
        if (4.0 * a[2] * a[0] > a[1] * a[1]) {
            *n = 0;
        } else {
            *n = 2;
            root[0] = (-a[1] + sqrt(a[1] * a[1] - 4.0 * a[2] * a[0])) / (2.0 * a[2]);
            root[1] = (-a[1] - sqrt(a[1] * a[1] - 4.0 * a[2] * a[0])) / (2.0 * a[2]);
        }
    }
}

static double orc_cube_root(double x)
{
    if (x == 0.0) return 0.0;
    if (x > 0.0) return pow(x, 1.0 / 3.0);
    return -pow(-x, 1.0 / 3.0);
}

static void orc_cube_zero(double* a, int* n, double* root)
{
    if (a[3] == 0.0) {
        orc_square_zero(a, n, root);
    } else {
        double p = ((-1.0 / 3.0) * (a[2] / a[3]) * (a[2] / a[3]) + a[1] / a[3]) / 3.0;
        double q = ((2.0 / 27.0) * (a[2] / a[3]) * (a[2] / a[3]) * (a[2] / a[3]) -
                    (1.0 / 3.0) * (a[2] / a[3]) * (a[1] / a[3]) + a[0] / a[3]) / 2.0;
        if (q * q + p * p * p >= 0.0) {
            *n = 1;
            root[0] = orc_cube_root(-q + sqrt(q * q + p * p * p)) +
                      orc_cube_root(-q - sqrt(q * q + p * p * p)) - a[2] / (3.0 * a[3]);
        } else {
            double phi = acos(-q / sqrt(-p * p * p));
            *n = 3;
            root[0] = 2.0 * sqrt(-p) * cos(phi / 3.0) - a[2] / (3.0 * a[3]);
            root[1] = -2.0 * sqrt(-p) * cos(phi / 3.0 + ORC_PI / 3.0) - a[2] / (3.0 * a[3]);
            root[2] = -2.0 * sqrt(-p) * cos(phi / 3.0 - ORC_PI / 3.0) - a[2] / (3.0 * a[3]);
        }
    }
}

static double orc_smallest_root(double* p)
{
    int n;
    double root[3], s = 1000.0;
    orc_cube_zero(p, &n, root);
    for (int i = 0; i < n; i++)
        if (root[i] > 0.0 && root[i] < s) s = root[i];
    return s;
}

/* Remap::init(param, W_eq, H_eq, W_cam, H_cam) (Remap.cpp:147-249); the
 * reference's "src" is the equirect canvas and its "dst" the camera. */
static void orc_remap_init(orc_remap* r, const orc_photo_param* p0, int eqW, int eqH, int camW,
                           int camH)
{
    (void)eqH;
    memset(r, 0, sizeof(*r));
    r->srcTX = eqW / 2;
    r->srcTY = eqH / 2;
    orc_photo_param p = *p0;
    if (p.imageType == 0 || p.imageType == 1) {
        p.cropX = 0;
        p.cropY = 0;
        p.cropWidth = camW;
        p.cropHeight = camH;
    }
    if (p.imageType == 0) r->type = 0;
    else if (p.imageType == 1) r->type = 3;
    else r->type = 2;
    r->destTX = p.cropX + p.cropWidth / 2;
    r->destTY = p.cropY + p.cropHeight / 2;
    r->horizontal = p.shiftX != 0 ? p.shiftX : 0;
    r->vertical = p.shiftY != 0 ? p.shiftY : 0;
    double a = ORC_DEG_TO_RAD(p.hfov);
    double b = ORC_DEG_TO_RAD(360);
    orc_set_matrix(-ORC_DEG_TO_RAD(p.pitch), 0.0, -ORC_DEG_TO_RAD(p.roll), r->mt, 0);
    r->distance = ((double)eqW) / b;
    if (r->type == 0)
        r->scale[0] = ((double)p.cropWidth) / (2.0 * tan(a / 2.0)) / r->distance;
    else
        r->scale[0] = ((double)p.cropWidth) / a / r->distance;
    r->scale[1] = r->scale[0];
    r->shear[0] = p.shearX / p.cropHeight;
    r->shear[1] = p.shearY / p.cropWidth;
    r->rot[0] = r->distance * ORC_PI;
    r->rot[1] = -p.yaw * r->distance * ORC_PI / 180.0;
    double rp[3][5];
    for (int i = 0; i < 3; i++) {
        rp[i][0] = 1.0;
        rp[i][4] = 1000.0;
        for (int k = 1; k < 4; k++) rp[i][k] = 0.0;
    }
    if (p.beta != 0.0 || p.gamma != 0.0 || p.alpha != 0.0) {
        for (int i = 0; i < 3; i++) {
            rp[i][3] = p.alpha;
            rp[i][2] = p.beta;
            rp[i][1] = p.gamma;
            rp[i][0] = 1.0 - (p.alpha + p.beta + p.gamma);
        }
    }
    for (int i = 0; i < 3; i++) {
        double t[4];
        for (int k = 0; k < 4; k++) {
            t[k] = 0.0;
            if (rp[i][k] != 0.0) t[k] = (k + 1) * rp[i][k];
        }
        rp[i][4] = orc_smallest_root(t);
    }
    for (int i = 0; i < 4; i++) r->rad[i] = rp[0][i];
    r->rad[5] = rp[0][4];
    r->rad[4] = ((double)(p.cropWidth < p.cropHeight ? p.cropWidth : p.cropHeight)) / 2.0;
}

/* Remap::remapImage (Remap.cpp:356-514). */
static void orc_remap_image(const orc_remap* m, double* xd, double* yd, double xs, double ys)
{
    xs -= m->srcTX - 0.5;
    ys -= m->srcTY - 0.5;
    double x = xs + m->rot[1];
    while (x < -m->rot[0]) x += 2 * m->rot[0];
    while (x > m->rot[0]) x -= 2 * m->rot[0];
    double y = ys;
    double phi = x / m->distance;
    double theta = -y / m->distance + ORC_PI / 2;
    if (theta < 0) {
        theta = -theta;
        phi += ORC_PI;
    }
    if (theta > ORC_PI) {
        theta = ORC_PI - (theta - ORC_PI);
        phi += ORC_PI;
    }
    double v0 = sin(theta) * sin(phi);
    double v1 = cos(theta);
    double v2 = sin(theta) * cos(phi);
    double v[3];
    for (int i = 0; i < 3; i++)
        v[i] = m->mt[0][i] * v0 + m->mt[1][i] * v1 + m->mt[2][i] * v2;
    double r = sqrt(v[0] * v[0] + v[1] * v[1]);
    if (r == 0.0)
        theta = 0.0;
    else
        theta = m->distance * atan2(r, v[2]) / r;
    x = theta * v[0];
    y = theta * v[1];
    if (m->type == 0) {
        double rr = sqrt(x * x + y * y);
        double th = rr / m->distance, rho;
        if (th >= ORC_PI / 2.0) rho = 1.6e16;
        else if (th == 0.0) rho = 1.0;
        else rho = tan(th) / th;
        x = rho * x;
        y = rho * y;
    }
    double xx = x * m->scale[0], yy = y * m->scale[1];
    x = xx;
    y = yy;
    double rt = (sqrt(x * x + y * y)) / m->rad[4], sc;
    if (rt < m->rad[5])
        sc = ((m->rad[3] * rt + m->rad[2]) * rt + m->rad[1]) * rt + m->rad[0];
    else
        sc = 1000.0;
    x = x * sc;
    y = y * sc;
    if (m->vertical != 0.0) y = y + m->vertical;
    if (m->horizontal != 0.0) x = x + m->horizontal;
    if (m->shear[0] != 0 || m->shear[1] != 0) {
        double nx = x + m->shear[0] * y;
        double ny = y + m->shear[1] * x;
        x = nx;
        y = ny;
    }
    x += m->destTX - 0.5;
    y += m->destTY - 0.5;
    *xd = x;
    *yd = y;
}

void orc_remap_point(const orc_photo_param* p, int src_w, int src_h, int dst_w, int dst_h,
                     double x, double y, double* sx, double* sy)
{
    orc_remap m;
    orc_remap_init(&m, p, dst_w, dst_h, src_w, src_h);
    orc_remap_image(&m, sx, sy, x, y);
}

/* getReprojectMapAndMask (ZReproject.cpp:132-219). */
void orc_reproject_map_and_mask(const orc_photo_param* p, int src_w, int src_h, int dst_w,
                                int dst_h, double* map_xy, uint8_t* mask, int nthreads)
{
    orc_remap m;
    orc_remap_init(&m, p, dst_w, dst_h, src_w, src_h);
    int circ = p->imageType == 3 || p->imageType == 2;
    double cx = 0, cy = 0, sq = 0;
    if (circ) {
        if (p->circleR > 0) {
            cx = p->circleX;
            cy = p->circleY;
            sq = (double)p->circleR * p->circleR;
        } else {
            cx = p->cropX + p->cropWidth / 2;
            cy = p->cropY + p->cropHeight / 2;
            sq = p->cropWidth > p->cropHeight ? p->cropWidth * p->cropWidth * 0.25
                                              : p->cropHeight * p->cropHeight * 0.25;
        }
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int h = 0; h < dst_h; h++) {
        for (int w = 0; w < dst_w; w++) {
            double sx, sy;
            orc_remap_image(&m, &sx, &sy, w, h);
            int ok;
            if (!circ)
                ok = sx >= 0 && sy >= 0 && sx < src_w && sy < src_h;
            else
                ok = sx >= p->cropX && sy >= p->cropY && sx < p->cropX + p->cropWidth &&
                     sy < p->cropY + p->cropHeight &&
                     (sx - cx) * (sx - cx) + (sy - cy) * (sy - cy) < sq;
            size_t o = (size_t)h * dst_w + w;
            mask[o] = ok ? 255 : 0;
            map_xy[o * 2] = ok ? sx : -1;
            map_xy[o * 2 + 1] = ok ? sy : -1;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CPUPanoramaRender::render, multiband branch (RicohUtil.cpp:1561-1578).     */
int orc_render(orc_blender* b, int n, const uint8_t* const* srcs, const int* src_w,
               const int* src_h, const float* const* xmaps, const float* const* ymaps,
               const uint8_t* luts, uint8_t* dst, int nthreads)
{
    if (!b || b->n != n)
        return 0;
    size_t P = (size_t)b->rows * b->cols;
    int16_t* rep = (int16_t*)malloc(sizeof(int16_t) * P * 3 * n);
    for (int i = 0; i < n; i++) {
        const uint8_t* s = srcs[i];
        uint8_t* corrected = NULL;
        if (luts) {
            corrected = (uint8_t*)malloc((size_t)src_w[i] * src_h[i] * 3);
            orc_transform(s, corrected, src_w[i], src_h[i], (size_t)src_w[i] * 3, 3,
                          luts + (size_t)i * 768);
            s = corrected;
        }
        orc_reproject_to16s(s, src_w[i], src_h[i], (size_t)src_w[i] * 3, 3, xmaps[i], ymaps[i],
                            b->cols, b->rows, rep + P * 3 * i, nthreads);
        free(corrected);
    }
    int ok = orc_blender_blend(b, rep, dst, nthreads);
    free(rep);
    return ok;
}
