/*
 * yuv.c -- CPU restatement of the encoder / decoder hand-off colour
 * conversions of CudaAccel/ImageProc.cu: cvtBGR32ToYUV420P and
 * cvtBGR32ToNV12 (:44-159 with the 2x2 block of :70-104) and their inverses
 * cvtYUV420PToBGR32 / cvtNV12ToBGR32 (:161-266).  TEST INFRASTRUCTURE ONLY
 * (part of the oracle library; see oracle.h).
 *
 * BT.601 limited range in 20-bit fixed point (the coefficients are OpenCV's
 * published cvtColor table).  Per 2x2 block: Y of each pixel; U and V from
 * the LEFT column only, top and bottom averaged with one extra shift (the
 * right column does not enter the chroma, :93-101).  Inverse: per 2x2 block
 * one (U, V); Y clamped at 16 from below before scaling (:172-193).
 *
 * Pinned by tests/test_oracle_yuv.py: black, white and grays against the
 * limited-range closed forms, every channel against the float BT.601 matrix
 * within one level, the left-column chroma, and the round trip.
 */
#include <stdint.h>
#include <string.h>

#include "oracle.h"

enum {
    CRY = 269484, CGY = 528482, CBY = 102760,
    CRU = -155188, CGU = -305135, CBU = 460324,
    CGV = -385875, CBV = -74448,
    CY = 1220542, CUB = 2116026, CUG = -409993, CVG = -852492, CVR = 1673527,
    SHIFT = 20
};

static int c255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

/* One 2x2 block: top / bottom rows of BGRx pixels (x, x + 1). */
static void block_to_yuv(const uint8_t* t, const uint8_t* b, uint8_t* yt, uint8_t* yb, uint8_t* u,
                         uint8_t* v)
{
    const int half = 1 << (SHIFT - 1), y16 = 16 << SHIFT, c128 = 128 << SHIFT;
    const uint8_t* px[4] = {t, t + 4, b, b + 4};
    int yv[4];
    for (int k = 0; k < 4; k++)
        yv[k] = CRY * px[k][2] + CGY * px[k][1] + CBY * px[k][0] + half + y16;
    yt[0] = (uint8_t)c255(yv[0] >> SHIFT);
    yt[1] = (uint8_t)c255(yv[1] >> SHIFT);
    yb[0] = (uint8_t)c255(yv[2] >> SHIFT);
    yb[1] = (uint8_t)c255(yv[3] >> SHIFT);
    /* chroma of the left column; V's red coefficient equals U's blue one */
    const int u0 = CRU * t[2] + CGU * t[1] + CBU * t[0] + half + c128;
    const int v0 = CBU * t[2] + CGV * t[1] + CBV * t[0] + half + c128;
    const int u1 = CRU * b[2] + CGU * b[1] + CBU * b[0] + half + c128;
    const int v1 = CBU * b[2] + CGV * b[1] + CBV * b[0] + half + c128;
    *u = (uint8_t)c255((u0 + u1) >> (SHIFT + 1));
    *v = (uint8_t)c255((v0 + v1) >> (SHIFT + 1));
}

int orc_bgr32_to_yuv(const uint8_t* bgr, size_t step, int w, int h, int nv12, uint8_t* y,
                     uint8_t* u, uint8_t* v)
{
    if (w <= 0 || h <= 0 || (w & 1) || (h & 1)) return 0;
    for (int r = 0; r < h / 2; r++)
        for (int c = 0; c < w / 2; c++) {
            const uint8_t* t = bgr + (size_t)(2 * r) * step + 8 * (size_t)c;
            uint8_t uu, vv;
            block_to_yuv(t, t + step, y + (size_t)(2 * r) * w + 2 * c, y + (size_t)(2 * r + 1) * w + 2 * c,
                         &uu, &vv);
            if (nv12) {
                u[(size_t)r * w + 2 * c] = uu;   /* interleaved UV plane, w bytes per row */
                u[(size_t)r * w + 2 * c + 1] = vv;
            } else {
                u[(size_t)r * (w / 2) + c] = uu;
                v[(size_t)r * (w / 2) + c] = vv;
            }
        }
    return 1;
}

static void yuv_to_block(const uint8_t* yt, const uint8_t* yb, int uval, int vval, uint8_t* t, uint8_t* b)
{
    const int u = uval - 128, v = vval - 128, half = 1 << (SHIFT - 1);
    const int ruv = half + CVR * v, guv = half + CVG * v + CUG * u, buv = half + CUB * u;
    const uint8_t* ys[4] = {yt, yt + 1, yb, yb + 1};
    uint8_t* out[4] = {t, t + 4, b, b + 4};
    for (int k = 0; k < 4; k++) {
        const int yy = (ys[k][0] > 16 ? ys[k][0] - 16 : 0) * CY;
        out[k][2] = (uint8_t)c255((yy + ruv) >> SHIFT);
        out[k][1] = (uint8_t)c255((yy + guv) >> SHIFT);
        out[k][0] = (uint8_t)c255((yy + buv) >> SHIFT);
    }
}

/* The inverse; channel 3 of the output is left as it is. */
int orc_yuv_to_bgr32(const uint8_t* y, const uint8_t* u, const uint8_t* v, int w, int h, int nv12,
                     uint8_t* bgr, size_t step)
{
    if (w <= 0 || h <= 0 || (w & 1) || (h & 1)) return 0;
    for (int r = 0; r < h / 2; r++)
        for (int c = 0; c < w / 2; c++) {
            const int uu = nv12 ? u[(size_t)r * w + 2 * c] : u[(size_t)r * (w / 2) + c];
            const int vv = nv12 ? u[(size_t)r * w + 2 * c + 1] : v[(size_t)r * (w / 2) + c];
            uint8_t* t = bgr + (size_t)(2 * r) * step + 8 * (size_t)c;
            yuv_to_block(y + (size_t)(2 * r) * w + 2 * c, y + (size_t)(2 * r + 1) * w + 2 * c, uu, vv, t,
                         t + step);
        }
    return 1;
}
