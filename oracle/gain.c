/*
 * gain.c -- CPU restatement of the per-frame gain estimation from overlap
 * means (TEST INFRASTRUCTURE ONLY, part of the oracle library; see oracle.h).
 *
 *   orc_gray_bgr        cv::cvtColor(BGR2GRAY) on 8-bit pixels: OpenCV 3.0's
 *                       fixed point (R2Y 4899, G2Y 9617, B2Y 1868, shift 14,
 *                       rounding constant folded in), the conversion both
 *                       estimators apply first (Compensate.cpp:629-631,
 *                       ExposureColorCorrect.cpp:41-43)
 *   orc_cv_solve        cv::solve(A, b, x) with the default DECOMP_LU: the
 *                       determinant forms OpenCV 3.0 uses for 1x1, 2x2 and
 *                       3x3 systems, Gaussian elimination with partial
 *                       pivoting (hal LU, eps = 100 DBL_EPSILON) above that
 *   orc_gain_estimate   getTransformsMeanApproxSiftPanoPaper
 *                       (Compensate.cpp:501-563) + rescale (:565-575), and
 *                       getTintTransformsMeanApproxMimicSiftPanoPaper
 *                       (TintAdjust.cpp:123-215) on the same images, then
 *                       getExposureColorOptimizeLUTs
 *                       (ExposureColorOptimize.cpp:1425-1443: B = e b, G = e,
 *                       R = e r, getLUT :1217-1223)
 *
 * cv::mean(img, mask) is sum * (1 / count) over the masked pixels (0 when
 * none); cv::countNonZero is exact.  All sums of 8-bit values are exact
 * integers.  Parity pinning (tests/test_oracle_gain.py): gray closed forms;
 * the solver against numpy's LAPACK solve and by its residual; the gains by
 * the gradient of the least-squares objective the normal equations come
 * from (zero at the solution), the rescale and anchor rules, the degenerate
 * cases (no overlap, empty masks) and the LUT closed form.  The OpenCV
 * boundary (cvtColor constants, solve's small-matrix forms) is parity
 * unpinned beyond those checks: no reference test holds vectors for it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int orc_gray_bgr(int b, int g, int r) { return (1868 * b + 9617 * g + 4899 * r + 8192) >> 14; }

static double det2(const double* m) { return m[0] * m[3] - m[1] * m[2]; }

static double det3(const double* m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* A: n x n row-major (overwritten for n > 3), b: n, x: n.  Returns 1 on
 * success, 0 when the system is singular (x untouched then). */
int orc_cv_solve(int n, double* A, const double* b, double* x)
{
    if (n == 1) {
        double d = A[0];
        if (d == 0.) return 0;
        x[0] = b[0] / d;
        return 1;
    }
    if (n == 2) {
        double d = det2(A);
        if (d == 0.) return 0;
        d = 1. / d;
        double t = (b[0] * A[3] - b[1] * A[1]) * d;
        x[1] = (b[1] * A[0] - b[0] * A[2]) * d;
        x[0] = t;
        return 1;
    }
    if (n == 3) {
        double d = det3(A);
        if (d == 0.) return 0;
        d = 1. / d;
        double t0 = ((A[4] * A[8] - A[5] * A[7]) * b[0] + (A[2] * A[7] - A[1] * A[8]) * b[1] +
                     (A[1] * A[5] - A[2] * A[4]) * b[2]) * d;
        double t1 = ((A[5] * A[6] - A[3] * A[8]) * b[0] + (A[0] * A[8] - A[2] * A[6]) * b[1] +
                     (A[2] * A[3] - A[0] * A[5]) * b[2]) * d;
        double t2 = ((A[3] * A[7] - A[4] * A[6]) * b[0] + (A[1] * A[6] - A[0] * A[7]) * b[1] +
                     (A[0] * A[4] - A[1] * A[3]) * b[2]) * d;
        x[0] = t0;
        x[1] = t1;
        x[2] = t2;
        return 1;
    }
    /* Gaussian elimination with partial pivoting; the pivot's reciprocal is
     * kept on the diagonal and the back substitution multiplies by it. */
    double* y = (double*)malloc(sizeof(double) * n);
    memcpy(y, b, sizeof(double) * n);
    const double eps = 2.220446049250313e-16 * 100;
    for (int i = 0; i < n; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (fabs(A[j * n + i]) > fabs(A[k * n + i])) k = j;
        if (fabs(A[k * n + i]) < eps) {
            free(y);
            return 0;
        }
        if (k != i) {
            for (int j = i; j < n; j++) {
                double t = A[i * n + j];
                A[i * n + j] = A[k * n + j];
                A[k * n + j] = t;
            }
            double t = y[i];
            y[i] = y[k];
            y[k] = t;
        }
        double d = -1 / A[i * n + i];
        for (int j = i + 1; j < n; j++) {
            double alpha = A[j * n + i] * d;
            for (int c = i + 1; c < n; c++) A[j * n + c] += alpha * A[i * n + c];
            y[j] += alpha * y[i];
        }
        A[i * n + i] = -d;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int c = i + 1; c < n; c++) s -= A[i * n + c] * y[c];
        y[i] = s * A[i * n + i];
    }
    memcpy(x, y, sizeof(double) * n);
    free(y);
    return 1;
}

/* One system of the overlap-means family: sum over (i, j) of
 * N_ij ((g_i I_ij - g_j I_ji)^2 sn + (1 - g_i)^2 sg); gains = 1 on failure. */
static void solve_means(int n, const double* N, const double* I, double sn, double sg, double* g)
{
    double* A = (double*)calloc((size_t)n * n, sizeof(double));
    double* b = (double*)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            A[i * n + i] += N[i * n + j] * (I[i * n + j] * I[i * n + j] * sn + sg);
            A[j * n + j] += N[i * n + j] * (I[j * n + i] * I[j * n + i] * sn);
            A[i * n + j] -= 2 * N[i * n + j] * (I[i * n + j] * I[j * n + i] * sn);
            b[i] += N[i * n + j] * sg;
        }
    if (!orc_cv_solve(n, A, b, g))
        for (int i = 0; i < n; i++) g[i] = 1;
    free(A);
    free(b);
}

void orc_gain_sums(int n, int rows, int cols, const uint8_t* const* bgr,
                   const uint8_t* const* masks, int64_t* sums)
{
    /* sums[(i * n + j) * 5 + f]: f = count, gray, B, G, R of image i over
     * mask_i & mask_j (mask_i alone for j = i) */
    memset(sums, 0, sizeof(int64_t) * (size_t)n * n * 5);
    for (int i = 0; i < n; i++)
        for (size_t p = 0; p < (size_t)rows * cols; p++) {
            if (!masks[i][p]) continue;
            const uint8_t* px = bgr[i] + 3 * p;
            const int gy = orc_gray_bgr(px[0], px[1], px[2]);
            for (int j = 0; j < n; j++) {
                if (!masks[j][p]) continue;
                int64_t* s = sums + ((size_t)i * n + j) * 5;
                s[0] += 1;
                s[1] += gy;
                s[2] += px[0];
                s[3] += px[1];
                s[4] += px[2];
            }
        }
}

int orc_gain_from_sums(int n, const int64_t* sums, int tint, double* gains, double* rg,
                       double* bg, int* max_index, uint8_t* luts)
{
    double* N = (double*)calloc((size_t)n * n, sizeof(double));
    double* I = (double*)calloc((size_t)n * n, sizeof(double));
    double* R = (double*)calloc((size_t)n * n, sizeof(double));
    double* B = (double*)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const int64_t* s = sums + ((size_t)i * n + j) * 5;
            const double cnt = (double)s[0];
            const double inv = s[0] ? 1. / cnt : 0.;
            N[i * n + j] = cnt;
            I[i * n + j] = (double)s[1] * inv;
            const double mb = (double)s[2] * inv, mg = (double)s[3] * inv, mr = (double)s[4] * inv;
            /* TintAdjust.cpp:137-155: ratios of the means; 0 off-diagonal without overlap */
            if (i == j || s[0]) {
                R[i * n + j] = mr / mg;
                B[i * n + j] = mb / mg;
            }
        }
    solve_means(n, N, I, 0.01, 100, gains);
    /* Compensate.cpp:549-560: the image with the largest mean is the anchor */
    double max_mean = -1;
    int mi = -1;
    for (int i = 0; i < n; i++)
        if (I[i * n + i] > max_mean) {
            max_mean = I[i * n + i];
            mi = i;
        }
    *max_index = mi;
    const double kscale = 1.0 / gains[mi];
    for (int i = 0; i < n; i++) gains[i] = kscale * gains[i];
    if (tint) {
        solve_means(n, N, R, 1, 0.05, rg);
        solve_means(n, N, B, 1, 0.05, bg);
    } else {
        for (int i = 0; i < n; i++) rg[i] = bg[i] = 1;
    }
    if (luts)
        for (int i = 0; i < n; i++) {
            orc_get_lut(luts + (size_t)i * 768, gains[i] * bg[i]);
            orc_get_lut(luts + (size_t)i * 768 + 256, gains[i]);
            orc_get_lut(luts + (size_t)i * 768 + 512, gains[i] * rg[i]);
        }
    free(N);
    free(I);
    free(R);
    free(B);
    return 1;
}

int orc_gain_estimate(int n, int rows, int cols, const uint8_t* const* bgr,
                      const uint8_t* const* masks, int tint, double* gains, double* rg, double* bg,
                      int* max_index, uint8_t* luts)
{
    int64_t* sums = (int64_t*)malloc(sizeof(int64_t) * (size_t)n * n * 5);
    orc_gain_sums(n, rows, cols, bgr, masks, sums);
    int ok = orc_gain_from_sums(n, sums, tint, gains, rg, bg, max_index, luts);
    free(sums);
    return ok;
}
