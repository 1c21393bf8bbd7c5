// kernels_gain.hip -- per-frame gain estimation from overlap means on gfx950.
//
// k_gain_stats   per camera i and every camera j whose original mask covers a
//                pixel of mask_i: count and sums of gray, B, G, R of camera
//                i's 8-bit reprojected sample (the bilinear of the render,
//                before any LUT) -- the inputs of cv::countNonZero and cv::mean
//                in getTransformsMeanApproxSiftPanoPaper (Compensate.cpp:
//                501-527) and getTintTransformsMeanApproxMimicSiftPanoPaper
//                (TintAdjust.cpp:123-157).  Exact integer sums.
// k_gain_solve   the normal equations of both (Compensate.cpp:529-547,
//                TintAdjust.cpp:160-210), cv::solve, the anchor rescale
//                (Compensate.cpp:549-575) and getExposureColorOptimizeLUTs
//                (ExposureColorOptimize.cpp:1425-1443), in FP64 in the
//                reference's operation order, writing the render's LUTs on
//                the device.
#include <hip/hip_runtime.h>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"
#include "pano_launch.hpp"

namespace pano {

namespace {

// cv::cvtColor(BGR2GRAY) for 8-bit pixels (OpenCV 3.0 fixed point, shift 14).
__device__ __forceinline__ uint32_t gray_bgr(uint32_t b, uint32_t g, uint32_t r)
{
    return (1868u * b + 9617u * g + 4899u * r + 8192u) >> 14;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

constexpr int kStatW = 2 * kFTW, kStatH = 2 * kFTH;   // level-0 core of a fused tile

}  // namespace

// One fused level-0 tile of camera i (the remap's jobs): its core pixels
// (2 kFTW x 2 kFTH at level 0, each canvas pixel of the camera's box once).
__global__ __launch_bounds__(256) void k_gain_stats(const Seg* __restrict__ segs, int nseg,
                                                    SrcTable src)
{
    __shared__ uint32_t part[4][kGainMaxCams * 5];
    const Seg& S = seg_of(segs, nseg, blockIdx.x);
    const RenderDesc R = ldc(S.rd);
    const DownJob J = ldc((const DownJob*)S.jobs + (blockIdx.x - S.begin));
    const RemapCam cam = ldc(R.cams + J.src);
    const Plane P0 = ldc(R.planes + cam.plane);
    const uint8_t* sp = src.p[S.src_base + cam.src];
    const uint32_t sstep = src.step[S.src_base + cam.src];
    const float fsw = (float)cam.sw, fsh = (float)cam.sh;
    const int n = R.ncams;
    const int cwid = 2 * J.nx, chgt = 2 * J.ny;
    // core pixel (r, c) is map storage (my + 2 + r, mx + 2 + c), canvas (2X + c, 2Y + r)
    const char* mbase = (const char*)(cam.map + ((size_t)(J.my + 2) * cam.mpitch + J.mx + 2));
    const uint32_t mrow = (uint32_t)cam.mpitch * 8u;
    uint32_t abg[kGainMaxCams], arg[kGainMaxCams], acnt[kGainMaxCams];
#pragma unroll
    for (int j = 0; j < kGainMaxCams; j++) abg[j] = arg[j] = acnt[j] = 0;
    uint32_t seen = 0;
    // per thread kStatW kStatH / 256 = 16 pixels: every packed 16-bit sum stays below 2^16.
    // kGB pixels at a time with their map, cover and tap loads in flight together.
    constexpr int kPerT = kStatW * kStatH / 256, kGB = 4;
    static_assert(kPerT * 256 == kStatW * kStatH && kPerT % kGB == 0, "whole pixel groups per thread");
#pragma unroll 1
    for (int k0 = 0; k0 < kPerT; k0 += kGB) {
        float2 m[kGB];
        uint32_t cov[kGB], px[kGB][4];
        bool in[kGB];
#pragma unroll
        for (int b = 0; b < kGB; b++) {
            const int e = threadIdx.x + 256 * (k0 + b);
            const int r = e / kStatW, c = e - (e / kStatW) * kStatW;
            const int vx = 2 * J.X + c, vy = 2 * J.Y + r;
            in[b] = r < chgt && c < cwid && vx < P0.cw && vy < P0.ch;
            m[b] = in[b] ? ldg_f2((const float2*)(mbase + __umul24((uint32_t)r, mrow) + 8u * (uint32_t)c))
                         : make_float2(0.f, 0.f);
            cov[b] = in[b] ? ldg(R.cover + __umul24((uint32_t)vy, (uint32_t)P0.cw) + vx) : 0u;
        }
#pragma unroll
        for (int b = 0; b < kGB; b++) {
            in[b] = in[b] && (__float_as_uint(m[b].y) >> 31);   // inside camera i's mask
            const float x = __builtin_fabsf(m[b].x), y = __builtin_fabsf(m[b].y);
            px[b][0] = px[b][1] = px[b][2] = px[b][3] = 0;
            if (in[b] && x < fsw && y < fsh) load_taps(sp, sstep, cam.sw, cam.sh, x, y, px[b]);
        }
#pragma unroll
        for (int b = 0; b < kGB; b++) {
            if (!in[b]) continue;
            const float x = __builtin_fabsf(m[b].x), y = __builtin_fabsf(m[b].y);
            uint32_t v = 0;
            if (x < fsw && y < fsh) {
                bool tie = false;
                v = bilinear_fast(__builtin_amdgcn_fractf(x), __builtin_amdgcn_fractf(y), px[b], tie);
                if (tie) v = bilinear_combine((double)x, (double)y, px[b], nullptr);
            }
            const uint32_t bb = v & 255u, g = (v >> 8) & 255u, rr = (v >> 16) & 255u;
            const uint32_t bg = bb | (g << 16), rg = rr | (gray_bgr(bb, g, rr) << 16);
            seen |= cov[b];
#pragma unroll
            for (int j = 0; j < kGainMaxCams; j++) {
                if (j >= n) break;   // uniform: the rig's camera count
                const uint32_t bit = (cov[b] >> j) & 1u;
                abg[j] += bit * bg;
                arg[j] += bit * rg;
                acnt[j] += bit;
            }
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // cameras touched anywhere in the wave (uniform)
    uint32_t wseen = seen;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) wseen |= __shfl_xor(wseen, m, 64);
#pragma unroll
    for (int j = 0; j < kGainMaxCams; j++) {
        uint32_t f[5] = {0, 0, 0, 0, 0};
        if ((wseen >> j) & 1u) {
            f[0] = wave_sum(acnt[j]);
            f[1] = wave_sum(arg[j] >> 16);
            f[2] = wave_sum(abg[j] & 0xffffu);
            f[3] = wave_sum(abg[j] >> 16);
            f[4] = wave_sum(arg[j] & 0xffffu);
        }
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 5; k++) part[wave][j * 5 + k] = f[k];
    }
    __syncthreads();
    if (threadIdx.x < n * 5) {
        const uint32_t t = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                           part[3][threadIdx.x];
        // one of kGainSlots copies per block: the blocks of a camera do not all
        // meet at the same addresses
        const size_t slot = (size_t)(blockIdx.x % kGainSlots) * n * n * 5;
        if (t) atomicAdd(R.gstats + slot + ((size_t)J.src * n) * 5 + threadIdx.x, (unsigned long long)t);
    }
}

namespace {

__device__ double det2(const double* m) { return m[0] * m[3] - m[1] * m[2]; }

__device__ double det3(const double* m)
{
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// cv::solve(A, b, x) with DECOMP_LU (OpenCV 3.0): determinant forms for n <= 3,
// Gaussian elimination with partial pivoting (eps 100 DBL_EPSILON) above.
// A is overwritten; y is scratch of n doubles.  Returns false when singular.
__device__ bool cv_solve(int n, double* A, const double* b, double* x, double* y)
{
    if (n == 1) {
        const double d = A[0];
        if (d == 0.) return false;
        x[0] = b[0] / d;
        return true;
    }
    if (n == 2) {
        double d = det2(A);
        if (d == 0.) return false;
        d = 1. / d;
        const double t = (b[0] * A[3] - b[1] * A[1]) * d;
        x[1] = (b[1] * A[0] - b[0] * A[2]) * d;
        x[0] = t;
        return true;
    }
    if (n == 3) {
        double d = det3(A);
        if (d == 0.) return false;
        d = 1. / d;
        const double t0 = ((A[4] * A[8] - A[5] * A[7]) * b[0] + (A[2] * A[7] - A[1] * A[8]) * b[1] +
                           (A[1] * A[5] - A[2] * A[4]) * b[2]) * d;
        const double t1 = ((A[5] * A[6] - A[3] * A[8]) * b[0] + (A[0] * A[8] - A[2] * A[6]) * b[1] +
                           (A[2] * A[3] - A[0] * A[5]) * b[2]) * d;
        const double t2 = ((A[3] * A[7] - A[4] * A[6]) * b[0] + (A[1] * A[6] - A[0] * A[7]) * b[1] +
                           (A[0] * A[4] - A[1] * A[3]) * b[2]) * d;
        x[0] = t0;
        x[1] = t1;
        x[2] = t2;
        return true;
    }
    for (int i = 0; i < n; i++) y[i] = b[i];
    const double eps = 2.220446049250313e-16 * 100;
    for (int i = 0; i < n; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (fabs(A[j * n + i]) > fabs(A[k * n + i])) k = j;
        if (fabs(A[k * n + i]) < eps) return false;
        if (k != i) {
            for (int j = i; j < n; j++) {
                const double t = A[i * n + j];
                A[i * n + j] = A[k * n + j];
                A[k * n + j] = t;
            }
            const double t = y[i];
            y[i] = y[k];
            y[k] = t;
        }
        const double d = -1 / A[i * n + i];
        for (int j = i + 1; j < n; j++) {
            const double alpha = A[j * n + i] * d;
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
    for (int i = 0; i < n; i++) x[i] = y[i];
    return true;
}

// The normal equations of sum_ij N_ij ((g_i I_ij - g_j I_ji)^2 sn + (1 - g_i)^2 sg),
// solved; gains = 1 when singular.
__device__ void solve_means(int n, const double* N, const double* I, double sn, double sg,
                            double* A, double* b, double* y, double* g)
{
    for (int k = 0; k < n * n; k++) A[k] = 0;
    for (int k = 0; k < n; k++) b[k] = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            A[i * n + i] += N[i * n + j] * (I[i * n + j] * I[i * n + j] * sn + sg);
            A[j * n + j] += N[i * n + j] * (I[j * n + i] * I[j * n + i] * sn);
            A[i * n + j] -= 2 * N[i * n + j] * (I[i * n + j] * I[j * n + i] * sn);
            b[i] += N[i * n + j] * sg;
        }
    if (!cv_solve(n, A, b, g, y))
        for (int i = 0; i < n; i++) g[i] = 1;
}

}  // namespace

__global__ __launch_bounds__(256) void k_gain_solve(const RenderDesc* __restrict__ rd, int flags)
{
    constexpr int M = kGainMaxCams;
    __shared__ double N[M * M], I[M * M], Rr[M * M], Bb[M * M];
    // one set of normal equations per solving thread: exposure, red / green, blue / green
    __shared__ double A[3][M * M], b[3][M], y[3][M], sol[3][M];
    __shared__ double fac[3 * M];
    __shared__ unsigned long long sums[M * M * 5];
    const RenderDesc R = ldc(rd);
    const int n = R.ncams;
    const int t0 = threadIdx.x;
    for (int t = t0; t < n * n * 5; t += blockDim.x) {
        unsigned long long v = 0;
        for (int k = 0; k < kGainSlots; k++) v += R.gstats[(size_t)k * n * n * 5 + t];
        sums[t] = v;
    }
    __syncthreads();
    // overlap counts and means, one (i, j) entry per thread
    for (int t = t0; t < n * n; t += blockDim.x) {
        const int i = t / n, j = t - (t / n) * n;
        const unsigned long long* s = sums + (size_t)t * 5;
        const double cnt = (double)s[0];
        const double inv = s[0] ? 1. / cnt : 0.;
        N[t] = cnt;
        I[t] = (double)s[1] * inv;
        const double mb = (double)s[2] * inv, mg = (double)s[3] * inv, mr = (double)s[4] * inv;
        Rr[t] = (i == j || s[0]) ? mr / mg : 0.;
        Bb[t] = (i == j || s[0]) ? mb / mg : 0.;
    }
    __syncthreads();
    const bool tint = flags & kGainTint;
    // one solve per wavefront, so the three run side by side on separate SIMDs
    if (t0 == 0) solve_means(n, N, I, 0.01, 100, A[0], b[0], y[0], sol[0]);
    if (t0 == 64 && tint) solve_means(n, N, Rr, 1, 0.05, A[1], b[1], y[1], sol[1]);
    if (t0 == 128 && tint) solve_means(n, N, Bb, 1, 0.05, A[2], b[2], y[2], sol[2]);
    __syncthreads();
    if (t0 == 0) {
        double* g = sol[0];
        double* rg = sol[1];
        double* bg = sol[2];
        double max_mean = -1;
        int mi = -1;
        for (int i = 0; i < n; i++)
            if (I[i * n + i] > max_mean) {
                max_mean = I[i * n + i];
                mi = i;
            }
        const double kscale = 1.0 / g[mi];
        for (int i = 0; i < n; i++) g[i] = kscale * g[i];
        if (!tint)
            for (int i = 0; i < n; i++) rg[i] = bg[i] = 1;
        for (int i = 0; i < n; i++) {
            stg(R.gains + i, g[i]);
            stg(R.gains + n + i, rg[i]);
            stg(R.gains + 2 * n + i, bg[i]);
            fac[3 * i] = g[i] * bg[i];
            fac[3 * i + 1] = g[i];
            fac[3 * i + 2] = g[i] * rg[i];
        }
        stg(R.gains + 3 * n, (double)mi);
    }
    __syncthreads();
    // getLUT (ExposureColorOptimize.cpp:1217-1223): saturate_cast<uchar>(k i)
    uint8_t* luts = (uint8_t*)R.luts;
    for (int e = threadIdx.x; e < n * 768; e += 256) {
        const int t = e >> 8, v = e & 255;
        stg(luts + e, (uint8_t)sat_u8(__double2int_rn(fac[t] * (double)v)));
    }
    if (threadIdx.x == 0) stg(luts + n * 768, (uint8_t)1);   // LUTs active for this frame
}

hipError_t launch_gain(const Seg* segs, int nseg, int blocks, const RenderDesc* rd, int flags,
                       const SrcTable& src, hipStream_t s)
{
    if (blocks <= 0 || nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gain_stats, dim3(blocks), dim3(256), 0, s, segs, nseg, src);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gain_solve, dim3(1), dim3(256), 0, s, rd, flags);
    return hipGetLastError();
}

}  // namespace pano
