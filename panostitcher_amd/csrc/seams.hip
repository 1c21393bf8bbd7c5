// seams.hip -- getNonIntersectingMasks (Blend/Mask.cpp:376-397) on gfx950.
//
// Camera i >= 1 takes the pixels where the L1 distance transform of its mask
// exceeds that of the union of the cameras before it; those pixels leave the
// earlier cameras.  The distance is cv::distanceTransform(CV_DIST_L1, 3) of
// OpenCV 3.0: its two-pass 3x3 chamfer with costs {1, 2} is the exact L1
// distance to the nearest zero pixel, its frame of INT_MAX >> 2 (fixed point,
// 16 fractional bits) adds the cap 8192 + distance to the outside, and the
// float result is exactly min(d_zero, 8192 + d_frame) (integers below 2^24).
// Here the L1 transform is separable: per row the distance to the nearest zero
// of the row (a block-wide prefix max / suffix min of zero positions), then per
// column min over rows of g + |dy| (segmented down / up sweeps); the cap last.
#include <hip/hip_runtime.h>

#include <vector>

#include "pano_dev.hpp"
#include "pano_kernel_common.hpp"

namespace pano {

namespace {

constexpr int kDtT = 256;           // threads of a row block
constexpr int kInf = 1 << 28;       // no zero in the row

// Row pass: g = distance to the nearest zero of the row (kInf if none).
__global__ __launch_bounds__(kDtT) void k_dt_rows(const uint8_t* __restrict__ m, int pitch, int w,
                                                  int32_t* __restrict__ g)
{
    __shared__ int lastz[kDtT], nextz[kDtT];
    const int y = blockIdx.x, t = threadIdx.x;
    const uint8_t* row = m + (size_t)y * pitch;
    const int chunk = (w + kDtT - 1) / kDtT;
    const int x0 = min(t * chunk, w), x1 = min(x0 + chunk, w);
    int lz = -1, nz = 1 << 30;
    for (int x = x0; x < x1; x++)
        if (!ldg(row + x)) {
            lz = x;
            if (nz == 1 << 30) nz = x;
        }
    lastz[t] = lz;
    nextz[t] = nz;
    __syncthreads();
    // inclusive prefix max of lastz, suffix min of nextz (Hillis-Steele)
    for (int s = 1; s < kDtT; s <<= 1) {
        const int a = t >= s ? lastz[t - s] : -1;
        const int b = t + s < kDtT ? nextz[t + s] : 1 << 30;
        __syncthreads();
        lastz[t] = max(lastz[t], a);
        nextz[t] = min(nextz[t], b);
        __syncthreads();
    }
    int prev = t > 0 ? lastz[t - 1] : -1;             // last zero before this chunk
    const int next = t + 1 < kDtT ? nextz[t + 1] : 1 << 30;   // first zero after it
    // forward then backward over the chunk
    int32_t* out = g + (size_t)y * pitch;
    for (int x = x0; x < x1; x++) {
        if (!ldg(row + x)) prev = x;
        out[x] = prev >= 0 ? x - prev : kInf;
    }
    int nx = next;
    for (int x = x1 - 1; x >= x0; x--) {
        if (!ldg(row + x)) nx = x;
        if (nx < (1 << 30)) out[x] = min(out[x], nx - x);
    }
}

// The same row pass with the row staged through LDS (rows up to kDtRowMax
// pixels): the mask row is read and the distances written with consecutive
// threads on consecutive pixels; the per-thread chunks run in LDS.
constexpr int kDtRowMax = 8192;
__global__ __launch_bounds__(kDtT) void k_dt_rows_lds(const uint8_t* __restrict__ m, int pitch, int w,
                                                      int32_t* __restrict__ g)
{
    __shared__ uint8_t row[kDtRowMax];
    __shared__ int32_t dist[kDtRowMax];
    __shared__ int lastz[kDtT], nextz[kDtT];
    const int y = blockIdx.x, t = threadIdx.x;
    const uint8_t* src = m + (size_t)y * pitch;
    for (int x = t; x < w; x += kDtT) row[x] = ldg(src + x);
    __syncthreads();
    const int chunk = (w + kDtT - 1) / kDtT;
    const int x0 = min(t * chunk, w), x1 = min(x0 + chunk, w);
    int lz = -1, nz = 1 << 30;
    for (int x = x0; x < x1; x++)
        if (!row[x]) {
            lz = x;
            if (nz == 1 << 30) nz = x;
        }
    lastz[t] = lz;
    nextz[t] = nz;
    __syncthreads();
    for (int s = 1; s < kDtT; s <<= 1) {
        const int a = t >= s ? lastz[t - s] : -1;
        const int b = t + s < kDtT ? nextz[t + s] : 1 << 30;
        __syncthreads();
        lastz[t] = max(lastz[t], a);
        nextz[t] = min(nextz[t], b);
        __syncthreads();
    }
    int prev = t > 0 ? lastz[t - 1] : -1;
    int nx = t + 1 < kDtT ? nextz[t + 1] : 1 << 30;
    for (int x = x0; x < x1; x++) {
        if (!row[x]) prev = x;
        dist[x] = prev >= 0 ? x - prev : kInf;
    }
    for (int x = x1 - 1; x >= x0; x--) {
        if (!row[x]) nx = x;
        if (nx < (1 << 30)) dist[x] = min(dist[x], nx - x);
    }
    __syncthreads();
    int32_t* out = g + (size_t)y * pitch;
    for (int x = t; x < w; x += kDtT) out[x] = dist[x];
}

// Column pass and cap: d = min(min_y' g(y') + |y - y'|, 8192 + distance to the
// frame).  A block holds 64 columns and kColSeg row segments; each thread sweeps
// its segment down and up, and the segments' boundary values (the minimum of
// g(y') + distance to the segment's first / last row over the segments above /
// below) are combined through LDS between the sweeps.
constexpr int kColSeg = 16;
__global__ __launch_bounds__(64 * kColSeg) void k_dt_cols(const int32_t* __restrict__ g, int pitch, int w, int h,
                                                          int32_t* __restrict__ d)
{
    __shared__ int down[kColSeg][64], up[kColSeg][64];
    const int cx = threadIdx.x & 63, sg = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + cx;
    const int seg = (h + kColSeg - 1) / kColSeg;
    const int y0 = min(sg * seg, h), y1 = min(y0 + seg, h);
    // local minima of g(y') + (y1 - 1 - y') (toward the segment's last row) and
    // g(y') + (y' - y0) (toward its first row)
    int ld = kInf, lu = kInf;
    if (x < w)
        for (int y = y0; y < y1; y++) {
            const int v = ldg(g + (size_t)y * pitch + x);
            ld = min(ld + 1, v);
            lu = min(lu, v + (y - y0));
        }
    down[sg][cx] = ld;
    up[sg][cx] = lu;
    __syncthreads();
    if (x >= w) return;
    // the best value entering this segment from above / below
    int in_d = kInf, in_u = kInf;
    for (int k = 0; k < sg; k++) {
        const int yk1 = min(min(k * seg, h) + seg, h);   // segment k's end
        in_d = min(in_d, down[k][cx] + (y0 - yk1) + 1);
    }
    for (int k = sg + 1; k < kColSeg; k++) {
        const int yk0 = min(k * seg, h);
        if (yk0 >= h) break;
        in_u = min(in_u, up[k][cx] + (yk0 - y1) + 1);
    }
    int f = in_d;   // value at row y0 - 1 seen from above, plus 1 per row
    for (int y = y0; y < y1; y++) {
        f = min(ldg(g + (size_t)y * pitch + x), f == kInf ? kInf : f);
        d[(size_t)y * pitch + x] = f;
        f = f + 1;
    }
    int b = in_u;
    for (int y = y1 - 1; y >= y0; y--) {
        const int v = min(d[(size_t)y * pitch + x], b);
        const int frame = 8192 + min(min(x + 1, w - x), min(y + 1, h - y));
        d[(size_t)y * pitch + x] = min(v, frame);
        b = v + 1;
    }
}

// Camera i's unique pixels: 255 where its distance exceeds the union's, those
// pixels cleared from the cameras before it and added to the union.
__global__ __launch_bounds__(256) void k_seam_step(const int32_t* __restrict__ di, const int32_t* __restrict__ du,
                                                   uint8_t* __restrict__ out, size_t plane, int i, int pitch,
                                                   int w, int h, uint8_t* __restrict__ acc)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const size_t q = (size_t)y * pitch + x;
    const bool u = di[q] > du[q];
    out[i * plane + q] = u ? 255 : 0;
    if (u) {
        for (int j = 0; j < i; j++) out[j * plane + q] = 0;
        acc[q] = 255;
    }
}

hipError_t dt(const uint8_t* m, int pitch, int w, int h, int32_t* g, int32_t* d, hipStream_t s)
{
    if (w <= kDtRowMax)
        hipLaunchKernelGGL(k_dt_rows_lds, dim3(h), dim3(kDtT), 0, s, m, pitch, w, g);
    else
        hipLaunchKernelGGL(k_dt_rows, dim3(h), dim3(kDtT), 0, s, m, pitch, w, g);
    hipLaunchKernelGGL(k_dt_cols, dim3((w + 63) / 64), dim3(64 * kColSeg), 0, s, g, pitch, w, h, d);
    return hipGetLastError();
}

struct Buf {
    void* p = nullptr;
    ~Buf()
    {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t n) { return hipMalloc(&p, n ? n : 1) == hipSuccess; }
};

}  // namespace

// Device masks in, device unique masks out, in place: n planes of rows x cols
// bytes (pitch cols) at M; synchronous.
hipError_t device_seams_inplace(int n, int rows, int cols, uint8_t* M)
{
    const int pitch = cols;
    const size_t plane = (size_t)rows * pitch;
    Buf acc, g, du, di;
    if (!acc.alloc(plane) || !g.alloc(plane * 4) || !du.alloc(plane * 4) || !di.alloc(plane * 4))
        return hipErrorOutOfMemory;
    hipError_t e = hipMemcpy(acc.p, M, plane, hipMemcpyDeviceToDevice);
    const dim3 grid((cols + 63) / 64, (rows + 3) / 4);
    for (int i = 1; i < n && e == hipSuccess; i++) {
        e = dt((const uint8_t*)acc.p, pitch, cols, rows, (int32_t*)g.p, (int32_t*)du.p, nullptr);
        if (e == hipSuccess) e = dt(M + plane * i, pitch, cols, rows, (int32_t*)g.p, (int32_t*)di.p, nullptr);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_seam_step, grid, dim3(256), 0, nullptr, (const int32_t*)di.p,
                               (const int32_t*)du.p, M, plane, i, pitch, cols, rows, (uint8_t*)acc.p);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
}

// Host masks in, host unique masks out (n planes of rows x cols); synchronous.
// Returns hipSuccess or the first error.
hipError_t device_nonintersecting_masks(int n, int rows, int cols, const uint8_t* const* masks,
                                        const size_t* steps, uint8_t* out)
{
    // planes packed (pitch = cols) so that host copies are one contiguous block
    const int pitch = cols;
    const size_t plane = (size_t)rows * pitch;
    Buf dm;
    if (!dm.alloc(plane * n)) return hipErrorOutOfMemory;
    uint8_t* M = (uint8_t*)dm.p;
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; i++)
        e = steps[i] == (size_t)cols
                ? hipMemcpy(M + plane * i, masks[i], plane, hipMemcpyHostToDevice)
                : hipMemcpy2D(M + plane * i, pitch, masks[i], steps[i], cols, rows, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = device_seams_inplace(n, rows, cols, M);
    if (e == hipSuccess) e = hipMemcpy(out, M, plane * n, hipMemcpyDeviceToHost);
    return e;
}

}  // namespace pano
