// maps.cpp -- equirect -> camera coordinate maps (getReprojectMap32FAndMask,
// Warp/ZReproject.cpp:132-240, over the PanoTools transform chain of
// Warp/Remap.cpp:147-514), on the host (rows in parallel) or on the device
// (cudaGenerateReprojectMapAndMask, CudaAccel/CudaReproject.cu:266-385): the
// same FP64 chain, one point per canvas pixel.  The reference names the
// equirect canvas "src" and the camera "dst" inside Remap (ZReproject.h:66-71);
// here canvas and camera.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/panohip.h"

namespace pano {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
inline double deg2rad(double d) { return d * 2.0 * kPi / 360.0; }

struct CameraChain {
    bool rectilinear = false;
    double canvasTX = 0, canvasTY = 0, camTX = 0, camTY = 0;
    double scale = 0, shear0 = 0, shear1 = 0, halfTurn = 0, yawShift = 0;
    double rad[6] = {0, 0, 0, 0, 0, 0};
    double mt[3][3];
    double distance = 0, horizontal = 0, vertical = 0;
};

void mul33(const double a[3][3], const double b[3][3], double r[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) r[i][k] = a[i][0] * b[0][k] + a[i][1] * b[1][k] + a[i][2] * b[2][k];
}

// Rotation about x by a, then (identity) y by b, z by c, as SetMatrix (Remap.cpp:22-43, cl = 0).
void rotation(double a, double b, double c, double m[3][3])
{
    double mx[3][3] = {{1.0, 0.0, 0.0}, {0.0, cos(a), sin(a)}, {0.0, 0.0, 0.0}};
    mx[2][1] = -mx[1][2];
    mx[2][2] = mx[1][1];
    double my[3][3] = {{cos(b), 0.0, -sin(b)}, {0.0, 1.0, 0.0}, {0.0, 0.0, 0.0}};
    my[2][0] = -my[0][2];
    my[2][2] = my[0][0];
    double mz[3][3] = {{cos(c), sin(c), 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}};
    mz[1][0] = -mz[0][1];
    mz[1][1] = mz[0][0];
    double t[3][3];
    mul33(mx, mz, t);
    mul33(t, my, m);
}

// Smallest positive root of the cubic p[0] + p[1] x + p[2] x^2 + p[3] x^3
// (1000 if none), as smallestRoot / cubeZero / squareZero (Remap.cpp:45-135).
double smallest_positive_root(const double* a)
{
    double roots[3];
    int n = 0;
    if (a[3] == 0.0) {
        if (a[2] == 0.0) {
            if (a[1] == 0.0) {
                if (a[0] == 0.0) { n = 1; roots[0] = 0.0; }
            } else {
                n = 1;
                roots[0] = -a[0] / a[1];
            }
        } else if (!(4.0 * a[2] * a[0] > a[1] * a[1])) {
            n = 2;
            roots[0] = (-a[1] + sqrt(a[1] * a[1] - 4.0 * a[2] * a[0])) / (2.0 * a[2]);
            roots[1] = (-a[1] - sqrt(a[1] * a[1] - 4.0 * a[2] * a[0])) / (2.0 * a[2]);
        }
    } else {
        double p = ((-1.0 / 3.0) * (a[2] / a[3]) * (a[2] / a[3]) + a[1] / a[3]) / 3.0;
        double q = ((2.0 / 27.0) * (a[2] / a[3]) * (a[2] / a[3]) * (a[2] / a[3]) -
                    (1.0 / 3.0) * (a[2] / a[3]) * (a[1] / a[3]) + a[0] / a[3]) / 2.0;
        auto cbrt_s = [](double x) { return x == 0.0 ? 0.0 : (x > 0.0 ? pow(x, 1.0 / 3.0) : -pow(-x, 1.0 / 3.0)); };
        if (q * q + p * p * p >= 0.0) {
            n = 1;
            roots[0] = cbrt_s(-q + sqrt(q * q + p * p * p)) + cbrt_s(-q - sqrt(q * q + p * p * p)) -
                       a[2] / (3.0 * a[3]);
        } else {
            double phi = acos(-q / sqrt(-p * p * p));
            n = 3;
            roots[0] = 2.0 * sqrt(-p) * cos(phi / 3.0) - a[2] / (3.0 * a[3]);
            roots[1] = -2.0 * sqrt(-p) * cos(phi / 3.0 + kPi / 3.0) - a[2] / (3.0 * a[3]);
            roots[2] = -2.0 * sqrt(-p) * cos(phi / 3.0 - kPi / 3.0) - a[2] / (3.0 * a[3]);
        }
    }
    double best = 1000.0;
    for (int i = 0; i < n; i++)
        if (roots[i] > 0.0 && roots[i] < best) best = roots[i];
    return best;
}

// Remap::init (Remap.cpp:147-249).
void chain_init(CameraChain& c, const pano_photo_param& p0, int canvasW, int canvasH, int camW,
                int camH)
{
    pano_photo_param p = p0;
    c.canvasTX = canvasW / 2;
    c.canvasTY = canvasH / 2;
    if (p.imageType == 0 || p.imageType == 1) {
        p.cropX = 0;
        p.cropY = 0;
        p.cropWidth = camW;
        p.cropHeight = camH;
    }
    c.rectilinear = p.imageType == 0;
    c.camTX = p.cropX + p.cropWidth / 2;
    c.camTY = p.cropY + p.cropHeight / 2;
    c.horizontal = p.shiftX != 0 ? p.shiftX : 0;
    c.vertical = p.shiftY != 0 ? p.shiftY : 0;
    double fov = deg2rad(p.hfov), turn = deg2rad(360);
    rotation(-deg2rad(p.pitch), 0.0, -deg2rad(p.roll), c.mt);
    c.distance = ((double)canvasW) / turn;
    if (c.rectilinear)
        c.scale = ((double)p.cropWidth) / (2.0 * tan(fov / 2.0)) / c.distance;
    else
        c.scale = ((double)p.cropWidth) / fov / c.distance;
    c.shear0 = p.shearX / p.cropHeight;
    c.shear1 = p.shearY / p.cropWidth;
    c.halfTurn = c.distance * kPi;
    c.yawShift = -p.yaw * c.distance * kPi / 180.0;
    double radial[4] = {1.0, 0.0, 0.0, 0.0};
    if (p.beta != 0.0 || p.gamma != 0.0 || p.alpha != 0.0) {
        radial[3] = p.alpha;
        radial[2] = p.beta;
        radial[1] = p.gamma;
        radial[0] = 1.0 - (p.alpha + p.beta + p.gamma);
    }
    double deriv[4];
    for (int k = 0; k < 4; k++) deriv[k] = radial[k] != 0.0 ? (k + 1) * radial[k] : 0.0;
    for (int k = 0; k < 4; k++) c.rad[k] = radial[k];
    c.rad[5] = smallest_positive_root(deriv);
    c.rad[4] = ((double)(p.cropWidth < p.cropHeight ? p.cropWidth : p.cropHeight)) / 2.0;
}

// Remap::remapImage (Remap.cpp:356-514): canvas pixel -> camera coordinates.
__host__ __device__ void chain_apply(const CameraChain& c, double cx, double cy, double& sx, double& sy)
{
    cx -= c.canvasTX - 0.5;
    cy -= c.canvasTY - 0.5;
    double x = cx + c.yawShift;
    while (x < -c.halfTurn) x += 2 * c.halfTurn;
    while (x > c.halfTurn) x -= 2 * c.halfTurn;
    double y = cy;
    double phi = x / c.distance;
    double theta = -y / c.distance + kPi / 2;
    if (theta < 0) {
        theta = -theta;
        phi += kPi;
    }
    if (theta > kPi) {
        theta = kPi - (theta - kPi);
        phi += kPi;
    }
    double a0 = sin(theta) * sin(phi), a1 = cos(theta), a2 = sin(theta) * cos(phi);
    double v[3];
    for (int i = 0; i < 3; i++) v[i] = c.mt[0][i] * a0 + c.mt[1][i] * a1 + c.mt[2][i] * a2;
    double r = sqrt(v[0] * v[0] + v[1] * v[1]);
    theta = r == 0.0 ? 0.0 : c.distance * atan2(r, v[2]) / r;
    x = theta * v[0];
    y = theta * v[1];
    if (c.rectilinear) {
        double rr = sqrt(x * x + y * y);
        double th = rr / c.distance;
        double rho = th >= kPi / 2.0 ? 1.6e16 : (th == 0.0 ? 1.0 : tan(th) / th);
        x = rho * x;
        y = rho * y;
    }
    x = x * c.scale;
    y = y * c.scale;
    double rt = (sqrt(x * x + y * y)) / c.rad[4];
    double s = rt < c.rad[5] ? ((c.rad[3] * rt + c.rad[2]) * rt + c.rad[1]) * rt + c.rad[0] : 1000.0;
    x = x * s;
    y = y * s;
    if (c.vertical != 0.0) y = y + c.vertical;
    if (c.horizontal != 0.0) x = x + c.horizontal;
    if (c.shear0 != 0 || c.shear1 != 0) {
        double nx = x + c.shear0 * y;
        double ny = y + c.shear1 * x;
        x = nx;
        y = ny;
    }
    sx = x + (c.camTX - 0.5);
    sy = y + (c.camTY - 0.5);
}

// The chain and the validity test of one camera (getReprojectMapAndMask,
// ZReproject.cpp:152 and :173-199): inside the frame, or for circular / drum
// fisheyes inside the crop rectangle and the circle.
struct MapSetup {
    CameraChain c;
    bool circ = false;
    double cx = 0, cy = 0, r2 = 0;
    int cropX = 0, cropY = 0, cropW = 0, cropH = 0, src_w = 0, src_h = 0;
};

MapSetup map_setup(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h)
{
    MapSetup m;
    chain_init(m.c, p, dst_w, dst_h, src_w, src_h);
    m.circ = p.imageType == 3 || p.imageType == 2;
    m.cropX = p.cropX;
    m.cropY = p.cropY;
    m.cropW = p.cropWidth;
    m.cropH = p.cropHeight;
    m.src_w = src_w;
    m.src_h = src_h;
    if (m.circ) {
        if (p.circleR > 0) {
            m.cx = p.circleX;
            m.cy = p.circleY;
            m.r2 = (double)p.circleR * p.circleR;
        } else {
            m.cx = p.cropX + p.cropWidth / 2;
            m.cy = p.cropY + p.cropHeight / 2;
            m.r2 = p.cropWidth > p.cropHeight ? p.cropWidth * p.cropWidth * 0.25 : p.cropHeight * p.cropHeight * 0.25;
        }
    }
    return m;
}

// Canvas pixel (w, h): the map point (float, -1 where invalid) and the mask byte.
__host__ __device__ void map_point(const MapSetup& m, int w, int h, float& x, float& y, uint8_t& mask)
{
    double sx, sy;
    chain_apply(m.c, w, h, sx, sy);
    bool ok;
    if (!m.circ)
        ok = sx >= 0 && sy >= 0 && sx < m.src_w && sy < m.src_h;
    else
        ok = sx >= m.cropX && sy >= m.cropY && sx < m.cropX + m.cropW && sy < m.cropY + m.cropH &&
             (sx - m.cx) * (sx - m.cx) + (sy - m.cy) * (sy - m.cy) < m.r2;
    mask = ok ? 255 : 0;
    x = ok ? (float)sx : -1.f;
    y = ok ? (float)sy : -1.f;
}

__global__ __launch_bounds__(256) void k_generate_map(MapSetup m, int dst_w, int dst_h, float* __restrict__ xmap,
                                                      float* __restrict__ ymap, size_t map_step,
                                                      uint8_t* __restrict__ mask, size_t mask_step)
{
    const int w = blockIdx.x * 64 + (threadIdx.x & 63), h = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (w >= dst_w || h >= dst_h) return;
    float x, y;
    uint8_t k;
    map_point(m, w, h, x, y, k);
    ((float*)((char*)xmap + (size_t)h * map_step))[w] = x;
    ((float*)((char*)ymap + (size_t)h * map_step))[w] = y;
    mask[(size_t)h * mask_step + w] = k;
}

}  // namespace

int generate_map(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h,
                 float* xmap, float* ymap, uint8_t* mask)
{
    if (src_w <= 0 || src_h <= 0 || dst_w <= 0 || dst_h <= 0 || !xmap || !ymap || !mask)
        return PANO_ERR_ARG;
    const MapSetup m = map_setup(p, src_w, src_h, dst_w, dst_h);
    int hw = (int)std::thread::hardware_concurrency();
    int nt = std::max(1, std::min(hw > 0 ? hw : 1, 32));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (int h = t; h < dst_h; h += nt)
                for (int w = 0; w < dst_w; w++) {
                    size_t o = (size_t)h * dst_w + w;
                    map_point(m, w, h, xmap[o], ymap[o], mask[o]);
                }
        });
    for (auto& x : th) x.join();
    return PANO_OK;
}

// The same on the device: xmap / ymap / mask are device planes with the given
// row steps (bytes).
int generate_map_device(const pano_photo_param& p, int src_w, int src_h, int dst_w, int dst_h, float* xmap,
                        float* ymap, size_t map_step, uint8_t* mask, size_t mask_step, hipStream_t s)
{
    if (src_w <= 0 || src_h <= 0 || dst_w <= 0 || dst_h <= 0 || !xmap || !ymap || !mask ||
        map_step < (size_t)dst_w * 4 || map_step % 4 || mask_step < (size_t)dst_w)
        return PANO_ERR_ARG;
    const MapSetup m = map_setup(p, src_w, src_h, dst_w, dst_h);
    hipLaunchKernelGGL(k_generate_map, dim3((dst_w + 63) / 64, (dst_h + 3) / 4), dim3(256), 0, s, m, dst_w, dst_h,
                       xmap, ymap, map_step, mask, mask_step);
    return hipGetLastError() == hipSuccess ? PANO_OK : PANO_ERR_HIP;
}

}  // namespace pano
