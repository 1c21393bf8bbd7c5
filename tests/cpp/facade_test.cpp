// facade_test.cpp -- the C++ API (include/panostitcher/CudaInterface.hpp) used the
// way PanoStitcher's tasks use CudaAccel: a rig file loaded with loadPhotoParams,
// maps and masks generated on the device, frames reprojected and blended with
// CudaTilingMultibandBlendFast (the reference's per-camera flow, RicohUtil.cpp:
// 1301-1313), and the same rig through CudaPanoramaRender2.  The two must give
// the same canvas byte for byte (each is checked against the oracle by the
// Python parity tests); error conventions are checked as the reference has them.
//
//   facade_test host <rigs dir>   host-only checks (no GPU call)
//   facade_test gpu <rigs dir>    the full flow on the GPU
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "panostitcher/CudaInterface.hpp"

using namespace pano;

static int failures = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static uint32_t lcg(uint32_t& s)
{
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

static void host_checks(const std::string& rigs)
{
    std::vector<PhotoParam> ps;
    EXPECT(loadPhotoParams(rigs + "/ptgui.pts", ps));
    EXPECT(ps.size() == 3);
    EXPECT(ps.size() == 3 && ps[0].imageType == 3 && ps[0].yaw == -90.5 && ps[0].cropWidth == 1900);
    std::vector<PhotoParam> xs;
    EXPECT(loadPhotoParams(rigs + "/per_camera.xml", xs) && xs.size() == 2);
    EXPECT(!loadPhotoParams(rigs + "/no_extension", xs) && xs.empty());
    const std::string out = "facade_test_rig.xml";
    exportPhotoParamToXML(out, ps);
    std::vector<PhotoParam> back;
    loadPhotoParamFromXML(out, back);
    EXPECT(back.size() == ps.size() && back[1].yaw == 89.75 && back[1].imageType == 1);
    std::remove(out.c_str());
    PhotoParam q = ps[2];
    rotateCamera(q, 0, 0, 0);
    EXPECT(q.yaw > 9.999999 && q.yaw < 10.000001);
    EXPECT(CudaTilingMultibandBlendFast::estimateMemorySize(7680, 3840, 8) > 0);
    Mat m(4, 6, kType8UC4);
    EXPECT(m.step == 24 && m.elemSize() == 4 && !m.empty() && m.size() == Size(6, 4));
    bool threw = false;
    try {
        check(PANO_ERR_ARG, "probe");
    } catch (const Exception& e) {
        threw = e.status == PANO_ERR_ARG;
    }
    EXPECT(threw);
}

static void gpu_checks(const std::string& rigs)
{
    const std::string rig = rigs + "/ptgui.pts";
    std::vector<PhotoParam> params;
    EXPECT(loadPhotoParams(rig, params));
    const int n = (int)params.size();
    const Size src(1920, 1920), dst(1024, 512);
    const int levels = 5;

    // frames and gains
    std::vector<Mat> frames(n);
    uint32_t seed = 12345;
    for (Mat& f : frames) {
        f.create(src.height, src.width, kType8UC4);
        for (int y = 0; y < f.rows; y++)
            for (int x = 0; x < f.cols; x++) {
                uint8_t* p = f.ptr(y) + 4 * x;
                const uint32_t v = lcg(seed);
                p[0] = (uint8_t)(v + x / 7);
                p[1] = (uint8_t)((v >> 8) + y / 5);
                p[2] = (uint8_t)((v >> 16) + (x + y) / 9);
                p[3] = 0;
            }
    }
    std::vector<std::vector<std::vector<unsigned char>>> luts(n, std::vector<std::vector<unsigned char>>(3));
    for (int i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) {
            luts[i][c].resize(256);
            check(pano_get_lut(luts[i][c].data(), 0.9 + 0.05 * i + 0.03 * c), "getLUT");
        }

    // the reference's per-camera flow: maps and masks, cudaTransform,
    // cudaReprojectTo16S, CudaTilingMultibandBlendFast::blend
    std::vector<GpuMat> xmaps, ymaps, dmasks;
    cudaGenerateReprojectMapsAndMasks(params, src, dst, xmaps, ymaps, dmasks);
    std::vector<Mat> masks(n);
    for (int i = 0; i < n; i++) dmasks[i].download(masks[i]);
    CudaTilingMultibandBlendFast blender;
    EXPECT(blender.prepare(masks, levels, 2));
    Stream stream(true);
    std::vector<GpuMat> gsrc(n), adj(n), reproj(n);
    for (int i = 0; i < n; i++) {
        gsrc[i].upload(frames[i], stream);
        cudaTransform(gsrc[i], adj[i], luts[i], stream);
        cudaReprojectTo16S(adj[i], reproj[i], xmaps[i], ymaps[i], stream);
    }
    GpuMat blended;
    blender.blend(reproj, blended, stream);
    stream.waitForCompletion();
    Mat a;
    blended.download(a);

    // begin / tile / end: the same canvas
    GpuMat tiled;
    blender.begin();
    for (int i = 0; i < n; i++) blender.tile(reproj[i], i, stream);
    blender.end(tiled, stream);
    stream.waitForCompletion();
    Mat t;
    tiled.download(t);

    // the whole rig through CudaPanoramaRender2 (fused path)
    CudaPanoramaRender2 render;
    EXPECT(render.prepare(rig, 1, levels, src, dst));
    EXPECT(render.getNumImages() == n);
    GpuMat rendered;
    EXPECT(render.render(frames, rendered, luts));
    Mat b;
    rendered.download(b);

    long diff_blend_render = 0, diff_tile = 0, nonzero = 0;
    for (int y = 0; y < dst.height; y++)
        for (int x = 0; x < 4 * dst.width; x++) {
            diff_blend_render += a.ptr(y)[x] != b.ptr(y)[x];
            diff_tile += a.ptr(y)[x] != t.ptr(y)[x];
            nonzero += a.ptr(y)[x] != 0;
        }
    std::printf("blend vs render: %ld differing bytes; tile vs blend: %ld; non-zero bytes %ld\n", diff_blend_render,
                diff_tile, nonzero);
    EXPECT(diff_blend_render == 0);
    EXPECT(diff_tile == 0);
    EXPECT(nonzero > dst.width * dst.height);

    // the linear mode of the render
    CudaPanoramaRender2 linear;
    EXPECT(linear.prepare(rig, 0, 8, src, dst));
    GpuMat lin;
    EXPECT(linear.render(frames, lin));

    // encoder hand-off round trip keeps the canvas' size
    GpuMat yp, uv, back;
    cvtBGR32ToNV12(rendered, yp, uv, stream);
    cvtNV12ToBGR32(yp, uv, back, stream);
    stream.waitForCompletion();
    EXPECT(back.size() == dst && uv.rows == dst.height / 2);

    // encoder planes downloaded asynchronously into pooled page-locked buffers
    // (PinnedMemoryPool, as the stitch-to-disk task does)
    PinnedMemoryPool ypool, uvpool;
    EXPECT(ypool.init(dst.height, dst.width, kType8UC1) && uvpool.init(dst.height / 2, dst.width, kType8UC1));
    const uint8_t* first = nullptr;
    {
        HostMem hy, huv;
        EXPECT(ypool.get(hy) && uvpool.get(huv));
        first = hy.data;
        HostMem other;
        EXPECT(ypool.get(other) && other.data != hy.data);   // hy is held: a second buffer
        download(yp, hy, stream);
        download(uv, huv, stream);
        stream.waitForCompletion();
        Mat ys, uvs;
        yp.download(ys);
        uv.download(uvs);
        long d = 0;
        for (int y = 0; y < ys.rows; y++)
            d += std::memcmp(ys.ptr(y), hy.createMatHeader().ptr(y), ys.cols) != 0;
        for (int y = 0; y < uvs.rows; y++)
            d += std::memcmp(uvs.ptr(y), huv.createMatHeader().ptr(y), uvs.cols) != 0;
        EXPECT(d == 0);
    }
    HostMem again;
    EXPECT(ypool.get(again) && again.data == first && ypool.size() == 2);   // released: reused

    // error conventions
    std::vector<Mat> bad = masks;
    bad[1] = Mat(10, 10, kType8UC1);
    CudaTilingMultibandBlendFast b2;
    EXPECT(!b2.prepare(bad, levels, 2));                 // mismatched masks: false
    GpuMat untouched;
    b2.blend(reproj, untouched);                         // not prepared: silent return
    EXPECT(untouched.empty());
    bool threw = false;
    try {
        cudaReprojectTo16S(reproj[0], untouched, xmaps[0], ymaps[0]);   // 16SC4 source
    } catch (const Exception& e) {
        threw = e.status == PANO_ERR_ARG;
    }
    EXPECT(threw);
    threw = false;
    try {
        std::vector<GpuMat> few(reproj.begin(), reproj.begin() + 1);
        blender.blend(few, untouched);
    } catch (const Exception&) {
        threw = true;
    }
    EXPECT(threw);
    std::vector<Mat> wrong = frames;
    wrong[0] = Mat(8, 8, kType8UC4);
    EXPECT(!render.render(wrong, rendered));             // wrong frame size: false
    EXPECT(!CudaPanoramaRender2().prepare(rig, 1, levels, src, Size(1000, 400)));   // not 2:1
    EXPECT(!CudaPanoramaRender2().prepare(rigs + "/missing.pts", 1, levels, src, dst));
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s host|gpu <rigs dir>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1], rigs = argv[2];
    try {
        host_checks(rigs);
        if (mode == "gpu") gpu_checks(rigs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 1;
    }
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("facade %s checks passed\n", mode.c_str());
    return 0;
}
