// host_prepare.cpp -- rig preparation on the host (the reference prepares on
// the CPU as well: CudaMultibandBlend.cpp:549-642 computes the seams with
// cv::distanceTransform on host Mats before uploading the pyramids).
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>

#include "../../include/panohip.h"
#include "pano_host.hpp"

namespace pano {

namespace {

int wrap_idx(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    p %= n;
    return p < 0 ? p + n : p;
}

int reflect_idx(int p, int n)
{
    if ((unsigned)p < (unsigned)n) return p;
    if (n == 1) return 0;
    do {
        p = p < 0 ? -p : 2 * n - 2 - p;
    } while ((unsigned)p >= (unsigned)n);
    return p;
}

// Smallest cyclic interval [x0, x0 + len) of [0, n) containing every set flag.
void cyclic_interval(const std::vector<uint8_t>& f, int n, int& x0, int& len)
{
    int count = 0;
    for (int i = 0; i < n; i++) count += f[i] != 0;
    if (count == 0) { x0 = 0; len = 0; return; }
    if (count == n) { x0 = 0; len = n; return; }
    // the longest run of empty columns (cyclic) is the complement
    int best = -1, best_start = 0;
    int start = -1;
    // find a set column to start scanning after it
    int s = 0;
    while (!f[s]) s++;
    int run = 0;
    for (int k = 1; k <= n; k++) {
        int i = (s + k) % n;
        if (!f[i]) {
            if (run == 0) start = i;
            run++;
        } else {
            if (run > best) { best = run; best_start = start; }
            run = 0;
        }
    }
    if (best <= 0) { x0 = 0; len = n; return; }
    x0 = (best_start + best) % n;
    len = n - best;
}

int round_pitch(int w) { return w <= 0 ? 0 : (w + 63) / 64 * 64; }

}  // namespace

// OpenCV 3.0 distanceTransform_3x3 (L1 metrics {1, 2}): fixed point with 16
// fractional bits; a one-pixel frame at INT_MAX >> 2 makes the outside count as
// far from any zero; forward pass over upper/left neighbours, backward pass
// over lower/right ones; float(t) * 2^-16.
void distance_transform_l1(const uint8_t* mask, size_t step, int w, int h, float* dst)
{
    const int HV = 1 << 16, DIAG = 2 << 16, INIT = INT_MAX >> 2;
    const float scale = 1.f / (1 << 16);
    const int tw = w + 2;
    std::vector<int> t((size_t)tw * (h + 2));
    for (int j = 0; j < tw; j++) {
        t[j] = INIT;
        t[(size_t)(h + 1) * tw + j] = INIT;
    }
    for (int i = 0; i < h; i++) {
        const uint8_t* s = mask + (size_t)i * step;
        int* r = t.data() + (size_t)(i + 1) * tw + 1;
        r[-1] = r[w] = INIT;
        const int* up = r - tw;
        for (int j = 0; j < w; j++) {
            if (!s[j]) {
                r[j] = 0;
            } else {
                int v = std::min(up[j - 1] + DIAG, up[j] + HV);
                v = std::min(v, up[j + 1] + DIAG);
                v = std::min(v, r[j - 1] + HV);
                r[j] = v;
            }
        }
    }
    for (int i = h - 1; i >= 0; i--) {
        int* r = t.data() + (size_t)(i + 1) * tw + 1;
        const int* dn = r + tw;
        float* d = dst + (size_t)i * w;
        for (int j = w - 1; j >= 0; j--) {
            int v = r[j];
            if (v > HV) {
                v = std::min(v, dn[j + 1] + DIAG);
                v = std::min(v, dn[j] + HV);
                v = std::min(v, dn[j - 1] + DIAG);
                v = std::min(v, r[j + 1] + HV);
                r[j] = v;
            }
            d[j] = (float)v * scale;
        }
    }
}

// getNonIntersectingMasks (Mask.cpp:376-397): camera i >= 1 takes the pixels
// where its own L1 distance to its mask edge exceeds the distance of the union
// of the cameras before it; those pixels are removed from all earlier cameras.
void nonintersecting_masks(int n, int rows, int cols, const uint8_t* const* masks,
                           const size_t* steps, uint8_t* out)
{
    const size_t P = (size_t)rows * cols;
    for (int i = 0; i < n; i++)
        for (int y = 0; y < rows; y++)
            memcpy(out + P * i + (size_t)y * cols, masks[i] + (size_t)y * steps[i], cols);
    if (n <= 1) return;
    std::vector<float> cur((size_t)n * P);
    parallel_for(n - 1, [&](int k) {
        distance_transform_l1(out + P * (k + 1), cols, cols, rows, cur.data() + P * (k + 1));
    });
    std::vector<uint8_t> acc(out, out + P);
    std::vector<float> dist(P);
    for (int i = 1; i < n; i++) {
        distance_transform_l1(acc.data(), cols, cols, rows, dist.data());
        uint8_t* ui = out + P * i;
        const float* ci = cur.data() + P * i;
        for (size_t p = 0; p < P; p++) ui[p] = ci[p] > dist[p] ? 255 : 0;
        for (int j = 0; j < i; j++) {
            uint8_t* uj = out + P * j;
            for (size_t p = 0; p < P; p++)
                if (ui[p]) uj[p] = 0;
        }
        for (size_t p = 0; p < P; p++) acc[p] |= ui[p];
    }
}

// getTrueNumLevels (MultiBandBlend.cpp:325-337).
int true_num_levels(int width, int height, int max_levels, int min_length)
{
    int num = max_levels < 0 ? 16 : (max_levels > 16 ? 16 : max_levels);
    int side = min_length < 0 ? 2 : (min_length > 64 ? 64 : min_length);
    int lw = (int)(log((double)width / side) / log(2.0));
    int lh = (int)(log((double)height / side) / log(2.0));
    return std::min(num, std::min(lw, lh));
}

void camera_boxes(int L, const int* lw, const int* lh, std::vector<uint8_t> colf, std::vector<uint8_t> rowf,
                  Geom* geom)
{
    for (int j = 0; j <= L; j++) {
        if (j > 0) {
            std::vector<uint8_t> c2(lw[j], 0), r2(lh[j], 0);
            for (int q = 0; q < lw[j]; q++)
                for (int b = -2; b <= 2 && !c2[q]; b++) c2[q] = colf[wrap_idx(2 * q + b, lw[j - 1])];
            for (int q = 0; q < lh[j]; q++)
                for (int a = -2; a <= 2 && !r2[q]; a++) r2[q] = rowf[reflect_idx(2 * q + a, lh[j - 1])];
            colf.swap(c2);
            rowf.swap(r2);
        }
        Geom& g = geom[j];
        g.cw = lw[j];
        g.ch = lh[j];
        int y0 = -1, y1 = -1;
        for (int y = 0; y < g.ch; y++)
            if (rowf[y]) {
                if (y0 < 0) y0 = y;
                y1 = y;
            }
        int x0, len;
        cyclic_interval(colf, g.cw, x0, len);
        if (y0 < 0 || len == 0) {
            g.x0 = g.y0 = g.bw = g.bh = 0;
        } else {
            g.x0 = x0;
            g.bw = len;
            g.y0 = y0;
            g.bh = y1 - y0 + 1;
        }
        g.pitch = round_pitch(g.bw);
    }
}

void work_lists(Prepared& out, const TileValidFn& tile_has_valid, const CamHitsFn& cam_hits,
                const TileOwnersFn& tile_owners)
{
    const int n = out.n, L = out.levels, rows = out.rows, cols = out.cols;
    const bool u8 = out.kind == kPixU8, fused0 = out.fused0;
    const int dtw = u8 ? out.dtw : kDownTW, dth = u8 ? out.dth : kDownTH;
    const int ttw = u8 ? 2 * kQTW : kTileW, tth = u8 ? 2 * kQTH : kTileH;
    out.down.assign(L, std::vector<DownJob>());
    for (int j = 0; j < L; j++) {
        for (int i = 0; i < n; i++) {
            const Geom& gd = out.geom[(size_t)i * (L + 1) + j + 1];
            if (gd.empty()) continue;
            int segs[2][2] = {{gd.x0, std::min(gd.x0 + gd.bw, gd.cw)}, {0, gd.x0 + gd.bw - gd.cw}};
            for (int sgi = 0; sgi < 2; sgi++) {
                int a = segs[sgi][0], b = segs[sgi][1];
                for (int X = a; X < b; X += dtw)
                    for (int Y = gd.y0; Y < gd.y0 + gd.bh; Y += dth) {
                        DownJob d{};
                        d.src = (fused0 && j == 0) ? i : i * (L + 1) + j;
                        d.dst = i * (L + 1) + j + 1;
                        d.X = X;
                        d.Y = Y;
                        d.nx = std::min(dtw, b - X);
                        d.ny = std::min(dth, gd.y0 + gd.bh - Y);
                        // outputs whose alpha is zero stay zero: skip empty tiles
                        if (!tile_has_valid(i, j + 1, d.X, d.Y, d.nx, d.ny)) continue;
                        out.down[j].push_back(d);
                    }
            }
        }
    }
    out.blend.assign(L + 1, std::vector<TileJob>());
    out.camlist.clear();
    out.max_tile_cams = 0;
    for (int j = 1; j <= L; j++) {
        int tw = (out.lw[j] + ttw - 1) / ttw, th = (out.lh[j] + tth - 1) / tth;
        std::vector<std::vector<int>> cams((size_t)tw * th);
        for (int i = 0; i < n; i++) {
            std::vector<uint8_t> hit((size_t)tw * th, 0);
            cam_hits(j, i, hit);
            for (size_t t = 0; t < hit.size(); t++)
                if (hit[t]) cams[t].push_back(i * (L + 1) + j);
        }
        for (int ty = 0; ty < th; ty++)
            for (int tx = 0; tx < tw; tx++) {
                TileJob t{};
                t.X = tx * ttw;
                t.Y = ty * tth;
                t.nx = std::min(ttw, out.lw[j] - t.X);
                t.ny = std::min(tth, out.lh[j] - t.Y);
                t.r = 0;
                const std::vector<int>& c = cams[(size_t)ty * tw + tx];
                t.off = (int)out.camlist.size();
                t.cnt = (int)c.size();
                t.flags = j == L ? kTileTop : 0;
                out.camlist.insert(out.camlist.end(), c.begin(), c.end());
                out.max_tile_cams = std::max(out.max_tile_cams, t.cnt);
                out.blend[j].push_back(t);
            }
    }
    {
        int tw = (cols + ttw - 1) / ttw, th = (rows + tth - 1) / tth;
        out.fin.clear();
        out.fin_owners.clear();
        for (int ty = 0; ty < th; ty++)
            for (int tx = 0; tx < tw; tx++) {
                TileJob t{};
                t.X = tx * ttw;
                t.Y = ty * tth;
                t.nx = std::min(ttw, cols - t.X);
                t.ny = std::min(tth, rows - t.Y);
                t.off = (int)out.fin_owners.size();
                if (u8) tile_owners(t, out.fin_owners);
                t.cnt = (int)out.fin_owners.size() - t.off;
                out.fin.push_back(t);
            }
    }
    out.remap.clear();
    for (int i = 0; i < n; i++) {
        const Geom& g = out.geom[(size_t)i * (L + 1)];
        for (int ly = 0; ly < g.bh; ly += kRemapTH)
            for (int lx = 0; lx < g.bw; lx += kRemapTW) {
                RemapJob r{};
                r.cam = i;
                r.lx = lx;
                r.ly = ly;
                r.nx = std::min(kRemapTW, g.bw - lx);
                r.ny = std::min(kRemapTH, g.bh - ly);
                out.remap.push_back(r);
            }
    }
}

int prepare_rig(int n, int rows, int cols, const uint8_t* const* masks, const size_t* steps,
                const uint8_t* const* support0, int max_levels, int min_length, bool full_boxes,
                int kind, bool fused0, Prepared& out, const uint8_t* unique, int shape)
{
    if (shape < 0 || shape >= kDownShapes) return PANO_ERR_ARG;
    if (n <= 0 || n > 255 || rows <= 0 || cols <= 0 || !masks || !steps) return PANO_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!masks[i] || steps[i] < (size_t)cols) return PANO_ERR_ARG;
    out = Prepared();
    out.kind = kind;
    out.fused0 = fused0;
    out.shape = kind == kPixU8 ? shape : 0;
    out.dtw = kDownTileW[out.shape];
    out.dth = kDownTileH[out.shape];
    out.n = n;
    out.rows = rows;
    out.cols = cols;
    const size_t P = (size_t)rows * cols;

    // seams and owners
    out.unique.resize(P * n);
    if (unique)
        std::copy(unique, unique + P * n, out.unique.begin());
    else
        nonintersecting_masks(n, rows, cols, masks, steps, out.unique.data());
    out.label.assign(P, kNoOwner);
    for (int i = 0; i < n; i++) {
        const uint8_t* u = out.unique.data() + P * i;
        for (size_t p = 0; p < P; p++)
            if (u[p]) out.label[p] = (uint8_t)i;
    }
    // coverage (MultiBandBlend.cpp:961-964)
    size_t covered = 0;
    for (int y = 0; y < rows; y++)
        for (int x = 0; x < cols; x++) {
            bool c = false;
            for (int i = 0; i < n && !c; i++) c = masks[i][(size_t)y * steps[i] + x] != 0;
            covered += c;
        }
    out.full = covered == P;

    const int L = true_num_levels(cols, rows, max_levels, min_length);
    if (L < 0) return PANO_ERR_ARG;
    out.levels = L;
    out.lw.resize(L + 1);
    out.lh.resize(L + 1);
    out.lw[0] = cols;
    out.lh[0] = rows;
    for (int j = 1; j <= L; j++) {
        out.lw[j] = (out.lw[j - 1] + 1) / 2;
        out.lh[j] = (out.lh[j - 1] + 1) / 2;
    }

    // per-camera support projections -> boxes per level
    out.geom.assign((size_t)n * (L + 1), Geom());
    parallel_for(n, [&](int i) {
        std::vector<uint8_t> colf(cols, 0), rowf(rows, 0);
        if (full_boxes) {
            std::fill(colf.begin(), colf.end(), 1);
            std::fill(rowf.begin(), rowf.end(), 1);
        } else {
            for (int y = 0; y < rows; y++) {
                const uint8_t* m = masks[i] + (size_t)y * steps[i];
                const uint8_t* s = support0 ? support0[i] + (size_t)y * cols : nullptr;
                bool any = false;
                for (int x = 0; x < cols; x++) {
                    if (m[x] || (s && s[x])) {
                        colf[x] = 1;
                        any = true;
                    }
                }
                rowf[y] = any;
            }
        }
        camera_boxes(L, out.lw.data(), out.lh.data(), std::move(colf), std::move(rowf),
                     out.geom.data() + (size_t)i * (L + 1));
    });

    // validity (alpha != 0) per level on the boxes, MultiBandBlend.cpp:945-959:
    // level 0 is the mask; level j+1 is non-zero where any level-j pixel of the
    // 5x5 footprint is.  Used to drop work tiles whose outputs are all zero.
    std::vector<std::vector<uint8_t>> valid((size_t)n * (L + 1));
    parallel_for(n, [&](int i) {
        const Geom& g0 = out.geom[(size_t)i * (L + 1)];
        std::vector<uint8_t>& v0 = valid[(size_t)i * (L + 1)];
        v0.assign(g0.pixels(), 0);
        for (int ly = 0; ly < g0.bh; ly++)
            for (int lx = 0; lx < g0.bw; lx++) {
                int X = (g0.x0 + lx) % g0.cw, Y = g0.y0 + ly;
                v0[(size_t)ly * g0.pitch + lx] = masks[i][(size_t)Y * steps[i] + X] != 0;
            }
        for (int j = 1; j <= L; j++) {
            const Geom& gs = out.geom[(size_t)i * (L + 1) + j - 1];
            const Geom& gd = out.geom[(size_t)i * (L + 1) + j];
            const std::vector<uint8_t>& prev = valid[(size_t)i * (L + 1) + j - 1];
            std::vector<uint8_t>& cur = valid[(size_t)i * (L + 1) + j];
            cur.assign(gd.pixels(), 0);
            auto at = [&](int X, int Y) -> int {
                int ly = Y - gs.y0;
                if ((unsigned)ly >= (unsigned)gs.bh) return 0;
                int lx = X - gs.x0;
                if (lx < 0) lx += gs.cw;
                if ((unsigned)lx >= (unsigned)gs.bw) return 0;
                return prev[(size_t)ly * gs.pitch + lx];
            };
            // rows: any over the 5 source rows of a horizontally dilated row
            int r0 = std::max(0, 2 * gd.y0 - 2), r1 = std::min(gs.ch - 1, 2 * (gd.y0 + gd.bh - 1) + 2);
            std::vector<uint8_t> hrow((size_t)std::max(r1 - r0 + 1, 0) * std::max(gd.bw, 1), 0);
            for (int r = r0; r <= r1 && gd.bw > 0; r++)
                for (int lx = 0; lx < gd.bw; lx++) {
                    int X = (gd.x0 + lx) % gd.cw;
                    uint8_t a = 0;
                    for (int b = -2; b <= 2 && !a; b++) a = (uint8_t)at(wrap_idx(2 * X + b, gs.cw), r);
                    hrow[(size_t)(r - r0) * gd.bw + lx] = a;
                }
            for (int ly = 0; ly < gd.bh; ly++) {
                int Y = gd.y0 + ly;
                for (int lx = 0; lx < gd.bw; lx++) {
                    uint8_t a = 0;
                    for (int k = -2; k <= 2 && !a; k++)
                        a = hrow[(size_t)(reflect_idx(2 * Y + k, gs.ch) - r0) * gd.bw + lx];
                    cur[(size_t)ly * gd.pitch + lx] = a;
                }
            }
        }
    });
    auto tile_has_valid = [&](int i, int j, int X, int Y, int nx, int ny) {
        const Geom& g = out.geom[(size_t)i * (L + 1) + j];
        const std::vector<uint8_t>& v = valid[(size_t)i * (L + 1) + j];
        for (int y = Y; y < Y + ny; y++)
            for (int x = X; x < X + nx; x++) {
                int ly = y - g.y0, lx = x - g.x0;
                if (lx < 0) lx += g.cw;
                if ((unsigned)ly < (unsigned)g.bh && (unsigned)lx < (unsigned)g.bw &&
                    v[(size_t)ly * g.pitch + lx])
                    return true;
            }
        return false;
    };

    // blend weights: Gaussian pyramid of 256 * unique_i (MultiBandBlend.cpp:936-943,
    // pyrDown with FixPtCast<short, 8>, WRAP columns, REFLECT_101 rows), on the boxes.
    out.weight.assign((size_t)n * (L + 1), std::vector<uint16_t>());
    parallel_for(n, [&](int i) {
        const Geom& g0 = out.geom[(size_t)i * (L + 1)];
        std::vector<uint16_t> prev(g0.pixels(), 0);
        const uint8_t* u = out.unique.data() + P * i;
        for (int ly = 0; ly < g0.bh; ly++)
            for (int lx = 0; lx < g0.bw; lx++) {
                int X = (g0.x0 + lx) % g0.cw, Y = g0.y0 + ly;
                prev[(size_t)ly * g0.pitch + lx] = u[(size_t)Y * cols + X] ? 256 : 0;
            }
        for (int j = 1; j <= L; j++) {
            const Geom& gs = out.geom[(size_t)i * (L + 1) + j - 1];
            const Geom& gd = out.geom[(size_t)i * (L + 1) + j];
            auto fetch = [&](int X, int Y) -> int {
                int ly = Y - gs.y0;
                if ((unsigned)ly >= (unsigned)gs.bh) return 0;
                int lx = X - gs.x0;
                if (lx < 0) lx += gs.cw;
                if ((unsigned)lx >= (unsigned)gs.bw) return 0;
                return prev[(size_t)ly * gs.pitch + lx];
            };
            std::vector<uint16_t> cur(gd.pixels(), 0);
            static const int k[5] = {1, 4, 6, 4, 1};
            // horizontal pass on the source rows the output needs, then vertical
            int r0 = std::max(0, 2 * gd.y0 - 2), r1 = std::min(gs.ch - 1, 2 * (gd.y0 + gd.bh - 1) + 2);
            int nr = r1 - r0 + 1;
            std::vector<int> hrow((size_t)std::max(nr, 0) * std::max(gd.bw, 1), 0);
            for (int r = r0; r <= r1 && gd.bw > 0; r++)
                for (int lx = 0; lx < gd.bw; lx++) {
                    int X = (gd.x0 + lx) % gd.cw;
                    int s = 0;
                    for (int b = 0; b < 5; b++) s += k[b] * fetch(wrap_idx(2 * X + b - 2, gs.cw), r);
                    hrow[(size_t)(r - r0) * gd.bw + lx] = s;
                }
            for (int ly = 0; ly < gd.bh; ly++) {
                int Y = gd.y0 + ly;
                for (int lx = 0; lx < gd.bw; lx++) {
                    int s = 0;
                    for (int a = 0; a < 5; a++) {
                        int r = reflect_idx(2 * Y + a - 2, gs.ch);
                        s += k[a] * hrow[(size_t)(r - r0) * gd.bw + lx];
                    }
                    cur[(size_t)ly * gd.pitch + lx] = (uint16_t)((s + 128) >> 8);
                }
            }
            out.weight[(size_t)i * (L + 1) + j] = cur;
            prev.swap(cur);
        }
    });

    // sum of weights per level (MultiBandBlend.cpp:970-980) when not fully covered
    out.wsum.assign(L + 1, std::vector<uint16_t>());
    if (!out.full) {
        for (int j = 1; j <= L; j++) {
            std::vector<uint16_t>& ws = out.wsum[j];
            ws.assign((size_t)out.lw[j] * out.lh[j], 0);
            for (int i = 0; i < n; i++) {
                const Geom& g = out.geom[(size_t)i * (L + 1) + j];
                const std::vector<uint16_t>& w = out.weight[(size_t)i * (L + 1) + j];
                for (int ly = 0; ly < g.bh; ly++)
                    for (int lx = 0; lx < g.bw; lx++) {
                        int X = (g.x0 + lx) % g.cw, Y = g.y0 + ly;
                        ws[(size_t)Y * out.lw[j] + X] += w[(size_t)ly * g.pitch + lx];
                    }
            }
        }
    }

    // work lists
    const bool u8 = kind == kPixU8;
    const int ttw = u8 ? 2 * kQTW : kTileW, tth = u8 ? 2 * kQTH : kTileH;
    auto cam_hits = [&](int j, int i, std::vector<uint8_t>& hit) {
        const int tw = (out.lw[j] + ttw - 1) / ttw;
        const Geom& g = out.geom[(size_t)i * (L + 1) + j];
        const std::vector<uint16_t>& w = out.weight[(size_t)i * (L + 1) + j];
        for (int ly = 0; ly < g.bh; ly++)
            for (int lx = 0; lx < g.bw; lx++)
                if (w[(size_t)ly * g.pitch + lx]) {
                    int X = (g.x0 + lx) % g.cw, Y = g.y0 + ly;
                    hit[(size_t)(Y / tth) * tw + X / ttw] = 1;
                }
    };
    auto tile_owners = [&](const TileJob& t, std::vector<int>& owners) {
        uint8_t seen[256] = {};
        for (int y = t.Y; y < t.Y + t.ny; y++)
            for (int x = t.X; x < t.X + t.nx; x++) {
                uint8_t o = out.label[(size_t)y * cols + x];
                if (o != kNoOwner && !seen[o]) {
                    seen[o] = 1;
                    owners.push_back(o);
                }
            }
    };
    work_lists(out, tile_has_valid, cam_hits, tile_owners);
    return PANO_OK;
}

}  // namespace pano
