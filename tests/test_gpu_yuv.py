"""GPU parity of the encoder / decoder hand-off: the standalone conversions
(cvtBGR32ToNV12 / ToYUV420P and inverses) against the oracle, and renders
written straight as NV12 / YUV420P planes against the oracle's conversion of
the same render's BGRx canvas.  All integer: bit-exact."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import panostitcher_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from panostitcher_amd import synth  # noqa: E402
from tests.gpu_util import from_dev, note, to_dev  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    P.lib()
    yield


def padded(h, w, pad, device="cuda"):
    """A uint8 (h, w) view with `pad` extra bytes per row."""
    big = torch.full((h, w + pad), 77, dtype=torch.uint8, device=device)
    return big[:, :w]


@pytest.mark.parametrize("h,w", [(38, 70), (2, 2), (64, 256)])
def test_conversions_match_oracle(h, w):
    rng = np.random.default_rng(h * w)
    img = rng.integers(0, 256, size=(h, w, 4)).astype(np.uint8)
    note(f"yuv conversions {w}x{h}")
    src = to_dev(img)
    y, u, v = P.cvt_bgr32_to_yuv420p(src, padded(h, w, 10), padded(h // 2, w // 2, 3), padded(h // 2, w // 2, 5))
    ry, ru, rv = O.bgr32_to_yuv420p(img)
    np.testing.assert_array_equal(from_dev(y), ry)
    np.testing.assert_array_equal(from_dev(u), ru)
    np.testing.assert_array_equal(from_dev(v), rv)
    y2, uv = P.cvt_bgr32_to_nv12(src, padded(h, w, 2), padded(h // 2, w, 4))
    ry2, ruv = O.bgr32_to_nv12(img)
    np.testing.assert_array_equal(from_dev(y2), ry2)
    np.testing.assert_array_equal(from_dev(uv), ruv)
    # inverses (channel 3 written as 0)
    back = from_dev(P.cvt_yuv420p_to_bgr32(y, u, v))
    np.testing.assert_array_equal(back, O.yuv420p_to_bgr32(ry, ru, rv))
    back2 = from_dev(P.cvt_nv12_to_bgr32(y2, uv))
    np.testing.assert_array_equal(back2, O.nv12_to_bgr32(ry2, ruv))


def render_planes(wl, r, dsrc, luts, fmt):
    h, w = wl.dst_h, wl.dst_w
    y = padded(h, w, 32)
    if fmt == "nv12":
        uv = padded(h // 2, w, 16)
        assert r.render_to(dsrc, P.Output.nv12(y, uv), luts)
        return from_dev(y), from_dev(uv)
    u, v = padded(h // 2, w // 2, 8), padded(h // 2, w // 2, 24)
    assert r.render_to(dsrc, P.Output.yuv420p(y, u, v), luts)
    return from_dev(y), from_dev(u), from_dev(v)


def new_render(wl, xs, ys, ms):
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys, ms,
                          wl.levels)
    return r


@pytest.mark.parametrize("name,scale", [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125)])
def test_render_writes_encoder_planes(name, scale):
    wl = synth.workload(name, scale)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    r = new_render(wl, xs, ys, ms)
    assert r.fused
    dsrc = [to_dev(synth.bgrx(s)) for s in srcs]
    canvas = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    assert r.render(dsrc, canvas, luts)
    bgrx = from_dev(canvas)
    # the canvas itself is checked against the oracle by the render tests
    note(f"render {name}@{scale} to NV12 / YUV420P")
    got = render_planes(wl, r, dsrc, luts, "nv12")
    for g, ref in zip(got, O.bgr32_to_nv12(bgrx)):
        np.testing.assert_array_equal(g, ref)
    got = render_planes(wl, r, dsrc, luts, "yuv420p")
    for g, ref in zip(got, O.bgr32_to_yuv420p(bgrx)):
        np.testing.assert_array_equal(g, ref)


def test_render_16bit_path_writes_encoder_planes():
    # maps sampling outside their masks keep the render on 16-bit planes; the
    # planes then come from a resident canvas and the standalone conversion
    wl = synth.workload("rect6", 0.2)
    xs, ys, ms = synth.maps_and_masks(wl)
    ms = [m.copy() for m in ms]
    rng = np.random.default_rng(5)
    for m in ms:
        ys_, xs_ = np.nonzero(m)
        pick = rng.choice(len(ys_), size=len(ys_) // 50, replace=False)
        m[ys_[pick], xs_[pick]] = 0
    r = new_render(wl, xs, ys, ms)
    assert not r.fused
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    ref = ob.render(srcs, xs, ys, luts, nthreads=8)
    note("render 16-bit path to NV12")
    dsrc = [to_dev(synth.bgrx(s)) for s in srcs]
    got = render_planes(wl, r, dsrc, luts, "nv12")
    for g, rr in zip(got, O.bgr32_to_nv12(synth.bgrx(ref))):
        np.testing.assert_array_equal(g, rr)


def test_batch_mixed_outputs_and_graph():
    cases = [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125)]
    renders, srcs, luts, bgrx = [], [], [], []
    for name, scale in cases:
        wl = synth.workload(name, scale)
        xs, ys, ms = synth.maps_and_masks(wl)
        r = new_render(wl, xs, ys, ms)
        d = [to_dev(synth.bgrx(s)) for s in synth.frames(wl)]
        lt = synth.luts(wl)
        c = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
        assert r.render(d, c, lt)
        renders.append((wl, r))
        srcs.append(d)
        luts.append(lt)
        bgrx.append(from_dev(c))
    b = P.Batch([r for _, r in renders])
    (w0, _), (w1, _), (w2, _) = renders
    c0 = torch.zeros((w0.dst_h, w0.dst_w, 4), dtype=torch.uint8, device="cuda")
    y1, uv1 = padded(w1.dst_h, w1.dst_w, 4), padded(w1.dst_h // 2, w1.dst_w, 6)
    y2, u2, v2 = (padded(w2.dst_h, w2.dst_w, 2), padded(w2.dst_h // 2, w2.dst_w // 2, 1),
                  padded(w2.dst_h // 2, w2.dst_w // 2, 3))
    outs = [P.Output.bgrx(c0), P.Output.nv12(y1, uv1), P.Output.yuv420p(y2, u2, v2)]

    def check(label):
        note(f"batch mixed outputs: {label}")
        np.testing.assert_array_equal(from_dev(c0), bgrx[0])
        for g, ref in zip((y1, uv1), O.bgr32_to_nv12(bgrx[1])):
            np.testing.assert_array_equal(from_dev(g), ref)
        for g, ref in zip((y2, u2, v2), O.bgr32_to_yuv420p(bgrx[2])):
            np.testing.assert_array_equal(from_dev(g), ref)

    b.render_to(srcs, outs, luts)
    check("eager")
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    b.capture_to(srcs, outs, luts, side)
    for t in (c0, y1, uv1, y2, u2, v2):
        t.fill_(0)
    torch.cuda.synchronize()
    for _ in range(2):
        b.replay(luts, side)
        side.synchronize()
        check("graph replay")
