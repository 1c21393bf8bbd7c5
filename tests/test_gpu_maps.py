"""Map generation on the device (cudaGenerateReprojectMapAndMask) against the
host generator (bit-exact against the oracle's FP64 chain in the CPU suite).
The device's FP64 sin / cos / atan2 / tan may differ from the host library's
in the last bits, so the float maps are expected identical almost everywhere
and within 1e-3 px wherever both are valid, and the masks identical except
where a point lies within rounding distance of the validity boundary.  A
render prepared from camera parameters (device maps) then matches the oracle
run on the same device maps."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import panostitcher_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from panostitcher_amd import synth  # noqa: E402
from tests.gpu_util import assert_parity, from_dev, note, to_dev  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    P.lib()
    yield


@pytest.mark.parametrize("name,scale", [("fisheye2", 0.25), ("rect6", 0.2), ("mixed8", 0.1), ("fisheye4", 0.08)])
def test_device_maps_match_host(name, scale):
    wl = synth.workload(name, scale)
    for cam in wl.cameras:
        note(f"device map {name}@{scale} {cam.width}x{cam.height} -> {wl.dst_w}x{wl.dst_h}")
        hx, hy, hm = P.generate_map(cam.param(), cam.width, cam.height, wl.dst_w, wl.dst_h)
        dx, dy, dm = (from_dev(t) for t in P.generate_map_device(cam.param(), cam.width, cam.height,
                                                                  wl.dst_w, wl.dst_h))
        flips = int((dm != hm).sum())
        assert flips <= max(2, hm.size // 100000), f"{flips} mask flips"
        both = (dm > 0) & (hm > 0)
        assert np.abs(dx[both] - hx[both]).max() <= 1e-3 and np.abs(dy[both] - hy[both]).max() <= 1e-3
        same = ((dx == hx) & (dy == hy)).mean()
        assert same >= 0.9999, same
        assert ((dx == -1) == (dm == 0)).all() and ((dy == -1) == (dm == 0)).all()


def test_render_prepared_from_params_matches_oracle():
    wl = synth.workload("rect6", 0.2)
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    r = P.PanoramaRender()
    assert r.prepare([c.param() for c in wl.cameras], [(c.width, c.height) for c in wl.cameras],
                     (wl.dst_w, wl.dst_h), wl.levels)
    # the oracle on the same (device) maps and masks
    maps = [P.generate_map_device(c.param(), c.width, c.height, wl.dst_w, wl.dst_h) for c in wl.cameras]
    xs = [from_dev(m[0]) for m in maps]
    ys = [from_dev(m[1]) for m in maps]
    ms = [from_dev(m[2]) for m in maps]
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    ref = ob.render(srcs, xs, ys, luts, nthreads=8)
    note(f"render prepared from params {wl.dst_w}x{wl.dst_h}")
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    assert r.render([to_dev(synth.bgrx(s)) for s in srcs], out, luts)
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label="render from params")


def test_reproject_params_is_map_then_reproject():
    wl = synth.workload("fisheye2", 0.25)
    cam = wl.cameras[0]
    src = to_dev(synth.bgrx(synth.frames(wl)[0]))
    note("cudaReproject(src, dst, mask, dstSize, param)")
    x, y, m = P.generate_map_device(cam.param(), cam.width, cam.height, wl.dst_w, wl.dst_h)
    d8, mk = P.cuda_reproject_params(src, cam.param(), (wl.dst_w, wl.dst_h), with_mask=True)
    np.testing.assert_array_equal(from_dev(mk), from_dev(m))
    np.testing.assert_array_equal(from_dev(d8), from_dev(P.cuda_reproject(src, x, y)))
    d16 = P.cuda_reproject_params(src, cam.param(), (wl.dst_w, wl.dst_h), to16s=True)
    np.testing.assert_array_equal(from_dev(d16), from_dev(P.cuda_reproject_to16s(src, x, y)))


def test_weighted_accumulate_matches_linear_render():
    # accumulating every camera with the free function, then rounding, is the
    # linear render (whose oracle parity is tested with the linear mode)
    wl = synth.workload("rect6", 0.125)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = synth.frames(wl)
    w = O.linear_weights(np.stack(ms), 4)
    ref = O.linear_render(srcs, xs, ys, w, None)
    note("cudaReprojectWeightedAccumulateTo32F over rect6")
    acc = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.float32, device="cuda")
    for i in range(wl.n):
        P.cuda_reproject_weighted_accumulate(to_dev(synth.bgrx(srcs[i])), acc, to_dev(xs[i]), to_dev(ys[i]),
                                             to_dev(w[i]))
    a = from_dev(acc)
    got = np.clip(np.rint(a[..., :3]), 0, 255).astype(np.uint8)   # numpy rint: half to even, as convertTo
    np.testing.assert_array_equal(got, ref)
    assert not a[..., 3].any()


def test_render_prepared_from_rig_file():
    # CPUPanoramaRender::prepare(path, ...) loads the rig with loadPhotoParams;
    # the PTGui fixture holds three cameras (two fisheyes, one rectilinear)
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "rigs", "ptgui.pts")
    params = P.load_photo_params(path)
    assert len(params) == 3
    sw, sh, dw, dh, levels = 1920, 1920, 1024, 512, 5
    rng = np.random.default_rng(11)
    srcs = [rng.integers(0, 256, (sh, sw, 3), dtype=np.uint8) for _ in params]
    r = P.PanoramaRender()
    note("render prepared from ptgui.pts 3 x 1920x1920 -> 1024x512")
    assert r.prepare(path, (sw, sh), (dw, dh), levels)
    assert not P.PanoramaRender().prepare(path + ".missing", (sw, sh), (dw, dh), levels)
    maps = [P.generate_map_device(q, sw, sh, dw, dh) for q in params]
    xs = [from_dev(m[0]) for m in maps]
    ys = [from_dev(m[1]) for m in maps]
    ms = [from_dev(m[2]) for m in maps]
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), levels)
    ref = ob.render(srcs, xs, ys, None, nthreads=8)
    out = torch.zeros((dh, dw, 4), dtype=torch.uint8, device="cuda")
    assert r.render([to_dev(synth.bgrx(s)) for s in srcs], out)
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label="render from rig file")
