"""Parity at the workloads' full sizes: the 8-camera 7680x3840 rig (L=9) of the
batched configuration, the 6-camera 3840x1920 rig with gains (L=7), and the
4-fisheye 7680x3840 rig (L=10) with per-frame gains, each against the oracle
on the same inputs (bit-exact), plus the batched form and the NV12 output of
the 8K rig through size-independent checks (equal to render() alone, equal to
the conversion of the canvas)."""
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


def setup(name):
    wl = synth.workload(name, 1.0)
    xs, ys, ms = synth.maps_and_masks(wl)
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
    return wl, xs, ys, ms, r


@pytest.mark.parametrize("name", ["mixed8", "rect6"])
def test_full_size_render_matches_oracle(name):
    wl, xs, ys, ms, r = setup(name)
    assert r.fused
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    assert ob.num_levels == r.num_levels == wl.levels
    note(f"full size {name} {wl.dst_w}x{wl.dst_h} n={wl.n} L={wl.levels}: oracle")
    ref = ob.render(srcs, xs, ys, luts, nthreads=16)
    note(f"full size {name}: device")
    dsrc = [to_dev(synth.bgrx(s)) for s in srcs]
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    assert r.render(dsrc, out, luts)
    got = from_dev(out)
    assert_parity(got[..., :3], ref, exact=False, label=f"{name} full")
    assert_parity(got[..., :3], ref, exact=True, label=f"{name} full")
    if name == "mixed8":
        # batched with a second stream of the same rig, and written as NV12
        note("full size mixed8: batch of 2 and NV12 output")
        r2 = P.PanoramaRender()
        assert r2.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys, ms,
                               wl.levels)
        b = P.Batch([r, r2])
        outs = [torch.zeros_like(out), torch.zeros_like(out)]
        b.render([dsrc, [t.clone() for t in dsrc]], outs, [luts, luts])
        for o in outs:
            np.testing.assert_array_equal(from_dev(o), got)
        y = torch.empty((wl.dst_h, wl.dst_w), dtype=torch.uint8, device="cuda")
        uv = torch.empty((wl.dst_h // 2, wl.dst_w), dtype=torch.uint8, device="cuda")
        assert r.render_to(dsrc, P.Output.nv12(y, uv), luts)
        ry, ruv = O.bgr32_to_nv12(got)
        np.testing.assert_array_equal(from_dev(y), ry)
        np.testing.assert_array_equal(from_dev(uv), ruv)


def test_full_size_autogain_matches_oracle():
    wl, xs, ys, ms, r = setup("fisheye4")
    assert r.fused and r.num_levels == 10
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    rng = np.random.default_rng(29)
    k = rng.uniform(0.85, 1.15, size=(wl.n, 3))
    srcs = [np.clip(np.rint(s * k[i]), 0, 255).astype(np.uint8) for i, s in enumerate(synth.frames(wl))]
    note(f"full size fisheye4 {wl.dst_w}x{wl.dst_h} L=10 autogain: oracle")
    rep = [O.reproject_to16s(s, x, y, nthreads=16).astype(np.uint8) for s, x, y in zip(srcs, xs, ys)]
    g, rg, bg, anchor, luts = O.gain_from_sums(O.gain_sums(rep, ms), True)
    ref = ob.render(srcs, xs, ys, luts, nthreads=16)
    note("full size fisheye4 autogain: device")
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    assert r.render_autogain([to_dev(synth.bgrx(s)) for s in srcs], out, tint=True)
    dg, drg, dbg, danchor = r.gains()
    np.testing.assert_array_equal(dg, g)
    assert danchor == anchor
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label="fisheye4 full autogain")
