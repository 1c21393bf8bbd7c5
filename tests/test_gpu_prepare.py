"""The fused render's prepare built on the device (boxes, weight pyramids,
weight sums, owners, cover bits, work lists, margined map storage) against
the same prepare built on the host: the rig must be the same, so the renders
are bit-identical and the work lists give the same per-stage byte counts and
the same device footprint.  The host prepare's renders are bit-exact against
the oracle in the parity tests; the device prepare is the default path there
as well."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import panostitcher_amd as P  # noqa: E402
from panostitcher_amd import synth  # noqa: E402
from tests.gpu_util import from_dev, note, to_dev  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    P.lib()
    yield


def _render(r, wl, srcs, luts, autogain):
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    if autogain:
        assert r.render_autogain(srcs, out, tint=True)
    else:
        assert r.render(srcs, out, luts)
    return from_dev(out)


# mixed8: full coverage; rect6: partial coverage (weight sums); fisheye4 with
# per-frame gains (cover bits); fisheye2: two cameras, small canvas
@pytest.mark.parametrize("name,scale,autogain", [("mixed8", 0.25, False), ("rect6", 0.3, False),
                                                 ("fisheye4", 0.15, True), ("fisheye2", 0.25, False)])
def test_device_prepare_matches_host(name, scale, autogain):
    wl = synth.workload(name, scale)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = [to_dev(synth.bgrx(f)) for f in synth.frames(wl)]
    luts = synth.luts(wl)
    sizes = [(c.width, c.height) for c in wl.cameras]
    res = []
    for on in (False, True):
        note(f"prepare {name}@{scale} on_device={on} {wl.dst_w}x{wl.dst_h} n={wl.n}")
        r = P.PanoramaRender()
        r.set_prepare_on_device(on)
        assert r.prepare_maps(sizes, (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
        assert r.fused
        img = _render(r, wl, srcs, luts, autogain)
        res.append((img, [r.model_bytes(k) for k in range(5)], r.mem_bytes(), r.prepare_ms()))
    note(f"prepare ms host {res[0][3]} device {res[1][3]}")
    np.testing.assert_array_equal(res[1][0], res[0][0])
    assert res[1][1] == res[0][1]
    assert res[1][2] == res[0][2]


def test_device_prepare_from_params_matches_host():
    wl = synth.workload("mixed8", 0.2)
    srcs = [to_dev(synth.bgrx(f)) for f in synth.frames(wl)]
    params = [c.param() for c in wl.cameras]
    sizes = [(c.width, c.height) for c in wl.cameras]
    imgs = []
    for on in (False, True):
        note(f"prepare from params on_device={on}")
        r = P.PanoramaRender()
        r.set_prepare_on_device(on)
        assert r.prepare(params, sizes, (wl.dst_w, wl.dst_h), wl.levels)
        imgs.append(_render(r, wl, srcs, None, False))
    np.testing.assert_array_equal(imgs[1], imgs[0])


def test_device_prepare_repeated_and_batched():
    # a render prepared twice on the device, and two renders in one batch
    wl = synth.workload("rect6", 0.25)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = [to_dev(synth.bgrx(f)) for f in synth.frames(wl)]
    sizes = [(c.width, c.height) for c in wl.cameras]
    r = P.PanoramaRender()
    r.set_prepare_on_device(False)
    assert r.prepare_maps(sizes, (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
    ref = _render(r, wl, srcs, None, False)
    d = P.PanoramaRender()
    for k in range(2):
        note(f"device prepare, round {k}")
        assert d.prepare_maps(sizes, (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
        np.testing.assert_array_equal(_render(d, wl, srcs, None, False), ref)


@pytest.mark.parametrize("case", ["empty_camera", "single_camera"])
def test_device_prepare_edge_rigs(case):
    # a camera whose mask is empty (its maps invalid everywhere), and a rig of
    # one camera (partial coverage)
    wl = synth.workload("mixed8" if case == "empty_camera" else "fisheye2", 0.15)
    xs, ys, ms = (list(v) for v in synth.maps_and_masks(wl))
    frames = synth.frames(wl)
    sizes = [(c.width, c.height) for c in wl.cameras]
    if case == "empty_camera":
        xs[3] = np.full_like(xs[3], -1)
        ys[3] = np.full_like(ys[3], -1)
        ms[3] = np.zeros_like(ms[3])
    else:
        xs, ys, ms, frames, sizes = xs[:1], ys[:1], ms[:1], frames[:1], sizes[:1]
    srcs = [to_dev(synth.bgrx(f)) for f in frames]
    imgs = []
    for on in (False, True):
        note(f"edge rig {case} on_device={on}")
        r = P.PanoramaRender()
        r.set_prepare_on_device(on)
        assert r.prepare_maps(sizes, (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
        out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
        assert r.render(srcs, out)
        imgs.append((from_dev(out), [r.model_bytes(k) for k in range(5)]))
    np.testing.assert_array_equal(imgs[1][0], imgs[0][0])
    assert imgs[1][1] == imgs[0][1]


@pytest.mark.parametrize("shape", [1, 2, 3, 4])
def test_tile_shapes_match_default(shape):
    # the other down tile shapes (64x32, 128x16, 32x32, 64x8) give the default shape's canvas bit
    # for bit (the default's canvas is bit-exact against the oracle), with the
    # tables built on the device and on the host alike
    wl = synth.workload("mixed8", 0.2)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = [to_dev(synth.bgrx(f)) for f in synth.frames(wl)]
    luts = synth.luts(wl)
    sizes = [(c.width, c.height) for c in wl.cameras]
    imgs = []
    for sh, on in ((0, True), (shape, True), (shape, False)):
        note(f"tile shape {sh} on_device={on}")
        r = P.PanoramaRender()
        r.set_prepare_on_device(on)
        r.set_tile_shape(sh)
        assert r.prepare_maps(sizes, (wl.dst_w, wl.dst_h), xs, ys, ms, wl.levels)
        assert r.tile_shape == sh
        imgs.append(_render(r, wl, srcs, luts, False))
    np.testing.assert_array_equal(imgs[1], imgs[0])
    np.testing.assert_array_equal(imgs[2], imgs[0])


def test_tile_autotune_and_batch_shapes():
    wl = synth.workload("rect6", 0.25)
    srcs = [to_dev(synth.bgrx(f)) for f in synth.frames(wl)]
    params = [c.param() for c in wl.cameras]
    sizes = [(c.width, c.height) for c in wl.cameras]
    ref = P.PanoramaRender()
    assert ref.prepare(params, sizes, (wl.dst_w, wl.dst_h), wl.levels)
    want = _render(ref, wl, srcs, None, False)
    r = P.PanoramaRender()
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    note("tile autotune rect6@0.25")
    best, ms = r.autotune_tile_shape(params, sizes, (wl.dst_w, wl.dst_h), wl.levels, srcs, out, iters=2)
    note(f"best shape {best}, ms {ms}")
    assert best in range(len(P.PanoramaRender.TILE_SHAPES)) and r.tile_shape == best and all(v > 0 for v in ms)
    np.testing.assert_array_equal(_render(r, wl, srcs, None, False), want)
    # renders of different shapes do not share a batched launch
    other = P.PanoramaRender()
    other.set_tile_shape((best + 1) % len(P.PanoramaRender.TILE_SHAPES))
    assert other.prepare(params, sizes, (wl.dst_w, wl.dst_h), wl.levels)
    b = P.Batch()
    with pytest.raises(P.PanoError):
        b.set([r, other])


def test_server_with_mixed_tile_shapes():
    # streams whose rigs were prepared with different down tile shapes go
    # through the serving loop in separate batches; every frame equals the
    # frame rendered alone
    rigs = []
    for name, scale, shape in (("rect6", 0.2, 0), ("rect6", 0.2, 3), ("fisheye2", 0.25, 4)):
        wl = synth.workload(name, scale)
        xs, ys, ms = synth.maps_and_masks(wl)
        r = P.PanoramaRender()
        r.set_tile_shape(shape)
        assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys, ms,
                              wl.levels)
        srcs = [to_dev(synth.bgrx(s)) for s in synth.frames(wl, 1)]
        alone = _render(r, wl, srcs, None, False)
        rigs.append((wl, r, srcs, alone))
    srv = P.Server(sum(r.mem_bytes() for _, r, _, _ in rigs) * 2, max_inflight=2)
    ids = [srv.admit(r) for _, r, _, _ in rigs]
    assert all(i is not None for i in ids)
    note("server with tile shapes 0, 3, 4")
    outs = []
    for rep in range(2):
        for k, (wl, r, srcs, _) in enumerate(rigs):
            o = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
            outs.append((k, o))
            srv.submit(ids[k], srcs, o, None, tag=rep)
    srv.drain()
    done = srv.poll(64)
    assert len(done) == len(outs) and all(d[3] == 0 for d in done)
    srv.close()
    for k, o in outs:
        np.testing.assert_array_equal(from_dev(o), rigs[k][3])
