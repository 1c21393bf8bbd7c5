"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same inputs.  The blend is integer and must match bit for bit; the remap is
FP64 evaluated in the reference's order and is expected to match exactly too
(the stated bar, max |diff| <= 1 with >= 99.9 % identical, is asserted as well)."""
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


def random_masks(rng, n, h, w):
    masks = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        cx = int(rng.integers(0, w))
        half = int(rng.integers(w // 4, w // 2 + 2))
        cols = np.arange(cx - half, cx + half) % w
        y0 = int(rng.integers(0, h // 3))
        y1 = int(rng.integers(2 * h // 3, h + 1))
        masks[i][y0:y1][:, cols] = 255
    return masks


def test_reproject_to16s_matches_oracle():
    rng = np.random.default_rng(1)
    sh, sw = 37, 53
    src = rng.integers(0, 256, size=(sh, sw, 3), dtype=np.uint8)
    dh, dw = 24, 40
    xm = (rng.random((dh, dw)) * (sw + 2) - 1).astype(np.float32)
    ym = (rng.random((dh, dw)) * (sh + 2) - 1).astype(np.float32)
    # edge cases: integer points, the last column (next-row read), the last pixel
    xm[0, :5] = [0, sw - 1, sw - 0.5, sw - 0.25, sw - 0.5]
    ym[0, :5] = [0, 3, 3, sh - 1, sh - 1]
    xm[1, :3] = [2.5, 3.5, sw]
    ym[1, :3] = [sh - 0.5, 1.0, 2]
    note("reproject 53x37 -> 40x24")
    ref = O.reproject_to16s(src, xm, ym)
    out = P.cuda_reproject_to16s(to_dev(synth.bgrx(src)), to_dev(xm), to_dev(ym))
    got = from_dev(out)
    np.testing.assert_array_equal(got[..., :3], ref)
    assert not got[..., 3].any()
    # cudaReproject: the same samples as bytes
    note("reproject 8UC4 53x37 -> 40x24")
    got8 = from_dev(P.cuda_reproject(to_dev(synth.bgrx(src)), to_dev(xm), to_dev(ym)))
    np.testing.assert_array_equal(got8[..., :3], ref)
    assert not got8[..., 3].any()


# rows up to 8192 pixels take the LDS-staged row pass, wider ones the chunked one
@pytest.mark.parametrize("seed,n,h,w", [(0, 3, 40, 96), (1, 4, 17, 2000), (2, 1, 9, 13), (3, 5, 64, 300),
                                        (4, 3, 1, 77), (5, 3, 6, 9000), (6, 2, 5, 8192)])
def test_nonintersecting_masks_device_matches_oracle(seed, n, h, w):
    rng = np.random.default_rng(seed)
    masks = random_masks(rng, n, h, w) if h > 3 else (rng.random((n, h, w)) < 0.7).astype(np.uint8) * 255
    if n > 2:
        masks[1][:, :] = 200        # no zero pixel: the distance is the 8192 frame cap
        masks[2][masks[2] == 0] = 0
        masks[2][(rng.random((h, w)) < 0.1)] = 0
    note(f"seams on the device seed={seed} n={n} {w}x{h}")
    got = P.nonintersecting_masks(masks)
    np.testing.assert_array_equal(got, O.nonintersecting_masks(masks))


def test_transform_matches_oracle():
    rng = np.random.default_rng(2)
    src = rng.integers(0, 256, size=(19, 33, 4), dtype=np.uint8)
    luts = np.stack([O.get_lut(0.9), O.get_lut(1.13), O.get_lut(1.02)])
    note("transform 33x19")
    got = from_dev(P.cuda_transform(to_dev(src), luts))
    np.testing.assert_array_equal(got[..., :3], O.transform(src[..., :3].copy(), luts))
    np.testing.assert_array_equal(got[..., 3], src[..., 3])


@pytest.mark.parametrize("seed,n,h,w,maxl,junk", [(0, 3, 48, 96, 5, False), (1, 2, 45, 90, 4, True),
                                                  (2, 4, 64, 128, 16, False),
                                                  (3, 3, 33, 70, 3, True)])
def test_blender_16s_matches_oracle(seed, n, h, w, maxl, junk):
    rng = np.random.default_rng(seed)
    masks = random_masks(rng, n, h, w)
    imgs = rng.integers(0, 256, size=(n, h, w, 3)).astype(np.int16)
    if junk:
        imgs[masks == 0] = rng.integers(-300, 300, size=int((masks == 0).sum() * 3 / 3))[:, None]
    else:
        imgs[masks == 0] = 0
    ob = O.Blender()
    assert ob.prepare(masks, maxl)
    ref = ob.blend(imgs)
    note(f"blender 16SC4 seed={seed} n={n} {w}x{h} L={ob.num_levels} junk={junk}")
    b = P.CudaTilingMultibandBlendFast()
    assert b.prepare(list(masks), maxl)
    assert b.num_levels == ob.num_levels
    for i, u in enumerate(b.get_unique_masks()):
        np.testing.assert_array_equal(u, ob.unique_masks()[i])
    dev = []
    for i in range(n):
        a = np.zeros((h, w, 4), np.int16)
        a[..., :3] = imgs[i]
        a[..., 3] = 77  # channel 3 of the caller image is ignored
        dev.append(to_dev(a))
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    assert b.blend(dev, out)
    got = from_dev(out)
    assert_parity(got[..., :3], ref, exact=True, label="blend16s")
    assert not got[..., 3].any()
    # begin / tile / end with the buffers reused
    out2 = torch.zeros_like(out)
    assert b.begin()
    scratch = torch.empty_like(dev[0])
    for i in range(n):
        scratch.copy_(dev[i])
        b.tile(scratch, i)
    b.end(out2)
    assert_parity(from_dev(out2)[..., :3], ref, exact=True, label="tile/end")


@pytest.mark.parametrize("seed,n,h,w,maxl,full,kind", [(0, 3, 48, 96, 5, False, "s16"),
                                                       (1, 2, 45, 90, 4, True, "u8"),
                                                       (2, 4, 64, 128, 16, False, "u8"),
                                                       (3, 3, 33, 70, 3, True, "s16")])
def test_blender_custom_masks_matches_oracle(seed, n, h, w, maxl, full, kind):
    rng = np.random.default_rng(seed + 100)
    masks = random_masks(rng, n, h, w)
    custom = random_masks(rng, n, h, w)
    if full:
        custom[n - 1][:, :] = 255
    imgs = rng.integers(0, 256, size=(n, h, w, 3)).astype(np.int16)
    ob = O.Blender()
    assert ob.prepare(masks, maxl)
    ref = ob.blend_custom(imgs, custom)
    note(f"blender custom masks seed={seed} n={n} {w}x{h} L={ob.num_levels} {kind}")
    b = P.CudaTilingMultibandBlendFast()
    assert b.prepare(list(masks), maxl)
    if kind == "u8":
        dev = [to_dev(synth.bgrx(imgs[i].astype(np.uint8))) for i in range(n)]
    else:
        dev = [to_dev(np.concatenate([imgs[i], np.full((h, w, 1), 5, np.int16)], axis=-1))
               for i in range(n)]
    dmasks = [to_dev(custom[i]) for i in range(n)]
    out = torch.full((h, w, 4), 7, dtype=torch.uint8, device="cuda")
    assert b.blend_with_masks(dev, dmasks, out)
    got = from_dev(out)
    assert_parity(got[..., :3], ref, exact=True, label="blend custom masks")
    assert not got[..., 3].any()
    # the prepared blend still gives its own result afterwards
    out2 = torch.zeros_like(out)
    assert b.blend(dev, out2)
    assert_parity(from_dev(out2)[..., :3], ob.blend(imgs), exact=True, label="blend after custom")


def test_blender_8u_input_matches_oracle():
    rng = np.random.default_rng(7)
    n, h, w = 3, 40, 80
    masks = random_masks(rng, n, h, w)
    imgs = rng.integers(0, 256, size=(n, h, w, 3)).astype(np.uint8)
    ob = O.Blender()
    ob.prepare(masks, 5)
    ref = ob.blend(imgs.astype(np.int16))
    note("blender 8UC4 80x40")
    b = P.CudaTilingMultibandBlendFast()
    assert b.prepare(list(masks), 5)
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    assert b.blend([to_dev(synth.bgrx(imgs[i])) for i in range(n)], out)
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label="blend8u")


def run_render(wl, nthreads=8, frame_index=0):
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = synth.frames(wl, frame_index)
    luts = synth.luts(wl, frame_index)
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    ref = ob.render(srcs, xs, ys, luts, nthreads=nthreads)
    note(f"render {wl.name} {wl.dst_w}x{wl.dst_h} n={wl.n} L={ob.num_levels} luts={luts is not None}")
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                          xs, ys, ms, wl.levels)
    assert r.num_levels == ob.num_levels
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    dsrc = [to_dev(synth.bgrx(s)) for s in srcs]
    assert r.render(dsrc, out, luts)
    got = from_dev(out)
    return got, ref, r, dsrc, luts


@pytest.mark.parametrize("name,scale", [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125),
                                        ("fisheye2", 1.0)])
def test_render_matches_oracle(name, scale):
    wl = synth.workload(name, scale)
    got, ref, *_ = run_render(wl)
    assert_parity(got[..., :3], ref, exact=False, label=f"{name}@{scale}")
    assert_parity(got[..., :3], ref, exact=True, label=f"{name}@{scale}")
    assert not got[..., 3].any()


def test_render_repeatable_and_pitched_output():
    wl = synth.workload("rect6", 0.2)
    got, ref, r, dsrc, luts = run_render(wl)
    big = torch.zeros((wl.dst_h, wl.dst_w + 16, 4), dtype=torch.uint8, device="cuda")
    view = big[:, :wl.dst_w]
    for _ in range(2):
        assert r.render(dsrc, view, luts)
    np.testing.assert_array_equal(from_dev(big)[:, :wl.dst_w], got)
    assert not from_dev(big)[:, wl.dst_w:].any()


def test_render_inconsistent_maps_use_16bit_planes():
    # A map that samples where the mask is 0 makes the reference's pyramid leave
    # [0, 255] (MultiBandBlend.cpp:732-768 divides by the masked alpha only); the
    # render then keeps 16-bit planes and must still match the oracle.
    wl = synth.workload("rect6", 0.2)
    xs, ys, ms = synth.maps_and_masks(wl)
    ms = [m.copy() for m in ms]
    rng = np.random.default_rng(5)
    for m in ms:
        ys_, xs_ = np.nonzero(m)
        pick = rng.choice(len(ys_), size=len(ys_) // 50, replace=False)
        m[ys_[pick], xs_[pick]] = 0
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    ref = ob.render(srcs, xs, ys, luts, nthreads=8)
    note(f"render inconsistent {wl.dst_w}x{wl.dst_h}")
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                          xs, ys, ms, wl.levels)
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    assert r.render([to_dev(synth.bgrx(s)) for s in srcs], out, luts)
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label="inconsistent")


@pytest.mark.parametrize("q", [2, 4])
def test_render_snapped_maps_hit_rounding_ties(q):
    # Map points on a 1/q pixel grid make the exact bilinear sum a half-integer
    # often (a cvRound tie); those pixels take the FP64 reference sequence.
    wl = synth.workload("rect6", 0.2)
    xs, ys, ms = synth.maps_and_masks(wl)
    xs = [np.where(x >= 0, np.floor(x * q) / q, x).astype(np.float32) for x in xs]
    ys = [np.where(y >= 0, np.floor(y * q) / q, y).astype(np.float32) for y in ys]
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    ref = ob.render(srcs, xs, ys, luts, nthreads=8)
    note(f"render snapped maps 1/{q} {wl.dst_w}x{wl.dst_h}")
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                          xs, ys, ms, wl.levels)
    assert r.fused
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    dsrc = [to_dev(synth.bgrx(s)) for s in srcs]
    assert r.render(dsrc, out, luts)
    assert_parity(from_dev(out)[..., :3], ref, exact=True, label=f"snapped 1/{q}")
    # the same frame without LUTs
    ref2 = ob.render(srcs, xs, ys, None, nthreads=8)
    assert r.render(dsrc, out, None)
    assert_parity(from_dev(out)[..., :3], ref2, exact=True, label=f"snapped 1/{q} no LUT")


@pytest.mark.parametrize("name,scale,tint", [("fisheye2", 0.25, True), ("rect6", 0.2, False),
                                             ("rect6", 0.2, True), ("mixed8", 0.125, True)])
def test_render_autogain_matches_oracle(name, scale, tint):
    # Per-frame gains from overlap means, estimated and applied on the device,
    # against the oracle's estimate on its own 8-bit reprojections.
    wl = synth.workload(name, scale)
    xs, ys, ms = synth.maps_and_masks(wl)
    r = P.PanoramaRender()
    assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                          xs, ys, ms, wl.levels)
    assert r.fused
    ob = O.Blender()
    assert ob.prepare(np.stack(ms), wl.levels)
    out = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(17)
    for frame in range(2):
        k = rng.uniform(0.8, 1.2, size=(wl.n, 3)) if tint else rng.uniform(0.8, 1.2, size=(wl.n, 1))
        srcs = [np.clip(np.rint(s * k[i]), 0, 255).astype(np.uint8)
                for i, s in enumerate(synth.frames(wl, frame))]
        rep = [O.reproject_to16s(s, x, y, nthreads=8).astype(np.uint8)
               for s, x, y in zip(srcs, xs, ys)]
        sums = O.gain_sums(rep, ms)
        g, rg, bg, anchor, luts = O.gain_from_sums(sums, tint)
        ref = ob.render(srcs, xs, ys, luts, nthreads=8)
        note(f"autogain {name}@{scale} tint={tint} frame={frame} n={wl.n}")
        assert r.render_autogain([to_dev(synth.bgrx(s)) for s in srcs], out, tint=tint)
        np.testing.assert_array_equal(r.gain_sums(), sums)
        dg, drg, dbg, danchor = r.gains()
        assert danchor == anchor
        np.testing.assert_array_equal(dg, g)
        np.testing.assert_array_equal(drg, rg)
        np.testing.assert_array_equal(dbg, bg)
        assert_parity(from_dev(out)[..., :3], ref, exact=True, label=f"autogain {name} f{frame}")
    # a plain render afterwards applies no LUTs again
    srcs = synth.frames(wl, 0)
    assert r.render([to_dev(synth.bgrx(s)) for s in srcs], out, None)
    assert_parity(from_dev(out)[..., :3], ob.render(srcs, xs, ys, None, nthreads=8), exact=True,
                  label="plain after autogain")


def test_batch_matches_alone_and_oracle():
    # Renders of different rigs and sizes stitched in one launch per stage give
    # the same canvases as each render alone, and the oracle's.
    cases = [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125)]
    renders, srcs, luts, refs, alone = [], [], [], [], []
    for name, scale in cases:
        wl = synth.workload(name, scale)
        got, ref, r, dsrc, lt = run_render(wl)
        assert r.fused
        renders.append(r)
        srcs.append(dsrc)
        luts.append(lt)
        refs.append(ref)
        alone.append(got)
    note("batch of 3 renders")
    b = P.Batch(renders)
    outs = [torch.zeros_like(torch.from_numpy(a)).cuda() for a in alone]
    for _ in range(2):
        b.render(srcs, outs, luts)
    for k, o in enumerate(outs):
        got = from_dev(o)
        np.testing.assert_array_equal(got, alone[k])
        assert_parity(got[..., :3], refs[k], exact=True, label=f"batch[{k}]")
    # a different composition and order
    b.set([renders[2], renders[0]])
    o2 = [torch.zeros_like(outs[2]), torch.zeros_like(outs[0])]
    b.render([srcs[2], srcs[0]], o2, [luts[2], luts[0]])
    np.testing.assert_array_equal(from_dev(o2[0]), alone[2])
    np.testing.assert_array_equal(from_dev(o2[1]), alone[0])
    # every block -> tile order of every stage gives the same canvases, and so
    # does the autotuner's choice
    b.set(renders)
    for sched in ([1, 1, 1, 1], [1, 0, 1, 0], [0, 1, 0, 1]):
        note(f"batch block order {sched}")
        b.set_sched(sched)
        assert b.sched() == sched
        for o in outs:
            o.zero_()
        b.render(srcs, outs, luts)
        for k, o in enumerate(outs):
            np.testing.assert_array_equal(from_dev(o), alone[k])
    note("batch autotune")
    chosen, ms = b.autotune(srcs, outs, luts, iters=2)
    assert all(c in (0, 1) for c in chosen) and ms.shape == (2, 4) and (ms[:, 0] > 0).all()
    for o in outs:
        o.zero_()
    b.render(srcs, outs, luts)
    for k, o in enumerate(outs):
        np.testing.assert_array_equal(from_dev(o), alone[k])


def test_server_results_same_alone_and_in_the_mix():
    # Streams of different rigs and sizes through the serving loop: every frame
    # equals the same frame rendered alone, and frames of several streams ran
    # together in one batch.
    cases = [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125), ("rect6", 0.2)]
    rigs = []
    for name, scale in cases:
        wl = synth.workload(name, scale)
        xs, ys, ms = synth.maps_and_masks(wl)
        r = P.PanoramaRender()
        assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                              xs, ys, ms, wl.levels)
        frames = []
        for f in range(3):
            srcs = [to_dev(synth.bgrx(s)) for s in synth.frames(wl, f)]
            lt = synth.luts(wl, f)
            alone = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
            assert r.render(srcs, alone, lt)
            frames.append((srcs, lt, from_dev(alone)))
        rigs.append((wl, r, frames))
    total = sum(r.mem_bytes() for _, r, _ in rigs)
    # admission: the budget holds all rigs but the largest, which fits once the
    # second largest leaves
    by_size = sorted(range(len(rigs)), key=lambda k: rigs[k][1].mem_bytes())
    big, second = by_size[-1], by_size[-2]
    budget = total - rigs[second][1].mem_bytes()
    srv = P.Server(budget, max_inflight=2)
    ids = {}
    for k, (_, r, _) in enumerate(rigs):
        if k != big:
            ids[k] = srv.admit(r)
            assert ids[k] is not None
    assert srv.admit(rigs[big][1]) is None
    note("server: mix of 3 streams, then 4 after a retire")
    outs = {}
    rng = np.random.default_rng(3)
    order = [(k, f) for f in range(3) for k in ids]
    rng.shuffle(order)
    for k, f in order:
        wl, r, frames = rigs[k]
        o = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
        outs[(k, f)] = o
        srv.submit(ids[k], frames[f][0], o, frames[f][1], tag=f)
    srv.drain()
    done = srv.poll(64)
    assert len(done) == len(order) and all(d[3] == 0 for d in done)
    assert max(d[4] for d in done) > 1
    # the second largest leaves and the largest is admitted in its place
    srv.retire(ids.pop(second))
    ids[big] = srv.admit(rigs[big][1])
    assert ids[big] is not None and len(ids) == 3
    for f in range(3):
        for k in ids:
            wl, r, frames = rigs[k]
            o = torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
            outs[(k, f, "b")] = o
            srv.submit(ids[k], frames[f][0], o, frames[f][1], tag=10 + f)
    srv.drain()
    done = srv.poll(64)
    assert len(done) == 3 * len(ids) and all(d[3] == 0 for d in done)
    srv.close()
    for key, o in outs.items():
        k, f = key[0], key[1]
        np.testing.assert_array_equal(from_dev(o), rigs[k][2][f][2])


def test_batch_graph_replays_match_render_alone():
    # The captured per-frame graph, replayed on frames refreshed in place and
    # with new LUT contents, gives each render's result alone; torch work and
    # synchronisation run between the replays.
    cases = [("fisheye2", 0.25), ("rect6", 0.25), ("mixed8", 0.125)]
    renders, wls, maps = [], [], []
    for name, scale in cases:
        wl = synth.workload(name, scale)
        xs, ys, ms = synth.maps_and_masks(wl)
        r = P.PanoramaRender()
        assert r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                              xs, ys, ms, wl.levels)
        renders.append(r)
        wls.append(wl)
    b = P.Batch(renders)
    bufs = [[to_dev(synth.bgrx(s)) for s in synth.frames(wl, 0)] for wl in wls]
    outs = [torch.zeros((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda") for wl in wls]
    note("batch graph capture")
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    b.capture(bufs, outs, [synth.luts(wl, 0) for wl in wls], stream=side)
    for f in range(3):
        for k, wl in enumerate(wls):
            for t, s in zip(bufs[k], synth.frames(wl, f)):
                t.copy_(to_dev(synth.bgrx(s)))
            outs[k].zero_()
        torch.cuda.synchronize()
        luts = [synth.luts(wl, f) for wl in wls]
        b.replay(luts, stream=side)
        torch.cuda.synchronize()
        for k, (wl, r) in enumerate(zip(wls, renders)):
            alone = torch.zeros_like(outs[k])
            assert r.render([to_dev(synth.bgrx(s)) for s in synth.frames(wl, f)], alone, luts[k])
            np.testing.assert_array_equal(from_dev(outs[k]), from_dev(alone))
