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
