"""GPU parity of the float-weight blender (CudaTilingMultibandBlendFast32F)
through the C-ABI against the oracle's restatement.  Both evaluate the same
float operations in the same order (explicit fused multiply-adds, correctly
rounded divisions), so the 8-bit output is expected to match bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import panostitcher_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.gpu_util import assert_parity, from_dev, note, to_dev  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    P.lib()
    yield


def partial_masks(rng, n, h, w):
    masks = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        cx = int(rng.integers(0, w))
        half = int(rng.integers(w // 4, w // 2 + 2))
        cols = np.arange(cx - half, cx + half) % w
        y0 = int(rng.integers(0, h // 3))
        y1 = int(rng.integers(2 * h // 3, h + 1))
        masks[i][y0:y1][:, cols] = 255
    return masks


def full_masks(rng, n, h, w):
    """Overlapping column bands (wrapping) that together cover the canvas."""
    masks = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        a = i * w // n - int(rng.integers(1, w // (2 * n) + 2))
        b = (i + 1) * w // n + int(rng.integers(1, w // (2 * n) + 2))
        masks[i][:, np.arange(a, b) % w] = 255
    assert masks.any(axis=0).all()
    return masks


@pytest.mark.parametrize("seed,n,h,w,maxl,full,kind", [
    (0, 3, 48, 96, 5, False, "u8"),
    (1, 2, 45, 90, 4, True, "u8"),
    (2, 4, 64, 128, 7, True, "u8"),
    (3, 3, 33, 70, 3, False, "f32"),
    (4, 5, 97, 203, 16, False, "u8"),
    (5, 2, 2, 9, 3, True, "f32"),
])
def test_blender32f_matches_oracle(seed, n, h, w, maxl, full, kind):
    rng = np.random.default_rng(seed)
    masks = (full_masks if full else partial_masks)(rng, n, h, w)
    if kind == "u8":
        imgs = rng.integers(0, 256, size=(n, h, w, 4)).astype(np.uint8)
    else:
        imgs = ((rng.random((n, h, w, 4)) - 0.1) * 300).astype(np.float32)
    ref, L = O.blend32f(masks, imgs, maxl)
    note(f"blender32f seed={seed} n={n} {w}x{h} L={L} full={full} {kind}")
    b = P.CudaTilingMultibandBlendFast32F()
    assert b.prepare(list(masks), maxl)
    assert b.num_levels == L
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    assert b.blend([to_dev(imgs[i]) for i in range(n)], out)
    assert_parity(from_dev(out), ref, exact=True, label="blend32f")
    # a second blend through the same prepared object gives the same bytes
    out.fill_(9)
    assert b.blend([to_dev(imgs[i]) for i in range(n)], out)
    assert_parity(from_dev(out), ref, exact=True, label="blend32f again")


def test_blender32f_pitched_images_and_output():
    rng = np.random.default_rng(11)
    n, h, w = 3, 40, 72
    masks = partial_masks(rng, n, h, w)
    imgs = rng.integers(0, 256, size=(n, h, w, 4)).astype(np.uint8)
    ref, _ = O.blend32f(masks, imgs, 4)
    note("blender32f pitched 72x40")
    b = P.CudaTilingMultibandBlendFast32F()
    assert b.prepare(list(masks), 4)
    big = [to_dev(np.pad(imgs[i], ((0, 0), (0, 24), (0, 0)))) for i in range(n)]
    views = [t[:, :w] for t in big]
    obig = torch.zeros((h, w + 40, 4), dtype=torch.uint8, device="cuda")
    assert b.blend(views, obig[:, :w])
    got = from_dev(obig)
    assert_parity(got[:, :w], ref, exact=True, label="blend32f pitched")
    assert not got[:, w:].any()


def test_blender32f_rejects_bad_calls():
    b = P.CudaTilingMultibandBlendFast32F()
    out = torch.zeros((8, 16, 4), dtype=torch.uint8, device="cuda")
    img = torch.zeros((8, 16, 4), dtype=torch.uint8, device="cuda")
    assert not b.blend([img], out)  # not prepared
    assert not b.prepare([], 3)
    assert not b.prepare([np.full((1, 16), 255, np.uint8)], 3)  # no level fits a 1-row canvas
    assert b.prepare([np.full((8, 16), 255, np.uint8)], 3)
    with pytest.raises(P.PanoError):
        b.blend([img, img], out)  # wrong count
    with pytest.raises(ValueError):
        b.blend([img.to(torch.int16)], out)
