"""Pins the oracle's linear-blend mode (oracle/linear.c): Gaussian taps against
the normalised Gaussian, the blur against a brute-force 2-D sum over a numpy
reflect-padded image, the weights' partition of unity, and the render
against float64 sums of the same bilinear samples."""
import numpy as np
import pytest

from oracle import oracle as O
from panostitcher_amd import synth


@pytest.mark.parametrize("r", [1, 2, 5, 16, 64])
def test_kernel_is_the_rounded_gaussian(r):
    n, sigma = 2 * r + 1, r / 3.0
    K = O.gaussian_kernel_q8(n, sigma)
    x = np.arange(n) - r
    g = np.exp(-x * x / (2 * sigma * sigma))
    g /= g.sum()
    assert np.abs(K - 256 * g).max() <= 0.5 + 1e-4
    np.testing.assert_array_equal(K, K[::-1])
    if r == 1:
        # sigma 1/3: the tails are exp(-4.5) of the centre
        np.testing.assert_array_equal(K, [3, 250, 3])


@pytest.mark.parametrize("h,w,r", [(12, 17, 1), (9, 30, 3), (40, 33, 6), (5, 6, 4)])
def test_blur_is_the_separable_fixed_point_sum(h, w, r):
    rng = np.random.default_rng(h * w + r)
    src = (rng.random((h, w)) < 0.5).astype(np.uint8) * 255
    src[rng.random((h, w)) < 0.1] = 77
    n = 2 * r + 1
    K = O.gaussian_kernel_q8(n, r / 3.0).astype(np.int64)
    # numpy's "reflect" padding is OpenCV's REFLECT_101 (edge not repeated)
    pad = np.pad(src.astype(np.int64), r, mode="reflect")
    ref = np.zeros((h, w), np.int64)
    for dy in range(n):
        for dx in range(n):
            ref += K[dy] * K[dx] * pad[dy:dy + h, dx:dx + w]
    ref = np.clip((ref + (1 << 15)) >> 16, 0, 255)
    np.testing.assert_array_equal(O.gaussian_blur_u8(src, n, r / 3.0), ref)


def band_masks(n, h, w, overlap):
    m = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        cols = np.arange(i * w // n - overlap, (i + 1) * w // n + overlap) % w
        m[i][:, cols] = 255
    return m


def test_weights_partition_of_unity():
    masks = band_masks(3, 24, 60, 6)
    masks[1, :3] = 0
    wf = O.linear_weights(masks, 4)
    s = wf.sum(axis=0)
    covered = (wf > 0).any(axis=0)
    assert np.abs(s[covered] - 1).max() < 1e-6
    assert not wf[:, ~covered].any()
    assert (wf[masks == 0] == 0).all()
    # a pixel that only one blurred seam reaches has weight exactly 1
    lone = ((wf > 0).sum(axis=0) == 1)
    assert lone.any() and (s[lone] == 1.0).all()
    wi = O.linear_weights(masks, 4, as_int=True)
    si = wi.sum(axis=0)
    # calcWeights adds 1 to every truncated share: the sum is 2^16 up to the count
    assert si.min() >= 3 and (si[covered] >= 65536).all() and (si[covered] <= 65536 + 3).all()


def test_single_camera_render_is_its_remap():
    wl = synth.workload("fisheye2", 0.125)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = synth.frames(wl)
    full = np.full((1,) + ms[0].shape, 255, np.uint8)
    w = O.linear_weights(full, 3)
    assert (w == 1).all()
    out = O.linear_render(srcs[:1], xs[:1], ys[:1], w)
    ref = O.reproject_to16s(srcs[0], xs[0], ys[0])
    np.testing.assert_array_equal(out, ref.astype(np.uint8))


def test_render_matches_float64_sum():
    wl = synth.workload("rect6", 0.125)
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = synth.frames(wl)
    luts = synth.luts(wl)
    w = O.linear_weights(np.stack(ms), 5)
    out = O.linear_render(srcs, xs, ys, w, luts)
    acc = np.zeros(out.shape, np.float64)
    for i in range(wl.n):
        adj = O.transform(srcs[i][..., :3].copy(), luts[i]) if luts is not None else srcs[i][..., :3]
        samp = O.reproject_to16s(np.ascontiguousarray(adj), xs[i], ys[i]).astype(np.float64)
        acc += samp * w[i][..., None].astype(np.float64)
    d = np.abs(out.astype(np.float64) - acc)
    # float accumulation of <= 6 terms of <= 255: within 1/2 + 2^-14 of the exact sum
    assert d.max() <= 0.5 + 2 ** -14
    assert (np.abs(out.astype(int) - np.rint(acc).astype(int)) <= 1).all()


def test_blend32s_closed_form():
    rng = np.random.default_rng(3)
    imgs = rng.integers(0, 256, size=(2, 5, 7, 4)).astype(np.uint8)
    w = np.zeros((2, 5, 7), np.int32)
    w[0] = 65536 + 1
    w[1] = 1
    out = O.linear_blend32s(imgs, w)
    ref = (imgs[0].astype(np.int64) * 65537 + imgs[1]) >> 16
    np.testing.assert_array_equal(out, np.clip(ref, 0, 255).astype(np.uint8))
