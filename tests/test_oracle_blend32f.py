"""Pins the oracle's float-weight blender (oracle/blend32f.c), the restatement of
CudaTilingMultibandBlendFast32F.  Its float pyramid filters are checked against
float64 convolutions written from their definition (5-tap binomial kernels,
rows REFLECT_101, columns WRAP); the blend against invariants and against the
integer blender of the CPU path."""
import numpy as np
import pytest

from oracle import oracle as O

K = np.array([1, 4, 6, 4, 1], np.float64) / 16.0


def reflect101(p, n):
    if n == 1:
        return 0
    while p < 0 or p >= n:
        p = -p if p < 0 else 2 * n - 2 - p
    return p


def down_f64(a):
    h, w = a.shape[:2]
    out = np.zeros(((h + 1) // 2, (w + 1) // 2) + a.shape[2:])
    for y in range(out.shape[0]):
        for x in range(out.shape[1]):
            s = 0
            for i in range(5):
                for j in range(5):
                    s = s + K[i] * K[j] * a[reflect101(2 * y - 2 + i, h), (2 * x - 2 + j) % w]
            out[y, x] = s
    return out


def up_f64(a, shape):
    h, w = a.shape[:2]
    dh, dw = shape
    out = np.zeros((dh, dw) + a.shape[2:])
    for y in range(dh):
        for x in range(dw):
            s = 0
            # zero insertion then the 5x5 kernel times 4
            for i in range(5):
                for j in range(5):
                    yy, xx = y - 2 + i, x - 2 + j
                    if yy % 2 or xx % 2:
                        continue
                    s = s + 4 * K[i] * K[j] * a[reflect101(yy // 2, h), (xx // 2) % w]
            out[y, x] = s
    return out


@pytest.mark.parametrize("h,w,cn", [(9, 13, 4), (1, 5, 1), (2, 3, 4), (17, 32, 1)])
def test_pyrdown32f_is_the_binomial_filter(h, w, cn):
    rng = np.random.default_rng(h * 31 + w)
    a = (rng.random((h, w, cn)) * 255).astype(np.float32)
    got = O.pyrdown32f(a if cn > 1 else a[..., 0])
    ref = down_f64(a.astype(np.float64))
    np.testing.assert_allclose(got.reshape(ref.shape), ref, rtol=2e-6, atol=2e-4)


@pytest.mark.parametrize("h,w,dh,dw", [(5, 7, 9, 13), (5, 7, 10, 14), (1, 3, 2, 5), (8, 16, 16, 32)])
def test_pyrup32f_is_zero_insertion_filter(h, w, dh, dw):
    rng = np.random.default_rng(h + 7 * w)
    a = ((rng.random((h, w, 4)) - 0.3) * 300).astype(np.float32)
    got = O.pyrup32f(a, (dh, dw))
    ref = up_f64(a.astype(np.float64), (dh, dw))
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-4)


def test_constant_images_are_fixed_points():
    a = np.full((6, 10, 4), 37.5, np.float32)
    np.testing.assert_array_equal(O.pyrdown32f(a), np.full((3, 5, 4), 37.5, np.float32))
    np.testing.assert_array_equal(O.pyrup32f(a, (12, 20)), np.full((12, 20, 4), 37.5, np.float32))


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


def test_single_camera_reproduces_its_image():
    rng = np.random.default_rng(1)
    h, w = 32, 64
    img = rng.integers(0, 256, size=(1, h, w, 4)).astype(np.uint8)
    img[..., 3] = 0
    out, L = O.blend32f(np.full((1, h, w), 255, np.uint8), img, 5)
    assert L == 4
    d = np.abs(out.astype(int) - img[0].astype(int))
    assert d.max() <= 1 and (d == 0).mean() > 0.999


@pytest.mark.parametrize("seed,n,h,w,maxl", [(0, 3, 48, 96, 5), (1, 2, 45, 90, 4), (2, 4, 64, 128, 7)])
def test_float_and_integer_blends_agree(seed, n, h, w, maxl):
    rng = np.random.default_rng(seed)
    masks = random_masks(rng, n, h, w)
    imgs = rng.integers(0, 256, size=(n, h, w, 4)).astype(np.uint8)
    imgs[..., 3] = 0
    imgs[masks == 0] = 0
    out, L = O.blend32f(masks, imgs, maxl)
    ob = O.Blender()
    ob.prepare(masks, maxl)
    ref = ob.blend(imgs[..., :3].astype(np.int16))
    assert L == ob.num_levels
    covered = masks.any(axis=0)
    assert not out[~covered].any()
    d = np.abs(out[..., :3].astype(int) - ref.astype(int))[covered]
    # the same pyramid in float and in the CPU path's integers (with its 255/256
    # normalisation): close everywhere, identical on most pixels
    assert d.max() <= 6 and (d <= 1).mean() > 0.8 and d.mean() < 1.0
