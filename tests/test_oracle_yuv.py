"""Pins the oracle's encoder / decoder hand-off conversions (oracle/yuv.c,
CudaAccel/ImageProc.cu) against the BT.601 limited-range definitions: closed
forms for black, white and grays, the float matrix within one level, the
left-column chroma of each 2x2 block, NV12 = interleaved YUV420P, and the
round trip through the inverse."""
import numpy as np
import pytest

from oracle import oracle as O


def bgrx(img3):
    out = np.zeros(img3.shape[:2] + (4,), np.uint8)
    out[..., :3] = img3
    return out


def test_grays_closed_form():
    g = np.arange(256, dtype=np.int64)
    img = np.zeros((2, 512, 4), np.uint8)
    img[:, 0::2, :3] = g[:, None]
    img[:, 1::2, :3] = g[:, None]
    y, u, v = O.bgr32_to_yuv420p(img)
    # 0.299 + 0.587 + 0.114 scaled by 219/255 in 20-bit fixed point: 900726
    np.testing.assert_array_equal(y[0, 0::2], (g * 900726 + (1 << 19) + (16 << 20)) >> 20)
    assert y[0, 0] == 16 and y[0, -1] == 235
    assert (u == 128).all() and (v == 128).all()


def float_bt601(b, g, r):
    """The three-decimal BT.601 studio-swing matrix (8-bit RGB in, Y 16..235)."""
    b, g, r = (np.asarray(a, np.float64) for a in (b, g, r))
    y = 16 + 0.257 * r + 0.504 * g + 0.098 * b
    u = 128 - 0.148 * r - 0.291 * g + 0.439 * b
    v = 128 + 0.439 * r - 0.368 * g - 0.071 * b
    return y, u, v


@pytest.mark.parametrize("seed", range(3))
def test_matches_float_matrix_rounded(seed):
    # the fixed-point coefficients are the matrix times 2^20 (to within one
    # unit), so every output is the matrix value rounded half up: |diff| <= 1/2
    rng = np.random.default_rng(seed)
    h, w = 16, 24
    img = rng.integers(0, 256, size=(h, w, 4)).astype(np.uint8)
    y, u, v = O.bgr32_to_yuv420p(img)
    fy, _, _ = float_bt601(img[..., 0], img[..., 1], img[..., 2])
    assert np.abs(y - fy).max() <= 0.5 + 1e-3
    # chroma: mean of the left column's top and bottom pixel of each block
    left = img[:, 0::2, :3].astype(np.float64)
    m = (left[0::2] + left[1::2]) / 2
    _, fu, fv = float_bt601(m[..., 0], m[..., 1], m[..., 2])
    assert np.abs(u - fu).max() <= 0.5 + 1e-3 and np.abs(v - fv).max() <= 0.5 + 1e-3
    # the inverse matrix (1.164, 2.018, -0.391, -0.813, 1.596) the same way
    back = O.yuv420p_to_bgr32(y, u, v).astype(np.float64)
    yy = np.maximum(y.astype(np.float64) - 16, 0)
    uu = np.repeat(np.repeat(u.astype(np.float64) - 128, 2, 0), 2, 1)
    vv = np.repeat(np.repeat(v.astype(np.float64) - 128, 2, 0), 2, 1)
    fb = np.clip(1.164 * yy + 2.018 * uu, 0, 255)
    fg = np.clip(1.164 * yy - 0.391 * uu - 0.813 * vv, 0, 255)
    fr = np.clip(1.164 * yy + 1.596 * vv, 0, 255)
    for c, f in ((0, fb), (1, fg), (2, fr)):
        assert np.abs(back[..., c] - f).max() <= 0.5 + 1e-3


def test_right_column_does_not_enter_chroma():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, size=(8, 8, 4)).astype(np.uint8)
    _, u0, v0 = O.bgr32_to_yuv420p(img)
    img2 = img.copy()
    img2[:, 1::2] = rng.integers(0, 256, size=img2[:, 1::2].shape)
    y2, u2, v2 = O.bgr32_to_yuv420p(img2)
    np.testing.assert_array_equal(u2, u0)
    np.testing.assert_array_equal(v2, v0)


def test_nv12_is_interleaved_yuv420p():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, size=(10, 14, 4)).astype(np.uint8)
    y, u, v = O.bgr32_to_yuv420p(img)
    y2, uv = O.bgr32_to_nv12(img)
    np.testing.assert_array_equal(y2, y)
    np.testing.assert_array_equal(uv[:, 0::2], u)
    np.testing.assert_array_equal(uv[:, 1::2], v)
    np.testing.assert_array_equal(O.nv12_to_bgr32(y2, uv), O.yuv420p_to_bgr32(y, u, v))


def test_inverse_closed_forms_and_round_trip():
    y = np.array([[16, 235], [16, 235]], np.uint8)
    u = np.array([[128]], np.uint8)
    out = O.yuv420p_to_bgr32(y, u, u.copy())
    np.testing.assert_array_equal(out[:, 0, :3], 0)
    np.testing.assert_array_equal(out[:, 1, :3], 255)
    # blocks of one colour, away from the gamut edges, come back within 2 levels
    rng = np.random.default_rng(6)
    cols = rng.integers(40, 216, size=(6, 9, 3))
    img = bgrx(np.repeat(np.repeat(cols, 2, axis=0), 2, axis=1).astype(np.uint8))
    back = O.yuv420p_to_bgr32(*O.bgr32_to_yuv420p(img))
    assert np.abs(back[..., :3].astype(int) - img[..., :3]).max() <= 2


def test_odd_sizes_rejected():
    with pytest.raises(AssertionError):
        O.bgr32_to_yuv420p(np.zeros((3, 4, 4), np.uint8))
