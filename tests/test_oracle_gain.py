"""Pins the oracle's per-frame gain estimation (oracle/gain.c).

The estimator's normal equations come from the least-squares error of the
overlap-means model, e(g) = sum_ij N_ij ((g_i I_ij - g_j I_ji)^2 sn + (1 - g_i)^2 sg)
(Compensate.cpp:501-547, TintAdjust.cpp:160-210).  The tests check the oracle's
gains by the gradient of that error, evaluated by central differences of e
itself, and its solver against LAPACK."""
import numpy as np
import pytest

from oracle import oracle as O


def test_gray_closed_forms():
    for v in range(256):
        assert O.gray_bgr(v, v, v) == v
    rng = np.random.default_rng(0)
    for b, g, r in rng.integers(0, 256, size=(2000, 3)):
        exact = 0.114 * b + 0.587 * g + 0.299 * r
        assert abs(O.gray_bgr(b, g, r) - exact) < 0.53
    # monotone in every channel
    for c in range(3):
        px = [0, 0, 0]
        prev = -1
        for v in range(256):
            px[c] = v
            y = O.gray_bgr(*px)
            assert y >= prev
            prev = y


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8])
def test_solve_matches_lapack(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        A = rng.normal(size=(n, n)) + np.eye(n) * n
        b = rng.normal(size=n)
        ok, x = O.cv_solve(A, b)
        assert ok
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-11, atol=1e-12)
        assert np.abs(A @ x - b).max() < 1e-11 * max(1.0, np.abs(b).max())


def test_solve_pivoting_and_singular():
    # a zero leading pivot needs a row swap
    A = np.array([[0., 2, 1, 0], [1, 1, 0, 0], [0, 1, 3, 1], [2, 0, 1, 4]])
    b = np.array([1., 2, 3, 4])
    ok, x = O.cv_solve(A, b)
    assert ok
    np.testing.assert_allclose(A @ x, b, atol=1e-12)
    for n in (1, 2, 3, 4):
        ok, _ = O.cv_solve(np.zeros((n, n)), np.ones(n))
        assert not ok
    ok, _ = O.cv_solve(np.array([[1., 2], [2, 4]]), np.ones(2))
    assert not ok


def scene(rng, n, h, w, gains=None):
    """n views of one smooth scene with overlapping cyclic column masks."""
    yy, xx = np.mgrid[0:h, 0:w]
    base = np.stack([60 + 50 * np.sin(xx / 7.0 + c) + 30 * np.cos(yy / 5.0 - c) for c in range(3)],
                    axis=-1)
    base += rng.uniform(-10, 10, size=base.shape)
    imgs, masks = [], []
    for i in range(n):
        m = np.zeros((h, w), np.uint8)
        c0 = int(i * w / n)
        cols = np.arange(c0 - w // (2 * n) - 3, c0 + w // (2 * n) + 4) % w
        m[:, cols] = 255
        k = 1.0 if gains is None else gains[i]
        img = np.clip(np.rint(base * k), 0, 255).astype(np.uint8)
        img[m == 0] = 0
        imgs.append(img)
        masks.append(m)
    return imgs, masks


def overlap_means(imgs, masks, gray: bool):
    n = len(imgs)
    N = np.zeros((n, n))
    I = np.zeros((n, n, 3 if not gray else 1))
    for i in range(n):
        g = np.vectorize(O.gray_bgr)(imgs[i][..., 0], imgs[i][..., 1], imgs[i][..., 2]) if gray else None
        for j in range(n):
            sel = (masks[i] > 0) & (masks[j] > 0)
            N[i, j] = sel.sum()
            if sel.any():
                I[i, j] = g[sel].mean() if gray else imgs[i][sel].mean(axis=0)
    return N, I


def error(g, N, I, sn, sg):
    n = len(g)
    e = 0.0
    for i in range(n):
        for j in range(n):
            e += N[i, j] * ((g[i] * I[i, j] - g[j] * I[j, i]) ** 2 * sn + (1 - g[i]) ** 2 * sg)
    return e


def grad(g, N, I, sn, sg, h=1e-4):
    out = np.zeros(len(g))
    for k in range(len(g)):
        d = np.zeros(len(g))
        d[k] = h
        out[k] = (error(g + d, N, I, sn, sg) - error(g - d, N, I, sn, sg)) / (2 * h)
    return out


def stationary_scale(r, N, I, sn, sg):
    """The oracle returns g / g[anchor]; recover the scale s at which s r is the
    minimiser: e is quadratic, so its gradient is affine in s."""
    g0, g1 = grad(0 * r, N, I, sn, sg), grad(r, N, I, sn, sg)
    d = g1 - g0
    s = -(g0 @ d) / (d @ d)
    return s, grad(s * r, N, I, sn, sg), np.abs(g0).max()


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (4, 3), (6, 4)])
def test_gains_are_the_least_squares_solution(n, seed):
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.8, 1.2, size=n)
    imgs, masks = scene(rng, n, 24, 96, k)
    gains, rg, bg, mi, luts = O.gain_estimate(imgs, masks, tint=True)
    N, Ig = overlap_means(imgs, masks, gray=True)
    Ig = Ig[..., 0]
    # anchor: the image with the largest mean; rescaled to gain 1
    assert mi == int(np.argmax(np.diag(Ig)))
    assert gains[mi] == 1.0
    s, gr, scale = stationary_scale(gains, N, Ig, 0.01, 100.0)
    assert np.abs(gr).max() < 1e-7 * scale
    # the tint systems: R/G and B/G ratios of the channel means
    _, Ic = overlap_means(imgs, masks, gray=False)
    with np.errstate(invalid="ignore", divide="ignore"):
        R = np.where(N > 0, Ic[..., 2] / Ic[..., 1], 0.0)
        B = np.where(N > 0, Ic[..., 0] / Ic[..., 1], 0.0)
    for sol, M in ((rg, R), (bg, B)):
        gsol = grad(sol, N, M, 1.0, 0.05)
        assert np.abs(gsol).max() < 1e-7 * np.abs(grad(0 * sol, N, M, 1.0, 0.05)).max()
    # LUTs: getExposureColorOptimizeLUTs (B: e b, G: e, R: e r)
    for i in range(n):
        np.testing.assert_array_equal(luts[i, 0], O.get_lut(gains[i] * bg[i]))
        np.testing.assert_array_equal(luts[i, 1], O.get_lut(gains[i]))
        np.testing.assert_array_equal(luts[i, 2], O.get_lut(gains[i] * rg[i]))


def test_gains_undo_exposure_differences():
    rng = np.random.default_rng(9)
    k = np.array([1.0, 0.85, 1.15, 0.92])
    imgs, masks = scene(rng, 4, 32, 128, k)
    gains, *_ = O.gain_estimate(imgs, masks, tint=False)
    # darker cameras get larger gains, and the corrected exposures g_i k_i are
    # closer together than k_i (the prior 1/sg^2 = 100 keeps the gains near 1)
    assert list(np.argsort(gains)) == list(np.argsort(-k))
    assert np.ptp(gains * k) < 0.7 * np.ptp(k)


def test_no_overlap_and_empty_masks():
    rng = np.random.default_rng(3)
    h, w = 16, 40
    imgs, masks = [], []
    for i in range(3):
        m = np.zeros((h, w), np.uint8)
        m[:, i * 13:(i + 1) * 13] = 255
        img = rng.integers(20, 230, size=(h, w, 3)).astype(np.uint8)
        imgs.append(img)
        masks.append(m)
    gains, rg, bg, mi, _ = O.gain_estimate(imgs, masks, tint=True)
    np.testing.assert_allclose(gains, 1.0, atol=1e-12)
    np.testing.assert_allclose(rg, 1.0, atol=1e-12)
    np.testing.assert_allclose(bg, 1.0, atol=1e-12)
    # no pixels at all: the systems are singular and the gains stay 1
    empty = [np.zeros((h, w), np.uint8)] * 3
    gains, rg, bg, mi, luts = O.gain_estimate(imgs, empty, tint=False)
    assert list(gains) == [1.0, 1.0, 1.0] and mi == 0
    np.testing.assert_array_equal(luts[0, 1], np.arange(256))


def test_sums_match_brute_force():
    rng = np.random.default_rng(5)
    imgs, masks = scene(rng, 3, 12, 30)
    s = O.gain_sums(imgs, masks)
    for i in range(3):
        for j in range(3):
            sel = (masks[i] > 0) & (masks[j] > 0)
            px = imgs[i][sel].astype(np.int64)
            assert s[i, j, 0] == sel.sum()
            assert s[i, j, 1] == sum(O.gray_bgr(*p) for p in px)
            np.testing.assert_array_equal(s[i, j, 2:], px.sum(axis=0))
