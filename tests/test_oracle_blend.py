"""Pins the oracle's seams (Blend/Mask.cpp:376-397) and multiband blender
(Blend/MultiBandBlend.cpp:906-1031) against brute force, an independent numpy
restatement, and closed-form invariants."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import np_ref


def rect_mask(h, w, y0, y1, x0, x1):
    m = np.zeros((h, w), np.uint8)
    m[y0:y1, x0:x1] = 255
    return m


@pytest.mark.parametrize("seed", range(6))
def test_distance_transform_bruteforce(seed):
    rng = np.random.default_rng(seed)
    h, w = rng.integers(3, 25, size=2)
    m = (rng.random((h, w)) < 0.85).astype(np.uint8) * 255
    np.testing.assert_array_equal(O.distance_transform_l1(m), np_ref.l1_distance(m))


def test_distance_transform_no_zero_pixel_cap():
    # No zero pixel: distance is 8192 + the L1 distance to the outside frame.
    m = np.full((5, 7), 255, np.uint8)
    d = O.distance_transform_l1(m)
    assert d[0, 0] == 8193 and d[2, 3] == 8195 and d[4, 6] == 8193
    np.testing.assert_array_equal(d, np_ref.l1_distance(m))
    assert not O.distance_transform_l1(np.zeros((4, 4), np.uint8)).any()


def test_seams_two_overlapping_rects():
    # Two bands overlapping in columns 10..19: the seam splits the overlap at
    # equal L1 distance; ties (currDist == dist) stay with the earlier camera.
    h, w = 6, 30
    m0 = rect_mask(h, w, 0, h, 0, 20)
    m1 = rect_mask(h, w, 0, h, 10, 30)
    u = O.nonintersecting_masks(np.stack([m0, m1]))
    # camera 1 owns x where dist-to-zero-of-m1 (x - 9) > dist-to-zero-of-m0 (20 - x)
    cols1 = [x for x in range(w) if u[1, 0, x]]
    assert cols1 == list(range(15, 30))
    assert not (u[0] & u[1]).any()
    np.testing.assert_array_equal((u[0] | u[1]) != 0, (m0 | m1) != 0)


@pytest.mark.parametrize("seed", range(5))
def test_seams_match_bruteforce_and_partition(seed):
    rng = np.random.default_rng(seed)
    h, w, n = 16, 24, rng.integers(2, 5)
    masks = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        y0, x0 = rng.integers(0, h - 4), rng.integers(0, w - 4)
        masks[i, y0:y0 + rng.integers(4, h), x0:x0 + rng.integers(4, w)] = 255
    u = O.nonintersecting_masks(masks)
    np.testing.assert_array_equal(u, np_ref.unique_masks(masks))
    for i in range(n):
        assert not (u[i] & ~masks[i]).any()                  # unique_i inside mask_i
        for j in range(i):
            assert not (u[i] & u[j]).any()                   # disjoint
    np.testing.assert_array_equal(u.any(0), masks.any(0))    # same coverage


def random_case(rng, n, h, w, junk_outside=False):
    masks = np.zeros((n, h, w), np.uint8)
    for i in range(n):
        cx = int(rng.integers(0, w))
        half = int(rng.integers(w // 4, w // 2 + 2))
        cols = (np.arange(cx - half, cx + half) % w)
        y0 = int(rng.integers(0, h // 3))
        y1 = int(rng.integers(2 * h // 3, h + 1))
        masks[i][y0:y1][:, cols] = 255
    imgs = rng.integers(0, 256, size=(n, h, w, 3)).astype(np.int16)
    if not junk_outside:
        imgs[masks == 0] = 0
    return masks, imgs


@pytest.mark.parametrize("seed,n,h,w,maxl", [(0, 2, 32, 64, 5), (1, 3, 24, 48, 3),
                                             (2, 3, 27, 45, 4), (3, 4, 33, 66, 9),
                                             (4, 1, 16, 32, 2), (5, 3, 31, 17, 3)])
def test_blend_matches_numpy_restatement(seed, n, h, w, maxl):
    rng = np.random.default_rng(seed)
    masks, imgs = random_case(rng, n, h, w, junk_outside=(seed % 2 == 1))
    b = O.Blender()
    assert b.prepare(masks, maxl)
    got = b.blend(imgs)
    ref = np_ref.multiband(imgs, masks, maxl)
    np.testing.assert_array_equal(got, ref)
    got_mt = b.blend(imgs, nthreads=4)
    np.testing.assert_array_equal(got_mt, got)


def test_blend_full_cover_variant():
    rng = np.random.default_rng(11)
    h, w = 32, 64
    m0 = np.zeros((h, w), np.uint8); m0[:, :40] = 255
    m1 = np.zeros((h, w), np.uint8); m1[:, 30:] = 255; m1[:, :4] = 255
    masks = np.stack([m0, m1])
    imgs = rng.integers(0, 256, size=(2, h, w, 3)).astype(np.int16)
    imgs[masks == 0] = 0
    b = O.Blender(); b.prepare(masks, 16)
    assert b.full_mask
    np.testing.assert_array_equal(b.blend(imgs), np_ref.multiband(imgs, masks, 16))


@pytest.mark.parametrize("h,w", [(32, 64), (27, 45), (48, 96)])
def test_single_full_camera_reproduces_input(h, w):
    # Laplacian collapse telescopes exactly in integers, whatever calcDstImage's
    # 255/256 scaling does to the Gaussian levels.
    rng = np.random.default_rng(h)
    img = rng.integers(0, 256, size=(1, h, w, 3)).astype(np.int16)
    b = O.Blender(); b.prepare(np.full((1, h, w), 255, np.uint8), 16)
    assert b.full_mask
    np.testing.assert_array_equal(b.blend(img), img[0].astype(np.uint8))


@pytest.mark.parametrize("seed,n,h,w,maxl,full", [(0, 2, 32, 64, 5, False), (1, 3, 24, 48, 3, True),
                                                  (2, 3, 27, 45, 4, False), (3, 4, 33, 66, 9, True)])
def test_blend_custom_masks_matches_numpy_restatement(seed, n, h, w, maxl, full):
    rng = np.random.default_rng(seed + 40)
    masks, imgs = random_case(rng, n, h, w, junk_outside=(seed % 2 == 1))
    custom, _ = random_case(rng, n, h, w)
    if full:
        custom[0][:, :] = 255
    b = O.Blender()
    assert b.prepare(masks, maxl)
    got = b.blend_custom(imgs, custom)
    np.testing.assert_array_equal(got, np_ref.multiband(imgs, masks, maxl, weight_masks=custom))
    if not full:
        assert not got[~custom.any(0)].any()


def test_blend_custom_with_the_seams_is_the_blend():
    # When the prepare masks leave the canvas partly uncovered, the prepared
    # blend divides by the seam weights' sums too: with the seams as custom
    # masks the two blends are the same computation.
    rng = np.random.default_rng(9)
    masks, imgs = random_case(rng, 3, 40, 80)
    masks[:, :3, :] = 0
    b = O.Blender(); b.prepare(masks, 5)
    assert not b.full_mask
    np.testing.assert_array_equal(b.blend_custom(imgs, b.unique_masks()), b.blend(imgs))


def test_blend_zero_outside_coverage_and_permutation():
    rng = np.random.default_rng(5)
    masks, imgs = random_case(rng, 3, 32, 64)
    masks[:, :4, :] = 0
    imgs[masks == 0] = 0
    b = O.Blender(); b.prepare(masks, 5)
    out = b.blend(imgs)
    assert not out[~masks.any(0)].any()
    # Permuting cameras while keeping each camera's seam region: integer
    # accumulation is order independent.  Build masks that are already
    # non-intersecting so that the seam labels do not depend on the order.
    u = b.unique_masks()
    b2 = O.Blender(); b2.prepare(u, 5)
    b3 = O.Blender(); b3.prepare(u[::-1].copy(), 5)
    np.testing.assert_array_equal(b2.blend(imgs), b3.blend(imgs[::-1].copy()))


def test_weight_and_alpha_pyramids():
    h, w = 32, 64
    m = rect_mask(h, w, 4, 28, 10, 50)
    b = O.Blender(); b.prepare(np.stack([m, 255 - m]), 4)
    # weights of the two cameras at level 0 are 256 on their unique regions
    w0 = b.weight(0, 0); w1 = b.weight(1, 0)
    np.testing.assert_array_equal(w0 + w1, np.full((h, w), 256))
    # alpha at level 1 is the 5x5 sum of 256 * valid (interior: 65536)
    a = b.alpha(0, 1)
    assert a[8, 15] == 65536 and a[0, 0] == 0
    np.testing.assert_array_equal(a, np_ref.pyr_down_sum(np.where(m > 0, 256, 0)))
