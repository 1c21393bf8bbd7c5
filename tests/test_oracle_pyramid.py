"""Pins the oracle's pyramid primitives (Blend/Pyramid.cpp) and borders
against independent formulations (numpy convolutions, known border values)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import np_ref


def test_border_known_values():
    R, W = O.REFLECT_101, O.WRAP
    # REFLECT_101: gfedcb|abcdefgh|gfedcba
    assert [O.border(p, 8, R) for p in (-3, -2, -1, 0, 7, 8, 9, 10)] == [3, 2, 1, 0, 7, 6, 5, 4]
    assert [O.border(p, 1, R) for p in (-2, -1, 0, 1, 2)] == [0, 0, 0, 0, 0]
    assert [O.border(p, 2, R) for p in (-2, -1, 2, 3)] == [0, 1, 0, 1]
    # WRAP: cdefgh|abcdefgh|abcdefg
    assert [O.border(p, 8, W) for p in (-9, -8, -3, -1, 8, 9, 17)] == [7, 0, 5, 7, 0, 1, 1]
    for n in (1, 2, 3, 5, 7):
        for p in range(-3 * n, 4 * n):
            assert O.border(p, n, W) == p % n


@pytest.mark.parametrize("h,w,cn", [(3, 3, 1), (4, 4, 3), (5, 7, 3), (8, 6, 1), (9, 13, 3),
                                    (16, 31, 3), (33, 64, 3), (3, 20, 3), (20, 3, 1)])
def test_pyrdown_matches_convolution(h, w, cn):
    rng = np.random.default_rng(h * 100 + w)
    img = rng.integers(-300, 300, size=(h, w, cn) if cn > 1 else (h, w), dtype=np.int16)
    ref = np_ref.pyr_down_sum(img)
    got = O.pyrdown_to32s(img)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(O.pyrdown_s16(img), ((ref + 128) >> 8).astype(np.int16))


def test_pyrdown_constant_alpha_closed_form():
    # pyrDown of a constant 256 plane is 256 * 256 = 65536 everywhere (sum of weights 256).
    img = np.full((10, 12), 256, np.int16)
    np.testing.assert_array_equal(O.pyrdown_to32s(img), np.full((5, 6), 65536))
    np.testing.assert_array_equal(O.pyrdown_s16(img), np.full((5, 6), 256))


@pytest.mark.parametrize("sh,sw,dh,dw,cn", [(1, 1, 2, 2, 1), (1, 1, 1, 1, 3), (2, 2, 4, 4, 3),
                                            (2, 3, 3, 5, 3), (3, 3, 5, 6, 1), (4, 7, 8, 13, 3),
                                            (5, 5, 10, 10, 3), (8, 16, 15, 31, 3)])
def test_pyrup_matches_zero_insertion(sh, sw, dh, dw, cn):
    rng = np.random.default_rng(sh * 1000 + sw * 10 + dh)
    img = rng.integers(-255, 256, size=(sh, sw, cn) if cn > 1 else (sh, sw), dtype=np.int16)
    ref = np_ref.pyr_up(img, (dh, dw))
    np.testing.assert_array_equal(O.pyrup_s16(img, (dh, dw)), ref.astype(np.int16))
    big = img.astype(np.int32) * 1000
    np.testing.assert_array_equal(O.pyrup_s32(big, (dh, dw)), np_ref.pyr_up(big, (dh, dw)))


def test_pyrup_constant_is_identity():
    img = np.full((4, 6, 3), 77, np.int16)
    np.testing.assert_array_equal(O.pyrup_s16(img, (8, 12)), np.full((8, 12, 3), 77))
    np.testing.assert_array_equal(O.pyrup_s16(img, (7, 11)), np.full((7, 11, 3), 77))


def test_pyrup_reflect_horizontal_border():
    # With REFLECT_101 in both directions the right neighbour of the last column is
    # column w-2; check one hand-computed value: src row [1, 2, 4] (h = 1).
    img = np.array([[1, 2, 4]], np.int16)
    out = O.pyrup_s16(img, (2, 6), hb=O.REFLECT_101, vb=O.REFLECT_101)
    # row of height 1: vertical taps are 1 + 6 + 1 = 8 times the row; even col 4: (2 + 24 + 2)*8
    assert out[0, 4] == ((2 + 6 * 4 + 2) * 8 + 32) >> 6
    assert out[0, 5] == ((4 * (4 + 2)) * 8 + 32) >> 6


@pytest.mark.parametrize("w,h,maxl,expect", [(2048, 1024, 5, 5), (3840, 1920, 7, 7),
                                             (7680, 3840, 9, 9), (7680, 3840, 10, 10),
                                             (7680, 3840, 16, 10), (100, 50, 16, 4),
                                             (64, 32, 16, 4), (8, 4, 16, 1), (4, 2, 16, 0)])
def test_true_num_levels(w, h, maxl, expect):
    assert O.true_num_levels(w, h, maxl, 2) == expect
