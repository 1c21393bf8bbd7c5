"""Pins the oracle's reprojection (Warp/ZReproject.cpp:101-130, 448-567), LUT
gains (Blend/ExposureColorOptimize.cpp:1217-1223, Blend/GainAdjust.cpp:295-346)
and the equirect camera chain (Warp/Remap.cpp:147-514) with closed forms."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def src_img(h=5, w=4, cn=3, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(h, w, cn), dtype=np.uint8)


def remap_one(img, x, y):
    xm = np.array([[x]], np.float32)
    ym = np.array([[y]], np.float32)
    return O.reproject_to16s(img, xm, ym)[0, 0]


def test_bilinear_integer_points_copy():
    img = src_img()
    h, w = img.shape[:2]
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    out = O.reproject_to16s(img, xs, ys)
    np.testing.assert_array_equal(out, img.astype(np.int16))


def test_bilinear_half_pixel_ties_round_to_even():
    img = np.zeros((3, 4, 3), np.uint8)
    img[1, 1] = (1, 2, 10)
    img[1, 2] = (2, 3, 13)
    # x = 1.5: (p + q) / 2 -> 1.5 -> 2, 2.5 -> 2, 11.5 -> 12
    assert remap_one(img, 1.5, 1.0).tolist() == [2, 2, 12]
    # quarter point: 0.75 * 1 + 0.25 * 2 = 1.25 -> 1
    assert remap_one(img, 1.25, 1.0).tolist()[0] == 1


def test_bilinear_invalid_points_are_zero():
    img = src_img()
    h, w = img.shape[:2]
    for x, y in [(-1, 0), (-0.5, 1), (w, 0), (0, h), (w + 3, h + 3), (-1, -1)]:
        assert not remap_one(img, x, y).any()


def test_bilinear_last_column_reads_next_pixel_in_memory():
    # The reference steps its pointer past x0 (it never uses its clamped x1):
    # at x0 = w - 1 the right sample is the first pixel of the next row.
    img = src_img(h=3, w=4, seed=3)
    got = remap_one(img, 3.5, 0.0)
    expect = np.rint((img[0, 3].astype(float) + img[1, 0]) / 2)
    np.testing.assert_array_equal(got, expect)
    # last pixel of the buffer: the sample itself (the reference reads past the end)
    got = remap_one(img, 3.5, 2.0)
    np.testing.assert_array_equal(got, img[2, 3])
    # vertical clamp: y1 = min(y0 + 1, h - 1)
    got = remap_one(img, 1.0, 2.5)
    np.testing.assert_array_equal(got, img[2, 1])


def test_bilinear_map64_agrees_with_float_maps():
    rng = np.random.default_rng(9)
    img = src_img(h=40, w=50, seed=9)
    xm = (rng.random((20, 30)) * 52 - 1).astype(np.float32)
    ym = (rng.random((20, 30)) * 42 - 1).astype(np.float32)
    a = O.reproject_to16s(img, xm, ym)
    b = O.reproject_to16s_map64(img, np.stack([xm, ym], -1).astype(np.float64))
    np.testing.assert_array_equal(a, b)
    # closed form on interior points: weights (1-fx, fx) x (1-fy, fy), rint
    for (y, x) in [(3, 4), (10, 17), (19, 29)]:
        fx, fy = float(xm[y, x]), float(ym[y, x])
        if not (0 <= fx < 49 and 0 <= fy < 39):
            continue
        x0, y0 = int(fx), int(fy)
        ax, ay = fx - x0, fy - y0
        p = img.astype(float)
        v = (p[y0, x0] * (1 - ax) * (1 - ay) + p[y0, x0 + 1] * ax * (1 - ay) +
             p[y0 + 1, x0] * (1 - ax) * ay + p[y0 + 1, x0 + 1] * ax * ay)
        assert np.abs(a[y, x] - v).max() <= 0.5 + 1e-9


def test_lut_closed_form_ties_and_saturation():
    lut = O.get_lut(0.5)
    assert lut[1] == 0 and lut[3] == 2 and lut[5] == 2 and lut[7] == 4 and lut[255] == 128
    lut = O.get_lut(2.0)
    assert lut[127] == 254 and lut[128] == 255 and lut[255] == 255
    lut = O.get_lut(1.0)
    np.testing.assert_array_equal(lut, np.arange(256))
    k = 1.1
    np.testing.assert_array_equal(O.get_lut(k), np.clip(np.rint(k * np.arange(256)), 0, 255))


def test_transform_applies_per_channel():
    img = src_img(h=6, w=7, cn=3, seed=4)
    luts = np.stack([O.get_lut(0.9), O.get_lut(1.0), O.get_lut(1.2)])
    out = O.transform(img, luts)
    for c in range(3):
        np.testing.assert_array_equal(out[..., c], luts[c][img[..., c]])


W, H = 2048, 1024


def test_remap_axis_maps_to_image_centre():
    for t in (1, 3):   # full-frame and circular fisheye
        p = O.photo_param(imageType=t, hfov=195.0, cropWidth=1920, cropHeight=1920,
                          cropX=0, cropY=0)
        sx, sy = O.remap_point(p, 1920, 1920, W, H, W / 2 - 0.5, H / 2 - 0.5)
        assert abs(sx - 959.5) < 1e-9 and abs(sy - 959.5) < 1e-9
    p = O.photo_param(imageType=0, hfov=100.0)
    sx, sy = O.remap_point(p, 1920, 1080, W, H, W / 2 - 0.5, H / 2 - 0.5)
    assert abs(sx - 959.5) < 1e-9 and abs(sy - 539.5) < 1e-9


def test_remap_yaw_is_a_column_shift():
    p0 = O.photo_param(imageType=0, hfov=100.0)
    p1 = O.photo_param(imageType=0, hfov=100.0, yaw=60.0)
    for x, y in [(1000.0, 500.0), (1300.5, 400.0), (900.0, 700.25)]:
        a = O.remap_point(p0, 1920, 1080, W, H, x, y)
        b = O.remap_point(p1, 1920, 1080, W, H, x + W * 60 / 360, y)
        assert abs(a[0] - b[0]) < 1e-6 and abs(a[1] - b[1]) < 1e-6


def test_remap_equidistant_and_rectilinear_radii():
    # equidistant fisheye: r = cropWidth * angle / hfov
    p = O.photo_param(imageType=1, hfov=180.0)
    x = W / 2 - 0.5 + W * 45 / 360
    sx, sy = O.remap_point(p, 1000, 1000, W, H, x, H / 2 - 0.5)
    assert abs((sx - 499.5) - 1000 * 45 / 180) < 1e-6 and abs(sy - 499.5) < 1e-6
    # rectilinear: r = f tan(angle), f = cropWidth / (2 tan(hfov / 2))
    p = O.photo_param(imageType=0, hfov=90.0)
    sx, sy = O.remap_point(p, 1000, 800, W, H, W / 2 - 0.5 + W * 30 / 360, H / 2 - 0.5)
    assert abs((sx - 499.5) - 500 * math.tan(math.radians(30))) < 1e-6
    # latitude: 20 degrees up on the optical meridian -> straight up in the image
    sx, sy = O.remap_point(p, 1000, 800, W, H, W / 2 - 0.5, H / 2 - 0.5 - H * 20 / 180)
    assert abs(sx - 499.5) < 1e-6 and abs((399.5 - sy) - 500 * math.tan(math.radians(20))) < 1e-6


def test_remap_pitch_moves_the_axis():
    p = O.photo_param(imageType=0, hfov=90.0, pitch=10.0)
    ups = O.remap_point(p, 1000, 800, W, H, W / 2 - 0.5, H / 2 - 0.5 - H * 10 / 180)
    dns = O.remap_point(p, 1000, 800, W, H, W / 2 - 0.5, H / 2 - 0.5 + H * 10 / 180)
    hit = [q for q in (ups, dns) if abs(q[0] - 499.5) < 1e-6 and abs(q[1] - 399.5) < 1e-6]
    assert len(hit) == 1


def test_remap_radial_polynomial_numbers():
    # Remap.cpp:465-477: the camera-plane radius is scaled by
    # a t^3 + b t^2 + c t + d, t = radius / (min(crop) / 2), d = 1 - (a + b + c).
    # Equidistant fisheye, hfov 180, crop 1000: longitude 45 deg -> radius 250,
    # t = 0.5; with a = 0.1, b = -0.05, c = 0.02 the factor is 0.0125 - 0.0125 +
    # 0.01 + 0.93 = 0.94 (a and c swapped would give 0.97)
    p = O.photo_param(imageType=1, hfov=180.0, alpha=0.1, beta=-0.05, gamma=0.02)
    sx, sy = O.remap_point(p, 1000, 1000, W, H, W / 2 - 0.5 + W * 45 / 360, H / 2 - 0.5)
    assert abs((sx - 499.5) - 250 * 0.94) < 1e-6 and abs(sy - 499.5) < 1e-6
    # t = 1 (longitude 90 deg, radius 500) is a fixed point of any (a, b, c)
    sx, sy = O.remap_point(p, 1000, 1000, W, H, W / 2 - 0.5 + W * 90 / 360, H / 2 - 0.5)
    assert abs((sx - 499.5) - 500) < 1e-6


def test_remap_shift_and_shear_numbers():
    # Remap.cpp:479-505: after the radial step y += shiftY, x += shiftX, then
    # x += (shearX / cropHeight) y and y += (shearY / cropWidth) x (both from the
    # unsheared point), around the crop centre
    base = O.photo_param(imageType=0, hfov=90.0)
    moved = O.photo_param(imageType=0, hfov=90.0, shiftX=12.5, shiftY=-7.25)
    for x, y in [(W / 2 - 0.5, H / 2 - 0.5), (W / 2 - 0.5 + 40, H / 2 - 0.5 - 30)]:
        a = O.remap_point(base, 1000, 800, W, H, x, y)
        b = O.remap_point(moved, 1000, 800, W, H, x, y)
        assert abs(b[0] - a[0] - 12.5) < 1e-9 and abs(b[1] - a[1] + 7.25) < 1e-9
    # 20 deg up on the optical meridian: (0, -500 tan 20) relative to the centre;
    # shearX = 80 (g = 0.1 with cropHeight 800) moves x by 0.1 y, y stays; shearY
    # moves y by x / cropWidth * shearY = 0
    sheared = O.photo_param(imageType=0, hfov=90.0, shearX=80.0, shearY=50.0)
    y_rel = -500 * math.tan(math.radians(20))
    sx, sy = O.remap_point(sheared, 1000, 800, W, H, W / 2 - 0.5, H / 2 - 0.5 - H * 20 / 180)
    assert abs((sx - 499.5) - 0.1 * y_rel) < 1e-6 and abs((sy - 399.5) - y_rel) < 1e-6
    # on the equator 30 deg right: (500 tan 30, 0): x stays, y moves by 0.05 x
    x_rel = 500 * math.tan(math.radians(30))
    sx, sy = O.remap_point(sheared, 1000, 800, W, H, W / 2 - 0.5 + W * 30 / 360, H / 2 - 0.5)
    assert abs((sx - 499.5) - x_rel) < 1e-6 and abs((sy - 399.5) - 0.05 * x_rel) < 1e-6


def test_remap_roll_turns_the_image():
    # roll 90 deg: a point 30 deg right on the equator lands on the image's
    # vertical axis at the same distance from the centre
    p = O.photo_param(imageType=0, hfov=90.0, roll=90.0)
    sx, sy = O.remap_point(p, 1000, 1000, W, H, W / 2 - 0.5 + W * 30 / 360, H / 2 - 0.5)
    assert abs(sx - 499.5) < 1e-6 and abs(abs(sy - 499.5) - 500 * math.tan(math.radians(30))) < 1e-6


def test_map_and_mask_validity():
    p = O.photo_param(imageType=3, hfov=195.0, cropWidth=200, cropHeight=200, circleX=100,
                      circleY=100, circleR=100)
    m, mask = O.reproject_map_and_mask(p, 200, 200, 256, 128)
    valid = mask > 0
    assert valid.any() and (~valid).any()
    assert (m[~valid] == -1).all()
    d2 = (m[valid][:, 0] - 100) ** 2 + (m[valid][:, 1] - 100) ** 2
    assert (d2 < 100 ** 2).all()
    p = O.photo_param(imageType=0, hfov=100.0)
    m, mask = O.reproject_map_and_mask(p, 192, 108, 256, 128)
    valid = mask > 0
    assert (m[valid][:, 0] >= 0).all() and (m[valid][:, 0] < 192).all()
    assert (m[valid][:, 1] >= 0).all() and (m[valid][:, 1] < 108).all()
    # a 100 degree camera covers 100/360 of the equator
    frac = valid[64].mean()
    assert abs(frac - 100 / 360) < 0.01
