"""The product's host map generator (getReprojectMap32FAndMask) against the
oracle's FP64 restatement: masks bit-exact, float32 maps equal to the
oracle's doubles rounded to float."""
import numpy as np
import pytest

import panostitcher_amd as P
from oracle import oracle as O
from panostitcher_amd import synth


@pytest.mark.parametrize("name,scale", [("fisheye2", 0.25), ("rect6", 0.2), ("mixed8", 0.08),
                                        ("fisheye4", 0.06)])
def test_maps_match_oracle(name, scale):
    wl = synth.workload(name, scale)
    for cam in wl.cameras:
        x, y, m = P.generate_map(cam.param(), cam.width, cam.height, wl.dst_w, wl.dst_h)
        op = O.photo_param(**cam.param().as_dict())
        mm, mk = O.reproject_map_and_mask(op, cam.width, cam.height, wl.dst_w, wl.dst_h)
        np.testing.assert_array_equal(m, mk)
        np.testing.assert_array_equal(x, mm[..., 0].astype(np.float32))
        np.testing.assert_array_equal(y, mm[..., 1].astype(np.float32))
        assert 0.02 < (m > 0).mean() < 0.9


def test_exposure_luts_match_get_lut():
    e, r, b = [0.9, 1.1], [1.05, 0.95], [0.97, 1.2]
    luts = P.exposure_luts(e, r, b)
    for i in range(2):
        np.testing.assert_array_equal(luts[i, 0], O.get_lut(e[i] * b[i]))
        np.testing.assert_array_equal(luts[i, 1], O.get_lut(e[i]))
        np.testing.assert_array_equal(luts[i, 2], O.get_lut(e[i] * r[i]))
