"""Rig files (Warp/PhotoParam.cpp): the XML layouts, PTGui .pts, export and
camera rotation, through the C-ABI.

The fixtures under tests/golden/rigs/ are hand-written; every expected value
below is read off the fixture text by the rules of PhotoParam.cpp (cited per
case), so a dropped field, a swapped key or a wrong truncation fails here.
The rotation is pinned by an independent construction: setRotationRM is
R = Ry(yaw) Rx(pitch) Rz(roll) with right-handed axis rotations
(Warp/Rotation.h:82-87), built here from the elementary matrices."""
import math
import os

import numpy as np
import pytest

import panostitcher_amd as P

RIGS = os.path.join(os.path.dirname(__file__), "golden", "rigs")
FIELDS = [n for n, _ in P.PhotoParam._fields_]


def rig(name):
    return os.path.join(RIGS, name)


def check(q, **want):
    """All fields of q: the ones given equal, the rest zero."""
    for f in FIELDS:
        assert getattr(q, f) == pytest.approx(want.get(f, 0), abs=0, rel=1e-15), f


def test_per_camera_layout():
    # PhotoParam.cpp:273-504; the VIDEO element is ignored when POSITION exists
    ps = P.load_photo_param_from_xml(rig("per_camera.xml"))
    assert len(ps) == 2
    # CROPLEFT 10.7 -> 10 (double to int), CROPRIGHT 1930.2 - 10 -> 1920,
    # CircleR 955.9 -> 955, "  4 " read as 4
    check(ps[0], imageType=3, hfov=195.5, alpha=0.012, beta=-0.034, gamma=0.0055, shiftX=12.5,
          shiftY=-3.25, shearY=1e-3, cropX=10, cropWidth=1920, cropY=4, cropHeight=1920,
          circleX=970, circleY=964, circleR=955, yaw=-90.25, pitch=1.5, roll=-0.75)
    check(ps[1], imageType=0, cropMode=1, hfov=100, vfov=67.5, cropWidth=1920, cropHeight=1080,
          yaw=90, pitch=-2, roll=0.5)


def test_shared_lens_layout():
    # PhotoParam.cpp:17-271: each Root's GlobalInfo lens, the crop from
    # GlobalInfo (first Root) or from each VIDEO element (second Root)
    ps = P.load_photo_param_from_xml(rig("shared_lens.xml"))
    assert len(ps) == 4
    lens = dict(imageType=2, cropMode=1, hfov=190, vfov=150, cropX=100, cropWidth=1280, cropY=20,
                cropHeight=1280, alpha=0.5, beta=-0.25, gamma=0.125, shiftX=2, shiftY=-1,
                shearX=0.001, shearY=-0.002)
    check(ps[0], **lens)
    check(ps[1], **lens, yaw=180, pitch=5, roll=-5)
    check(ps[2], hfov=90, cropWidth=2560, cropHeight=1440, yaw=45)
    check(ps[3], hfov=90, cropX=8, cropWidth=2544, cropY=4, cropHeight=1432, yaw=-45, pitch=1, roll=2)


def test_pts():
    # PhotoParam.cpp:681-754: the first full 'o' line is the lens (not a
    # camera); f2 -> circular (3), f3 -> full frame (1), f4 -> 0; the f0 line is
    # dropped because its value text after C splits into five parts at the
    # commas (the file name holds one); comments, 'p', 'm', 'k' and a CR-only
    # line are skipped
    ps = P.load_photo_param_from_pts(rig("ptgui.pts"))
    assert len(ps) == 3
    lens = dict(hfov=195.5, alpha=0.012, beta=-0.034, gamma=0.0055, shiftX=12.5, shiftY=-3.25,
                shearY=0.001)
    check(ps[0], imageType=3, yaw=-90.5, roll=0.25, pitch=1.5, cropX=10, cropY=5, cropWidth=1900,
          cropHeight=1910, **lens)
    check(ps[1], imageType=1, yaw=89.75, roll=-0.5, pitch=-2, **lens)
    check(ps[2], imageType=0, yaw=10, roll=1, pitch=2, **lens)


def test_dispatch_by_extension(tmp_path):
    # loadPhotoParams (PhotoParam.cpp:756-770): "pts" exactly, anything else XML
    assert len(P.load_photo_params(rig("ptgui.pts"))) == 3
    assert len(P.load_photo_params(rig("per_camera.xml"))) == 2
    upper = tmp_path / "rig.PTS"
    upper.write_bytes(open(rig("ptgui.pts"), "rb").read())
    assert P.load_photo_params(str(upper)) == []          # read as XML: not XML
    with pytest.raises(P.PanoError):
        P.load_photo_params(str(tmp_path / "noext"))
    assert P.load_photo_params(str(tmp_path / "missing.xml")) == []


def test_pts_without_lens_line_or_header(tmp_path):
    f = tmp_path / "a.pts"
    f.write_text("# ptGui project file\no f2 y1 r2 p3\n")      # no v a b c d e g t line
    assert P.load_photo_param_from_pts(str(f)) == []
    f.write_text("o f2 y0 r0 p0 v100 a0 b0 c0 d0 e0 g0 t0\no f2 y1 r2 p3\n")   # no header
    assert P.load_photo_param_from_pts(str(f)) == []
    f.write_text("# ptGui project file\no f2 y0 r0 p0 v100 a0 b0 c0 d0 e0 g0 t0\no f0 y1 r2 p3 C1,2,3\n")
    assert P.load_photo_param_from_pts(str(f)) == []            # three crop values: line dropped


def test_xml_errors(tmp_path):
    good = open(rig("per_camera.xml")).read()
    f = tmp_path / "r.xml"
    # a field that is not a number: the whole load fails (ticpp throws)
    f.write_text(good.replace("<HFOV>100</HFOV>", "<HFOV>wide</HFOV>"))
    with pytest.raises(P.PanoError, match="not a number"):
        P.load_photo_param_from_xml(str(f))
    f.write_text(good.replace("<HFOV>100</HFOV>", "<HFOV></HFOV>"))
    with pytest.raises(P.PanoError):
        P.load_photo_param_from_xml(str(f))
    # a missing element: the per-camera layout fails and so does the shared-lens one
    f.write_text(good.replace("<CircleR>0</CircleR>", ""))
    assert P.load_photo_param_from_xml(str(f)) == []
    # shared lens without the crop in one camera element
    sl = open(rig("shared_lens.xml")).read()
    f.write_text(sl.replace("<CROPLEFT>8</CROPLEFT>", ""))
    assert P.load_photo_param_from_xml(str(f)) == []
    # not well formed
    f.write_text(good.replace("</Root>", ""))
    assert P.load_photo_param_from_xml(str(f)) == []
    f.write_text("")
    assert P.load_photo_param_from_xml(str(f)) == []
    # pathological nesting is refused, not recursed into
    f.write_text("<Root>" + "<a>" * 100000 + "</a>" * 100000 + "</Root>")
    assert P.load_photo_param_from_xml(str(f)) == []
    # attributes with '>' inside quotes, self-closing elements, CDATA text
    f.write_text(good.replace("<POSITION>", '<POSITION id="a>b" />\n<POSITION>', 1)
                     .replace("<HFOV>195.5</HFOV>", "<HFOV><![CDATA[195.5]]></HFOV>"))
    assert P.load_photo_param_from_xml(str(f)) == []     # the empty POSITION lacks every field
    f.write_text(good.replace("<HFOV>195.5</HFOV>", "<HFOV><![CDATA[195.5]]></HFOV>"))
    assert P.load_photo_param_from_xml(str(f))[0].hfov == 195.5


def test_export_round_trip(tmp_path):
    ps = P.load_photo_param_from_xml(rig("per_camera.xml"))
    ps[0].yaw = 12.3456789
    ps[1].hfov = 1.0 / 3.0
    out = tmp_path / "e.xml"
    P.export_photo_param_to_xml(str(out), ps)
    text = out.read_text()
    # numbers as an ostream prints them: 6 significant digits
    assert "<YAW>12.3457</YAW>" in text and "<HFOV>0.333333</HFOV>" in text
    assert "<CROPRIGHT>1930</CROPRIGHT>" in text and "<ShearY>0.001</ShearY>" in text
    back = P.load_photo_params(str(out))
    assert len(back) == 2
    for a, b in zip(ps, back):
        for f in FIELDS:
            va, vb = getattr(a, f), getattr(b, f)
            assert vb == float(f"{va:.6g}"), f
    P.export_photo_param_to_xml(str(out), [])
    assert P.load_photo_params(str(out)) == []


def _axes(yaw, pitch, roll):
    """Ry(yaw) Rx(pitch) Rz(roll), right-handed elementary rotations."""
    cy, sy, cp, sp, cr, sr = (math.cos(yaw), math.sin(yaw), math.cos(pitch), math.sin(pitch),
                              math.cos(roll), math.sin(roll))
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    rz = np.array([[cr, -sr, 0], [sr, cr, 0], [0, 0, 1]])
    return ry @ rx @ rz


def test_rotate_cameras_composes_rotations():
    rng = np.random.default_rng(7)
    cams = [P.PhotoParam.make(yaw=y, pitch=p, roll=r) for y, p, r in
            zip(rng.uniform(-180, 180, 12), rng.uniform(-89, 89, 12), rng.uniform(-180, 180, 12))]
    cams.append(P.PhotoParam.make(yaw=30, pitch=90, roll=0))     # gimbal lock: second branch
    cams.append(P.PhotoParam.make(yaw=-60, pitch=-90, roll=0))
    before = [(c.yaw, c.pitch, c.roll) for c in cams]
    rot = (0.3, -0.2, 0.1)
    P.rotate_cameras(cams, *rot)
    d = math.pi / 180
    for (y0, p0, r0), c in zip(before, cams):
        want = _axes(*rot) @ _axes(y0 * d, p0 * d, r0 * d)
        np.testing.assert_allclose(_axes(c.yaw * d, c.pitch * d, c.roll * d), want, atol=1e-9)
        assert -90 <= c.pitch <= 90 + 1e-9 and -180 <= c.yaw <= 180 and -180 <= c.roll <= 180


def test_rotate_camera_closed_forms():
    c = P.PhotoParam.make(yaw=10, pitch=0, roll=0)
    P.rotate_camera(c, math.radians(20), 0, 0)
    assert (c.yaw, c.pitch, c.roll) == (pytest.approx(30, abs=1e-9), pytest.approx(0, abs=1e-12),
                                        pytest.approx(0, abs=1e-12))
    c = P.PhotoParam.make(yaw=-40, pitch=15, roll=-25)
    P.rotate_camera(c, 0, 0, 0)
    assert (c.yaw, c.pitch, c.roll) == (pytest.approx(-40, abs=1e-9), pytest.approx(15, abs=1e-9),
                                        pytest.approx(-25, abs=1e-9))
    c = P.PhotoParam.make(yaw=170, pitch=0, roll=0)
    P.rotate_camera(c, math.radians(20), 0, 0)
    assert c.yaw == pytest.approx(-170, abs=1e-9)                  # wraps through atan2


def test_rotate_photo_param_in_xml(tmp_path):
    out = tmp_path / "rot.xml"
    P.rotate_photo_param_in_xml(rig("per_camera.xml"), str(out), 0.5, 0.1, -0.2)
    ps = P.load_photo_param_from_xml(rig("per_camera.xml"))
    P.rotate_cameras(ps, 0.5, 0.1, -0.2)
    back = P.load_photo_params(str(out))
    for a, b in zip(ps, back):
        for f in ("yaw", "pitch", "roll", "hfov", "cropWidth", "circleR"):
            assert getattr(b, f) == float(f"{getattr(a, f):.6g}"), f
