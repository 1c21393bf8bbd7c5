"""Seeded synthetic rigs and frames for the stitch workloads.

There are no datasets or rig files offline, so every workload is generated:
camera parameters (PhotoParam fields), per-camera uint8 BGR content (smooth
value noise plus random axis-aligned rectangles, uniform 0-255), maps and
masks from the camera model, and per-channel gains turned into LUTs.

Workloads (sizes as in the task's benchmark table; ``scale`` shrinks canvas and
sources together for quick runs):
  fisheye2   2 x 1920^2 circular fisheye 195 deg, yaw 0/180 -> 2048x1024, L=5,
             masks cut to yaw sectors +-90 deg widened by 32 px, unit gains
  rect6      6 x 1920x1080 rectilinear 100 deg, yaw 0..300, pitch U(-5,5)
             -> 3840x1920, L=7, masks cut to a 10% sector overlap, gains U(0.85,1.15)
  mixed8     2 x 2560x1440 rect 90 + 4 x 1920x1080 rect 100 + 2 x 1280^2 fisheye 190,
             yaws every 45 deg -> 7680x3840, L=9
  fisheye4   4 x 3840x2160 fisheye 200 deg, yaw 0/90/180/270 -> 7680x3840, L=10
"""
from __future__ import annotations

import dataclasses

import numpy as np

SEED_BASE = 20261203


@dataclasses.dataclass
class Camera:
    kind: str            # "rect" | "fisheye"
    width: int
    height: int
    hfov: float
    yaw: float
    pitch: float = 0.0
    roll: float = 0.0

    def param(self):
        from panostitcher_amd import PhotoParam

        if self.kind == "rect":
            return PhotoParam.make(imageType=0, hfov=self.hfov, yaw=self.yaw, pitch=self.pitch,
                                   roll=self.roll, cropWidth=self.width, cropHeight=self.height)
        r = min(self.width, self.height) // 2
        return PhotoParam.make(imageType=3, hfov=self.hfov, yaw=self.yaw, pitch=self.pitch,
                               roll=self.roll, cropX=0, cropY=0, cropWidth=self.width,
                               cropHeight=self.height, circleX=self.width // 2,
                               circleY=self.height // 2, circleR=r)


@dataclasses.dataclass
class Workload:
    name: str
    cameras: list
    dst_w: int
    dst_h: int
    levels: int
    sector_overlap_px: int | None = None   # cut masks to yaw sectors widened by this many px
    gains: tuple | None = None             # (lo, hi) per-channel gain range, None = no LUTs
    seed: int = SEED_BASE

    @property
    def n(self):
        return len(self.cameras)

    @property
    def out_pixels(self):
        return self.dst_w * self.dst_h


def _even(v):
    v = max(2, int(round(v)))
    return v + (v & 1)


def workload(name: str, scale: float = 1.0, seed: int | None = None) -> Workload:
    s = scale
    rng = np.random.default_rng(SEED_BASE + 1000 * {"fisheye2": 1, "rect6": 2, "mixed8": 3,
                                                     "fisheye4": 5}[name] + (seed or 0))
    if name == "fisheye2":
        cams = [Camera("fisheye", _even(1920 * s), _even(1920 * s), 195.0, y) for y in (0.0, 180.0)]
        return Workload(name, cams, _even(2048 * s) * 1, _even(1024 * s), 5,
                        sector_overlap_px=max(1, int(round(32 * s))), seed=SEED_BASE + 1000)
    if name == "rect6":
        cams = [Camera("rect", _even(1920 * s), _even(1080 * s), 100.0, 60.0 * k,
                       pitch=float(rng.uniform(-5, 5))) for k in range(6)]
        dh = _even(1920 * s)
        # 10% overlap of each 60 deg sector: widen each side by 3 deg
        return Workload(name, cams, 2 * dh, dh, 7,
                        sector_overlap_px=max(1, int(round(2 * dh * 3.0 / 360.0))),
                        gains=(0.85, 1.15), seed=SEED_BASE + 2000)
    if name == "mixed8":
        kinds = [("rect", 2560, 1440, 90.0), ("rect", 1920, 1080, 100.0),
                 ("fisheye", 1280, 1280, 190.0), ("rect", 1920, 1080, 100.0),
                 ("rect", 2560, 1440, 90.0), ("rect", 1920, 1080, 100.0),
                 ("fisheye", 1280, 1280, 190.0), ("rect", 1920, 1080, 100.0)]
        cams = [Camera(k, _even(w * s), _even(h * s), f, 45.0 * i)
                for i, (k, w, h, f) in enumerate(kinds)]
        dh = _even(3840 * s)
        return Workload(name, cams, 2 * dh, dh, 9, seed=SEED_BASE + 3000)
    if name == "fisheye4":
        cams = [Camera("fisheye", _even(3840 * s), _even(2160 * s), 200.0, 90.0 * k)
                for k in range(4)]
        dh = _even(3840 * s)
        return Workload(name, cams, 2 * dh, dh, 10, gains=(0.9, 1.1), seed=SEED_BASE + 5000)
    raise ValueError(name)


SERVED_SIZES = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
SERVED_OUTPUTS = [(2048, 1024, 5), (3840, 1920, 7), (7680, 3840, 9)]


def served_rig(k: int) -> Workload:
    """The k-th seeded rig of the served mix: 2-8 cameras of one size from
    1280x720 to 3840x2160, fisheye (always for <= 3 cameras) or rectilinear
    wide enough to overlap its neighbours, yaws evenly spaced, output
    2048x1024 / 3840x1920 / 7680x3840 with 5 / 7 / 9 levels."""
    rng = np.random.default_rng(SEED_BASE + 4000 + 7 * k)
    n = int(rng.integers(2, 9))
    w, h = SERVED_SIZES[int(rng.integers(0, len(SERVED_SIZES)))]
    kind = "fisheye" if n <= 3 or rng.random() < 0.4 else "rect"
    if kind == "fisheye":
        w = h = min(w, h)
        fov = float(rng.uniform(190, 200))
    else:
        fov = float(min(120.0, max(360.0 / n * 1.2, 60.0)))
    cams = [Camera(kind, w, h, fov, 360.0 * i / n + float(rng.uniform(-3, 3)),
                   pitch=float(rng.uniform(-3, 3))) for i in range(n)]
    dw, dh, levels = SERVED_OUTPUTS[int(rng.integers(0, len(SERVED_OUTPUTS)))]
    return Workload(f"served{k}", cams, dw, dh, levels, seed=SEED_BASE + 4000 + 7 * k)


def value_noise(h: int, w: int, rng: np.random.Generator, octaves=(8, 32, 128)) -> np.ndarray:
    """Smooth noise in [0, 1): bilinear upsampling of random grids at several scales."""
    out = np.zeros((h, w), np.float32)
    total = 0.0
    for k, cell in enumerate(octaves):
        gh, gw = h // cell + 2, w // cell + 2
        g = rng.random((gh, gw)).astype(np.float32)
        ys = np.arange(h, dtype=np.float32) / cell
        xs = np.arange(w, dtype=np.float32) / cell
        y0 = ys.astype(np.int32)
        x0 = xs.astype(np.int32)
        fy = (ys - y0)[:, None]
        fx = (xs - x0)[None, :]
        a = g[y0][:, x0]
        b = g[y0][:, x0 + 1]
        c = g[y0 + 1][:, x0]
        d = g[y0 + 1][:, x0 + 1]
        amp = 0.5 ** (len(octaves) - 1 - k)
        out += amp * ((a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy)
        total += amp
    return out / total


def frame(h: int, w: int, seed: int, rects: int = 24) -> np.ndarray:
    """uint8 BGR frame: per-channel value noise plus random uniform rectangles."""
    rng = np.random.default_rng(seed)
    img = np.empty((h, w, 3), np.uint8)
    for c in range(3):
        img[..., c] = np.clip(value_noise(h, w, rng) * 256, 0, 255).astype(np.uint8)
    for _ in range(rects):
        y0, x0 = int(rng.integers(0, h)), int(rng.integers(0, w))
        rh, rw = int(rng.integers(h // 32 + 1, h // 4 + 2)), int(rng.integers(w // 32 + 1, w // 4 + 2))
        img[y0:y0 + rh, x0:x0 + rw] = rng.integers(0, 256, size=3, dtype=np.uint8)
    return img


def frames(wl: Workload, frame_index: int = 0):
    return [frame(c.height, c.width, wl.seed + 97 * i + 7919 * frame_index)
            for i, c in enumerate(wl.cameras)]


def bgrx(img: np.ndarray) -> np.ndarray:
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    out[..., :3] = img
    return out


def sector_cut(masks, xmaps, ymaps, wl: Workload):
    """Restrict each camera's mask to its yaw sector widened by sector_overlap_px
    on each side; the map is set to (-1, -1) where the mask is cut, as
    getReprojectMapAndMask does for invalid pixels."""
    if wl.sector_overlap_px is None:
        return masks, xmaps, ymaps
    n = wl.n
    W = wl.dst_w
    half = W / n / 2.0 + wl.sector_overlap_px
    cols = np.arange(W, dtype=np.float64) + 0.5
    out = ([], [], [])
    for i, cam in enumerate(wl.cameras):
        centre = (W / 2.0 + cam.yaw / 360.0 * W) % W
        d = np.abs(((cols - centre) + W / 2.0) % W - W / 2.0)
        keep = (d <= half)[None, :]
        m = np.where(keep, masks[i], 0).astype(np.uint8)
        x = np.where(m > 0, xmaps[i], -1).astype(np.float32)
        y = np.where(m > 0, ymaps[i], -1).astype(np.float32)
        out[0].append(m)
        out[1].append(x)
        out[2].append(y)
    return out


def maps_and_masks(wl: Workload):
    """(xmaps, ymaps, masks) from the product's map generator (host), with the
    workload's sector cut applied."""
    from panostitcher_amd import generate_map

    xs, ys, ms = [], [], []
    for cam in wl.cameras:
        x, y, m = generate_map(cam.param(), cam.width, cam.height, wl.dst_w, wl.dst_h)
        xs.append(x)
        ys.append(y)
        ms.append(m)
    ms, xs, ys = sector_cut(ms, xs, ys, wl)
    return xs, ys, ms


def luts(wl: Workload, frame_index: int = 0):
    """Per-camera 3x256 LUTs from gains U(lo, hi) per channel (getLUT), or None."""
    if wl.gains is None:
        return None
    from panostitcher_amd import get_lut

    rng = np.random.default_rng(wl.seed + 31 + frame_index)
    lo, hi = wl.gains
    out = np.zeros((wl.n, 3, 256), np.uint8)
    for i in range(wl.n):
        for c in range(3):
            out[i, c] = get_lut(float(rng.uniform(lo, hi)))
    return out
