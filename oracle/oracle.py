"""ctypes wrapper over oracle/liboracle.so -- the CPU restatement of the
reference's per-frame stitch path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py``, always as the checker (or as the
timed CPU baseline), never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

REFLECT_101 = 4
WRAP = 3


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        i = C.c_int
        L.orc_border.argtypes = [i, i, i]
        L.orc_border.restype = i
        for nm in ("orc_pyrdown_to32s_s16", "orc_pyrdown_s16", "orc_pyrup_s16", "orc_pyrup_s32"):
            getattr(L, nm).argtypes = [vp, i, i, i, i, vp, i, i, i, i, i]
            getattr(L, nm).restype = None
        L.orc_true_num_levels.argtypes = [i, i, i, i]
        L.orc_true_num_levels.restype = i
        L.orc_distance_transform_l1.argtypes = [vp, i, i, vp]
        L.orc_nonintersecting_masks.argtypes = [i, i, i, vp, vp]
        L.orc_blender_create.restype = vp
        L.orc_blender_destroy.argtypes = [vp]
        L.orc_blender_prepare.argtypes = [vp, i, i, i, vp, i, i]
        L.orc_blender_prepare.restype = i
        L.orc_blender_blend.argtypes = [vp, vp, vp, i]
        L.orc_blender_blend.restype = i
        L.orc_gaussian_kernel_q8.argtypes = [i, C.c_double, vp]
        L.orc_gaussian_kernel_q8.restype = None
        L.orc_gaussian_blur_u8.argtypes = [vp, i, i, i, C.c_double, vp]
        L.orc_gaussian_blur_u8.restype = None
        L.orc_linear_weights.argtypes = [i, i, i, vp, i, i, vp]
        L.orc_linear_weights.restype = i
        L.orc_linear_render.argtypes = [i, vp, vp, vp, vp, vp, vp, i, i, vp, vp, i]
        L.orc_linear_render.restype = i
        L.orc_linear_blend32s.argtypes = [i, i, i, i, vp, vp, vp]
        L.orc_linear_blend32s.restype = i
        L.orc_bgr32_to_yuv.argtypes = [vp, C.c_size_t, i, i, i, vp, vp, vp]
        L.orc_bgr32_to_yuv.restype = i
        L.orc_yuv_to_bgr32.argtypes = [vp, vp, vp, i, i, i, vp, C.c_size_t]
        L.orc_yuv_to_bgr32.restype = i
        L.orc_blender_blend_custom.argtypes = [vp, vp, vp, vp, i]
        L.orc_blender_blend_custom.restype = i
        L.orc_blender_num_levels.argtypes = [vp]
        L.orc_blender_num_levels.restype = i
        L.orc_blender_full_mask.argtypes = [vp]
        L.orc_blender_full_mask.restype = i
        L.orc_blender_unique_masks.argtypes = [vp, vp]
        L.orc_blender_level_size.argtypes = [vp, i, C.POINTER(i), C.POINTER(i)]
        L.orc_blender_level_size.restype = i
        L.orc_blender_weight.argtypes = [vp, i, i, vp]
        L.orc_blender_weight.restype = i
        L.orc_blender_alpha.argtypes = [vp, i, i, vp]
        L.orc_blender_alpha.restype = i
        L.orc_reproject_to16s.argtypes = [vp, i, i, C.c_size_t, i, vp, vp, i, i, vp, i]
        L.orc_reproject_to16s_map64.argtypes = [vp, i, i, C.c_size_t, i, vp, i, i, vp, i]
        L.orc_get_lut.argtypes = [vp, C.c_double]
        L.orc_transform.argtypes = [vp, vp, i, i, C.c_size_t, i, vp]
        L.orc_remap_point.argtypes = [vp, i, i, i, i, C.c_double, C.c_double,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_reproject_map_and_mask.argtypes = [vp, i, i, i, i, vp, vp, i]
        L.orc_render.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, i]
        L.orc_render.restype = i
        L.orc_gray_bgr.argtypes = [i, i, i]
        L.orc_gray_bgr.restype = i
        L.orc_cv_solve.argtypes = [i, vp, vp, vp]
        L.orc_cv_solve.restype = i
        L.orc_gain_sums.argtypes = [i, i, i, vp, vp, vp]
        L.orc_gain_from_sums.argtypes = [i, vp, i, vp, vp, vp, C.POINTER(i), vp]
        L.orc_gain_from_sums.restype = i
        L.orc_pyrdown32f.argtypes = [vp, i, i, i, vp]
        L.orc_pyrup32f.argtypes = [vp, i, i, i, vp, i, i]
        L.orc_blend32f.argtypes = [i, i, i, vp, vp, i, i, vp]
        L.orc_blend32f.restype = i
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def border(p: int, n: int, kind: int) -> int:
    return lib().orc_border(p, n, kind)


def _pyr(fn, src, dshape, dtype, hb, vb):
    src = np.ascontiguousarray(src)
    sh, sw = src.shape[:2]
    cn = 1 if src.ndim == 2 else src.shape[2]
    dh, dw = dshape
    dst = np.zeros((dh, dw, cn) if src.ndim == 3 else (dh, dw), dtype=dtype)
    fn(_p(src), sw, sh, sw * cn, cn, _p(dst), dw, dh, dw * cn, hb, vb)
    return dst


def pyrdown_to32s(src, hb=WRAP, vb=REFLECT_101):
    h, w = src.shape[:2]
    return _pyr(lib().orc_pyrdown_to32s_s16, src.astype(np.int16), ((h + 1) // 2, (w + 1) // 2),
                np.int32, hb, vb)


def pyrdown_s16(src, hb=WRAP, vb=REFLECT_101):
    h, w = src.shape[:2]
    return _pyr(lib().orc_pyrdown_s16, src.astype(np.int16), ((h + 1) // 2, (w + 1) // 2),
                np.int16, hb, vb)


def pyrup_s16(src, dshape, hb=WRAP, vb=REFLECT_101):
    return _pyr(lib().orc_pyrup_s16, src.astype(np.int16), dshape, np.int16, hb, vb)


def pyrup_s32(src, dshape, hb=WRAP, vb=REFLECT_101):
    return _pyr(lib().orc_pyrup_s32, src.astype(np.int32), dshape, np.int32, hb, vb)


def true_num_levels(w, h, max_levels, min_length=2):
    return lib().orc_true_num_levels(w, h, max_levels, min_length)


def distance_transform_l1(mask: np.ndarray) -> np.ndarray:
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    h, w = mask.shape
    out = np.zeros((h, w), np.float32)
    lib().orc_distance_transform_l1(_p(mask), w, h, _p(out))
    return out


def nonintersecting_masks(masks: np.ndarray) -> np.ndarray:
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    n, h, w = masks.shape
    out = np.zeros_like(masks)
    lib().orc_nonintersecting_masks(n, w, h, _p(masks), _p(out))
    return out


class Blender:
    """TilingMultibandBlendFast (Blend/ZBlend.h:114-143) on the CPU."""

    def __init__(self):
        self.h = lib().orc_blender_create()
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_blender_destroy(self.h)
            self.h = None

    def prepare(self, masks: np.ndarray, max_levels: int, min_length: int = 2) -> bool:
        masks = np.ascontiguousarray(masks, dtype=np.uint8)
        n, rows, cols = masks.shape
        self.n, self.rows, self.cols = n, rows, cols
        return bool(lib().orc_blender_prepare(self.h, n, rows, cols, _p(masks), max_levels,
                                              min_length))

    @property
    def num_levels(self):
        return lib().orc_blender_num_levels(self.h)

    @property
    def full_mask(self):
        return bool(lib().orc_blender_full_mask(self.h))

    def unique_masks(self):
        out = np.zeros((self.n, self.rows, self.cols), np.uint8)
        lib().orc_blender_unique_masks(self.h, _p(out))
        return out

    def level_size(self, level):
        w, h = C.c_int(), C.c_int()
        assert lib().orc_blender_level_size(self.h, level, C.byref(w), C.byref(h))
        return w.value, h.value

    def weight(self, cam, level):
        w, h = self.level_size(level)
        out = np.zeros((h, w), np.int16)
        assert lib().orc_blender_weight(self.h, cam, level, _p(out))
        return out

    def alpha(self, cam, level):
        w, h = self.level_size(level)
        out = np.zeros((h, w), np.int32)
        assert lib().orc_blender_alpha(self.h, cam, level, _p(out))
        return out

    def blend(self, images: np.ndarray, nthreads: int = 1) -> np.ndarray:
        images = np.ascontiguousarray(images, dtype=np.int16)
        assert images.shape == (self.n, self.rows, self.cols, 3)
        out = np.zeros((self.rows, self.cols, 3), np.uint8)
        assert lib().orc_blender_blend(self.h, _p(images), _p(out), nthreads)
        return out

    def blend_custom(self, images: np.ndarray, masks: np.ndarray, nthreads: int = 1) -> np.ndarray:
        """blend(images, masks, blendImage): weights from these masks."""
        images = np.ascontiguousarray(images, dtype=np.int16)
        masks = np.ascontiguousarray(masks, dtype=np.uint8)
        assert images.shape == (self.n, self.rows, self.cols, 3)
        assert masks.shape == (self.n, self.rows, self.cols)
        out = np.zeros((self.rows, self.cols, 3), np.uint8)
        assert lib().orc_blender_blend_custom(self.h, _p(images), _p(masks), _p(out), nthreads)
        return out

    def render(self, srcs, xmaps, ymaps, luts=None, nthreads=1) -> np.ndarray:
        """CPUPanoramaRender::render on BGR sources (list of HxWx3 uint8)."""
        n = len(srcs)
        srcs = [np.ascontiguousarray(s[..., :3]) for s in srcs]
        xmaps = [np.ascontiguousarray(x, dtype=np.float32) for x in xmaps]
        ymaps = [np.ascontiguousarray(y, dtype=np.float32) for y in ymaps]
        sp = (C.c_void_p * n)(*[s.ctypes.data for s in srcs])
        xp = (C.c_void_p * n)(*[x.ctypes.data for x in xmaps])
        yp = (C.c_void_p * n)(*[y.ctypes.data for y in ymaps])
        sw = np.array([s.shape[1] for s in srcs], np.int32)
        shh = np.array([s.shape[0] for s in srcs], np.int32)
        lp = None
        if luts is not None:
            luts = np.ascontiguousarray(luts, dtype=np.uint8)
            assert luts.shape == (n, 3, 256)
            lp = _p(luts)
        out = np.zeros((self.rows, self.cols, 3), np.uint8)
        assert lib().orc_render(self.h, n, sp, _p(sw), _p(shh), xp, yp, lp, _p(out), nthreads)
        return out


def reproject_to16s(src: np.ndarray, xmap: np.ndarray, ymap: np.ndarray, nthreads=1):
    src = np.ascontiguousarray(src, dtype=np.uint8)
    sh, sw, cn = src.shape
    xmap = np.ascontiguousarray(xmap, dtype=np.float32)
    ymap = np.ascontiguousarray(ymap, dtype=np.float32)
    dh, dw = xmap.shape
    out = np.zeros((dh, dw, cn), np.int16)
    lib().orc_reproject_to16s(_p(src), sw, sh, sw * cn, cn, _p(xmap), _p(ymap), dw, dh, _p(out),
                              nthreads)
    return out


def reproject_to16s_map64(src: np.ndarray, map_xy: np.ndarray, nthreads=1):
    src = np.ascontiguousarray(src, dtype=np.uint8)
    sh, sw, cn = src.shape
    map_xy = np.ascontiguousarray(map_xy, dtype=np.float64)
    dh, dw = map_xy.shape[:2]
    out = np.zeros((dh, dw, cn), np.int16)
    lib().orc_reproject_to16s_map64(_p(src), sw, sh, sw * cn, cn, _p(map_xy), dw, dh, _p(out),
                                    nthreads)
    return out


def get_lut(k: float) -> np.ndarray:
    out = np.zeros(256, np.uint8)
    lib().orc_get_lut(_p(out), float(k))
    return out


def transform(src: np.ndarray, lut3: np.ndarray) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    h, w, cn = src.shape
    lut3 = np.ascontiguousarray(lut3, dtype=np.uint8)
    out = np.zeros_like(src)
    lib().orc_transform(_p(src), _p(out), w, h, w * cn, cn, _p(lut3))
    return out


def gray_bgr(b: int, g: int, r: int) -> int:
    """cv::cvtColor(BGR2GRAY) of one 8-bit pixel."""
    return lib().orc_gray_bgr(int(b), int(g), int(r))


def cv_solve(A: np.ndarray, b: np.ndarray):
    """cv::solve(A, b, x, DECOMP_LU) -> (ok, x)."""
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(n)
    x = np.zeros(n, np.float64)
    ok = lib().orc_cv_solve(n, _p(A), _p(b), _p(x))
    return bool(ok), x


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def gain_sums(bgr, masks) -> np.ndarray:
    """(n, n, 5) int64: count, gray, B, G, R of image i over mask_i & mask_j."""
    bgr = [np.ascontiguousarray(a, dtype=np.uint8) for a in bgr]
    masks = [np.ascontiguousarray(m, dtype=np.uint8) for m in masks]
    n = len(bgr)
    rows, cols = masks[0].shape
    out = np.zeros((n, n, 5), np.int64)
    lib().orc_gain_sums(n, rows, cols, _ptrs(bgr), _ptrs(masks), _p(out))
    return out


def gain_from_sums(sums: np.ndarray, tint: bool = True):
    """-> (gains, rg, bg, max_index, luts (n, 3, 256))."""
    sums = np.ascontiguousarray(sums, dtype=np.int64)
    n = sums.shape[0]
    g, rg, bg = (np.zeros(n, np.float64) for _ in range(3))
    mi = C.c_int(0)
    luts = np.zeros((n, 3, 256), np.uint8)
    lib().orc_gain_from_sums(n, _p(sums), int(tint), _p(g), _p(rg), _p(bg), C.byref(mi), _p(luts))
    return g, rg, bg, mi.value, luts


def gain_estimate(bgr, masks, tint: bool = True):
    """Gains from overlap means of reprojected BGR images (see gain.c)."""
    return gain_from_sums(gain_sums(bgr, masks), tint)


def pyrdown32f(src: np.ndarray) -> np.ndarray:
    """pyrDown32FC1/32FC4 of the float blender on an (h, w) or (h, w, cn) float32 image."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    cn = 1 if src.ndim == 2 else src.shape[2]
    out = np.zeros(((h + 1) // 2, (w + 1) // 2) + src.shape[2:], np.float32)
    lib().orc_pyrdown32f(_p(src), w, h, cn, _p(out))
    return out


def pyrup32f(src: np.ndarray, dshape) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    cn = 1 if src.ndim == 2 else src.shape[2]
    dh, dw = dshape
    out = np.zeros((dh, dw) + src.shape[2:], np.float32)
    lib().orc_pyrup32f(_p(src), w, h, cn, _p(out), dw, dh)
    return out


def blend32f(masks: np.ndarray, images: np.ndarray, max_levels: int, min_length: int = 2):
    """CudaTilingMultibandBlendFast32F prepare + blend: masks (n, h, w) u8,
    images (n, h, w, 4) u8 or float32 -> (h, w, 4) u8 and the level count."""
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    images = np.ascontiguousarray(images, dtype=np.float32)
    n, h, w = masks.shape
    out = np.zeros((h, w, 4), np.uint8)
    L = lib().orc_blend32f(n, h, w, _p(masks), _p(images), max_levels, min_length, _p(out))
    return out, L


def gaussian_kernel_q8(n: int, sigma: float) -> np.ndarray:
    """getGaussianKernel(n, sigma, CV_32F) as 8-bit fixed-point taps."""
    K = np.zeros(n, np.int32)
    lib().orc_gaussian_kernel_q8(n, sigma, _p(K))
    return K


def gaussian_blur_u8(src: np.ndarray, n: int, sigma: float) -> np.ndarray:
    src = np.ascontiguousarray(src, np.uint8)
    h, w = src.shape
    out = np.zeros_like(src)
    lib().orc_gaussian_blur_u8(_p(src), w, h, n, sigma, _p(out))
    return out


def linear_weights(masks: np.ndarray, radius: int, as_int: bool = False) -> np.ndarray:
    """getWeightsLinearBlend32F (float) / getWeightsLinearBlend (int32, 2^16 units)."""
    masks = np.ascontiguousarray(masks, np.uint8)
    n, h, w = masks.shape
    out = np.zeros((n, h, w), np.int32 if as_int else np.float32)
    assert lib().orc_linear_weights(n, h, w, _p(masks), radius, int(as_int), _p(out))
    return out


def linear_render(srcs, xmaps, ymaps, weights, luts=None, nthreads=8) -> np.ndarray:
    """CPUPanoramaRender::render, linear branch, on BGR frames -> (h, w, 3) u8."""
    n = len(srcs)
    srcs = [np.ascontiguousarray(s[..., :3]) for s in srcs]
    xmaps = [np.ascontiguousarray(x, np.float32) for x in xmaps]
    ymaps = [np.ascontiguousarray(y, np.float32) for y in ymaps]
    weights = np.ascontiguousarray(weights, np.float32)
    dh, dw = xmaps[0].shape
    sw = (C.c_int * n)(*[s.shape[1] for s in srcs])
    sh = (C.c_int * n)(*[s.shape[0] for s in srcs])
    lp = None
    if luts is not None:
        luts = np.ascontiguousarray(luts, np.uint8)
        lp = _p(luts)
    out = np.zeros((dh, dw, 3), np.uint8)
    assert lib().orc_linear_render(n, _ptrs(srcs), sw, sh, _ptrs(xmaps), _ptrs(ymaps), _p(weights), dw, dh,
                                   lp, _p(out), nthreads)
    return out


def linear_blend32s(images: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """TilingLinearBlend::blend / CudaTilingLinearBlend::blend: images (n, h, w, cn) u8."""
    images = np.ascontiguousarray(images, np.uint8)
    weights = np.ascontiguousarray(weights, np.int32)
    n, h, w, cn = images.shape
    out = np.zeros((h, w, cn), np.uint8)
    assert lib().orc_linear_blend32s(n, h, w, cn, _p(images), _p(weights), _p(out))
    return out


def bgr32_to_yuv420p(bgrx: np.ndarray):
    """cvtBGR32ToYUV420P (ImageProc.cu:44-141): (h, w, 4) u8 -> y, u, v planes."""
    bgrx = np.ascontiguousarray(bgrx, np.uint8)
    h, w = bgrx.shape[:2]
    y = np.zeros((h, w), np.uint8)
    u = np.zeros((h // 2, w // 2), np.uint8)
    v = np.zeros((h // 2, w // 2), np.uint8)
    assert lib().orc_bgr32_to_yuv(_p(bgrx), w * 4, w, h, 0, _p(y), _p(u), _p(v))
    return y, u, v


def bgr32_to_nv12(bgrx: np.ndarray):
    """cvtBGR32ToNV12 (ImageProc.cu:123-159): (h, w, 4) u8 -> y, interleaved uv."""
    bgrx = np.ascontiguousarray(bgrx, np.uint8)
    h, w = bgrx.shape[:2]
    y = np.zeros((h, w), np.uint8)
    uv = np.zeros((h // 2, w), np.uint8)
    assert lib().orc_bgr32_to_yuv(_p(bgrx), w * 4, w, h, 1, _p(y), _p(uv), None)
    return y, uv


def yuv420p_to_bgr32(y, u, v):
    """cvtYUV420PToBGR32 (ImageProc.cu:197-242); channel 3 of the output is 0."""
    y, u, v = (np.ascontiguousarray(a, np.uint8) for a in (y, u, v))
    h, w = y.shape
    out = np.zeros((h, w, 4), np.uint8)
    assert lib().orc_yuv_to_bgr32(_p(y), _p(u), _p(v), w, h, 0, _p(out), w * 4)
    return out


def nv12_to_bgr32(y, uv):
    """cvtNV12ToBGR32 (ImageProc.cu:211-258); channel 3 of the output is 0."""
    y, uv = (np.ascontiguousarray(a, np.uint8) for a in (y, uv))
    h, w = y.shape
    out = np.zeros((h, w, 4), np.uint8)
    assert lib().orc_yuv_to_bgr32(_p(y), _p(uv), None, w, h, 1, _p(out), w * 4)
    return out


class PhotoParam(C.Structure):
    """PhotoParam (Warp/ZReproject.h:17-50), same field order."""
    _fields_ = [(n, C.c_int) for n in ("imageType", "cropMode", "cropX", "cropY", "cropWidth",
                                        "cropHeight", "circleX", "circleY", "circleR")] + \
               [(n, C.c_double) for n in ("hfov", "vfov", "alpha", "beta", "gamma", "shiftX",
                                           "shiftY", "shearX", "shearY", "yaw", "pitch", "roll")]


def photo_param(**kw) -> PhotoParam:
    p = PhotoParam()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def remap_point(param: PhotoParam, src_w, src_h, dst_w, dst_h, x, y):
    sx, sy = C.c_double(), C.c_double()
    lib().orc_remap_point(C.byref(param), src_w, src_h, dst_w, dst_h, float(x), float(y),
                          C.byref(sx), C.byref(sy))
    return sx.value, sy.value


def reproject_map_and_mask(param: PhotoParam, src_w, src_h, dst_w, dst_h, nthreads=1):
    m = np.zeros((dst_h, dst_w, 2), np.float64)
    mask = np.zeros((dst_h, dst_w), np.uint8)
    lib().orc_reproject_map_and_mask(C.byref(param), src_w, src_h, dst_w, dst_h, _p(m), _p(mask),
                                     nthreads)
    return m, mask
