"""panostitcher_amd -- MI355X (gfx950) per-frame panorama stitch path.

Python mirror of the reference's CudaAccel interface for this path, over the
C-ABI library ``libpanohip.so`` (include/panohip.h).  Names and argument order
follow CudaAccel/CudaInterface.h and Task/RicohUtil.h:

  * ``CudaTilingMultibandBlendFast``: prepare(masks, max_levels, min_length),
    blend(images, out), begin()/tile(image, index)/end(out), get_unique_masks(),
    calc_used_memory_size(), estimate_memory_size(w, h, n);
  * ``cuda_reproject_to16s(src, xmap, ymap, stream)``  (bilinear, CPU arithmetic);
  * ``cuda_transform(src, luts, stream)``;
  * ``CudaPanoramaRender2``-style ``PanoramaRender``: prepare(params | maps+masks)
    and render(srcs, luts) -> BGRx canvas.

Device memory is torch tensors on ``cuda`` (ROCm); streams are torch streams.
Errors: prepare() returns False like the reference; a render/blend on an
unprepared object returns False; invalid arguments raise ValueError (the
reference's CV_Assert throws).  The native library is required: importing the
wrappers without it raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpanohip.so")

PANO_OK = 0
PANO_ERR_BUDGET = -6
PANO_ERR_IO = -7
PANO_ERR_FORMAT = -8
PANO_8UC4 = 24
PANO_16SC4 = 27
PANO_32FC4 = 29
PANO_OUT_BGRX, PANO_OUT_NV12, PANO_OUT_YUV420P = 0, 1, 2

_lib = None


def build(arch: str = "gfx950", jobs: int = 8) -> str:
    """Compile libpanohip.so in-tree with hipcc (cross-compiles without a GPU), and
    the C++ API's test program (tests/cpp/facade_test) over it."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, f"-j{jobs}", f"ARCH={arch}", "all"], check=True)
    return LIB_PATH


class PanoError(RuntimeError):
    pass


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PanoError(f"{LIB_PATH} is missing: run panostitcher_amd.build() first")
        L = C.CDLL(LIB_PATH)
        vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
        i64 = C.c_int64
        pp = C.POINTER(C.c_void_p)
        psz = C.POINTER(C.c_size_t)
        sigs = {
            "pano_status_string": ([i], C.c_char_p),
            "pano_version": ([], i),
            "pano_blender_create": ([pp], i),
            "pano_blender_destroy": ([vp], None),
            "pano_blender_prepare": ([vp, i, i, i, pp, psz, i, i], i),
            "pano_blender_prepare_device": ([vp, i, i, i, pp, psz, i, i], i),
            "pano_blender_blend": ([vp, i, pp, psz, i, vp, sz, vp], i),
            "pano_blender_blend_custom": ([vp, i, pp, psz, i, pp, psz, vp, sz, vp], i),
            "pano_blender_begin": ([vp], i),
            "pano_blender_tile": ([vp, vp, sz, i, i, vp], i),
            "pano_blender_end": ([vp, vp, sz, vp], i),
            "pano_blender_unique_masks": ([vp, pp, psz], i),
            "pano_blender_num_levels": ([vp], i),
            "pano_blender_mem_bytes": ([vp], i64),
            "pano_reproject": ([vp, sz, i, i, vp, vp, sz, vp, sz, i, i, vp], i),
            "pano_linear_weights": ([i, i, i, pp, psz, i, i, pp, psz], i),
            "pano_nonintersecting_masks": ([i, i, i, pp, psz, pp, psz], i),
            "pano_render_prepare_ms": ([vp, vp], i),
            "pano_render_set_prepare_on_device": ([vp, i], i),
            "pano_render_set_tile_shape": ([vp, i], i),
            "pano_render_tile_shape": ([vp], i),
            "pano_render_autotune_tile_shape": ([vp, i, vp, vp, vp, i, i, i, pp, psz, vp, sz, i, vp, vp, vp], i),
            "pano_generate_map_device": ([vp, i, i, i, i, vp, vp, sz, vp, sz, vp], i),
            "pano_reproject_weighted_accumulate": ([vp, sz, i, i, vp, vp, sz, vp, sz, vp, sz, i, i, vp], i),
            "pano_reproject_params": ([vp, sz, i, i, vp, i, i, i, vp, sz, vp, sz, vp], i),
            "pano_linear_blender_create": ([pp], i),
            "pano_linear_blender_destroy": ([vp], None),
            "pano_linear_blender_prepare": ([vp, i, i, i, pp, psz, i], i),
            "pano_linear_blender_blend": ([vp, i, pp, psz, vp, sz, vp], i),
            "pano_linear_render_create": ([pp], i),
            "pano_linear_render_destroy": ([vp], None),
            "pano_linear_render_prepare_maps": ([vp, i, vp, vp, i, i, pp, pp, pp, i], i),
            "pano_linear_render_prepare": ([vp, i, vp, vp, vp, i, i, i], i),
            "pano_linear_render_render": ([vp, i, pp, psz, vp, vp, sz, vp], i),
            "pano_linear_render_weights": ([vp, i, vp], i),
            "pano_blender32f_create": ([pp], i),
            "pano_blender32f_destroy": ([vp], None),
            "pano_blender32f_prepare": ([vp, i, i, i, pp, psz, i, i], i),
            "pano_blender32f_blend": ([vp, i, pp, psz, i, vp, sz, vp], i),
            "pano_blender32f_num_levels": ([vp], i),
            "pano_estimate_mem_bytes": ([i, i, i], i64),
            "pano_reproject_to16s": ([vp, sz, i, i, vp, vp, sz, vp, sz, i, i, vp], i),
            "pano_transform_lut3": ([vp, sz, vp, sz, i, i, vp, vp], i),
            "pano_get_lut": ([vp, C.c_double], i),
            "pano_exposure_luts": ([i, vp, vp, vp, vp], i),
            "pano_generate_map": ([vp, i, i, i, i, vp, vp, vp], i),
            "pano_load_photo_params": ([C.c_char_p, vp, i], i),
            "pano_load_photo_params_xml": ([C.c_char_p, vp, i], i),
            "pano_load_photo_params_pts": ([C.c_char_p, vp, i], i),
            "pano_export_photo_params_xml": ([C.c_char_p, vp, i], i),
            "pano_rotate_cameras": ([vp, i, C.c_double, C.c_double, C.c_double], i),
            "pano_rotate_photo_params_xml": ([C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double], i),
            "pano_render_create": ([pp], i),
            "pano_render_destroy": ([vp], None),
            "pano_render_prepare": ([vp, i, vp, vp, vp, i, i, i], i),
            "pano_render_prepare_maps": ([vp, i, vp, vp, i, i, pp, pp, pp, i], i),
            "pano_render_render": ([vp, i, pp, psz, vp, vp, sz, vp], i),
            "pano_render_render_to": ([vp, i, pp, psz, vp, vp, vp], i),
            "pano_batch_render_to": ([vp, pp, psz, pp, vp, vp], i),
            "pano_batch_capture_to": ([vp, pp, psz, pp, vp, vp], i),
            "pano_bgr32_to_yuv420p": ([vp, sz, i, i, vp, sz, vp, sz, vp, sz, vp], i),
            "pano_bgr32_to_nv12": ([vp, sz, i, i, vp, sz, vp, sz, vp], i),
            "pano_yuv420p_to_bgr32": ([vp, sz, vp, sz, vp, sz, i, i, vp, sz, vp], i),
            "pano_nv12_to_bgr32": ([vp, sz, vp, sz, i, i, vp, sz, vp], i),
            "pano_render_num_images": ([vp], i),
            "pano_render_num_levels": ([vp], i),
            "pano_render_mem_bytes": ([vp], i64),
            "pano_render_model_bytes": ([vp, i], i64),
            "pano_render_set_timing": ([vp, i], i),
            "pano_render_stage_ms": ([vp, vp, C.POINTER(C.c_int)], i),
            "pano_render_fused": ([vp], i),
            "pano_render_render_autogain": ([vp, i, pp, psz, i, vp, sz, vp], i),
            "pano_render_gains": ([vp, vp, vp, vp, C.POINTER(C.c_int)], i),
            "pano_render_gain_sums": ([vp, vp], i),
            "pano_server_create": ([pp, i64, i], i),
            "pano_server_destroy": ([vp], None),
            "pano_server_admit": ([vp, vp, C.POINTER(C.c_int)], i),
            "pano_server_retire": ([vp, i], i),
            "pano_server_submit": ([vp, i, pp, psz, vp, vp, sz, C.c_uint64], i),
            "pano_server_poll": ([vp, vp, i, i], i),
            "pano_server_drain": ([vp], i),
            "pano_server_stats": ([vp, vp, vp, vp], i),
            "pano_batch_create": ([pp], i),
            "pano_batch_destroy": ([vp], None),
            "pano_batch_set": ([vp, i, pp], i),
            "pano_batch_render": ([vp, pp, psz, pp, pp, psz, vp], i),
            "pano_batch_set_timing": ([vp, i], i),
            "pano_batch_stage_ms": ([vp, vp, C.POINTER(C.c_int)], i),
            "pano_batch_capture": ([vp, pp, psz, pp, pp, psz, vp], i),
            "pano_batch_replay": ([vp, pp, vp], i),
            "pano_batch_set_sched": ([vp, vp], i),
            "pano_batch_get_sched": ([vp, vp], i),
            "pano_batch_autotune": ([vp, pp, psz, pp, pp, psz, vp, i, vp], i),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


PANO_GAIN_TINT = 1


def exported_symbols():
    return [
        "pano_status_string", "pano_version", "pano_blender_create", "pano_blender_destroy",
        "pano_blender_prepare", "pano_blender_prepare_device", "pano_blender_blend",
        "pano_blender_begin", "pano_blender_tile", "pano_blender_end",
        "pano_blender_unique_masks", "pano_blender_num_levels", "pano_blender_mem_bytes",
        "pano_estimate_mem_bytes", "pano_reproject_to16s", "pano_transform_lut3", "pano_get_lut",
        "pano_exposure_luts", "pano_generate_map", "pano_render_create", "pano_render_destroy",
        "pano_render_prepare", "pano_render_prepare_maps", "pano_render_render",
        "pano_render_num_images", "pano_render_num_levels", "pano_render_mem_bytes",
        "pano_render_model_bytes", "pano_render_set_timing", "pano_render_stage_ms",
        "pano_render_fused", "pano_render_render_autogain", "pano_render_gains",
        "pano_render_gain_sums", "pano_batch_create", "pano_batch_destroy", "pano_batch_set",
        "pano_batch_render", "pano_batch_set_timing", "pano_batch_stage_ms",
        "pano_batch_capture", "pano_batch_replay", "pano_batch_set_sched", "pano_batch_get_sched",
        "pano_batch_autotune",
        "pano_server_create", "pano_server_destroy", "pano_server_admit", "pano_server_retire",
        "pano_server_submit", "pano_server_poll", "pano_server_drain", "pano_server_stats",
        "pano_blender32f_create", "pano_blender32f_destroy", "pano_blender32f_prepare",
        "pano_blender32f_blend", "pano_blender32f_num_levels", "pano_reproject",
        "pano_blender_blend_custom", "pano_render_render_to", "pano_batch_render_to",
        "pano_batch_capture_to", "pano_bgr32_to_yuv420p", "pano_bgr32_to_nv12",
        "pano_yuv420p_to_bgr32", "pano_nv12_to_bgr32", "pano_linear_weights",
        "pano_linear_blender_create", "pano_linear_blender_destroy", "pano_linear_blender_prepare",
        "pano_linear_blender_blend", "pano_linear_render_create", "pano_linear_render_destroy",
        "pano_linear_render_prepare_maps", "pano_linear_render_render", "pano_linear_render_weights",
        "pano_nonintersecting_masks", "pano_render_prepare_ms", "pano_generate_map_device",
        "pano_reproject_weighted_accumulate", "pano_reproject_params",
        "pano_load_photo_params", "pano_load_photo_params_xml", "pano_load_photo_params_pts",
        "pano_export_photo_params_xml", "pano_rotate_cameras", "pano_rotate_photo_params_xml",
        "pano_render_set_prepare_on_device", "pano_linear_render_prepare", "pano_render_set_tile_shape",
        "pano_render_tile_shape", "pano_render_autotune_tile_shape",
    ]


def status_string(st: int) -> str:
    return lib().pano_status_string(st).decode()


def _check(st: int, what: str):
    if st != PANO_OK:
        raise PanoError(f"{what}: {status_string(st)} ({st})")


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[C.c_void_p(a) for a in arrs])


def _sizes(vals):
    return (C.c_size_t * len(vals))(*vals)


def _stream(stream):
    if stream is None:
        import torch

        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class Output(C.Structure):
    """pano_output: a render's destination.  Output.bgrx(canvas), Output.nv12(y, uv)
    or Output.yuv420p(y, u, v) over uint8 cuda tensors (rows may be padded)."""
    _fields_ = [("format", C.c_int), ("plane", C.c_void_p * 3), ("step", C.c_size_t * 3)]

    @classmethod
    def _make(cls, fmt, planes):
        o = cls()
        o.format = fmt
        for k, t in enumerate(planes):
            p, st = _dev_ptr_step(t)
            o.plane[k] = p
            o.step[k] = st
        o._tensors = list(planes)
        return o

    @classmethod
    def bgrx(cls, canvas):
        return cls._make(PANO_OUT_BGRX, [canvas])

    @classmethod
    def nv12(cls, y, uv):
        return cls._make(PANO_OUT_NV12, [y, uv])

    @classmethod
    def yuv420p(cls, y, u, v):
        return cls._make(PANO_OUT_YUV420P, [y, u, v])


def _dev_ptr_step(t):
    """(pointer, row step in bytes) of a torch tensor HxWxC (rows may be padded)."""
    assert t.is_cuda, "device tensor expected"
    return t.data_ptr(), t.stride(0) * t.element_size()


# ---------------------------------------------------------------- photo params --

class PhotoParam(C.Structure):
    """PhotoParam (Warp/ZReproject.h:17-50) as pano_photo_param."""
    _fields_ = [(n, C.c_int) for n in ("imageType", "cropMode", "cropX", "cropY", "cropWidth",
                                        "cropHeight", "circleX", "circleY", "circleR")] + \
               [(n, C.c_double) for n in ("hfov", "vfov", "alpha", "beta", "gamma", "shiftX",
                                           "shiftY", "shearX", "shearY", "yaw", "pitch", "roll")]

    @classmethod
    def make(cls, **kw):
        p = cls()
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# ---------------------------------------------------------------- rig files --

def _load_params(fn, path, what):
    f = getattr(lib(), fn)
    p = os.fsencode(path)
    n = f(p, None, 0)
    if n < 0:
        _check(n, what)
    if n == 0:
        return []
    arr = (PhotoParam * n)()
    m = f(p, arr, n)
    if m < 0:
        _check(m, what)
    return list(arr[:min(m, n)])


def load_photo_params(path) -> list:
    """loadPhotoParams (Warp/PhotoParam.cpp:756-770): ".pts" files as PTGui
    projects, others as XML; an empty list where the reference returns false.
    A path without an extension raises, as the reference reports an error."""
    return _load_params("pano_load_photo_params", path, "loadPhotoParams")


def load_photo_param_from_xml(path) -> list:
    """loadPhotoParamFromXML (PhotoParam.cpp:506-511); a non-numeric field raises."""
    return _load_params("pano_load_photo_params_xml", path, "loadPhotoParamFromXML")


def load_photo_param_from_pts(path) -> list:
    """loadPhotoParamFromPTS (PhotoParam.cpp:681-754)."""
    return _load_params("pano_load_photo_params_pts", path, "loadPhotoParamFromPTS")


def export_photo_param_to_xml(path, params) -> None:
    """exportPhotoParamToXML (PhotoParam.cpp:772-925)."""
    arr = (PhotoParam * len(params))(*params)
    _check(lib().pano_export_photo_params_xml(os.fsencode(path), arr, len(params)), "exportPhotoParamToXML")


def rotate_cameras(params, yaw: float, pitch: float, roll: float) -> None:
    """rotateCameras (PhotoParam.cpp:962-967), in place; the rotation in radians."""
    arr = (PhotoParam * len(params))(*params)
    _check(lib().pano_rotate_cameras(arr, len(params), yaw, pitch, roll), "rotateCameras")
    for k, q in enumerate(arr):
        C.memmove(C.addressof(params[k]), C.addressof(q), C.sizeof(PhotoParam))


def rotate_camera(param: PhotoParam, yaw: float, pitch: float, roll: float) -> None:
    """rotateCamera (PhotoParam.cpp:931-960), in place; the rotation in radians."""
    _check(lib().pano_rotate_cameras(C.byref(param), 1, yaw, pitch, roll), "rotateCamera")


def rotate_photo_param_in_xml(src, dst, yaw: float, pitch: float, roll: float) -> None:
    """rotatePhotoParamInXML (PhotoParam.cpp:969-975)."""
    _check(lib().pano_rotate_photo_params_xml(os.fsencode(src), os.fsencode(dst), yaw, pitch, roll),
           "rotatePhotoParamInXML")


def get_lut(k: float) -> np.ndarray:
    out = np.zeros(256, np.uint8)
    _check(lib().pano_get_lut(out.ctypes.data, float(k)), "getLUT")
    return out


def exposure_luts(exposures, red_ratios, blue_ratios) -> np.ndarray:
    e = np.ascontiguousarray(exposures, np.float64)
    r = np.ascontiguousarray(red_ratios, np.float64)
    b = np.ascontiguousarray(blue_ratios, np.float64)
    out = np.zeros((len(e), 3, 256), np.uint8)
    _check(lib().pano_exposure_luts(len(e), e.ctypes.data, r.ctypes.data, b.ctypes.data,
                                    out.ctypes.data), "getExposureColorOptimizeLUTs")
    return out


def generate_map_device(param: PhotoParam, src_w: int, src_h: int, dst_w: int, dst_h: int, stream=None):
    """cudaGenerateReprojectMapAndMask: (xmap, ymap, mask) cuda tensors of dst_h x dst_w."""
    import torch

    xm = torch.empty((dst_h, dst_w), dtype=torch.float32, device="cuda")
    ym = torch.empty_like(xm)
    mk = torch.empty((dst_h, dst_w), dtype=torch.uint8, device="cuda")
    _check(lib().pano_generate_map_device(C.byref(param), src_w, src_h, dst_w, dst_h,
                                          C.c_void_p(xm.data_ptr()), C.c_void_p(ym.data_ptr()), dst_w * 4,
                                          C.c_void_p(mk.data_ptr()), dst_w, _stream(stream)),
           "cudaGenerateReprojectMapAndMask")
    return xm, ym, mk


def generate_map(param: PhotoParam, src_w: int, src_h: int, dst_w: int, dst_h: int):
    """getReprojectMap32FAndMask on the host: (xmap, ymap, mask)."""
    xm = np.zeros((dst_h, dst_w), np.float32)
    ym = np.zeros((dst_h, dst_w), np.float32)
    mk = np.zeros((dst_h, dst_w), np.uint8)
    _check(lib().pano_generate_map(C.byref(param), src_w, src_h, dst_w, dst_h, xm.ctypes.data,
                                   ym.ctypes.data, mk.ctypes.data), "generate map")
    return xm, ym, mk


# ------------------------------------------------------------------- blender --

class CudaTilingMultibandBlendFast:
    """CudaTilingMultibandBlendFast (CudaAccel/CudaInterface.h:45-82)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pano_blender_create(C.byref(h)), "create blender")
        self._h = h
        self.num_images = 0
        self.rows = self.cols = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_blender_destroy(h)
            self._h = None

    def prepare(self, masks, max_levels: int, min_length: int = 2) -> bool:
        """masks: list of HxW uint8 host arrays (or torch cuda tensors)."""
        if len(masks) == 0:
            return False
        dev = hasattr(masks[0], "is_cuda") and masks[0].is_cuda
        rows, cols = masks[0].shape[:2]
        if dev:
            ptrs = [m.data_ptr() for m in masks]
            steps = [m.stride(0) for m in masks]
            st = lib().pano_blender_prepare_device(self._h, len(masks), rows, cols, _ptrs(ptrs),
                                                   _sizes(steps), max_levels, min_length)
        else:
            arrs = [np.ascontiguousarray(m, np.uint8) for m in masks]
            for a in arrs:
                if a.shape != (rows, cols):
                    return False
            st = lib().pano_blender_prepare(self._h, len(arrs), rows, cols,
                                            _ptrs([a.ctypes.data for a in arrs]),
                                            _sizes([a.strides[0] for a in arrs]), max_levels,
                                            min_length)
        if st == PANO_OK:
            self.num_images, self.rows, self.cols = len(masks), rows, cols
        return st == PANO_OK

    @property
    def num_levels(self):
        return lib().pano_blender_num_levels(self._h)

    def _type(self, t):
        import torch

        if t.dtype == torch.int16 and t.shape[-1] == 4:
            return PANO_16SC4, 8
        if t.dtype == torch.uint8 and t.shape[-1] == 4:
            return PANO_8UC4, 4
        raise ValueError("images must be HxWx4 int16 (CV_16SC4) or uint8 (CV_8UC4)")

    def blend(self, images, out, stream=None) -> bool:
        """images: list of HxWx4 cuda tensors; out: HxWx4 uint8 cuda tensor."""
        typ, _ = self._type(images[0])
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in images])
        op, ostep = _dev_ptr_step(out)
        st = lib().pano_blender_blend(self._h, len(images), _ptrs(ptrs), _sizes(steps), typ,
                                      C.c_void_p(op), ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "blend")
        return True

    def blend_with_masks(self, images, masks, out, stream=None) -> bool:
        """blend(images, masks, blendImage) (CudaInterface.h:55-56): weights from
        these masks (HxW uint8 cuda tensors) instead of the prepared seams."""
        typ, _ = self._type(images[0])
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in images])
        mptrs = [m.data_ptr() for m in masks]
        msteps = [m.stride(0) for m in masks]
        op, ostep = _dev_ptr_step(out)
        st = lib().pano_blender_blend_custom(self._h, len(images), _ptrs(ptrs), _sizes(steps), typ,
                                             _ptrs(mptrs), _sizes(msteps), C.c_void_p(op), ostep,
                                             _stream(stream))
        if st == -2:
            return False
        _check(st, "blend with masks")
        return True

    def begin(self):
        return lib().pano_blender_begin(self._h) == PANO_OK

    def tile(self, image, index: int, stream=None):
        typ, _ = self._type(image)
        p, s = _dev_ptr_step(image)
        _check(lib().pano_blender_tile(self._h, C.c_void_p(p), s, typ, index, _stream(stream)),
               "tile")

    def end(self, out, stream=None):
        p, s = _dev_ptr_step(out)
        _check(lib().pano_blender_end(self._h, C.c_void_p(p), s, _stream(stream)), "end")

    def get_unique_masks(self):
        outs = [np.zeros((self.rows, self.cols), np.uint8) for _ in range(self.num_images)]
        _check(lib().pano_blender_unique_masks(self._h, _ptrs([o.ctypes.data for o in outs]),
                                               _sizes([self.cols] * self.num_images)),
               "getUniqueMasks")
        return outs

    def calc_used_memory_size(self) -> int:
        return lib().pano_blender_mem_bytes(self._h)

    @staticmethod
    def estimate_memory_size(width: int, height: int, num_images: int) -> int:
        return lib().pano_estimate_mem_bytes(width, height, num_images)


def nonintersecting_masks(masks) -> np.ndarray:
    """getNonIntersectingMasks (Blend/Mask.cpp:376-397) on the device: host
    (n, h, w) uint8 masks -> host (n, h, w) unique masks."""
    masks = np.ascontiguousarray(masks, np.uint8)
    n, h, w = masks.shape
    out = np.zeros_like(masks)
    _check(lib().pano_nonintersecting_masks(n, h, w, _ptrs([m.ctypes.data for m in masks]), _sizes([w] * n),
                                            _ptrs([o.ctypes.data for o in out]), _sizes([w] * n)),
           "getNonIntersectingMasks")
    return out


def linear_weights(masks, radius: int, as_int: bool = False):
    """getWeightsLinearBlend32F / getWeightsLinearBlend on the device: masks host
    (n, h, w) uint8 -> (n, h, w) float32 (int32 in 2^16 units) cuda tensor."""
    import torch

    masks = np.ascontiguousarray(masks, np.uint8)
    n, h, w = masks.shape
    out = torch.empty((n, h, w), dtype=torch.int32 if as_int else torch.float32, device="cuda")
    _check(lib().pano_linear_weights(n, h, w, _ptrs([m.ctypes.data for m in masks]), _sizes([w] * n),
                                     int(radius), int(as_int), _ptrs([out[i].data_ptr() for i in range(n)]),
                                     _sizes([w * 4] * n)), "linear weights")
    return out


class CudaTilingLinearBlend:
    """CudaTilingLinearBlend (CudaAccel/CudaInterface.h:103-115)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pano_linear_blender_create(C.byref(h)), "create linear blender")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_linear_blender_destroy(h)
            self._h = None

    def prepare(self, masks, radius: int) -> bool:
        if len(masks) == 0:
            return False
        arrs = [np.ascontiguousarray(m, np.uint8) for m in masks]
        h, w = arrs[0].shape
        if any(a.shape != (h, w) for a in arrs):
            return False
        return lib().pano_linear_blender_prepare(self._h, len(arrs), h, w, _ptrs([a.ctypes.data for a in arrs]),
                                                 _sizes([w] * len(arrs)), int(radius)) == PANO_OK

    def blend(self, images, out, stream=None) -> bool:
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in images])
        op, ostep = _dev_ptr_step(out)
        st = lib().pano_linear_blender_blend(self._h, len(images), _ptrs(ptrs), _sizes(steps), C.c_void_p(op),
                                             ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "linear blend")
        return True


class LinearRender:
    """The render's linear-blend mode (highQualityBlend = 0, blendParam = radius)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pano_linear_render_create(C.byref(h)), "create linear render")
        self._h = h
        self.n = 0
        self.dst = (0, 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_linear_render_destroy(h)
            self._h = None

    def prepare_maps(self, src_sizes, dst_size, xmaps, ymaps, masks, radius: int) -> bool:
        n = len(src_sizes)
        dw, dh = dst_size
        xs = [np.ascontiguousarray(x, np.float32) for x in xmaps]
        ys = [np.ascontiguousarray(y, np.float32) for y in ymaps]
        ms = [np.ascontiguousarray(m, np.uint8) for m in masks]
        sw = (C.c_int * n)(*[s[0] for s in src_sizes])
        sh = (C.c_int * n)(*[s[1] for s in src_sizes])
        st = lib().pano_linear_render_prepare_maps(self._h, n, sw, sh, dw, dh,
                                                   _ptrs([a.ctypes.data for a in xs]),
                                                   _ptrs([a.ctypes.data for a in ys]),
                                                   _ptrs([a.ctypes.data for a in ms]), int(radius))
        if st == PANO_OK:
            self.n, self.dst = n, (dw, dh)
        return st == PANO_OK

    def prepare(self, params, src_sizes, dst_size, radius: int) -> bool:
        """From camera parameters (PhotoParams or a rig file path): maps and masks
        generated and prepared on the device.  src_sizes: one (w, h) per camera,
        or one for all."""
        if isinstance(params, (str, bytes, os.PathLike)):
            params = load_photo_params(params)
            if not params:
                return False
        n = len(params)
        if len(src_sizes) == 2 and all(isinstance(v, (int, np.integer)) for v in src_sizes):
            src_sizes = [tuple(src_sizes)] * n
        dw, dh = dst_size
        arr = (PhotoParam * n)(*params)
        sw = (C.c_int * n)(*[s[0] for s in src_sizes])
        sh = (C.c_int * n)(*[s[1] for s in src_sizes])
        st = lib().pano_linear_render_prepare(self._h, n, C.cast(arr, C.c_void_p), sw, sh, dw, dh, int(radius))
        if st == PANO_OK:
            self.n, self.dst = n, (dw, dh)
        return st == PANO_OK

    def render(self, srcs, out, luts=None, stream=None) -> bool:
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        op, ostep = _dev_ptr_step(out)
        lp = None
        if luts is not None:
            luts = np.ascontiguousarray(luts, np.uint8)
            assert luts.shape == (len(srcs), 3, 256)
            lp = C.c_void_p(luts.ctypes.data)
        st = lib().pano_linear_render_render(self._h, len(srcs), _ptrs(ptrs), _sizes(steps), lp, C.c_void_p(op),
                                             ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "linear render")
        return True

    def weights(self, cam: int) -> np.ndarray:
        dw, dh = self.dst
        out = np.zeros((dh, dw), np.float32)
        _check(lib().pano_linear_render_weights(self._h, cam, C.c_void_p(out.ctypes.data)), "weights")
        return out


class CudaTilingMultibandBlendFast32F:
    """CudaTilingMultibandBlendFast32F (CudaAccel/CudaInterface.h:84-101): the
    float-weight blender; blend() takes CV_8UC4 or CV_32FC4 images."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pano_blender32f_create(C.byref(h)), "create blender32f")
        self._h = h
        self.num_images = 0
        self.rows = self.cols = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_blender32f_destroy(h)
            self._h = None

    def prepare(self, masks, max_levels: int, min_length: int = 2) -> bool:
        """masks: list of HxW uint8 host arrays."""
        if len(masks) == 0:
            return False
        rows, cols = masks[0].shape[:2]
        arrs = [np.ascontiguousarray(m, np.uint8) for m in masks]
        if any(a.shape != (rows, cols) for a in arrs):
            return False
        st = lib().pano_blender32f_prepare(self._h, len(arrs), rows, cols,
                                           _ptrs([a.ctypes.data for a in arrs]),
                                           _sizes([a.strides[0] for a in arrs]), max_levels,
                                           min_length)
        if st == PANO_OK:
            self.num_images, self.rows, self.cols = len(masks), rows, cols
        return st == PANO_OK

    @property
    def num_levels(self):
        return lib().pano_blender32f_num_levels(self._h)

    def blend(self, images, out, stream=None) -> bool:
        """images: list of HxWx4 uint8 or float32 cuda tensors; out: HxWx4 uint8."""
        import torch

        t = images[0]
        if t.shape[-1] != 4 or t.dtype not in (torch.uint8, torch.float32):
            raise ValueError("images must be HxWx4 uint8 (CV_8UC4) or float32 (CV_32FC4)")
        typ = PANO_8UC4 if t.dtype == torch.uint8 else PANO_32FC4
        ptrs, steps = zip(*[_dev_ptr_step(x) for x in images])
        op, ostep = _dev_ptr_step(out)
        st = lib().pano_blender32f_blend(self._h, len(images), _ptrs(ptrs), _sizes(steps), typ,
                                         C.c_void_p(op), ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "blend")
        return True


# ------------------------------------------------------------- free functions --

def cuda_reproject_to16s(src, xmap, ymap, dst=None, stream=None):
    """cudaReprojectTo16S(src 8UC4, dst 16SC4, xmap, ymap) with bilinear sampling."""
    import torch

    sh, sw = src.shape[:2]
    dh, dw = xmap.shape
    if dst is None:
        dst = torch.empty((dh, dw, 4), dtype=torch.int16, device=src.device)
    sp, ss = _dev_ptr_step(src)
    dp, ds = _dev_ptr_step(dst)
    assert xmap.stride(0) == ymap.stride(0)
    _check(lib().pano_reproject_to16s(C.c_void_p(sp), ss, sw, sh, C.c_void_p(xmap.data_ptr()),
                                      C.c_void_p(ymap.data_ptr()), xmap.stride(0) * 4,
                                      C.c_void_p(dp), ds, dw, dh, _stream(stream)),
           "cudaReprojectTo16S")
    return dst


def cuda_reproject(src, xmap, ymap, dst=None, stream=None):
    """cudaReproject(src 8UC4, dst 8UC4, xmap, ymap) with bilinear sampling."""
    import torch

    sh, sw = src.shape[:2]
    dh, dw = xmap.shape
    if dst is None:
        dst = torch.empty((dh, dw, 4), dtype=torch.uint8, device=src.device)
    sp, ss = _dev_ptr_step(src)
    dp, ds = _dev_ptr_step(dst)
    assert xmap.stride(0) == ymap.stride(0)
    _check(lib().pano_reproject(C.c_void_p(sp), ss, sw, sh, C.c_void_p(xmap.data_ptr()),
                                C.c_void_p(ymap.data_ptr()), xmap.stride(0) * 4, C.c_void_p(dp), ds,
                                dw, dh, _stream(stream)),
           "cudaReproject")
    return dst


def _planes_yuv(h, w, nv12, device):
    import torch

    y = torch.empty((h, w), dtype=torch.uint8, device=device)
    if nv12:
        return y, torch.empty((h // 2, w), dtype=torch.uint8, device=device)
    return (y, torch.empty((h // 2, w // 2), dtype=torch.uint8, device=device),
            torch.empty((h // 2, w // 2), dtype=torch.uint8, device=device))


def cvt_bgr32_to_nv12(src, y=None, uv=None, stream=None):
    """cvtBGR32ToNV12(bgr32, y, uv) (CudaInterface.h:164-165)."""
    h, w = src.shape[:2]
    if y is None or uv is None:
        y, uv = _planes_yuv(h, w, True, src.device)
    sp, ss = _dev_ptr_step(src)
    _check(lib().pano_bgr32_to_nv12(C.c_void_p(sp), ss, w, h, C.c_void_p(y.data_ptr()), y.stride(0),
                                    C.c_void_p(uv.data_ptr()), uv.stride(0), _stream(stream)),
           "cvtBGR32ToNV12")
    return y, uv


def cvt_bgr32_to_yuv420p(src, y=None, u=None, v=None, stream=None):
    """cvtBGR32ToYUV420P(bgr32, y, u, v) (CudaInterface.h:161-162)."""
    h, w = src.shape[:2]
    if y is None or u is None or v is None:
        y, u, v = _planes_yuv(h, w, False, src.device)
    sp, ss = _dev_ptr_step(src)
    _check(lib().pano_bgr32_to_yuv420p(C.c_void_p(sp), ss, w, h, C.c_void_p(y.data_ptr()), y.stride(0),
                                       C.c_void_p(u.data_ptr()), u.stride(0), C.c_void_p(v.data_ptr()),
                                       v.stride(0), _stream(stream)), "cvtBGR32ToYUV420P")
    return y, u, v


def cvt_nv12_to_bgr32(y, uv, dst=None, stream=None):
    """cvtNV12ToBGR32(y, uv, bgr32) (CudaInterface.h:170-171); channel 3 = 0."""
    import torch

    h, w = y.shape
    if dst is None:
        dst = torch.empty((h, w, 4), dtype=torch.uint8, device=y.device)
    dp, ds = _dev_ptr_step(dst)
    _check(lib().pano_nv12_to_bgr32(C.c_void_p(y.data_ptr()), y.stride(0), C.c_void_p(uv.data_ptr()),
                                    uv.stride(0), w, h, C.c_void_p(dp), ds, _stream(stream)),
           "cvtNV12ToBGR32")
    return dst


def cvt_yuv420p_to_bgr32(y, u, v, dst=None, stream=None):
    """cvtYUV420PToBGR32(y, u, v, bgr32) (CudaInterface.h:167-168); channel 3 = 0."""
    import torch

    h, w = y.shape
    if dst is None:
        dst = torch.empty((h, w, 4), dtype=torch.uint8, device=y.device)
    dp, ds = _dev_ptr_step(dst)
    _check(lib().pano_yuv420p_to_bgr32(C.c_void_p(y.data_ptr()), y.stride(0), C.c_void_p(u.data_ptr()),
                                       u.stride(0), C.c_void_p(v.data_ptr()), v.stride(0), w, h,
                                       C.c_void_p(dp), ds, _stream(stream)), "cvtYUV420PToBGR32")
    return dst


def cuda_reproject_weighted_accumulate(src, accum, xmap, ymap, weight, stream=None):
    """cudaReprojectWeightedAccumulateTo32F: accum (HxWx4 float32 cuda) += sample x weight."""
    sh, sw = src.shape[:2]
    dh, dw = xmap.shape
    sp, ss = _dev_ptr_step(src)
    ap, ast = _dev_ptr_step(accum)
    _check(lib().pano_reproject_weighted_accumulate(C.c_void_p(sp), ss, sw, sh, C.c_void_p(xmap.data_ptr()),
                                                    C.c_void_p(ymap.data_ptr()), xmap.stride(0) * 4,
                                                    C.c_void_p(weight.data_ptr()), weight.stride(0) * 4,
                                                    C.c_void_p(ap), ast, dw, dh, _stream(stream)),
           "cudaReprojectWeightedAccumulateTo32F")
    return accum


def cuda_reproject_params(src, param, dst_size, to16s=False, with_mask=False, stream=None):
    """cudaReproject / cudaReprojectTo16S(src, dst, [mask,] dstSize, param)."""
    import torch

    sh, sw = src.shape[:2]
    dw, dh = dst_size
    dst = torch.empty((dh, dw, 4), dtype=torch.int16 if to16s else torch.uint8, device=src.device)
    mask = torch.empty((dh, dw), dtype=torch.uint8, device=src.device) if with_mask else None
    sp, ss = _dev_ptr_step(src)
    dp, ds = _dev_ptr_step(dst)
    _check(lib().pano_reproject_params(C.c_void_p(sp), ss, sw, sh, C.byref(param), dw, dh,
                                       PANO_16SC4 if to16s else PANO_8UC4, C.c_void_p(dp), ds,
                                       C.c_void_p(mask.data_ptr()) if with_mask else None, dw if with_mask else 0,
                                       _stream(stream)), "cudaReproject")
    return (dst, mask) if with_mask else dst


def cuda_transform(src, luts, dst=None, stream=None):
    """cudaTransform(src 8UC4, dst, luts 3x256)."""
    import torch

    h, w = src.shape[:2]
    if dst is None:
        dst = torch.empty_like(src)
    luts = np.ascontiguousarray(luts, np.uint8)
    assert luts.shape == (3, 256)
    sp, ss = _dev_ptr_step(src)
    dp, ds = _dev_ptr_step(dst)
    _check(lib().pano_transform_lut3(C.c_void_p(sp), ss, C.c_void_p(dp), ds, w, h,
                                     luts.ctypes.data, _stream(stream)), "cudaTransform")
    return dst


# -------------------------------------------------------------------- render --

class PanoramaRender:
    """CudaPanoramaRender2 (Task/RicohUtil.h:134-161), multiband branch: maps,
    seams and weights resident on the device; render() = LUT + remap + blend."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().pano_render_create(C.byref(h)), "create render")
        self._h = h
        self.num_images = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_render_destroy(h)
            self._h = None

    def prepare(self, params, src_sizes, dst_size, max_levels: int) -> bool:
        """params: PhotoParams, or a rig file path loaded by loadPhotoParams as
        CPUPanoramaRender::prepare does (RicohUtil.cpp:1437-1443: False when it
        yields no camera).  src_sizes: one (w, h) per camera, or one for all."""
        if isinstance(params, (str, bytes, os.PathLike)):
            params = load_photo_params(params)
            if not params:
                return False
        n = len(params)
        if len(src_sizes) == 2 and all(isinstance(v, (int, np.integer)) for v in src_sizes):
            src_sizes = [tuple(src_sizes)] * n
        arr = (PhotoParam * n)(*params)
        sw = np.array([s[0] for s in src_sizes], np.int32)
        shh = np.array([s[1] for s in src_sizes], np.int32)
        st = lib().pano_render_prepare(self._h, n, C.cast(arr, C.c_void_p), sw.ctypes.data,
                                       shh.ctypes.data, dst_size[0], dst_size[1], max_levels)
        if st == PANO_OK:
            self.num_images, self.dst_size = n, tuple(dst_size)
        return st == PANO_OK

    def prepare_maps(self, src_sizes, dst_size, xmaps, ymaps, masks, max_levels: int) -> bool:
        n = len(xmaps)
        xs = [np.ascontiguousarray(x, np.float32) for x in xmaps]
        ys = [np.ascontiguousarray(y, np.float32) for y in ymaps]
        ms = [np.ascontiguousarray(m, np.uint8) for m in masks]
        sw = np.array([s[0] for s in src_sizes], np.int32)
        shh = np.array([s[1] for s in src_sizes], np.int32)
        st = lib().pano_render_prepare_maps(self._h, n, sw.ctypes.data, shh.ctypes.data,
                                            dst_size[0], dst_size[1],
                                            _ptrs([x.ctypes.data for x in xs]),
                                            _ptrs([y.ctypes.data for y in ys]),
                                            _ptrs([m.ctypes.data for m in ms]), max_levels)
        if st == PANO_OK:
            self.num_images, self.dst_size = n, tuple(dst_size)
        return st == PANO_OK

    @property
    def num_levels(self):
        return lib().pano_render_num_levels(self._h)

    @property
    def fused(self) -> bool:
        return bool(lib().pano_render_fused(self._h))

    def mem_bytes(self) -> int:
        return lib().pano_render_mem_bytes(self._h)

    def model_bytes(self, stage: int = 4) -> int:
        return lib().pano_render_model_bytes(self._h, stage)

    STAGES = ("remap", "down", "blend", "final")

    def set_timing(self, on: bool):
        _check(lib().pano_render_set_timing(self._h, int(bool(on))), "set timing")

    def stage_ms(self):
        """Summed device milliseconds per stage since the last call, and frames."""
        ms = np.zeros(4, np.float64)
        nf = C.c_int()
        _check(lib().pano_render_stage_ms(self._h, ms.ctypes.data, C.byref(nf)), "stage times")
        return dict(zip(self.STAGES, ms.tolist())), nf.value

    def render(self, srcs, out, luts=None, stream=None) -> bool:
        """srcs: list of HxWx4 uint8 cuda tensors (BGRx); out: HxWx4 uint8 cuda tensor;
        luts: None or (n, 3, 256) uint8 host array."""
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        op, ostep = _dev_ptr_step(out)
        lp = None
        if luts is not None:
            luts = np.ascontiguousarray(luts, np.uint8)
            assert luts.shape == (len(srcs), 3, 256)
            lp = C.c_void_p(luts.ctypes.data)
        st = lib().pano_render_render(self._h, len(srcs), _ptrs(ptrs), _sizes(steps), lp,
                                      C.c_void_p(op), ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "render")
        return True

    TILE_SHAPES = ((64, 16), (64, 32), (128, 16), (32, 32), (64, 8))   # down tile (width, height) per shape

    def set_tile_shape(self, shape: int) -> None:
        """Down tile shape of the next prepare (an index of TILE_SHAPES)."""
        _check(lib().pano_render_set_tile_shape(self._h, int(shape)), "set_tile_shape")

    @property
    def tile_shape(self) -> int:
        return lib().pano_render_tile_shape(self._h)

    def autotune_tile_shape(self, params, src_sizes, dst_size, max_levels: int, srcs, out, iters: int = 5,
                            stream=None):
        """Prepares with each tile shape, times `iters` frames of srcs (device BGRx
        tensors) into out, keeps the fastest.  Returns (shape, ms per frame by shape)."""
        n = len(params)
        arr = (PhotoParam * n)(*params)
        sw = np.array([s_[0] for s_ in src_sizes], np.int32)
        shh = np.array([s_[1] for s_ in src_sizes], np.int32)
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        op, ostep = _dev_ptr_step(out)
        best = C.c_int()
        ms = np.zeros(len(self.TILE_SHAPES), np.float32)
        _check(lib().pano_render_autotune_tile_shape(self._h, n, C.cast(arr, C.c_void_p), sw.ctypes.data,
                                                     shh.ctypes.data, dst_size[0], dst_size[1], max_levels,
                                                     _ptrs(ptrs), _sizes(steps), C.c_void_p(op), ostep, int(iters),
                                                     C.byref(best), ms.ctypes.data, _stream(stream)),
               "autotune_tile_shape")
        self.num_images, self.dst_size = n, tuple(dst_size)
        return best.value, [round(float(v), 4) for v in ms]

    def set_prepare_on_device(self, on: bool = True) -> None:
        """Build the fused render's tables on the device (default) or the host."""
        _check(lib().pano_render_set_prepare_on_device(self._h, int(bool(on))), "set_prepare_on_device")

    def prepare_ms(self):
        """Wall time of the last prepare by phase (ms)."""
        ms = np.zeros(4, np.float64)
        _check(lib().pano_render_prepare_ms(self._h, C.c_void_p(ms.ctypes.data)), "prepare_ms")
        return dict(zip(("support", "seams", "tables", "build"), ms.round(1).tolist()))

    def render_to(self, srcs, output: "Output", luts=None, stream=None) -> bool:
        """render() into an Output (the canvas, or NV12 / YUV420P planes)."""
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        lp = None
        if luts is not None:
            luts = np.ascontiguousarray(luts, np.uint8)
            assert luts.shape == (len(srcs), 3, 256)
            lp = C.c_void_p(luts.ctypes.data)
        st = lib().pano_render_render_to(self._h, len(srcs), _ptrs(ptrs), _sizes(steps), lp,
                                         C.byref(output), _stream(stream))
        if st == -2:
            return False
        _check(st, "render_to")
        return True

    def render_autogain(self, srcs, out, tint=True, stream=None) -> bool:
        """render() with gains estimated from this frame's overlap means
        (GainCompensate / getTransformsMeanApproxSiftPanoPaper, optionally the
        tint ratios) and applied as LUTs, all on the device."""
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        op, ostep = _dev_ptr_step(out)
        st = lib().pano_render_render_autogain(self._h, len(srcs), _ptrs(ptrs), _sizes(steps),
                                               PANO_GAIN_TINT if tint else 0, C.c_void_p(op),
                                               ostep, _stream(stream))
        if st == -2:
            return False
        _check(st, "render_autogain")
        return True

    def gains(self):
        """The last estimate: (gains, red ratio gains, blue ratio gains, anchor)."""
        n = lib().pano_render_num_images(self._h)
        g, r, b = (np.zeros(n, np.float64) for _ in range(3))
        a = C.c_int()
        _check(lib().pano_render_gains(self._h, g.ctypes.data, r.ctypes.data, b.ctypes.data,
                                       C.byref(a)), "gains")
        return g, r, b, a.value

    def gain_sums(self) -> np.ndarray:
        """The last estimate's (n, n, 5) int64 sums: count, gray, B, G, R."""
        n = lib().pano_render_num_images(self._h)
        out = np.zeros((n, n, 5), np.int64)
        _check(lib().pano_render_gain_sums(self._h, out.ctypes.data), "gain sums")
        return out


class Batch:
    """Several prepared renders stitched with one launch per pyramid stage
    (pano_batch_*).  The renders must stay alive while the batch uses them."""

    def __init__(self, renders=None):
        h = C.c_void_p()
        _check(lib().pano_batch_create(C.byref(h)), "create batch")
        self._h = h
        self.renders = []
        if renders:
            self.set(renders)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pano_batch_destroy(h)
            self._h = None

    def set(self, renders):
        _check(lib().pano_batch_set(self._h, len(renders), _ptrs([r._h.value for r in renders])),
               "batch set")
        self.renders = list(renders)

    STAGES = PanoramaRender.STAGES

    def set_timing(self, on: bool):
        _check(lib().pano_batch_set_timing(self._h, int(bool(on))), "set timing")

    def stage_ms(self):
        ms = np.zeros(4, np.float64)
        nf = C.c_int()
        _check(lib().pano_batch_stage_ms(self._h, ms.ctypes.data, C.byref(nf)), "stage times")
        return dict(zip(self.STAGES, ms.tolist())), nf.value

    @staticmethod
    def _lut_array(luts_per_render, keep):
        if luts_per_render is None:
            return None
        arr = []
        for lt in luts_per_render:
            if lt is None:
                arr.append(None)
            else:
                a = np.ascontiguousarray(lt, np.uint8)
                keep.append(a)
                arr.append(a.ctypes.data)
        return (C.c_void_p * len(arr))(*[C.c_void_p(a) if a else None for a in arr])

    def render(self, srcs_per_render, outs, luts_per_render=None, stream=None):
        """srcs_per_render: list (per render) of lists of BGRx cuda tensors;
        outs: per render BGRx cuda tensor; luts_per_render: per render (n,3,256) or None."""
        flat = [t for srcs in srcs_per_render for t in srcs]
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in flat])
        optrs, osteps = zip(*[_dev_ptr_step(t) for t in outs])
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        _check(lib().pano_batch_render(self._h, _ptrs(ptrs), _sizes(steps), lp, _ptrs(optrs),
                                       _sizes(osteps), _stream(stream)), "batch render")
        return True

    def render_to(self, srcs_per_render, outputs, luts_per_render=None, stream=None):
        """render() with one Output per render."""
        flat = [t for srcs in srcs_per_render for t in srcs]
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in flat])
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        arr = (Output * len(outputs))(*outputs)
        _check(lib().pano_batch_render_to(self._h, _ptrs(ptrs), _sizes(steps), lp, arr,
                                          _stream(stream)), "batch render_to")
        return True

    def capture_to(self, srcs_per_render, outputs, luts_per_render=None, stream=None):
        flat = [t for srcs in srcs_per_render for t in srcs]
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in flat])
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        arr = (Output * len(outputs))(*outputs)
        _check(lib().pano_batch_capture_to(self._h, _ptrs(ptrs), _sizes(steps), lp, arr,
                                           _stream(stream)), "batch capture_to")
        self._graph_buffers = (flat, list(outputs))

    def capture(self, srcs_per_render, outs, luts_per_render=None, stream=None):
        """Records one render of these buffers as a HIP graph (pano_batch_capture);
        the tensors are kept alive for the replays."""
        flat = [t for srcs in srcs_per_render for t in srcs]
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in flat])
        optrs, osteps = zip(*[_dev_ptr_step(t) for t in outs])
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        _check(lib().pano_batch_capture(self._h, _ptrs(ptrs), _sizes(steps), lp, _ptrs(optrs),
                                        _sizes(osteps), _stream(stream)), "batch capture")
        self._graph_buffers = (flat, list(outs))

    def replay(self, luts_per_render=None, stream=None):
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        _check(lib().pano_batch_replay(self._h, lp, _stream(stream)), "batch replay")

    SCHED_MODES = 2

    def set_sched(self, sched4):
        a = np.ascontiguousarray(sched4, np.int32)
        _check(lib().pano_batch_set_sched(self._h, C.c_void_p(a.ctypes.data)), "set sched")

    def sched(self):
        a = np.zeros(4, np.int32)
        _check(lib().pano_batch_get_sched(self._h, C.c_void_p(a.ctypes.data)), "get sched")
        return a.tolist()

    def autotune(self, srcs_per_render, outs, luts_per_render=None, stream=None, iters=5):
        """Times every block order per stage on these frames and keeps the fastest
        (pano_batch_autotune).  Returns (chosen orders, ms[mode][stage])."""
        flat = [t for srcs in srcs_per_render for t in srcs]
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in flat])
        optrs, osteps = zip(*[_dev_ptr_step(t) for t in outs])
        keep = []
        lp = self._lut_array(luts_per_render, keep)
        ms = np.zeros((self.SCHED_MODES, 4), np.float64)
        _check(lib().pano_batch_autotune(self._h, _ptrs(ptrs), _sizes(steps), lp, _ptrs(optrs),
                                         _sizes(osteps), _stream(stream), int(iters),
                                         C.c_void_p(ms.ctypes.data)), "batch autotune")
        return self.sched(), ms


class ServerDone(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("latency_ms", C.c_double), ("stream_id", C.c_int),
                ("status", C.c_int), ("batch_frames", C.c_int), ("pad", C.c_int)]


class Server:
    """The serving loop (pano_server_*): prepared renders admitted as streams
    against an HBM budget, frames batched continuously across streams.  The
    tensors of a submitted frame are kept alive here until it is reported done."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 2):
        h = C.c_void_p()
        _check(lib().pano_server_create(C.byref(h), int(hbm_budget_bytes), int(max_inflight)),
               "server create")
        self._h = h
        self._keep = {}
        self._renders = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            lib().pano_server_destroy(h)
        self._h = None
        self._keep.clear()

    def __del__(self):
        self.close()

    def admit(self, render):
        """Stream id, or None when the render does not fit the budget."""
        sid = C.c_int()
        st = lib().pano_server_admit(self._h, render._h, C.byref(sid))
        if st == PANO_ERR_BUDGET:
            return None
        _check(st, "admit")
        self._renders[sid.value] = render
        return sid.value

    def retire(self, sid: int):
        _check(lib().pano_server_retire(self._h, int(sid)), "retire")
        self._renders.pop(sid, None)

    def submit(self, sid: int, srcs, out, luts=None, tag: int = 0):
        ptrs, steps = zip(*[_dev_ptr_step(t) for t in srcs])
        op, ostep = _dev_ptr_step(out)
        lp = None
        if luts is not None:
            luts = np.ascontiguousarray(luts, np.uint8)
            lp = C.c_void_p(luts.ctypes.data)
        key = (int(sid), int(tag))
        self._keep[key] = (list(srcs), out)
        st = lib().pano_server_submit(self._h, int(sid), _ptrs(ptrs), _sizes(steps), lp,
                                      C.c_void_p(op), ostep, C.c_uint64(int(tag)))
        if st != PANO_OK:
            self._keep.pop(key, None)
        _check(st, "submit")

    def poll(self, max_out: int = 256, wait_ms: int = 0):
        """Completed frames: list of (stream_id, tag, latency_ms, status, batch_frames)."""
        buf = (ServerDone * max_out)()
        k = lib().pano_server_poll(self._h, buf, max_out, int(wait_ms))
        if k < 0:
            _check(k, "poll")
        out = []
        for d in buf[:k]:
            self._keep.pop((d.stream_id, int(d.tag)), None)
            out.append((d.stream_id, int(d.tag), d.latency_ms, d.status, d.batch_frames))
        return out

    def drain(self):
        _check(lib().pano_server_drain(self._h), "drain")

    def stats(self):
        u, f, b = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().pano_server_stats(self._h, C.byref(u), C.byref(f), C.byref(b)), "stats")
        return {"used_bytes": u.value, "frames_done": f.value, "batches": b.value}

