"""ctypes bindings of include/retrofire_amd.h.

The HIP library must be built (`python -m retrofire_amd.build` or
`__graft_entry__.build()`); importing the bindings without it raises, there is
no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librfamd.so")


class RfCtx(C.Structure):
    _fields_ = [("depth_test", C.c_int32), ("face_cull", C.c_int32), ("color_write", C.c_int32),
                ("depth_write", C.c_int32), ("depth_sort", C.c_int32), ("clear_color", C.c_int32),
                ("color_clear", C.c_uint8 * 4), ("clear_depth", C.c_int32), ("depth_clear", C.c_float)]


class RfStats(C.Structure):
    _fields_ = [("calls", C.c_uint64), ("prims_i", C.c_uint64), ("prims_o", C.c_uint64),
                ("verts_i", C.c_uint64), ("verts_o", C.c_uint64), ("frags_i", C.c_uint64),
                ("frags_o", C.c_uint64), ("seconds", C.c_double), ("objs_i", C.c_uint64),
                ("objs_o", C.c_uint64)]


class RfTarget(C.Structure):
    _fields_ = [("color", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("color_stride", C.c_uint32), ("fmt", C.c_uint32), ("depth", C.POINTER(C.c_float)),
                ("depth_stride", C.c_uint32)]


class RfDraw(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("tris", C.POINTER(C.c_uint32)), ("n_tris", C.c_uint64),
                ("pos", C.POINTER(C.c_float)), ("attr", C.POINTER(C.c_float)), ("n_verts", C.c_uint64),
                ("mvp", C.c_float * 16), ("viewport", C.c_float * 16), ("flat_color", C.c_uint8 * 4),
                ("tex", C.POINTER(C.c_uint8)), ("tex_w", C.c_uint32), ("tex_h", C.c_uint32),
                ("tex_stride", C.c_uint32), ("prim", C.c_uint32)]


class RfMeshDraw(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("kind", C.c_uint32), ("mvp", C.c_float * 16),
                ("viewport", C.c_float * 16), ("flat_color", C.c_uint8 * 4), ("tex", C.c_void_p),
                ("first_tri", C.c_uint64), ("n_tris", C.c_uint64), ("cull", C.c_uint32)]


class RfFrame(C.Structure):
    _fields_ = [("target", C.c_void_p), ("ctx", RfCtx), ("draws", C.POINTER(RfMeshDraw)),
                ("n_draws", C.c_uint32)]


class RfRequest(C.Structure):
    _fields_ = [("id", C.c_uint64), ("draws", C.POINTER(RfMeshDraw)), ("n_draws", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("fmt", C.c_uint32),
                ("with_depth", C.c_int32), ("ctx", RfCtx)]


class RfResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("fmt", C.c_uint32), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("done_s", C.c_double), ("stats", RfStats)]


class RfTuneEntry(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("chunk", C.c_uint32), ("ms", C.c_float)]


# Every symbol declared in include/retrofire_amd.h, with its signature.
_vp = C.c_void_p
_st = C.c_int32
SIGNATURES = {
    "rf_device_open": (_st, [C.c_int, C.POINTER(_vp)]),
    "rf_device_close": (_st, [_vp]),
    "rf_status_str": (C.c_char_p, [_st]),
    "rf_device_bytes": (C.c_uint64, [_vp]),
    "rf_device_set_tile_width": (_st, [_vp, C.c_uint32]),
    "rf_device_set_raster_chunk": (_st, [_vp, C.c_uint32]),
    "rf_tile_height": (C.c_uint32, []),
    "rf_device_set_tile_height": (_st, [_vp, C.c_uint32]),
    "rf_render": (_st, [_vp, C.POINTER(RfDraw), C.POINTER(RfTarget), C.POINTER(RfCtx), C.POINTER(RfStats)]),
    "rf_render_noalloc": (_st, [_vp, C.POINTER(RfDraw), C.POINTER(RfTarget), C.POINTER(RfCtx),
                                C.POINTER(RfStats)]),
    "rf_mesh_create": (_st, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.c_uint64,
                             C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(_vp)]),
    "rf_mesh_create_edges": (_st, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_uint32, C.c_uint64,
                                   C.POINTER(C.c_uint32), C.c_uint64, C.POINTER(_vp)]),
    "rf_mesh_free": (_st, [_vp, _vp]),
    "rf_texture_view": (_st, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "rf_texture_create": (_st, [_vp, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, C.c_uint32,
                                C.POINTER(_vp)]),
    "rf_texture_free": (_st, [_vp, _vp]),
    "rf_framebuf_create": (_st, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(_vp)]),
    "rf_framebuf_free": (_st, [_vp, _vp]),
    "rf_framebuf_target": (_st, [_vp, C.POINTER(RfTarget)]),
    "rf_framebuf_read": (_st, [_vp, _vp, _vp, C.POINTER(C.c_float)]),
    "rf_batch_create": (_st, [_vp, C.POINTER(RfFrame), C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "rf_batch_run": (_st, [_vp, C.c_int]),
    "rf_batch_wait": (_st, [_vp, C.POINTER(RfStats)]),
    "rf_batch_free": (_st, [_vp]),
    "rf_batch_tile_width": (C.c_uint32, [_vp]),
    "rf_batch_raster_chunk": (C.c_uint32, [_vp]),
    "rf_batch_tile_height": (C.c_uint32, [_vp]),
    "rf_batch_tune_results": (C.c_uint32, [_vp, C.POINTER(RfTuneEntry), C.c_uint32]),
    "rf_batch_last_ms": (_st, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rf_batch_profile": (_st, [_vp, C.c_int, C.POINTER(C.c_float)]),
    "rf_serve_open": (_st, [_vp, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(_vp)]),
    "rf_serve_submit": (_st, [_vp, C.POINTER(RfRequest)]),
    "rf_serve_poll": (_st, [_vp, C.POINTER(RfResult), C.POINTER(_vp), C.c_uint32, C.POINTER(C.c_uint32),
                            C.c_uint32]),
    "rf_serve_close": (_st, [_vp]),
    "rf_now_s": (C.c_double, []),
}

_lib = None


def lib():
    """Loads librfamd.so, or raises: the render path has no CPU fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m retrofire_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue  # reported by missing_symbols()
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def missing_symbols():
    """Names declared in the header that the built library does not export."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is None]
