"""ctypes loader for the oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


class OracleCtx(C.Structure):
    _fields_ = [("depth_test", C.c_int32), ("face_cull", C.c_int32),
                ("color_write", C.c_int32), ("depth_write", C.c_int32), ("depth_sort", C.c_int32)]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        srcs = [os.path.join(ORACLE_DIR, f) for f in ("rf_oracle.hpp", "rf_oracle_capi.cpp")]
        if not os.path.exists(ORACLE_SO) or any(
                os.path.exists(s) and os.path.getmtime(s) > os.path.getmtime(ORACLE_SO) for s in srcs):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = C.CDLL(ORACLE_SO)
        f32p, u32p, u8p, i64p, u64p = (C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint8), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_uint64))
        L.oracle_render.restype = C.c_int
        L.oracle_render.argtypes = [C.c_int, u32p, C.c_uint64, f32p, f32p, C.c_uint64, f32p, f32p,
                                    u8p, u8p, C.c_uint32, C.c_uint32, C.c_uint32, u8p, C.c_uint64,
                                    C.c_uint64, C.c_uint64, C.c_int32, f32p, C.c_uint64,
                                    C.POINTER(OracleCtx), u64p]
        L.oracle_clip_edges_f32.restype = C.c_int64
        L.oracle_clip_edges_f32.argtypes = [f32p, C.c_int64, C.POINTER(C.c_int32), C.c_int32, f32p, C.c_int64]
        L.oracle_line_f32.restype = C.c_int64
        L.oracle_line_f32.argtypes = [f32p, f32p, i64p, f32p, C.c_int64]
        L.oracle_render_prims.restype = C.c_int
        L.oracle_render_prims.argtypes = [C.c_int] + L.oracle_render.argtypes
        L.oracle_render_hello_fp.restype = C.c_int
        L.oracle_render_hello_fp.argtypes = [u32p, C.c_uint64, f32p, f32p, C.c_uint64, f32p, f32p,
                                             u8p, C.c_uint64, C.c_uint64, u64p]
        L.oracle_tri_fill_f32.restype = C.c_int64
        L.oracle_tri_fill_f32.argtypes = [f32p, i64p, C.c_int64, f32p, C.c_int64, i64p]
        L.oracle_vary_f32.argtypes = [C.c_float, C.c_float, C.c_uint32, f32p]
        L.oracle_vary_to_f32.argtypes = [C.c_float, C.c_float, C.c_uint32, f32p]
        L.oracle_scanline_frags_f32.argtypes = [f32p, f32p, C.c_uint32, f32p]
        L.oracle_bbox_of.restype = None
        L.oracle_bbox_of.argtypes = [f32p, C.c_uint64, f32p]
        L.oracle_bbox_visibility.restype = C.c_int
        L.oracle_bbox_visibility.argtypes = [f32p, f32p]
        L.oracle_signed_dist.restype = C.c_float
        L.oracle_signed_dist.argtypes = [C.c_int, f32p]
        L.oracle_outcode.restype = C.c_int
        L.oracle_outcode.argtypes = [f32p]
        L.oracle_clip_tris_f32.restype = C.c_int64
        L.oracle_clip_tris_f32.argtypes = [f32p, C.c_int64, C.POINTER(C.c_int32), C.c_int32, f32p,
                                           C.c_int64]
        L.oracle_clip_plane_polygon_f32.restype = C.c_int64
        L.oracle_clip_plane_polygon_f32.argtypes = [C.c_int32, f32p, C.c_int64, f32p, C.c_int64]
        L.oracle_clip_histogram.restype = C.c_int64
        L.oracle_clip_histogram.argtypes = [u64p]
        L.oracle_sample.restype = C.c_int
        L.oracle_sample.argtypes = [C.c_int, u8p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                    C.c_float, u8p]
        L.oracle_time_frames.restype = C.c_double
        L.oracle_time_frames.argtypes = [C.c_int, u32p, C.c_uint64, f32p, f32p, C.c_uint64, f32p,
                                         f32p, u8p, u8p, C.c_uint32, C.c_uint32, C.c_uint32, u8p,
                                         C.c_uint64, C.c_uint64, C.c_int32, f32p,
                                         C.POINTER(OracleCtx), u8p, C.c_float, C.c_int32,
                                         C.POINTER(C.c_double), u64p]
        _lib = L
    return _lib


FMT_BPP = {0: 4, 1: 3, 2: 4, 3: 4, 4: 4, 5: 4, 6: 2, 7: 2}


def render(kind, tris, pos, attr, mvp, vp, color, depth=None, fmt=0, flat=(255, 255, 255, 255),
           tex=None, ctx=(1, 1, 1, 1), stats=None, prim=0):
    """One render() call on the oracle. `color` is an (h, w, bpp) uint8, (h, w)
    uint32 or (h, w) uint16 array updated in place; `depth` an (h, w) float32
    array or None. `prim` 0: `tris` holds index triples; 1: index pairs (edges)."""
    L = lib()
    tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    attr = np.ascontiguousarray(attr if attr is not None else np.zeros(1), dtype=np.float32).reshape(-1)
    mvp = np.ascontiguousarray(mvp, dtype=np.float32).reshape(-1)
    vp = np.ascontiguousarray(vp, dtype=np.float32).reshape(-1)
    flat = np.ascontiguousarray(flat, dtype=np.uint8)
    h, w = color.shape[0], color.shape[1]
    if tex is None:
        texa, tw, th = np.zeros(4, dtype=np.uint8), 0, 0
    else:
        texa = np.ascontiguousarray(tex, dtype=np.uint8)
        th, tw = texa.shape[0], texa.shape[1]
    if stats is None:
        stats = np.zeros(8, dtype=np.uint64)
    c = OracleCtx(*ctx)
    assert color.flags["C_CONTIGUOUS"]
    stride_bytes = color.strides[0]
    err = L.oracle_render_prims(prim, kind, _p(tris, C.c_uint32), tris.size // (2 if prim else 3), _p(pos, C.c_float),
                          _p(attr, C.c_float), pos.size // 3, _p(mvp, C.c_float), _p(vp, C.c_float),
                          _p(flat, C.c_uint8), _p(texa, C.c_uint8), tw, th, tw,
                          color.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, stride_bytes, fmt,
                          _p(depth, C.c_float), (w if depth is not None else 0), C.byref(c),
                          _p(stats, C.c_uint64))
    return err, stats


VISIBLE, CLIPPED, HIDDEN = 0, 1, 2


def bbox_of(pos):
    """BBox::of (scene.rs:37-39) of an (n, 3) float32 array: (low xyz, high xyz)."""
    p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    out = np.zeros(6, np.float32)
    lib().oracle_bbox_of(_p(p, C.c_float), p.size // 3, _p(out, C.c_float))
    return out


def bbox_visibility(bb6, mvp):
    """BBox::visibility (scene.rs:81-87): VISIBLE, CLIPPED or HIDDEN."""
    b = np.ascontiguousarray(bb6, dtype=np.float32).reshape(6)
    m = np.ascontiguousarray(mvp, dtype=np.float32).reshape(16)
    return lib().oracle_bbox_visibility(_p(b, C.c_float), _p(m, C.c_float))
