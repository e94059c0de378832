"""Resident meshes/textures/framebuffers, batched frames and the serving loop
over the C ABI (include/retrofire_amd.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from . import Context, Stats, check, default_device

BPP = {0: 4, 1: 3, 2: 4, 3: 4, 4: 4, 5: 4, 6: 2, 7: 2}


class Mesh:
    """A device-resident mesh (geom/mesh.rs:24-30): positions, an optional
    per-vertex attribute (n_attr floats) and index triples, or index pairs
    (an (n, 2) array: edges, drawn as lines)."""

    def __init__(self, pos, tris, attr=None, device=None):
        self.dev = device or default_device()
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        tris = np.asarray(tris)
        self.edges = tris.ndim == 2 and tris.shape[1] == 2
        tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1, 2 if self.edges else 3)
        n_attr = 0
        ap = None
        if attr is not None:
            attr = np.ascontiguousarray(attr, dtype=np.float32).reshape(pos.shape[0], -1)
            n_attr = attr.shape[1]
            ap = attr.ctypes.data_as(C.POINTER(C.c_float))
        self.n_verts, self.n_tris, self.n_attr = pos.shape[0], tris.shape[0], n_attr
        self.h = C.c_void_p()
        create = _ffi.lib().rf_mesh_create_edges if self.edges else _ffi.lib().rf_mesh_create
        check(create(self.dev.h, pos.ctypes.data_as(C.POINTER(C.c_float)), ap, n_attr, pos.shape[0],
                     tris.ctypes.data_as(C.POINTER(C.c_uint32)), tris.shape[0], C.byref(self.h)))

    def free(self):
        if self.h:
            _ffi.lib().rf_mesh_free(self.dev.h, self.h)
            self.h = C.c_void_p()


class Texture:
    def __init__(self, rgba: np.ndarray | None, device=None):
        self.dev = device or default_device()
        self.h = C.c_void_p()
        if rgba is None:
            return
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        check(_ffi.lib().rf_texture_create(self.dev.h, rgba.ctypes.data_as(C.POINTER(C.c_uint8)), rgba.shape[1],
                                           rgba.shape[0], rgba.shape[1], C.byref(self.h)))

    def view(self, x, y, w, h):
        """The sub-texture of the rectangle at (x, y), w x h texels, sharing
        this texture's memory (Atlas::get, tex.rs:157-172). Free it first."""
        v = Texture(None, self.dev)
        check(_ffi.lib().rf_texture_view(self.dev.h, self.h, x, y, w, h, C.byref(v.h)))
        return v

    def free(self):
        if self.h:
            _ffi.lib().rf_texture_free(self.dev.h, self.h)
            self.h = C.c_void_p()


class DeviceFramebuf:
    """A device-resident colour (+ optional depth) target."""

    def __init__(self, width, height, fmt=0, with_depth=True, device=None):
        self.dev = device or default_device()
        self.width, self.height, self.fmt, self.with_depth = width, height, fmt, with_depth
        self.h = C.c_void_p()
        check(_ffi.lib().rf_framebuf_create(self.dev.h, width, height, fmt, int(with_depth), C.byref(self.h)))

    def read(self):
        if self.fmt in (2, 3, 4, 5):
            col = np.empty((self.height, self.width), np.uint32)
        elif self.fmt in (6, 7):
            col = np.empty((self.height, self.width), np.uint16)
        else:
            col = np.empty((self.height, self.width, BPP[self.fmt]), np.uint8)
        dep = np.empty((self.height, self.width), np.float32) if self.with_depth else None
        check(_ffi.lib().rf_framebuf_read(self.dev.h, self.h, col.ctypes.data,
                                          dep.ctypes.data_as(C.POINTER(C.c_float)) if dep is not None else None))
        return col, dep

    def free(self):
        if self.h:
            _ffi.lib().rf_framebuf_free(self.dev.h, self.h)
            self.h = C.c_void_p()


def mesh_draw(mesh: Mesh, shd, mvp, viewport, first_tri=0, n_tris=0, texture: Texture | None = None,
              cull=False):
    """One render() call over a resident mesh: `shd` is a shader kind object
    (retrofire_amd.shader.*); textured kinds take a resident `texture`. With
    `cull`, the device skips the draw when the mesh's bounding box is outside
    the view frustum (BBox::visibility is Hidden, scene.rs:81-87)."""
    d = _ffi.RfMeshDraw()
    d.mesh = mesh.h.value
    d.kind = shd.kind
    m = np.ascontiguousarray(mvp, np.float32).reshape(16)
    v = np.ascontiguousarray(viewport, np.float32).reshape(16)
    for i in range(16):
        d.mvp[i] = m[i]
        d.viewport[i] = v[i]
    if shd.kind == 0:
        for i in range(4):
            d.flat_color[i] = shd.color[i]
    if shd.kind >= 2:
        assert texture is not None
        d.tex = texture.h.value
    d.first_tri = first_tri
    d.n_tris = n_tris
    d.cull = int(bool(cull))
    return d


class Batch:
    """Frames (target, ctx, draws) planned once and replayed as one HIP graph."""

    def __init__(self, frames, tile_w=0, device=None):
        self.dev = device or default_device()
        self._keep = []
        arr = (_ffi.RfFrame * len(frames))()
        for i, (fb, ctx, draws) in enumerate(frames):
            da = (_ffi.RfMeshDraw * max(1, len(draws)))(*draws)
            self._keep.append(da)
            arr[i].target = fb.h.value
            arr[i].ctx = ctx.to_c()
            arr[i].draws = da
            arr[i].n_draws = len(draws)
        self.n = len(frames)
        self.h = C.c_void_p()
        check(_ffi.lib().rf_batch_create(self.dev.h, arr, len(frames), tile_w, C.byref(self.h)))

    @property
    def tile_w(self):
        return int(_ffi.lib().rf_batch_tile_width(self.h))

    @property
    def tile_h(self):
        return int(_ffi.lib().rf_batch_tile_height(self.h))

    def tune_results(self):
        """The autotuner's timings: [(tile_w, tile_h, chunk, ms)] per candidate."""
        n = _ffi.lib().rf_batch_tune_results(self.h, None, 0)
        arr = (_ffi.RfTuneEntry * max(1, n))()
        _ffi.lib().rf_batch_tune_results(self.h, arr, n)
        return [(e.tile_w, e.tile_h, e.chunk, round(e.ms, 4)) for e in arr[:n]]

    @property
    def chunk(self):
        """The rasterizer's span chunk (6 or 8) this batch runs with."""
        return int(_ffi.lib().rf_batch_raster_chunk(self.h))

    def run(self, sync=True):
        check(_ffi.lib().rf_batch_run(self.h, int(sync)))

    def wait(self):
        st = (_ffi.RfStats * self.n)()
        check(_ffi.lib().rf_batch_wait(self.h, st))
        out = []
        for s in st:
            x = Stats()
            x.add(s)
            out.append(x)
        return out

    def last_ms(self):
        t, r = C.c_float(), C.c_float()
        check(_ffi.lib().rf_batch_last_ms(self.h, C.byref(t), C.byref(r)))
        return t.value

    def profile(self, iters=5):
        """Mean ms of vertex, setup, scan, bin, binsort, raster (eager run with
        HIP events on the batch stream)."""
        ms = (C.c_float * 6)()
        check(_ffi.lib().rf_batch_profile(self.h, iters, ms))
        return list(ms)

    def free(self):
        if self.h:
            _ffi.lib().rf_batch_free(self.h)
            self.h = C.c_void_p()


class Server:
    """The serving loop (rf_serve_*): submit requests, poll results."""

    def __init__(self, budget_bytes, max_batch=16, n_streams=2, device=None):
        self.dev = device or default_device()
        self.h = C.c_void_p()
        check(_ffi.lib().rf_serve_open(self.dev.h, budget_bytes, max_batch, n_streams, C.byref(self.h)))

    def submit(self, rid, draws, width, height, fmt=0, with_depth=True, ctx: Context | None = None):
        ctx = ctx or Context()
        r = _ffi.RfRequest()
        r.id = rid
        da = (_ffi.RfMeshDraw * max(1, len(draws)))(*draws)
        r.draws = da
        r.n_draws = len(draws)
        r.width, r.height, r.fmt, r.with_depth = width, height, fmt, int(with_depth)
        r.ctx = ctx.to_c()
        check(_ffi.lib().rf_serve_submit(self.h, C.byref(r)))

    def poll(self, max_results=64, timeout_ms=0, buffers=None):
        """Returns a list of (rf_result, pixels-or-None). `buffers` maps request
        id -> preallocated numpy array to receive the pixels."""
        res = (_ffi.RfResult * max_results)()
        n = C.c_uint32()
        ptrs = None
        if buffers is not None:
            # pixels are copied in completion order; one buffer per slot
            ptrs = (C.c_void_p * max_results)(*([b.ctypes.data for b in buffers[:max_results]] +
                                                [None] * (max_results - len(buffers[:max_results]))))
        check(_ffi.lib().rf_serve_poll(self.h, res, ptrs, max_results, C.byref(n), timeout_ms))
        return [res[i] for i in range(n.value)]

    def close(self):
        if self.h:
            check(_ffi.lib().rf_serve_close(self.h))
            self.h = C.c_void_p()
