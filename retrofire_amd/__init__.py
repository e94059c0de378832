"""retrofire_amd: retrofire's triangle rendering path on an MI355X.

Host-side mirror of the reference interface
`render(prims, verts, shader, uniform, to_screen, target, ctx)`
(core/src/render.rs:134-207), over the C ABI of include/retrofire_amd.h. The
built-in shader kinds are the types in `shader`; a target is a `Framebuf`
(colour + depth, target.rs:35-38) or a `Colorbuf` (colour only,
target.rs:41-44 / Buf2<Color3|Color4>). Errors raise `RenderError`, where the
reference panics. There is no CPU fallback: a missing HIP library raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _ffi

__all__ = ["Context", "Stats", "Framebuf", "Colorbuf", "RenderError", "render", "render_noalloc", "shader", "Device",
           "default_device", "PIX_RGBA8", "PIX_RGB8", "PIX_U32_RGBA", "PIX_U32_XRGB", "PIX_U32_ARGB",
           "PIX_U32_BGRA", "PIX_RGBA4444", "PIX_RGB565"]

PIX_RGBA8, PIX_RGB8, PIX_U32_RGBA, PIX_U32_XRGB = 0, 1, 2, 3
PIX_U32_ARGB, PIX_U32_BGRA, PIX_RGBA4444, PIX_RGB565 = 4, 5, 6, 7
_DT = {None: 0, "less": 1, "equal": 2, "greater": 3}
_FC = {None: 0, "back": 1, "front": 2}
_DS = {None: 0, "front_to_back": 1, "back_to_front": 2}


class RenderError(RuntimeError):
    def __init__(self, status: int):
        self.status = status
        super().__init__(f"rf status {status}: {_ffi.lib().rf_status_str(status).decode()}")


def check(status: int):
    if status != 0:
        raise RenderError(status)


@dataclass
class Stats:
    """Stats (stats.rs:16-40)."""
    calls: int = 0
    prims_i: int = 0
    prims_o: int = 0
    verts_i: int = 0
    verts_o: int = 0
    frags_i: int = 0
    frags_o: int = 0
    seconds: float = 0.0
    objs_i: int = 0
    objs_o: int = 0

    def add(self, s: "_ffi.RfStats"):
        for k in ("calls", "prims_i", "prims_o", "verts_i", "verts_o", "frags_i", "frags_o", "objs_i", "objs_o"):
            setattr(self, k, getattr(self, k) + int(getattr(s, k)))
        self.seconds += float(s.seconds)


@dataclass
class Context:
    """Context (ctx.rs:11-65); `depth_test` is None/"less"/"equal"/"greater",
    `face_cull` None/"back"/"front". `color_clear`/`depth_clear` are applied
    before drawing when `clear` is set (the frontend's Frame::clear,
    front/src/lib.rs:103-120); render() itself never clears."""
    color_clear: tuple | None = (0, 0, 0, 255)
    depth_clear: float | None = float("inf")
    face_cull: str | None = "back"
    depth_sort: str | None = None
    depth_test: str | None = "less"
    color_write: bool = True
    depth_write: bool = True
    clear: bool = False
    stats: Stats = field(default_factory=Stats)

    def to_c(self) -> _ffi.RfCtx:
        c = _ffi.RfCtx()
        c.depth_test = _DT[self.depth_test]
        c.face_cull = _FC[self.face_cull]
        c.color_write = int(self.color_write)
        c.depth_write = int(self.depth_write)
        c.depth_sort = _DS[self.depth_sort]
        c.clear_color = int(self.clear and self.color_clear is not None)
        cc = self.color_clear or (0, 0, 0, 0)
        for i in range(4):
            c.color_clear[i] = int(cc[i])
        c.clear_depth = int(self.clear and self.depth_clear is not None)
        c.depth_clear = float(self.depth_clear if self.depth_clear is not None else float("inf"))
        return c


class shader:
    """The built-in shader kinds (vertex: `vertex(mvp.apply(&v.pos), v.attrib)`)."""

    class Flat:
        kind = 0
        n_attr = 0

        def __init__(self, color=(255, 255, 255, 255)):
            self.color = tuple(int(c) for c in color)

    class Gouraud:
        kind = 1
        n_attr = 3

    class TexClamp:
        kind = 2
        n_attr = 2

        def __init__(self, texture: np.ndarray):
            self.texture = np.ascontiguousarray(texture, dtype=np.uint8)

    class TexRepeatPot:
        kind = 3
        n_attr = 2

        def __init__(self, texture: np.ndarray):
            self.texture = np.ascontiguousarray(texture, dtype=np.uint8)

    class TexOnce:
        """SamplerOnce (tex.rs:313-357): texture coordinates must stay in
        0..1; a texel outside the texture raises RenderError (RF_E_BOUNDS)."""
        kind = 4
        n_attr = 2

        def __init__(self, texture: np.ndarray):
            self.texture = np.asarray(texture, dtype=np.uint8)


class Colorbuf:
    """Colour-only target: `color` is (h, w, 4) uint8 (RGBA8), (h, w, 3) uint8
    (RGB8), (h, w) uint32 with fmt PIX_U32_RGBA (the default for uint32),
    PIX_U32_XRGB, PIX_U32_ARGB or PIX_U32_BGRA, or (h, w) uint16 with fmt
    PIX_RGBA4444 or PIX_RGB565 (which must be given)."""
    depth = None

    _ITEM = {PIX_RGBA8: (np.uint8, 3), PIX_RGB8: (np.uint8, 3), PIX_U32_RGBA: (np.uint32, 2),
             PIX_U32_XRGB: (np.uint32, 2), PIX_U32_ARGB: (np.uint32, 2), PIX_U32_BGRA: (np.uint32, 2),
             PIX_RGBA4444: (np.uint16, 2), PIX_RGB565: (np.uint16, 2)}

    def __init__(self, color: np.ndarray, fmt: int | None = None):
        self.color = color
        if fmt is None:
            if color.dtype == np.uint16:
                raise ValueError("a uint16 colour buffer needs fmt PIX_RGBA4444 or PIX_RGB565")
            fmt = PIX_RGB8 if color.ndim == 3 and color.shape[2] == 3 else (
                PIX_RGBA8 if color.ndim == 3 else PIX_U32_RGBA)
        want = self._ITEM.get(fmt)
        if want is None or color.dtype != want[0] or color.ndim != want[1]:
            raise ValueError(f"colour buffer of dtype {color.dtype}, {color.ndim} dims does not fit fmt {fmt}")
        self.fmt = fmt


class Framebuf(Colorbuf):
    """Colour + f32 depth (1/w) target."""

    def __init__(self, color: np.ndarray, depth: np.ndarray, fmt: int | None = None):
        super().__init__(color, fmt)
        self.depth = depth


class Device:
    def __init__(self, ordinal: int = 0):
        self.h = C.c_void_p()
        check(_ffi.lib().rf_device_open(ordinal, C.byref(self.h)))

    def close(self):
        if self.h:
            _ffi.lib().rf_device_close(self.h)
            self.h = C.c_void_p()

    @property
    def bytes(self) -> int:
        return int(_ffi.lib().rf_device_bytes(self.h))

    def set_tile_width(self, tile_w: int = 0):
        """Rasterizer tile width of render() on this device (16, 32, 64; 0 = default)."""
        check(_ffi.lib().rf_device_set_tile_width(self.h, tile_w))

    def set_tile_height(self, tile_h: int = 0):
        """Rasterizer tile height of render() and of batches with a forced tile
        width on this device (16, 32, 64 rows; 0 = default, 32)."""
        check(_ffi.lib().rf_device_set_tile_height(self.h, tile_h))

    def set_raster_chunk(self, chunk: int = 0):
        """Rasterizer span chunk of render() and of batches with a forced tile
        width on this device (6 or 8 pixels; 0 = default, 8)."""
        check(_ffi.lib().rf_device_set_raster_chunk(self.h, chunk))


_default = None


def default_device() -> Device:
    global _default
    if _default is None:
        _default = Device(0)
    return _default


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _rows(a: np.ndarray) -> bool:
    """Rows of packed elements at a positive row stride: a whole buffer or a
    Slice2/MutSlice2 view of one (buf.rs:570-580), e.g. a numpy slice
    `parent[y0:y1, x0:x1]` or a pitched surface."""
    if a.ndim == 2:
        px = a.itemsize
    elif a.ndim == 3 and a.strides[2] == a.itemsize:
        px = a.itemsize * a.shape[2]
    else:
        return False
    return a.strides[1] == px and a.strides[0] >= px * a.shape[1] and a.strides[0] % a.itemsize == 0


def make_target(target) -> _ffi.RfTarget:
    col = target.color
    if not _rows(col):
        raise ValueError("colour buffer rows must hold packed pixels at a positive row stride")
    t = _ffi.RfTarget()
    t.color = col.ctypes.data
    t.height, t.width = col.shape[0], col.shape[1]
    t.color_stride = col.strides[0]
    t.fmt = target.fmt
    if target.depth is not None:
        d = target.depth
        if not (d.dtype == np.float32 and d.ndim == 2 and _rows(d) and d.shape[:2] == col.shape[:2]):
            raise ValueError("depth buffer must be float32 rows of the colour buffer's size")
        t.depth = _fptr(d)
        t.depth_stride = d.strides[0] // 4
    return t


def make_draw(prims, verts, shd, uniform, to_screen, keep: list) -> _ffi.RfDraw:
    prims = np.asarray(prims)
    edges = prims.ndim == 2 and prims.shape[1] == 2
    tris = np.ascontiguousarray(prims, dtype=np.uint32).reshape(-1, 2 if edges else 3)
    pos, attr = verts if isinstance(verts, tuple) else (verts, None)
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    d = _ffi.RfDraw()
    d.kind = shd.kind
    d.prim = 1 if edges else 0
    d.tris = tris.ctypes.data_as(C.POINTER(C.c_uint32))
    d.n_tris = tris.shape[0]
    d.pos = _fptr(pos)
    d.n_verts = pos.shape[0]
    keep += [tris, pos]
    if shd.n_attr:
        a = np.ascontiguousarray(attr, dtype=np.float32).reshape(pos.shape[0], shd.n_attr)
        d.attr = _fptr(a)
        keep.append(a)
    m = np.ascontiguousarray(uniform, dtype=np.float32).reshape(16)
    v = np.ascontiguousarray(to_screen, dtype=np.float32).reshape(16)
    for i in range(16):
        d.mvp[i] = m[i]
        d.viewport[i] = v[i]
    if shd.kind == 0:
        for i in range(4):
            d.flat_color[i] = shd.color[i]
    if shd.kind >= 2:
        tex = shd.texture
        # a view into a larger texture (an atlas cell) keeps its row stride
        assert tex.ndim == 3 and tex.shape[2] == 4 and tex.strides[2] == 1 and tex.strides[1] == 4
        d.tex = tex.ctypes.data_as(C.POINTER(C.c_uint8))
        d.tex_h, d.tex_w = tex.shape[0], tex.shape[1]
        d.tex_stride = tex.strides[0] // 4
        keep.append(tex)
    return d


def render_noalloc(prims, verts, shd, uniform, to_screen, target, ctx: Context, device: Device | None = None):
    """render() through rf_render_noalloc: the same call and result, with
    every device buffer taken from the library's fixed 64 MiB array per GPU
    instead of device allocations (for callers that cannot allocate). A call
    whose buffers do not fit raises RenderError(-9)."""
    _render(prims, verts, shd, uniform, to_screen, target, ctx, device, "rf_render_noalloc")


def render(prims, verts, shd, uniform, to_screen, target, ctx: Context, device: Device | None = None):
    """render(prims, verts, shader, uniform, to_screen, target, ctx)
    (core/src/render.rs:134-207) on the GPU. `prims` is an (n, 3) index
    array of triangles (Tri<usize>) or an (n, 2) one of edges (Edge<usize>,
    drawn as lines); `verts` is positions (n, 3) or a tuple (positions, attributes);
    `uniform` is the mvp matrix of the built-in vertex shader; `to_screen` the
    viewport matrix. The target's arrays are updated in place; ctx.stats
    accumulates the call's statistics."""
    _render(prims, verts, shd, uniform, to_screen, target, ctx, device, "rf_render")


def _render(prims, verts, shd, uniform, to_screen, target, ctx, device, entry):
    if not isinstance(shd, (shader.Flat, shader.Gouraud, shader.TexClamp, shader.TexRepeatPot, shader.TexOnce)):
        raise RenderError(-5)
    dev = device or default_device()
    keep: list = []
    d = make_draw(prims, verts, shd, uniform, to_screen, keep)
    t = make_target(target)
    c = ctx.to_c()
    st = _ffi.RfStats()
    check(getattr(_ffi.lib(), entry)(dev.h, C.byref(d), C.byref(t), C.byref(c), C.byref(st)))
    ctx.stats.add(st)
