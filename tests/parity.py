"""Helpers shared by the GPU parity tests: run one render() call on the
oracle and on the HIP path from identical inputs and compare."""
import numpy as np

import oracle_lib as O
import retrofire_amd as rf

DT = {None: 0, "less": 1, "equal": 2, "greater": 3}

# Rasterizer configurations (tile width, span chunk, tile height) every
# parity case runs with: each width with each chunk at the default height,
# and the other heights (results must be identical for all of them).
TILE_CONFIGS = [(16, 8, 32), (32, 8, 32), (64, 8, 32), (16, 6, 32), (32, 6, 32), (64, 6, 32),
                (32, 8, 16), (32, 8, 64), (16, 6, 64), (64, 6, 16)]
TILE_IDS = [f"tw{w}c{c}h{h}" for (w, c, h) in TILE_CONFIGS]


def use_tile_config(cfg):
    """Sets the device's rasterizer configuration for rf_render, then restores
    the defaults (a generator for a pytest fixture)."""
    dev = rf.default_device()
    dev.set_tile_width(cfg[0])
    dev.set_raster_chunk(cfg[1])
    dev.set_tile_height(cfg[2])
    yield cfg
    dev.set_tile_width(0)
    dev.set_raster_chunk(0)
    dev.set_tile_height(0)
FC = {None: 0, "back": 1, "front": 2}
DS = {None: 0, "front_to_back": 1, "back_to_front": 2}


def make_target(w, h, fmt, with_depth, seed=0, clear=None):
    rng = np.random.default_rng(seed)
    if fmt in (rf.PIX_U32_RGBA, rf.PIX_U32_XRGB, rf.PIX_U32_ARGB, rf.PIX_U32_BGRA):
        col = rng.integers(0, 2 ** 32, size=(h, w), dtype=np.uint32)
    elif fmt in (rf.PIX_RGBA4444, rf.PIX_RGB565):
        col = rng.integers(0, 2 ** 16, size=(h, w), dtype=np.uint16)
    else:
        col = rng.integers(0, 256, size=(h, w, 3 if fmt == rf.PIX_RGB8 else 4), dtype=np.uint8)
    dep = None
    if with_depth:
        dep = rng.uniform(0.0, 0.02, size=(h, w)).astype(np.float32)
    return col, dep


def pack_clear(fmt, c):
    r, g, b, a = c
    if fmt == rf.PIX_RGBA8:
        return np.array([r, g, b, a], np.uint8)
    if fmt == rf.PIX_RGB8:
        return np.array([r, g, b], np.uint8)
    if fmt == rf.PIX_U32_RGBA:
        return np.uint32((r << 24) | (g << 16) | (b << 8) | a)
    if fmt == rf.PIX_U32_ARGB:
        return np.uint32((a << 24) | (r << 16) | (g << 8) | b)
    if fmt == rf.PIX_U32_BGRA:
        return np.uint32((b << 24) | (g << 16) | (r << 8) | a)
    if fmt == rf.PIX_RGBA4444:
        return np.uint16(((r >> 4) << 12) | ((g >> 4) << 8) | ((b >> 4) << 4) | (a >> 4))
    if fmt == rf.PIX_RGB565:
        return np.uint16(((r >> 3) << 11) | ((g >> 2) << 5) | (b >> 3))
    return np.uint32((r << 16) | (g << 8) | b)


def run_both(kind, tris, pos, attr, mvp, vp, w, h, fmt=rf.PIX_RGBA8, with_depth=True, flat=(255, 255, 255, 255),
             tex=None, ctx=None, seed=0, calls=1):
    """Renders `calls` times into identical targets on both paths; returns
    (oracle colour, oracle depth, oracle stats), (gpu colour, gpu depth, gpu stats)."""
    ctx = ctx or rf.Context()
    col, dep = make_target(w, h, fmt, with_depth, seed)
    ocol, odep = col.copy(), (dep.copy() if dep is not None else None)
    gcol, gdep = col.copy(), (dep.copy() if dep is not None else None)
    if ctx.clear:
        if ctx.color_clear is not None:
            ocol[...] = pack_clear(fmt, ctx.color_clear)
        if odep is not None and ctx.depth_clear is not None:
            odep[...] = np.float32(1.0) / np.float32(ctx.depth_clear)
    ost = np.zeros(8, np.uint64)
    octx = (DT[ctx.depth_test], FC[ctx.face_cull], int(ctx.color_write), int(ctx.depth_write), DS[ctx.depth_sort])
    if kind == 0:
        shd = rf.shader.Flat(flat)
    elif kind == 1:
        shd = rf.shader.Gouraud()
    elif kind == 2:
        shd = rf.shader.TexClamp(tex)
    elif kind == 3:
        shd = rf.shader.TexRepeatPot(tex)
    else:
        shd = rf.shader.TexOnce(tex)
    gctx = rf.Context(color_clear=ctx.color_clear, depth_clear=ctx.depth_clear, face_cull=ctx.face_cull,
                      depth_test=ctx.depth_test, color_write=ctx.color_write, depth_write=ctx.depth_write,
                      clear=ctx.clear, depth_sort=ctx.depth_sort)
    target = rf.Framebuf(gcol, gdep, fmt) if with_depth else rf.Colorbuf(gcol, fmt)
    verts = (pos, attr) if attr is not None else pos
    for i in range(calls):
        err, _ = O.render(kind, tris, pos, attr, mvp, vp, ocol, depth=odep, fmt=fmt, flat=flat, tex=tex,
                          ctx=octx, stats=ost, prim=1 if np.asarray(tris).shape[-1] == 2 else 0)
        assert err == 0, f"oracle error {err}"
        rf.render(tris, verts, shd, mvp, vp, target, gctx)
        gctx.clear = False  # a frame clears once
    gst = gctx.stats
    g = np.array([gst.calls, gst.prims_i, gst.prims_o, gst.verts_i, gst.verts_o, gst.frags_i, gst.frags_o, 0],
                 np.uint64)
    return (ocol, odep, ost), (gcol, gdep, g)


def assert_same(o, g, color_tol=0):
    ocol, odep, ost = o
    gcol, gdep, gst = g
    if odep is not None:
        ob, gb = odep.view(np.uint32), gdep.view(np.uint32)
        bad = np.argwhere(ob != gb)
        assert bad.shape[0] == 0, f"depth differs at {bad.shape[0]} px, first {bad[:5].tolist()}"
    if color_tol == 0:
        bad = np.argwhere(ocol != gcol)
        assert bad.shape[0] == 0, f"colour differs at {bad.shape[0]} entries, first {bad[:5].tolist()}"
    else:
        d = np.abs(ocol.astype(np.int64) - gcol.astype(np.int64))
        assert d.max() <= color_tol
    assert ost[:7].tolist() == gst[:7].tolist(), f"stats oracle {ost[:7].tolist()} gpu {gst[:7].tolist()}"
