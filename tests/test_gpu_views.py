"""Targets and textures that are views into larger buffers (Slice2 /
MutSlice2, buf.rs:570-580): a bottom-right sub-rectangle of a parent image,
as a pitched SDL surface or an atlas cell hands it over. The view's pixels
match the oracle rendering into a packed copy of the view, and every byte of
the parent outside the view is left as it was, with and without the fused
clear."""
import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from parity import DT, FC, pack_clear
from retrofire_amd import mat as M
from test_gpu_parity import random_soup

pytestmark = pytest.mark.gpu

PH, PW = 150, 262      # parent image
Y0, X0 = 41, 97        # the view: rows Y0.., columns X0.. (reaches the parent's last row and column)


def _parent(fmt, rng):
    if fmt in (rf.PIX_U32_RGBA, rf.PIX_U32_XRGB, rf.PIX_U32_ARGB, rf.PIX_U32_BGRA):
        return rng.integers(0, 2 ** 32, size=(PH, PW), dtype=np.uint32)
    if fmt in (rf.PIX_RGBA4444, rf.PIX_RGB565):
        return rng.integers(0, 2 ** 16, size=(PH, PW), dtype=np.uint16)
    return rng.integers(0, 256, size=(PH, PW, 3 if fmt == rf.PIX_RGB8 else 4), dtype=np.uint8)


@pytest.mark.parametrize("fmt", [rf.PIX_RGBA8, rf.PIX_RGB8, rf.PIX_U32_XRGB, rf.PIX_RGB565])
@pytest.mark.parametrize("clear", [False, True])
@pytest.mark.parametrize("with_depth", [False, True])
def test_render_into_bottom_right_view(fmt, clear, with_depth):
    rng = np.random.default_rng(31 * fmt + 2 * clear + with_depth)
    kind = 3
    w, h = PW - X0, PH - Y0
    pos, attr, tris, mvp, vp = random_soup(rng, 200, w, h, kind)
    tex = rng.integers(0, 256, size=(32, 64, 4), dtype=np.uint8)
    parent = _parent(fmt, rng)
    pdep = rng.uniform(0.0, 0.02, size=(PH, PW)).astype(np.float32) if with_depth else None
    before, dbefore = parent.copy(), (pdep.copy() if with_depth else None)
    view = parent[Y0:, X0:]
    dview = pdep[Y0:, X0:] if with_depth else None
    assert not view.flags["C_CONTIGUOUS"]

    # oracle: a packed copy of the view
    ocol = np.ascontiguousarray(view).copy()
    odep = np.ascontiguousarray(dview).copy() if with_depth else None
    ctx = rf.Context(clear=clear, color_clear=(12, 34, 56, 78), depth_clear=float("inf"))
    if clear:
        ocol[...] = pack_clear(fmt, ctx.color_clear)
        if with_depth:
            odep[...] = np.float32(0.0)
    ost = np.zeros(8, np.uint64)
    err, _ = O.render(kind, tris, pos, attr, mvp, vp, ocol, depth=odep, fmt=fmt, tex=tex,
                      ctx=(DT["less"] if with_depth else 0, FC["back"], 1, 1), stats=ost)
    assert err == 0

    target = rf.Framebuf(view, dview, fmt) if with_depth else rf.Colorbuf(view, fmt)
    gctx = rf.Context(clear=clear, color_clear=(12, 34, 56, 78), depth_clear=float("inf"),
                      depth_test="less" if with_depth else None)
    rf.render(tris, (pos, attr), rf.shader.TexRepeatPot(tex), mvp, vp, target, gctx)

    assert np.array_equal(view, ocol), "view pixels differ from the oracle"
    if with_depth:
        assert np.array_equal(dview.view(np.uint32), odep.view(np.uint32)), "view depth differs"
    assert gctx.stats.frags_o == int(ost[6]) and gctx.stats.frags_o > 1000
    # nothing outside the view changed
    outside = np.ones((PH, PW), bool)
    outside[Y0:, X0:] = False
    assert np.array_equal(parent[outside], before[outside]), "bytes outside the view were written"
    if with_depth:
        assert np.array_equal(pdep[outside].view(np.uint32), dbefore[outside].view(np.uint32))


@pytest.mark.parametrize("kind", [2, 3, 4])
def test_texture_view_at_parent_bottom_right(kind):
    """A texture that is the last cell of an atlas (its last row is the
    parent's last row): only the cell's texels are read, and the image equals
    the oracle's with a packed copy of the cell."""
    rng = np.random.default_rng(77 + kind)
    atlas = rng.integers(0, 256, size=(48, 64, 4), dtype=np.uint8)
    cell = atlas[32:48, 48:64]
    w, h = 96, 80
    pos = np.array([(-0.9, -0.9, 0), (0.9, -0.9, 0), (-0.9, 0.9, 0), (0.9, 0.9, 0)], np.float32)
    uv = np.array([(0.01, 0.01), (0.99, 0.01), (0.01, 0.99), (0.99, 0.99)], np.float32)
    tris = np.array([[0, 2, 1], [1, 2, 3]], np.uint32)
    vp = M.viewport(0, h, w, 0)
    ocol = np.zeros((h, w, 4), np.uint8)
    err, _ = O.render(kind, tris, pos, uv, M.identity(), vp, ocol, tex=np.ascontiguousarray(cell), ctx=(0, 0, 1, 1))
    assert err == 0
    shd = {2: rf.shader.TexClamp, 3: rf.shader.TexRepeatPot, 4: rf.shader.TexOnce}[kind](cell)
    if kind != 4:
        shd.texture = cell  # keep the view (the Clamp and RepeatPot wrappers pack it)
    gcol = np.zeros((h, w, 4), np.uint8)
    rf.render(tris, (pos, uv), shd, M.identity(), vp, rf.Colorbuf(gcol), rf.Context(face_cull=None, depth_test=None))
    assert np.array_equal(gcol, ocol)
