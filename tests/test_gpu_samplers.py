"""SamplerOnce (tex.rs:313-357) and atlas sub-textures (Atlas::get,
tex.rs:157-172) on the HIP path against the oracle: colour bytes, depth bits
and Stats identical; a texel outside the texture is an error on both."""
import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from parity import TILE_CONFIGS, TILE_IDS, assert_same, run_both, use_tile_config
from retrofire_amd import batch as RB
from retrofire_amd import mat as M
from test_gpu_parity import random_soup

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=TILE_CONFIGS, ids=TILE_IDS)
def tile_config(request):
    """Every case runs with each rasterizer configuration."""
    yield from use_tile_config(request.param)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sampler_once_soup(seed):
    rng = np.random.default_rng(900 + seed)
    w, h = 301, 197
    pos, attr, tris, mvp, vp = random_soup(rng, 300, w, h, 2)
    attr = rng.uniform(0.05, 0.95, size=attr.shape).astype(np.float32)  # SamplerOnce needs 0..1
    tex = rng.integers(0, 256, size=(24, 40, 4), dtype=np.uint8)
    o, g = run_both(4, tris, pos, attr, mvp, vp, w, h, tex=tex, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    assert g[2][6] > 10000


def test_sampler_once_out_of_range_is_an_error():
    w, h = 64, 48
    pos = np.array([(-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0)], np.float32)
    uv = np.array([(0, 0), (1.5, 0), (0, 1), (1.5, 1)], np.float32)
    tris = np.array([[0, 2, 1], [1, 2, 3]], np.uint32)
    tex = np.zeros((8, 8, 4), np.uint8)
    fb = np.zeros((h, w, 4), np.uint8)
    err, _ = O.render(4, tris, pos, uv, M.identity(), M.viewport(0, h, w, 0), fb.copy(), tex=tex, ctx=(0, 0, 1, 1))
    assert err == -2
    with pytest.raises(rf.RenderError) as e:
        rf.render(tris, (pos, uv), rf.shader.TexOnce(tex), M.identity(), M.viewport(0, h, w, 0), rf.Colorbuf(fb),
                  rf.Context(face_cull=None))
    assert e.value.status == -2
    # with colour writes off the shader still runs for every passing fragment
    with pytest.raises(rf.RenderError) as e:
        rf.render(tris, (pos, uv), rf.shader.TexOnce(tex), M.identity(), M.viewport(0, h, w, 0),
                  rf.Framebuf(fb, np.zeros((h, w), np.float32)), rf.Context(face_cull=None, color_write=False))
    assert e.value.status == -2


def _atlas():
    """A 64 x 48 atlas of 16 x 16 cells (Layout::Grid), each cell a distinct
    pattern."""
    rng = np.random.default_rng(4)
    return rng.integers(0, 256, size=(48, 64, 4), dtype=np.uint8)


def _cell(i, cols=4, cw=16, ch=16):
    """Atlas::rect for a Grid layout (tex.rs:146-155): top-left corner."""
    return (i % cols) * cw, (i // cols) * ch


def test_atlas_cells_plain_and_resident():
    """Quads textured with atlas cells through every sampler, from numpy views
    (plain path) and from resident texture views (batched path)."""
    atlas = _atlas()
    w, h = 200, 120
    vp = M.viewport(0, h, w, 0)
    rng = np.random.default_rng(5)
    quads = []
    for k in range(12):
        x0, y0 = rng.uniform(-1, 0.5, 2)
        s = rng.uniform(0.2, 0.6)
        pos = np.array([(x0, y0, 0), (x0 + s, y0, 0), (x0, y0 + s, 0), (x0 + s, y0 + s, 0)], np.float32)
        lo, hi = (0.02, 0.98) if k % 3 == 2 else (-0.3, 1.4)
        uv = np.array([(lo, lo), (hi, lo), (lo, hi), (hi, hi)], np.float32)
        quads.append((pos, uv, k % 12, [2, 3, 4][k % 3]))
    tris = np.array([[0, 2, 1], [1, 2, 3]], np.uint32)
    # oracle and plain GPU path, one render() call per quad
    ocol = np.zeros((h, w, 4), np.uint8)
    gcol = ocol.copy()
    ctx = rf.Context(face_cull=None, depth_test=None)
    shaders = {2: rf.shader.TexClamp, 3: rf.shader.TexRepeatPot, 4: rf.shader.TexOnce}
    for pos, uv, cell, kind in quads:
        cx, cy = _cell(cell)
        view = atlas[cy:cy + 16, cx:cx + 16]
        err, _ = O.render(kind, tris, pos, uv, M.identity(), vp, ocol, tex=view, ctx=(0, 0, 1, 1))
        assert err == 0
        rf.render(tris, (pos, uv), shaders[kind](view), M.identity(), vp, rf.Colorbuf(gcol), ctx)
    assert np.array_equal(ocol, gcol)
    # batched path: one resident atlas, one view per cell
    tex = RB.Texture(atlas)
    views = {}
    draws, meshes = [], []
    for pos, uv, cell, kind in quads:
        if cell not in views:
            views[cell] = tex.view(*_cell(cell), 16, 16)
        m = RB.Mesh(pos, tris, uv)
        meshes.append(m)
        draws.append(RB.mesh_draw(m, shaders[kind](atlas), M.identity(), vp, texture=views[cell]))
    fb = RB.DeviceFramebuf(w, h, 0, False)
    b = RB.Batch([(fb, rf.Context(clear=True, color_clear=(0, 0, 0, 0), face_cull=None, depth_test=None), draws)])
    b.run(sync=True)
    b.wait()
    assert np.array_equal(fb.read()[0], ocol)
    b.free()
    fb.free()
    for m in meshes:
        m.free()
    for v in views.values():
        v.free()
    tex.free()
