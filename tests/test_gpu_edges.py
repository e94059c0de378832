"""Edges (Edge<usize> primitives, prim.rs:41-60): edge-list clipping
(clip.rs:311-348) and raster::line (raster.rs:122-177) on the HIP path,
against the oracle on the same seeded inputs: colour bytes, depth bits and
Stats identical, at both rasterizer tile widths."""
import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from parity import assert_same, run_both, use_tile_config
from retrofire_amd import batch as RB
from retrofire_amd import mat as M
from retrofire_amd import scenes as S

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[(16, 8, 32), (32, 8, 32), (64, 8, 32), (32, 6, 16), (16, 8, 64)],
                ids=["tw16", "tw32", "tw64", "tw32c6h16", "tw16h64"])
def tile_config(request):
    yield from use_tile_config(request.param)


def edge_soup(rng, n, w, h, kind, spread=3.0):
    """Random edges in front of, across and behind the camera (near-plane
    and side clipping), of every slope."""
    pos = rng.uniform([-spread, -spread, -1.5], [spread, spread, 6], size=(2 * n, 3)).astype(np.float32)
    edges = np.arange(2 * n, dtype=np.uint32).reshape(n, 2)
    attr = None
    if kind == 1:
        attr = rng.uniform(0, 1, size=(2 * n, 3)).astype(np.float32)
    elif kind >= 2:
        attr = rng.uniform(-1.5, 2.5, size=(2 * n, 2)).astype(np.float32)
    mvp = M.then(M.translate3(0, 0, 2.0), M.perspective(1.2, np.float32(w) / np.float32(h), 0.1, 8.0))
    return pos, attr, edges, mvp, M.viewport(0, h, w, 0)


CTXS = [
    dict(depth_test="less", face_cull="back"),
    dict(depth_test=None, face_cull=None),
    dict(depth_test="greater", face_cull="front"),  # no edge is a backface: FaceCull::Front drops them all
    dict(depth_test="equal", face_cull=None),
    dict(depth_test="less", face_cull=None, color_write=False),
    dict(depth_test="less", face_cull=None, depth_write=False),
]


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("ci", range(len(CTXS)))
def test_edge_soup_contexts(kind, ci):
    rng = np.random.default_rng(700 + 10 * kind + ci)
    w, h = 301, 197
    pos, attr, edges, mvp, vp = edge_soup(rng, 400, w, h, kind)
    tex = None
    if kind >= 2:
        tex = rng.integers(0, 256, size=(32, 64 if kind == 3 else 37, 4), dtype=np.uint8)
    ctx = rf.Context(clear=(ci % 2 == 0), **CTXS[ci])
    o, g = run_both(kind, edges, pos, attr, mvp, vp, w, h, tex=tex, flat=(9, 200, 77, 255), ctx=ctx, seed=ci,
                    calls=2)
    assert_same(o, g)
    if ci == 2:
        assert g[2][2] == 0 and g[2][6] == 0
    else:
        assert g[2][2] > 100 and g[2][5] > 1000


def test_long_lines_across_many_tiles():
    """Lines spanning the whole 1920x1080 target at every angle: the stepped
    coordinate enters each tile through the exact fast-forward."""
    w, h = 1920, 1080
    n = 720
    t = np.arange(n, dtype=np.float64) / n * 2 * np.pi
    c = np.stack([np.cos(t), np.sin(t)], 1)
    a = (c * 0.99).astype(np.float32)
    b = (-c * np.array([0.97, 0.93])).astype(np.float32)
    pos = np.concatenate([np.c_[a, np.zeros(n)], np.c_[b, np.full(n, 0.5)]]).astype(np.float32)
    edges = np.stack([np.arange(n), np.arange(n) + n], 1).astype(np.uint32)
    attr = np.random.default_rng(3).uniform(0, 1, size=(2 * n, 3)).astype(np.float32)
    o, g = run_both(1, edges, pos, attr, M.identity(), M.viewport(0, h, w, 0), w, h,
                    ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    assert g[2][5] > 500000


def test_wireframe_over_triangles_one_launch():
    """A sphere and then its wireframe, as two draws of one batched frame:
    lines and triangles meet in the same tiles in submission order."""
    sc = S.sphere640()
    tris = sc.tris
    ed = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]).astype(np.uint32)
    col = np.random.default_rng(9).uniform(0, 1, size=(sc.pos.shape[0], 3)).astype(np.float32)
    w, h = sc.width, sc.height
    # oracle: two render() calls into one frame
    ocol = np.zeros((h, w, 4), np.uint8)
    ocol[...] = (0, 0, 0, 255)
    odep = np.zeros((h, w), np.float32)
    st = np.zeros(8, np.uint64)
    e1, _ = O.render(sc.kind, tris, sc.pos, sc.attr, sc.mvp, sc.viewport, ocol, depth=odep, tex=sc.tex,
                     flat=sc.flat, stats=st)
    e2, _ = O.render(1, ed, sc.pos, col, sc.mvp, sc.viewport, ocol, depth=odep, stats=st, prim=1)
    assert e1 == 0 and e2 == 0
    mesh = RB.Mesh(sc.pos, tris, sc.attr)
    wire = RB.Mesh(sc.pos, ed, col)
    shd = rf.shader.Flat(sc.flat) if sc.kind == 0 else rf.shader.Gouraud()
    fb = RB.DeviceFramebuf(w, h, 0, True)
    draws = [RB.mesh_draw(mesh, shd, sc.mvp, sc.viewport), RB.mesh_draw(wire, rf.shader.Gouraud(), sc.mvp, sc.viewport)]
    for tw in (16, 32, 64):
        b = RB.Batch([(fb, rf.Context(clear=True), draws)], tile_w=tw)
        b.run(sync=True)
        s = b.wait()[0]
        gcol, gdep = fb.read()
        assert np.array_equal(gcol, ocol)
        assert np.array_equal(gdep.view(np.uint32), odep.view(np.uint32))
        assert [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o] == \
            [int(x) for x in st[:7]]
        b.free()
    fb.free()
    mesh.free()
    wire.free()


def test_edge_special_cases():
    w, h = 64, 48
    vp = M.viewport(0, h, w, 0)
    I = M.identity()
    pos = np.array([
        (-0.5, -0.5, 0), (-0.5, -0.5, 0),        # zero length
        (-0.9, 0.1, 0), (0.9, 0.1, 0),           # horizontal
        (0.2, -0.9, 0), (0.2, 0.9, 0),           # vertical
        (-0.9, -0.9, 0), (0.9, 0.9, 0),          # 45 degrees
        (-3, 0.3, 0), (3, -0.2, 0),              # clipped both sides
        (0.1, 0.1, -3), (0.1, 0.1, 3),           # through the near and far planes
        (2, 2, 0), (3, 3, 0),                    # outside
        (-0.99, 0.5, 0.5), (0.99, 0.52, 0.5),   # almost horizontal, full width
    ], np.float32)
    edges = np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 2)
    for kind, attr in ((0, None), (1, np.random.default_rng(1).uniform(0, 1, (pos.shape[0], 3)).astype(np.float32))):
        for ctx in (rf.Context(clear=True, face_cull=None), rf.Context(clear=False, depth_test=None)):
            o, g = run_both(kind, edges, pos, attr, I, vp, w, h, ctx=ctx)
            assert_same(o, g)
            o, g = run_both(kind, edges, pos, attr, I, vp, w, h, ctx=ctx, with_depth=False)
            assert_same(o, g)


def test_edge_errors():
    fb = np.zeros((8, 8, 4), np.uint8)
    pos = np.array([(-0.5, -0.5, 0), (0.5, 0.5, 0)], np.float32)
    with pytest.raises(rf.RenderError) as e:
        rf.render(np.array([[0, 2]]), pos, rf.shader.Flat(), M.identity(), M.viewport(0, 0, 8, 8),
                  rf.Colorbuf(fb), rf.Context())
    assert e.value.status == -1
    # a wide line whose last pixel centre lands exactly on its end at the
    # bottom edge (screen (0.5, 6) to (4.5, 8), slope 1/2): row 8 is the
    # target's height, so the reference indexes past the buffer and panics
    pos = np.array([(-0.875, -0.5, 0), (0.125, -1.0, 0)], np.float32)
    ob = np.zeros((8, 8, 4), np.uint8)
    err, _ = O.render(0, np.array([[0, 1]], np.uint32), pos, None, M.identity(), M.viewport(0, 8, 8, 0), ob,
                      prim=1)
    assert err == -2
    with pytest.raises(rf.RenderError) as e:
        rf.render(np.array([[0, 1]]), pos, rf.shader.Flat(), M.identity(), M.viewport(0, 8, 8, 0),
                  rf.Colorbuf(fb), rf.Context(face_cull=None))
    assert e.value.status == -2


@pytest.mark.parametrize("fmt", [rf.PIX_RGB8, rf.PIX_U32_XRGB, rf.PIX_U32_BGRA, rf.PIX_RGBA4444, rf.PIX_RGB565])
def test_edges_pixel_formats(fmt):
    rng = np.random.default_rng(31)
    w, h = 157, 93
    pos, attr, edges, mvp, vp = edge_soup(rng, 250, w, h, 1)
    for ctx, with_depth in ((rf.Context(clear=True, face_cull=None), True), (rf.Context(clear=False), False)):
        o, g = run_both(1, edges, pos, attr, mvp, vp, w, h, fmt=fmt, with_depth=with_depth, ctx=ctx)
        assert_same(o, g)
