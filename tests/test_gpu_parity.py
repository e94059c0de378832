"""Parity of the HIP path with the oracle on the same seeded inputs: colour
bytes, depth bits and Stats must be identical (the tolerance the brief allows
for colour is +-1 LSB per 8-bit channel; the tests use 0)."""
import os

import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from parity import TILE_CONFIGS, TILE_IDS, assert_same, run_both, use_tile_config
from retrofire_amd import batch as RB
from retrofire_amd import mat as M
from retrofire_amd import scenes as S
from test_oracle_kats import GOLDEN, hello_tri_inputs, read_ppm, textured_quad_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=TILE_CONFIGS, ids=TILE_IDS)
def tile_config(request):
    """Every case runs with each rasterizer configuration."""
    yield from use_tile_config(request.param)


def test_textured_quad_golden_gpu():  # core/tests/rendering.rs:18-60
    verts, uvs, faces, tex, mvp, vp = textured_quad_inputs()
    fb = np.zeros((256, 256, 3), np.uint8)
    ctx = rf.Context()
    rf.render(faces, (verts, uvs), rf.shader.TexClamp(tex), mvp, vp, rf.Colorbuf(fb), ctx)
    np.testing.assert_array_equal(fb, read_ppm(os.path.join(GOLDEN, "textured_quad.ppm")))
    assert ctx.stats.frags_i == 256 * 256 and ctx.stats.prims_o == 2


def test_hello_tri_gpu():  # core/examples/hello_tri.rs:47-53
    verts, cols, faces, mvp, vp = hello_tri_inputs()
    fb = np.zeros((480, 640, 4), np.uint8)
    rf.render(faces, (verts, cols), rf.shader.Gouraud(), mvp, vp, rf.Colorbuf(fb), rf.Context())
    assert tuple(fb[240, 320]) == (114, 102, 128, 255)
    gold = read_ppm(os.path.join(GOLDEN, "triangle.ppm"))
    np.testing.assert_array_equal(fb[..., 3] == 255, gold.any(axis=2))


def _scene_parity(sc, ctx=None, with_depth=True, fmt=rf.PIX_RGBA8):
    ctx = ctx or rf.Context(clear=True)
    o, g = run_both(sc.kind, sc.tris, sc.pos, sc.attr, sc.mvp, sc.viewport, sc.width, sc.height, fmt=fmt,
                    with_depth=with_depth, tex=sc.tex, flat=sc.flat, ctx=ctx)
    assert_same(o, g)
    return o, g


def test_sphere640_parity():
    o, g = _scene_parity(S.sphere640())
    assert g[2][6] > 50000  # the visible hemisphere is drawn


def test_crates1080_parity():
    o, g = _scene_parity(S.crates1080())
    assert g[2][6] > 100000


def test_icospheres_full_parity():
    """The 10,000-icosphere scene at full size (3.2M triangles, 1920x1080)."""
    o, g = _scene_parity(S.icospheres1080())
    assert g[2][1] == 3200000 and g[2][6] > 1000000


def test_icospheres_parity_reduced():
    sc = S.icospheres1080(n=1500)
    o, g = _scene_parity(sc)
    assert g[2][6] > 10000


@pytest.mark.parametrize("fmt", [rf.PIX_RGBA8, rf.PIX_RGB8, rf.PIX_U32_RGBA, rf.PIX_U32_XRGB, rf.PIX_U32_ARGB,
                                 rf.PIX_U32_BGRA, rf.PIX_RGBA4444, rf.PIX_RGB565])
def test_pixel_formats(fmt):
    _scene_parity(S.sphere640(), fmt=fmt, ctx=rf.Context(clear=True))
    _scene_parity(S.sphere640(), fmt=fmt, ctx=rf.Context(clear=False), with_depth=False)


CTXS = [
    dict(depth_test="less", face_cull="back"),
    dict(depth_test=None, face_cull=None),
    dict(depth_test="greater", face_cull="front"),
    dict(depth_test="equal", face_cull=None),
    dict(depth_test="less", face_cull=None, color_write=False),
    dict(depth_test="less", face_cull=None, depth_write=False),
]


def random_soup(rng, n, w, h, kind):
    """Random overlapping triangles in front of and across the camera (some
    crossing the near plane, some outside the frustum)."""
    pos = rng.uniform([-3, -2, -1.5], [3, 2, 6], size=(3 * n, 3)).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    attr = None
    if kind == 1:
        attr = rng.uniform(0, 1, size=(3 * n, 3)).astype(np.float32)
    elif kind >= 2:
        attr = rng.uniform(-1.5, 2.5, size=(3 * n, 2)).astype(np.float32)
    mvp = M.then(M.translate3(0, 0, 2.0), M.perspective(1.2, np.float32(w) / np.float32(h), 0.1, 8.0))
    vp = M.viewport(0, h, w, 0)
    return pos, attr, tris, mvp, vp


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("ci", range(len(CTXS)))
def test_random_soup_contexts(kind, ci):
    rng = np.random.default_rng(100 * kind + ci)
    w, h = 301, 197
    pos, attr, tris, mvp, vp = random_soup(rng, 300, w, h, kind)
    tex = None
    if kind >= 2:
        tw = 64 if kind == 3 else 37
        tex = rng.integers(0, 256, size=(32, tw, 4), dtype=np.uint8)
    ctx = rf.Context(clear=(ci % 2 == 0), **CTXS[ci])
    o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, flat=(9, 200, 77, 255), ctx=ctx, seed=ci,
                    calls=2)
    assert_same(o, g)


def test_many_calls_same_target_ties():
    """Several render() calls into one target: equal depths keep the first
    drawn fragment (ctx.rs:86-89)."""
    sc = S.sphere640()
    o, g = run_both(sc.kind, sc.tris, sc.pos, sc.attr, sc.mvp, sc.viewport, sc.width, sc.height,
                    ctx=rf.Context(clear=True), calls=3)
    assert_same(o, g)


def test_viewport_offset_and_small_target():
    sc = S.crates1080(n_crates=300)
    w, h = 500, 300
    vp = M.viewport(10, h - 10, w - 10, 10)  # as the crates demo (crates.rs:53)
    o, g = run_both(sc.kind, sc.tris, sc.pos, sc.attr, sc.mvp, vp, w, h, tex=sc.tex, ctx=rf.Context(clear=True))
    assert_same(o, g)


def test_empty_and_degenerate():
    w, h = 64, 48
    vp = M.viewport(0, h, w, 0)
    mvp = M.then(M.translate3(0, 0, 3), M.perspective(1.0, 1.0, 0.1, 10.0))
    pos = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 1, 0], [1, 1, 0], [0.5, 0.2, 0], [-1, -1, 0],
                    [1, -1, 0], [0, 1, 0]], np.float32)
    tris = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8], [6, 8, 7]], np.uint32)
    o, g = run_both(0, tris, pos, None, mvp, vp, w, h, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    o, g = run_both(0, tris[:0], pos, None, mvp, vp, w, h, ctx=rf.Context(clear=True))
    assert_same(o, g)


def test_errors():
    fb = np.zeros((8, 8, 4), np.uint8)
    pos = np.zeros((3, 3), np.float32)
    with pytest.raises(rf.RenderError) as e:
        rf.render(np.array([[0, 1, 3]]), pos, rf.shader.Flat(), M.identity(), M.viewport(0, 0, 8, 8),
                  rf.Colorbuf(fb), rf.Context())
    assert e.value.status == -1
    tex = np.zeros((3, 5, 4), np.uint8)
    with pytest.raises(rf.RenderError) as e:
        rf.render(np.array([[0, 1, 2]]), (pos, np.zeros((3, 2), np.float32)), rf.shader.TexRepeatPot(tex),
                  M.identity(), M.viewport(0, 0, 8, 8), rf.Colorbuf(fb), rf.Context())
    assert e.value.status == -3
    with pytest.raises(rf.RenderError) as e:  # index errors also under a depth sort
        rf.render(np.array([[0, 1, 3]]), pos, rf.shader.Flat(), M.identity(), M.viewport(0, 0, 8, 8),
                  rf.Colorbuf(fb), rf.Context(depth_sort="back_to_front"))
    assert e.value.status == -1
    verts = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    with pytest.raises(rf.RenderError) as e:  # viewport larger than the target (target.rs:173 panics)
        rf.render(np.array([[0, 1, 2]]), verts, rf.shader.Flat(), M.identity(), M.viewport(0, 0, 16, 16),
                  rf.Colorbuf(fb), rf.Context(face_cull=None))
    assert e.value.status == -2


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_huge_triangles_long_spans(kind):
    """Full-screen perspective geometry at 2048x2048: rows thousands of steps
    below a triangle's top and spans thousands of pixels long, so every tile
    enters its spans through the exact fast-forward of the serial sums."""
    rng = np.random.default_rng(40 + kind)
    w = h = 2048
    n = 24
    pos = np.concatenate([rng.uniform([-30, -30, -2], [30, 30, 40], size=(3 * n, 3))]).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    attr = rng.uniform(0, 1, size=(3 * n, 3)).astype(np.float32) if kind == 1 else \
        rng.uniform(-3, 4, size=(3 * n, 2)).astype(np.float32)
    tex = rng.integers(0, 256, size=(64, 128, 4), dtype=np.uint8) if kind >= 2 else None
    mvp = M.then(M.translate3(0, 0, 1.0), M.perspective(0.7, 1.0, 0.1, 100.0))
    vp = M.viewport(0, h, w, 0)
    ctx = rf.Context(clear=True, face_cull=None)
    o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, ctx=ctx)
    assert_same(o, g)
    assert g[2][6] > w * h // 4


def test_tall_thin_fan():
    """Slivers spanning the full height (row counts >> 64) with tiny widths."""
    w, h = 300, 2000
    n = 60
    xs = np.linspace(-1, 1, n + 1).astype(np.float32)
    pos = []
    for i in range(n):
        pos += [[xs[i], -1.2, 0.5], [xs[i + 1], -1.2, 0.5], [(xs[i] + xs[i + 1]) / 2, 1.2, 0.3]]
    pos = np.array(pos, np.float32)
    cols = np.random.default_rng(3).uniform(0, 1, size=(pos.shape[0], 3)).astype(np.float32)
    tris = np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3)
    mvp = M.then(M.translate3(0, 0, 2.0), M.perspective(1.0, np.float32(w) / np.float32(h), 0.1, 10.0))
    o, g = run_both(1, tris, pos, cols, mvp, M.viewport(0, h, w, 0), w, h, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)


@pytest.mark.parametrize("n", [3000, 30000, 80000])
def test_many_groups_in_one_tile(n):
    """Many 64-triangle groups touching the same few tiles: the per-tile run
    lists (one run per group) are ordered in registers (<= 64 runs), in LDS
    (<= 1024) and in place in global memory (more); any misordering shows as
    a different winner of overlapping fragments."""
    rng = np.random.default_rng(n)
    w, h = 64, 48
    c = rng.uniform([20, 14, 0.2], [44, 34, 0.9], size=(n, 1, 3)).astype(np.float32)
    pos = (c + rng.uniform([-6, -6, -0.05], [6, 6, 0.05], size=(n, 3, 3)).astype(np.float32)).reshape(-1, 3)
    # screen-space coordinates through an orthographic mvp: x, y in pixels, z in (0, 1)
    mvp = M.then(M.scale3(2.0 / w, 2.0 / h, 1.0), M.translate3(-1.0, -1.0, 0.0))
    vp = M.viewport(0, h, w, 0)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    cols = rng.uniform(0, 1, size=(3 * n, 3)).astype(np.float32)
    o, g = run_both(1, tris, pos, cols, mvp, vp, w, h, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    assert g[2][6] > 500


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_extreme_attributes(kind):
    """Varyings far outside their usual range (huge, infinite, NaN, negative
    zero): the samplers' and the Gouraud shader's float-to-integer casts
    saturate and map NaN to 0 as Rust's `as` does."""
    rng = np.random.default_rng(40 + kind)
    w, h = 96, 64
    pos, attr, tris, mvp, vp = random_soup(rng, 200, w, h, kind)
    specials = np.array([1e20, -1e20, np.inf, -np.inf, np.nan, -0.0, 3e9, -3e9, 1e-30, 300.0, -300.0],
                        np.float32)
    pick = rng.integers(0, len(specials), size=attr.shape)
    mask = rng.uniform(size=attr.shape) < 0.3
    attr = np.where(mask, specials[pick], attr).astype(np.float32)
    tex = None
    if kind >= 2:
        tex = rng.integers(0, 256, size=(16, 32, 4), dtype=np.uint8)
    o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    assert g[2][6] > 0


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("sort", ["front_to_back", "back_to_front"])
@pytest.mark.parametrize("dt", [None, "less"])
def test_depth_sort(kind, sort, dt):
    """render() with depth_sort (render.rs:180-183, 209-219): clipped
    triangles (near-plane fans included) drawn in depth order."""
    rng = np.random.default_rng(50 + kind)
    w, h = 301, 197
    pos, attr, tris, mvp, vp = random_soup(rng, 300, w, h, kind)
    tex = None
    if kind >= 2:
        tex = rng.integers(0, 256, size=(32, 64 if kind == 3 else 37, 4), dtype=np.uint8)
    ctx = rf.Context(clear=True, depth_test=dt, face_cull=None, depth_sort=sort)
    o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, flat=(9, 200, 77, 255), ctx=ctx, calls=2)
    assert_same(o, g)
    assert g[2][2] > 200


def test_depth_sort_crates_and_edges():
    sc = S.crates1080(n_crates=200)
    for sort in ("front_to_back", "back_to_front"):
        ctx = rf.Context(clear=True, depth_test=None, depth_sort=sort)
        o, g = run_both(sc.kind, sc.tris, sc.pos, sc.attr, sc.mvp, sc.viewport, sc.width, sc.height, tex=sc.tex,
                        ctx=ctx)
        assert_same(o, g)
    # edges keep their order under a depth sort (every Render::depth is +inf)
    rng = np.random.default_rng(8)
    pos = rng.uniform(-1, 1, size=(200, 3)).astype(np.float32)
    edges = np.arange(200, dtype=np.uint32).reshape(100, 2)
    col = rng.uniform(0, 1, size=(200, 3)).astype(np.float32)
    o, g = run_both(1, edges, pos, col, M.identity(), M.viewport(0, 60, 80, 0), 80, 60,
                    ctx=rf.Context(clear=True, depth_test=None, depth_sort="back_to_front"))
    assert_same(o, g)


@pytest.mark.parametrize("fmt", [rf.PIX_RGB8, rf.PIX_U32_ARGB, rf.PIX_RGB565])
def test_depth_sort_pixel_formats(fmt):
    rng = np.random.default_rng(77)
    w, h = 211, 149
    pos, attr, tris, mvp, vp = random_soup(rng, 200, w, h, 1)
    ctx = rf.Context(clear=True, depth_test=None, face_cull=None, depth_sort="back_to_front")
    o, g = run_both(1, tris, pos, attr, mvp, vp, w, h, fmt=fmt, ctx=ctx)
    assert_same(o, g)
    o, g = run_both(1, tris, pos, attr, mvp, vp, w, h, fmt=fmt, with_depth=False,
                    ctx=rf.Context(clear=False, depth_sort="front_to_back"))
    assert_same(o, g)


@pytest.mark.parametrize("w,h", [(16384, 256), (200, 16384)])
def test_largest_target_sides(w, h):
    """Targets at the largest side (16,384 pixels): the widest and tallest
    tile grids, record x packed in 16 bits."""
    sc = S.sphere640()
    vp = M.viewport(0, 0, w, h)
    o, g = run_both(sc.kind, sc.tris, sc.pos, sc.attr, sc.mvp, vp, w, h, ctx=rf.Context(clear=True))
    assert_same(o, g)
    assert g[2][6] > 1000


@pytest.mark.parametrize("above", [40, 700])
@pytest.mark.parametrize("kind", [1, 3])
def test_rows_above_the_target_fold_onto_row_0(above, kind):
    """A viewport reaching `above` rows over the target: every scanline above
    it is drawn onto row 0 (`self.y as usize` saturates, raster.rs:106), in
    row order, each with its own span, depth test and stats. 700 rows is more
    than one rasterizer batch of edge records, so a triangle's folded rows go
    in several parts."""
    rng = np.random.default_rng(60 + above + kind)
    w, h = 257, 180
    pos, attr, tris, mvp, _ = random_soup(rng, 250, w, h, kind)
    tex = rng.integers(0, 256, size=(32, 64, 4), dtype=np.uint8) if kind == 3 else None
    vp = M.viewport(0, h, w, -above)
    for clear in (True, False):
        o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, ctx=rf.Context(clear=clear, face_cull=None))
        assert_same(o, g)
    assert g[2][5] > 10000


def test_rows_entirely_above_the_target():
    """Triangles whose every row lies above the target still draw their
    folded spans onto row 0."""
    w, h = 120, 40
    pos = np.array([[-0.9, 2.0, 0.5], [0.8, 2.5, 0.5], [0.1, 4.0, 0.5], [-0.5, 1.2, 0.2], [0.9, 1.4, 0.2],
                    [0.0, 3.0, 0.2]], np.float32)
    tris = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    cols = np.random.default_rng(1).uniform(0, 1, size=(6, 3)).astype(np.float32)
    mvp = M.scale3(1.0, 0.2, 1.0)  # y in NDC 0.24 .. 0.8, viewport maps NDC y 0 .. 1 to rows -200 .. -400
    vp = M.viewport(0, 0, w, -400)
    o, g = run_both(1, tris, pos, cols, mvp, vp, w, h, ctx=rf.Context(clear=True, face_cull=None))
    assert_same(o, g)
    assert g[2][6] > 0 and (g[0][1:] == (0, 0, 0, 255)).all()  # only row 0 is drawn


def test_w_zero_vertex():
    """A clip-space vertex (0, 0, 0, 0) lies on every frustum plane, so it is
    inside and unclipped (clip.rs:101-111); its screen position is 0/0 = NaN.
    The reference sorts the NaN y first (x86-64's NaN from 0/0 has its sign
    set, total_cmp), rows whose left edge is NaN become spans 0..x1 with no
    fragment (`as usize` of NaN is 0, `as u32` count 0) that still count in
    frags.i (raster.rs:104-107, target.rs:176)."""
    w, h = 160, 120
    # w' = z: the vertex at the origin maps to (0, 0, 0, 0)
    mvp = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 1, 0]], np.float32)
    rng = np.random.default_rng(5)
    pos = [[0, 0, 0], [0.5, 0.2, 1.0], [-0.3, 0.6, 1.0], [0.7, -0.4, 1.0], [0, 0, 0], [-0.6, -0.5, 1.0]]
    pos += rng.uniform([-0.8, -0.8, 0.9], [0.8, 0.8, 1.0], size=(30, 3)).tolist()
    pos = np.array(pos, np.float32)
    tris = np.array([[0, 1, 2], [1, 0, 3], [3, 4, 5], [5, 2, 0]] + [[6 + 3 * i, 7 + 3 * i, 8 + 3 * i] for i in range(10)],
                    np.uint32)
    cols = rng.uniform(0, 1, size=(pos.shape[0], 3)).astype(np.float32)
    vp = M.viewport(0, h, w, 0)
    for kind, attr in ((1, cols), (2, cols[:, :2].copy())):
        tex = rng.integers(0, 256, size=(8, 8, 4), dtype=np.uint8) if kind == 2 else None
        o, g = run_both(kind, tris, pos, attr, mvp, vp, w, h, tex=tex, ctx=rf.Context(clear=True, face_cull=None))
        assert_same(o, g)
        assert g[2][5] > g[2][6]  # the NaN rows count in frags.i without writing


_DRIFT = {}


def _drift_case():
    """Thin triangles on a 16384 x 16384 target whose long, nearly vertical
    edges at x ~ 16,000 step by just over half an ulp of x per row: every row
    step rounds up by one ulp, so after 16k rows the stepped edge runs pixels
    past its end vertex (raster.rs:84-113)."""
    if "o" not in _DRIFT:
        w = h = 16384
        verts, tris = [], []
        ulp = 2.0 ** -10  # x in [8192, 16384)
        for i, (x, slope) in enumerate([(16000.25, 0.51), (15800.75, -0.52), (9000.5, 0.55), (12000.1, -0.6)]):
            dl = slope * ulp
            top = (x, 0.3)
            bot = (x + dl * 16383.0, 16383.4)
            side = (x - 30.0 if slope > 0 else x + 30.0, 9000.2)
            verts += [top, bot, side]
            tris.append([3 * i, 3 * i + 1, 3 * i + 2])
        # screen coordinates through an orthographic mvp: x, y in pixels
        pos = np.array([(x, y, 0.5) for (x, y) in verts], np.float32)
        mvp = M.then(M.scale3(2.0 / w, 2.0 / h, 1.0), M.translate3(-1.0, -1.0, 0.0))
        vp = M.viewport(0, 0, w, h)
        tris = np.array(tris, np.uint32)
        cols = np.random.default_rng(9).uniform(0, 1, size=(pos.shape[0], 3)).astype(np.float32)
        ocol = np.zeros((h, w, 4), np.uint8)
        ocol[...] = (0, 0, 0, 255)
        ost = np.zeros(8, np.uint64)
        err, _ = O.render(1, tris, pos, cols, mvp, vp, ocol, ctx=(0, 0, 1, 1), stats=ost)
        assert err == 0
        _DRIFT.update(o=(ocol, ost), inputs=(tris, pos, cols, mvp, vp, w, h))
    return _DRIFT


def test_long_edges_drift_past_their_vertices():
    d = _drift_case()
    tris, pos, cols, mvp, vp, w, h = d["inputs"]
    ocol, ost = d["o"]
    gcol = np.zeros((h, w, 4), np.uint8)
    ctx = rf.Context(clear=True, face_cull=None, depth_test=None)
    rf.render(tris, (pos, cols), rf.shader.Gouraud(), mvp, vp, rf.Colorbuf(gcol), ctx)
    assert np.array_equal(gcol, ocol)
    assert (ctx.stats.frags_i, ctx.stats.frags_o) == (int(ost[5]), int(ost[6])) and ctx.stats.frags_o > 100000
