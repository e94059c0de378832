"""The batched path (one HIP-graph launch over many frames of uneven
resolution) and the serving loop give, frame by frame, exactly what the
plain per-frame path and the oracle give."""
import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from retrofire_amd import batch as RB
from retrofire_amd import mat as M
from retrofire_amd import scenes as S

pytestmark = pytest.mark.gpu

RES = [(256, 256), (512, 512), (1280, 720), (1920, 1080), (2048, 2048)]


def _persp(w, h):
    a = np.float32(w) / np.float32(h)
    return M.perspective(M.fov_diagonal_focal_ratio(90.0, a), a, 0.1, 1000.0)


@pytest.fixture(scope="module")
def world():
    cpos, cuv, ctris, tex = S.crates_world(400)
    ipos, iattr, itris, nv, nt = S.icospheres_world(300)
    return dict(cpos=cpos, cuv=cuv, ctris=ctris, tex=tex, ipos=ipos, iattr=iattr, itris=itris, nt=nt)


def mixed_frames(world, n, seed):
    """n frames, each a seeded mix of crates (10-100 % of the objects) and
    icospheres (10-100 %) with random camera poses and resolutions."""
    rng = S.Xorshift64(S.DEFAULT_SEED + 3 + seed)
    out = []
    for i in range(n):
        w, h = RES[rng.uniform_usize(0, 5)]
        fc = float(rng.uniform_f32(0.1, 1.0))
        fi = float(rng.uniform_f32(0.1, 1.0))
        yaw = float(rng.uniform_f32(0.0, 1.0))
        orbit = float(rng.uniform_f32(0.0, 1.0))
        nc = max(1, int(fc * world["ctris"].shape[0] / 12)) * 12
        ni = max(1, int(fi * world["itris"].shape[0] / world["nt"])) * world["nt"]
        vp = M.viewport(0, h, w, 0)
        mvp_c = M.then(S.first_person(0.0, 1.7, 0.0, yaw), _persp(w, h))
        view_i = M.then(M.rotate_y(orbit), M.translate3(0.0, 0.0, 2.5))
        mvp_i = M.then(view_i, M.perspective(np.float32(1.7320508), np.float32(w) / np.float32(h), 0.1, 100.0))
        out.append(dict(w=w, h=h, vp=vp, mvp_c=mvp_c, mvp_i=mvp_i, nc=nc, ni=ni))
    return out


def oracle_frame(world, f):
    col = np.zeros((f["h"], f["w"], 4), np.uint8)
    col[...] = (0, 0, 0, 255)
    dep = np.zeros((f["h"], f["w"]), np.float32)
    st = np.zeros(8, np.uint64)
    e1, _ = O.render(3, world["ctris"][: f["nc"]], world["cpos"], world["cuv"], f["mvp_c"], f["vp"], col,
                     depth=dep, tex=world["tex"], stats=st)
    e2, _ = O.render(1, world["itris"][: f["ni"]], world["ipos"], world["iattr"], f["mvp_i"], f["vp"], col, depth=dep,
                     stats=st)
    assert e1 == 0 and e2 == 0
    return col, dep, st


def test_batched_uneven_frames_match_oracle(world):
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    icos = RB.Mesh(world["ipos"], world["itris"], world["iattr"])
    tex = RB.Texture(world["tex"])
    frames = mixed_frames(world, 12, seed=0)
    fbs, spec = [], []
    for f in frames:
        fb = RB.DeviceFramebuf(f["w"], f["h"], 0, True)
        fbs.append(fb)
        draws = [RB.mesh_draw(crates, rf.shader.TexRepeatPot(world["tex"]), f["mvp_c"], f["vp"], n_tris=f["nc"],
                              texture=tex),
                 RB.mesh_draw(icos, rf.shader.Gouraud(), f["mvp_i"], f["vp"], n_tris=f["ni"])]
        spec.append((fb, rf.Context(clear=True), draws))
    dev = rf.default_device()
    # autotuned, then every forced tile width with each span chunk
    for tw, ch in ((0, 0), (16, 8), (32, 8), (64, 8), (16, 6), (32, 6), (64, 6)):
        dev.set_raster_chunk(ch)
        b = RB.Batch(spec, tile_w=tw)
        assert b.chunk in (6, 8) and (tw == 0 or b.chunk == ch)
        for rep in range(3):  # graph replays give the same result every time
            b.run(sync=True)
            stats = b.wait()
            for k, f in enumerate(frames):
                col, dep = fbs[k].read()
                ocol, odep, ost = oracle_frame(world, f)
                assert np.array_equal(col, ocol), f"frame {k} colour (tile_w {b.tile_w}, chunk {b.chunk}, replay {rep})"
                assert np.array_equal(dep.view(np.uint32), odep.view(np.uint32)), f"frame {k} depth"
                s = stats[k]
                assert [s.prims_o, s.verts_o, s.frags_i, s.frags_o] == [int(x) for x in ost[[2, 4, 5, 6]]]
                assert s.calls == 2
        b.free()
    dev.set_raster_chunk(0)
    for fb in fbs:
        fb.free()
    crates.free()
    icos.free()
    tex.free()


def test_serving_alone_equals_mixed(world):
    """Each served frame is identical whether it is rendered alone or batched
    with concurrent requests (continuous batching over several streams)."""
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    tex = RB.Texture(world["tex"])
    frames = mixed_frames(world, 24, seed=1)
    shd = rf.shader.TexRepeatPot(world["tex"])

    def req_draws(f):
        return [RB.mesh_draw(crates, shd, f["mvp_c"], f["vp"], n_tris=f["nc"], texture=tex)]

    def collect(srv, frames_, ids, deadline_s=60.0):
        """Polls until every id in `ids` is back (fails after a deadline)."""
        import time
        out = {}
        t_end = time.time() + deadline_s
        while len(out) < len(ids):
            assert time.time() < t_end, f"serving results missing: {sorted(set(ids) - set(out))}"
            bufs = [np.empty(2048 * 2048 * 4, np.uint8) for _ in range(4)]
            res = srv.poll(4, timeout_ms=500, buffers=bufs)
            for r, b in zip(res, bufs):
                f = frames_[r.id]
                assert r.status == 0, f"request {r.id} status {r.status}"
                out[r.id] = (b[: f["w"] * f["h"] * 4].reshape(f["h"], f["w"], 4).copy(), r)
        return out

    def run(frames_, max_batch, streams, together):
        srv = RB.Server(budget_bytes=1 << 30, max_batch=max_batch, n_streams=streams)
        out = {}
        try:
            if together:
                for i, f in enumerate(frames_):
                    srv.submit(i, req_draws(f), f["w"], f["h"])
                out = collect(srv, frames_, list(range(len(frames_))))
            else:
                for i, f in enumerate(frames_):
                    srv.submit(i, req_draws(f), f["w"], f["h"])
                    out.update(collect(srv, frames_, [i]))
        finally:
            srv.close()
        return out

    alone = run(frames, 1, 1, together=False)
    mixed = run(frames, 8, 3, together=True)
    for i, f in enumerate(frames):
        assert np.array_equal(alone[i][0], mixed[i][0]), f"request {i} differs alone vs mixed"
        assert alone[i][1].stats.frags_o == mixed[i][1].stats.frags_o
        col = np.zeros((f["h"], f["w"], 4), np.uint8)
        col[...] = (0, 0, 0, 255)
        dep = np.zeros((f["h"], f["w"]), np.float32)
        err, _ = O.render(3, world["ctris"][: f["nc"]], world["cpos"], world["cuv"], f["mvp_c"], f["vp"], col,
                          depth=dep, tex=world["tex"])
        assert np.array_equal(col, alone[i][0])
    crates.free()
    tex.free()


def test_serving_budget_rejects_oversized():
    m = RB.Mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.uint32))
    srv = RB.Server(budget_bytes=1 << 20, max_batch=4, n_streams=1)
    srv.submit(7, [RB.mesh_draw(m, rf.shader.Flat(), M.identity(), M.viewport(0, 0, 2048, 2048))], 2048, 2048)
    res = srv.poll(1, timeout_ms=5000)
    assert len(res) == 1 and res[0].id == 7 and res[0].status == -9
    srv.close()
    m.free()


def test_object_culling_matches_caller_side_culling():
    """Per-object draws with `cull` set: the device skips each object whose
    bounding box is Hidden, exactly as the crates demo skips it before calling
    render() (crates.rs:114-133). The oracle side does that culling itself,
    object by object, and renders the rest in order; images and every stat
    (objs included) agree. The first 20 draws ask for no culling."""
    cpos, cuv, ctri, models, tex_px = S.crates_objects(300)
    cube = RB.Mesh(cpos, ctri, cuv)
    tex = RB.Texture(tex_px)
    shd = rf.shader.TexRepeatPot(tex_px)
    bb = O.bbox_of(cpos)
    frames = []
    for k, (w, h, yaw) in enumerate([(1280, 720, 0.0), (640, 480, 0.3), (1920, 1080, 0.62)]):
        wtp = M.then(S.first_person(0.0, 1.7, 0.0, yaw), _persp(w, h))
        frames.append(dict(w=w, h=h, vp=M.viewport(0, h, w, 0), mvps=[M.then(m, wtp) for m in models]))
    fbs, spec = [], []
    for f in frames:
        fb = RB.DeviceFramebuf(f["w"], f["h"], 0, True)
        fbs.append(fb)
        draws = [RB.mesh_draw(cube, shd, mvp, f["vp"], texture=tex, cull=(i >= 20)) for i, mvp in enumerate(f["mvps"])]
        spec.append((fb, rf.Context(clear=True), draws))
    b = RB.Batch(spec)
    b.run(sync=True)
    stats = b.wait()
    for k, f in enumerate(frames):
        col = np.zeros((f["h"], f["w"], 4), np.uint8)
        col[...] = (0, 0, 0, 255)
        dep = np.zeros((f["h"], f["w"]), np.float32)
        st = np.zeros(8, np.uint64)
        calls = objs_o = 0
        for i, mvp in enumerate(f["mvps"]):
            if i >= 20 and O.bbox_visibility(bb, mvp) == O.HIDDEN:
                continue
            err, _ = O.render(3, ctri, cpos, cuv, mvp, f["vp"], col, depth=dep, tex=tex_px, stats=st)
            assert err == 0
            calls += 1
            objs_o += i >= 20
        assert 10 < objs_o < 200, objs_o  # the camera sees part of the field
        gcol, gdep = fbs[k].read()
        assert np.array_equal(gcol, col), f"frame {k} colour"
        assert np.array_equal(gdep.view(np.uint32), dep.view(np.uint32)), f"frame {k} depth"
        s = stats[k]
        assert [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o] == \
            [calls] + [int(x) for x in st[1:7]]
        assert (s.objs_i, s.objs_o) == (len(models) - 20, objs_o)
    # the serving loop culls the same way
    srv = RB.Server(budget_bytes=1 << 30, max_batch=4, n_streams=1)
    try:
        f = frames[0]
        draws = [RB.mesh_draw(cube, shd, mvp, f["vp"], texture=tex, cull=(i >= 20)) for i, mvp in enumerate(f["mvps"])]
        srv.submit(0, draws, f["w"], f["h"])
        buf = np.empty(f["w"] * f["h"] * 4, np.uint8)
        res = srv.poll(1, timeout_ms=30000, buffers=[buf])
        assert len(res) == 1 and res[0].status == 0
        assert np.array_equal(buf.reshape(f["h"], f["w"], 4), fbs[0].read()[0])
        assert (res[0].stats.calls, res[0].stats.objs_i, res[0].stats.objs_o, res[0].stats.frags_o) == \
            (stats[0].calls, stats[0].objs_i, stats[0].objs_o, stats[0].frags_o)
    finally:
        srv.close()
    b.free()
    for fb in fbs:
        fb.free()
    cube.free()
    tex.free()


def test_argument_checks():
    """Invalid tile widths, culling flags and shader kinds are rejected before
    any work is queued."""
    from retrofire_amd import _ffi
    import ctypes as C
    dev = rf.default_device()
    assert _ffi.lib().rf_device_set_tile_width(dev.h, 48) == -3
    assert _ffi.lib().rf_device_set_raster_chunk(dev.h, 7) == -3
    m = RB.Mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.uint32))
    fb = RB.DeviceFramebuf(16, 16)
    d = RB.mesh_draw(m, rf.shader.Flat(), M.identity(), M.viewport(0, 16, 16, 0))
    with pytest.raises(rf.RenderError) as e:
        RB.Batch([(fb, rf.Context(), [d])], tile_w=48)
    assert e.value.status == -3
    d.cull = 2
    with pytest.raises(rf.RenderError) as e:
        RB.Batch([(fb, rf.Context(), [d])])
    assert e.value.status == -3
    d.cull = 0
    d.kind = 9
    with pytest.raises(rf.RenderError) as e:
        RB.Batch([(fb, rf.Context(), [d])])
    assert e.value.status == -3
    fb.free()
    m.free()
