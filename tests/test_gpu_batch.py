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
    oracle = [oracle_frame(world, f) for f in frames]
    # autotuned, then every forced tile width with each span chunk, and other tile heights
    for tw, ch, th in ((0, 0, 0), (16, 8, 32), (32, 8, 32), (64, 8, 32), (16, 6, 32), (32, 6, 32), (64, 6, 32),
                       (32, 8, 16), (64, 6, 64), (16, 8, 64)):
        dev.set_raster_chunk(ch)
        dev.set_tile_height(th)
        b = RB.Batch(spec, tile_w=tw)
        assert b.chunk in (6, 8) and (tw == 0 or b.chunk == ch)
        assert b.tile_h in (16, 32, 64) and (tw == 0 or b.tile_h == th)
        if tw == 0:
            assert len(b.tune_results()) == 18
        for rep in range(3):  # graph replays give the same result every time
            b.run(sync=True)
            stats = b.wait()
            for k, f in enumerate(frames):
                col, dep = fbs[k].read()
                ocol, odep, ost = oracle[k]
                assert np.array_equal(col, ocol), \
                    f"frame {k} colour (tile {b.tile_w}x{b.tile_h}, chunk {b.chunk}, replay {rep})"
                assert np.array_equal(dep.view(np.uint32), odep.view(np.uint32)), f"frame {k} depth"
                s = stats[k]
                assert [s.prims_o, s.verts_o, s.frags_i, s.frags_o] == [int(x) for x in ost[[2, 4, 5, 6]]]
                assert s.calls == 2
        b.free()
    dev.set_raster_chunk(0)
    dev.set_tile_height(0)
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


def test_serving_readback_space_waits_for_polls(world):
    """Results wait in the server's pinned read-back block until they are
    polled. With a 256 MiB budget the block is 64 MiB: four 2048x2048 RGBA
    frames. Twelve such requests submitted before any poll: the server
    launches what fits, the rest start only after polls free space, and every
    frame still equals the plain path."""
    import time
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    tex = RB.Texture(world["tex"])
    shd = rf.shader.TexRepeatPot(world["tex"])
    w = h = 2048
    vp = M.viewport(0, h, w, 0)
    mvps = [M.then(S.first_person(0.0, 1.7, 0.0, k / 12.0), _persp(w, h)) for k in range(12)]
    srv = RB.Server(budget_bytes=256 << 20, max_batch=4, n_streams=2)
    got = {}
    try:
        for k in range(12):
            srv.submit(k, [RB.mesh_draw(crates, shd, mvps[k], vp, texture=tex)], w, h)
        time.sleep(0.5)  # the server fills its read-back block and waits
        t_poll = time.monotonic()
        t_end = t_poll + 60.0
        while len(got) < 12:
            assert time.monotonic() < t_end, f"results missing: {sorted(set(range(12)) - set(got))}"
            bufs = [np.empty(w * h * 4, np.uint8) for _ in range(2)]
            for r, b in zip(srv.poll(2, timeout_ms=500, buffers=bufs), bufs):
                assert r.status == 0, f"request {r.id} status {r.status}"
                got[r.id] = (r, b.reshape(h, w, 4).copy())
    finally:
        srv.close()
    late = sum(1 for r, _ in got.values() if r.start_s >= t_poll)
    assert late >= 8, f"only {late} of 12 requests started after the first poll"
    for k in range(12):
        col = np.zeros((h, w, 4), np.uint8)
        dep = np.zeros((h, w), np.float32)
        rf.render(world["ctris"], (world["cpos"], world["cuv"]), shd, mvps[k], vp, rf.Framebuf(col, dep),
                  rf.Context(clear=True))
        assert np.array_equal(got[k][1], col), f"request {k}: served != plain"
    crates.free()
    tex.free()


def test_serving_frame_larger_than_readback_block(world):
    """A frame larger than the whole pinned read-back block (64 MiB with a
    256 MiB budget) gets a pinned buffer of its own and equals the plain
    path; a small frame served beside it still uses the block."""
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    tex = RB.Texture(world["tex"])
    shd = rf.shader.TexRepeatPot(world["tex"])
    sizes = [(4096, 4112), (256, 256)]  # 4096 * 4112 * 4 B = 64.25 MiB
    srv = RB.Server(budget_bytes=256 << 20, max_batch=2, n_streams=1)
    got = {}
    try:
        for k, (w, h) in enumerate(sizes):
            mvp = M.then(S.first_person(0.0, 1.7, 0.0, 0.1 * k), _persp(w, h))
            srv.submit(k, [RB.mesh_draw(crates, shd, mvp, M.viewport(0, h, w, 0), texture=tex)], w, h)
        import time
        t_end = time.monotonic() + 60.0
        while len(got) < 2:
            assert time.monotonic() < t_end, "results missing"
            bufs = [np.empty(4096 * 4112 * 4, np.uint8)]
            for r, b in zip(srv.poll(1, timeout_ms=500, buffers=bufs), bufs):
                assert r.status == 0, f"request {r.id} status {r.status}"
                w, h = sizes[r.id]
                got[r.id] = b[: w * h * 4].reshape(h, w, 4).copy()
    finally:
        srv.close()
    for k, (w, h) in enumerate(sizes):
        mvp = M.then(S.first_person(0.0, 1.7, 0.0, 0.1 * k), _persp(w, h))
        col = np.zeros((h, w, 4), np.uint8)
        dep = np.zeros((h, w), np.float32)
        rf.render(world["ctris"], (world["cpos"], world["cuv"]), shd, mvp, M.viewport(0, h, w, 0),
                  rf.Framebuf(col, dep), rf.Context(clear=True))
        assert np.array_equal(got[k], col), f"request {k} ({w}x{h}): served != plain"
    crates.free()
    tex.free()


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


@pytest.mark.parametrize("sort", ["front_to_back", "back_to_front"])
@pytest.mark.parametrize("dt", [None, "less"])
def test_depth_sorted_frames_in_batches(sort, dt):
    """Batched frames whose context asks for a depth sort (render.rs:180-183,
    209-219): every render() call sorts its own clip fans (near-plane fans
    included), each frame holds several sorted draws of every shader kind,
    edges keep their order, frames without a sort sit between them, and some
    draws are culled whole; every frame, replayed, equals the oracle."""
    from test_gpu_parity import random_soup
    rng = np.random.default_rng(70 + (sort == "back_to_front") + 2 * (dt is None))
    tex = rng.integers(0, 256, size=(32, 64, 4), dtype=np.uint8)
    dtex = RB.Texture(tex)
    shaders = {0: rf.shader.Flat((9, 200, 77, 255)), 1: rf.shader.Gouraud(), 2: rf.shader.TexClamp(tex),
               3: rf.shader.TexRepeatPot(tex)}
    meshes, frames, oracle = [], [], []
    for fi in range(6):
        w, h = [(301, 197), (256, 256), (128, 96)][fi % 3]
        sorted_frame = fi != 2
        ctx = rf.Context(clear=True, depth_test=dt, face_cull=None, depth_sort=sort if sorted_frame else None)
        fb = RB.DeviceFramebuf(w, h, 0, True)
        ocol = np.zeros((h, w, 4), np.uint8)
        ocol[...] = (0, 0, 0, 255)
        odep = np.zeros((h, w), np.float32)
        ost = np.zeros(8, np.uint64)
        draws = []
        for k, kind in enumerate([1, 3, 0, 2, 1]):
            pos, attr, tris, mvp, vp = random_soup(rng, 60 + 30 * k, w, h, kind)
            if k == 4:  # far off to the side: culled whole by its bounding box
                mvp = M.then(M.translate3(40.0, 0.0, 0.0), mvp)
            mesh = RB.Mesh(pos, tris, attr)
            meshes.append(mesh)
            draws.append(RB.mesh_draw(mesh, shaders[kind], mvp, vp, texture=dtex if kind >= 2 else None,
                                      cull=(k == 4)))
            if not (k == 4 and O.bbox_visibility(O.bbox_of(pos), mvp) == O.HIDDEN):
                e, _ = O.render(kind, tris, pos, attr, mvp, vp, ocol, depth=odep, flat=(9, 200, 77, 255),
                                tex=tex if kind >= 2 else None,
                                ctx=(DT_[dt], 0, 1, 1, DS_[sort] if sorted_frame else 0), stats=ost)
                assert e == 0
        # an edge draw keeps its order under the sort
        epos = rng.uniform(-1, 1, size=(80, 3)).astype(np.float32)
        ecol = rng.uniform(0, 1, size=(80, 3)).astype(np.float32)
        edges = np.arange(80, dtype=np.uint32).reshape(40, 2)
        emesh = RB.Mesh(epos, edges, ecol)
        meshes.append(emesh)
        draws.append(RB.mesh_draw(emesh, rf.shader.Gouraud(), M.identity(), M.viewport(0, h, w, 0)))
        e, _ = O.render(1, edges, epos, ecol, M.identity(), M.viewport(0, h, w, 0), ocol, depth=odep,
                        ctx=(DT_[dt], 0, 1, 1, DS_[sort] if sorted_frame else 0), stats=ost, prim=1)
        assert e == 0
        frames.append((fb, ctx, draws))
        oracle.append((ocol, odep, ost))
    b = RB.Batch(frames)
    for rep in range(3):
        b.run(sync=True)
        stats = b.wait()
        for k, (fb, _, _) in enumerate(frames):
            ocol, odep, ost = oracle[k]
            gcol, gdep = fb.read()
            assert np.array_equal(gcol, ocol), f"frame {k} colour (replay {rep})"
            assert np.array_equal(gdep.view(np.uint32), odep.view(np.uint32)), f"frame {k} depth"
            s = stats[k]
            assert [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o] == \
                [int(x) for x in ost[:7]], f"frame {k} stats"
    b.free()
    for fb, _, _ in frames:
        fb.free()
    for m in meshes:
        m.free()
    dtex.free()


@pytest.mark.parametrize("sort", ["front_to_back", "back_to_front"])
@pytest.mark.parametrize("dt", [None, "less"])
def test_depth_sorted_served_frames(sort, dt):
    """Served requests whose context asks for a depth sort (render.rs:180-183,
    209-219), beside requests without one: each sorted request is launched
    alone through its stream's sort pass; every frame (colour and stats)
    equals the oracle."""
    import time
    from test_gpu_parity import random_soup
    rng = np.random.default_rng(170 + (sort == "back_to_front") + 2 * (dt is None))
    tex = rng.integers(0, 256, size=(32, 64, 4), dtype=np.uint8)
    dtex = RB.Texture(tex)
    shaders = {0: rf.shader.Flat((9, 200, 77, 255)), 1: rf.shader.Gouraud(), 2: rf.shader.TexClamp(tex),
               3: rf.shader.TexRepeatPot(tex)}
    meshes, reqs, oracle = [], [], []
    for fi in range(8):
        w, h = [(301, 197), (256, 256), (128, 96)][fi % 3]
        sorted_frame = fi % 4 != 2
        ctx = rf.Context(clear=True, depth_test=dt, face_cull=None, depth_sort=sort if sorted_frame else None)
        ocol = np.zeros((h, w, 4), np.uint8)
        ocol[...] = (0, 0, 0, 255)
        odep = np.zeros((h, w), np.float32)
        ost = np.zeros(8, np.uint64)
        draws = []
        for k, kind in enumerate([1, 3, 0, 2]):
            pos, attr, tris, mvp, vp = random_soup(rng, 60 + 30 * k, w, h, kind)
            mesh = RB.Mesh(pos, tris, attr)
            meshes.append(mesh)
            draws.append(RB.mesh_draw(mesh, shaders[kind], mvp, vp, texture=dtex if kind >= 2 else None))
            e, _ = O.render(kind, tris, pos, attr, mvp, vp, ocol, depth=odep, flat=(9, 200, 77, 255),
                            tex=tex if kind >= 2 else None,
                            ctx=(DT_[dt], 0, 1, 1, DS_[sort] if sorted_frame else 0), stats=ost)
            assert e == 0
        # an edge draw keeps its order under the sort
        epos = rng.uniform(-1, 1, size=(80, 3)).astype(np.float32)
        ecol = rng.uniform(0, 1, size=(80, 3)).astype(np.float32)
        edges = np.arange(80, dtype=np.uint32).reshape(40, 2)
        emesh = RB.Mesh(epos, edges, ecol)
        meshes.append(emesh)
        draws.append(RB.mesh_draw(emesh, rf.shader.Gouraud(), M.identity(), M.viewport(0, h, w, 0)))
        e, _ = O.render(1, edges, epos, ecol, M.identity(), M.viewport(0, h, w, 0), ocol, depth=odep,
                        ctx=(DT_[dt], 0, 1, 1, DS_[sort] if sorted_frame else 0), stats=ost, prim=1)
        assert e == 0
        reqs.append((w, h, ctx, draws))
        oracle.append((ocol, ost))
    srv = RB.Server(budget_bytes=256 << 20, max_batch=4, n_streams=2)
    got = {}
    try:
        for k, (w, h, ctx, draws) in enumerate(reqs):
            srv.submit(k, draws, w, h, ctx=ctx)
        t_end = time.monotonic() + 60.0
        while len(got) < len(reqs):
            assert time.monotonic() < t_end, f"results missing: {sorted(set(range(len(reqs))) - set(got))}"
            bufs = [np.empty(301 * 256 * 4, np.uint8) for _ in range(4)]
            for r, b in zip(srv.poll(4, timeout_ms=500, buffers=bufs), bufs):
                assert r.status == 0, f"request {r.id} status {r.status}"
                w, h = reqs[r.id][0], reqs[r.id][1]
                got[r.id] = (r, b[: w * h * 4].reshape(h, w, 4).copy())
    finally:
        srv.close()
    for k in range(len(reqs)):
        r, col = got[k]
        ocol, ost = oracle[k]
        assert np.array_equal(col, ocol), f"request {k} colour"
        s = r.stats
        assert [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o] == \
            [int(x) for x in ost[:7]], f"request {k} stats"
    for m in meshes:
        m.free()
    dtex.free()


_FMT_ARRAYS = {rf.PIX_RGBA8: (np.uint8, 4), rf.PIX_RGB8: (np.uint8, 3), rf.PIX_U32_RGBA: (np.uint32, 0),
               rf.PIX_U32_XRGB: (np.uint32, 0), rf.PIX_U32_ARGB: (np.uint32, 0), rf.PIX_U32_BGRA: (np.uint32, 0),
               rf.PIX_RGBA4444: (np.uint16, 0), rf.PIX_RGB565: (np.uint16, 0)}


@pytest.mark.parametrize("fmt", sorted(_FMT_ARRAYS))
def test_pixel_formats_batched_and_served(world, fmt):
    """Every pixel format through the batched and the serving path, with and
    without depth, two draws per frame (textured crates, then Gouraud
    icospheres): each frame equals the plain path, whose formats are checked
    against the oracle."""
    import time
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    icos = RB.Mesh(world["ipos"], world["itris"], world["iattr"])
    tex = RB.Texture(world["tex"])
    shd = rf.shader.TexRepeatPot(world["tex"])
    dt, ch = _FMT_ARRAYS[fmt]
    frames = []
    for k, (w, h, with_depth) in enumerate([(320, 200, True), (256, 160, False)]):
        mvp_c = M.then(S.first_person(0.0, 1.7, 0.0, 0.15 * k), _persp(w, h))
        view_i = M.then(M.rotate_y(0.3 * k), M.translate3(0.0, 0.0, 2.5))
        mvp_i = M.then(view_i, M.perspective(np.float32(1.7320508), np.float32(w) / np.float32(h), 0.1, 100.0))
        vp = M.viewport(0, h, w, 0)
        col = np.zeros((h, w, ch) if ch else (h, w), dt)
        dep = np.zeros((h, w), np.float32) if with_depth else None
        tgt = rf.Framebuf(col, dep, fmt) if with_depth else rf.Colorbuf(col, fmt)
        ctx = rf.Context(clear=True)
        rf.render(world["ctris"], (world["cpos"], world["cuv"]), shd, mvp_c, vp, tgt, ctx)
        ctx.clear = False
        rf.render(world["itris"], (world["ipos"], world["iattr"]), rf.shader.Gouraud(), mvp_i, vp, tgt, ctx)
        draws = [RB.mesh_draw(crates, shd, mvp_c, vp, texture=tex), RB.mesh_draw(icos, rf.shader.Gouraud(), mvp_i, vp)]
        frames.append((w, h, with_depth, draws, col, dep))
    # batched
    fbs = [RB.DeviceFramebuf(w, h, fmt, with_depth) for (w, h, with_depth, _, _, _) in frames]
    b = RB.Batch([(fb, rf.Context(clear=True), f[3]) for fb, f in zip(fbs, frames)])
    b.run(sync=True)
    b.wait()
    for k, (fb, f) in enumerate(zip(fbs, frames)):
        gcol, gdep = fb.read()
        assert np.array_equal(gcol, f[4]), f"batched frame {k} colour"
        if f[2]:
            assert np.array_equal(gdep.view(np.uint32), f[5].view(np.uint32)), f"batched frame {k} depth"
    b.free()
    for fb in fbs:
        fb.free()
    # served
    srv = RB.Server(budget_bytes=256 << 20, max_batch=4, n_streams=2)
    got = {}
    try:
        for k, (w, h, with_depth, draws, _, _) in enumerate(frames):
            srv.submit(k, draws, w, h, fmt=fmt, with_depth=with_depth)
        t_end = time.monotonic() + 60.0
        while len(got) < len(frames):
            assert time.monotonic() < t_end, "results missing"
            bufs = [np.empty(320 * 200 * 4, np.uint8) for _ in range(2)]
            for r, buf in zip(srv.poll(2, timeout_ms=500, buffers=bufs), bufs):
                assert r.status == 0, f"request {r.id} status {r.status}"
                got[r.id] = buf
    finally:
        srv.close()
    for k, f in enumerate(frames):
        want = f[4].tobytes()
        assert got[k][: len(want)].tobytes() == want, f"served frame {k} colour"
    crates.free()
    icos.free()
    tex.free()


_CTX_VARIANTS = [
    dict(depth_test="less", face_cull="back", color_clear=(10, 20, 30, 255), depth_clear=50.0),
    dict(depth_test=None, face_cull=None),
    dict(depth_test="greater", face_cull="front", depth_clear=0.5),
    dict(depth_test="equal", face_cull=None),
    dict(depth_test="less", face_cull=None, color_write=False, color_clear=(1, 2, 3, 4)),
    dict(depth_test="less", face_cull=None, depth_write=False),
]


def test_contexts_batched_and_served(world):
    """Each Context setting (depth tests, face culling, colour or depth writes
    off, clear values) on its own frame of one batch and of one served
    stream: every frame equals the plain path."""
    import time
    crates = RB.Mesh(world["cpos"], world["ctris"], world["cuv"])
    icos = RB.Mesh(world["ipos"], world["itris"], world["iattr"])
    tex = RB.Texture(world["tex"])
    shd = rf.shader.TexRepeatPot(world["tex"])
    w, h = 256, 160
    vp = M.viewport(0, h, w, 0)
    frames = []
    for k, cv in enumerate(_CTX_VARIANTS):
        mvp_c = M.then(S.first_person(0.0, 1.7, 0.0, 0.1 * k), _persp(w, h))
        view_i = M.then(M.rotate_y(0.2 * k), M.translate3(0.0, 0.0, 2.5))
        mvp_i = M.then(view_i, M.perspective(np.float32(1.7320508), np.float32(w) / np.float32(h), 0.1, 100.0))
        col = np.zeros((h, w, 4), np.uint8)
        dep = np.zeros((h, w), np.float32)
        ctx = rf.Context(clear=True, **cv)
        rf.render(world["ctris"], (world["cpos"], world["cuv"]), shd, mvp_c, vp, rf.Framebuf(col, dep), ctx)
        ctx.clear = False
        rf.render(world["itris"], (world["ipos"], world["iattr"]), rf.shader.Gouraud(), mvp_i, vp,
                  rf.Framebuf(col, dep), ctx)
        draws = [RB.mesh_draw(crates, shd, mvp_c, vp, texture=tex), RB.mesh_draw(icos, rf.shader.Gouraud(), mvp_i, vp)]
        frames.append((rf.Context(clear=True, **cv), draws, col, dep))
    fbs = [RB.DeviceFramebuf(w, h, 0, True) for _ in frames]
    b = RB.Batch([(fb, f[0], f[1]) for fb, f in zip(fbs, frames)])
    b.run(sync=True)
    b.wait()
    for k, (fb, f) in enumerate(zip(fbs, frames)):
        gcol, gdep = fb.read()
        assert np.array_equal(gcol, f[2]), f"batched frame {k} ({_CTX_VARIANTS[k]}) colour"
        assert np.array_equal(gdep.view(np.uint32), f[3].view(np.uint32)), f"batched frame {k} depth"
    b.free()
    for fb in fbs:
        fb.free()
    srv = RB.Server(budget_bytes=256 << 20, max_batch=4, n_streams=2)
    got = {}
    try:
        for k, (ctx, draws, _, _) in enumerate(frames):
            srv.submit(k, draws, w, h, ctx=ctx)
        t_end = time.monotonic() + 60.0
        while len(got) < len(frames):
            assert time.monotonic() < t_end, "results missing"
            bufs = [np.empty(w * h * 4, np.uint8) for _ in range(3)]
            for r, buf in zip(srv.poll(3, timeout_ms=500, buffers=bufs), bufs):
                assert r.status == 0, f"request {r.id} status {r.status}"
                got[r.id] = buf.reshape(h, w, 4)
    finally:
        srv.close()
    for k, f in enumerate(frames):
        assert np.array_equal(got[k], f[2]), f"served frame {k} ({_CTX_VARIANTS[k]}) colour"
    crates.free()
    icos.free()
    tex.free()


DT_ = {None: 0, "less": 1, "equal": 2, "greater": 3}
DS_ = {None: 0, "front_to_back": 1, "back_to_front": 2}
