"""The batched and served workloads at their stated sizes.

Batched: 64 frames in one HIP-graph launch, resolutions uniform over
256x256 .. 2048x2048, each frame a seeded mix of the full crates scene (1,000
textured crates) and the full icosphere scene (10,000 spheres, 3.2M
triangles) at 10-100 % of their objects, with random camera poses. Every frame
equals, bit for bit, the plain per-frame path (one synchronous rf_render per
draw, in order, the first clearing: the render() calls Batch::render makes,
batch.rs:127-146, render.rs:134-207), and four seeded frames equal the oracle.

Served: an open-loop Poisson stream of 320 requests over 16 resident meshes
and the five resolutions, under a memory budget that admits only a few frames
at once. Every request comes back (none rejected or lost), some wait for
memory, every result equals the same request rendered alone through
rf_render, and a sample equals the oracle.
"""
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
import retrofire_amd as rf
from retrofire_amd import batch as RB
from retrofire_amd import scenes as S

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full_world():
    cpos, cuv, ctris, tex = S.crates_world(1000)
    ipos, iattr, itris, nv, nt = S.icospheres_world(10000)
    return dict(cpos=cpos, cuv=cuv, ctris=ctris, tex=tex, ipos=ipos, iattr=iattr, itris=itris, nt=nt)


def plain_frame(world, f):
    """The frame through the plain drop-in path: the crates draw (clearing
    first), then the icosphere draw, each one synchronous rf_render."""
    col = np.zeros((f["h"], f["w"], 4), np.uint8)
    dep = np.zeros((f["h"], f["w"]), np.float32)
    ctx = rf.Context(clear=True)
    rf.render(world["ctris"][: f["nc"]], (world["cpos"], world["cuv"]), rf.shader.TexRepeatPot(world["tex"]),
              f["mvp_c"], f["vp"], rf.Framebuf(col, dep), ctx)
    ctx.clear = False
    rf.render(world["itris"][: f["ni"]], (world["ipos"], world["iattr"]), rf.shader.Gouraud(), f["mvp_i"], f["vp"],
              rf.Framebuf(col, dep), ctx)
    return col, dep, ctx.stats


def oracle_frame(world, f):
    col = np.zeros((f["h"], f["w"], 4), np.uint8)
    col[...] = (0, 0, 0, 255)
    dep = np.zeros((f["h"], f["w"]), np.float32)
    st = np.zeros(8, np.uint64)
    e1, _ = O.render(3, world["ctris"][: f["nc"]], world["cpos"], world["cuv"], f["mvp_c"], f["vp"], col,
                     depth=dep, tex=world["tex"], stats=st)
    e2, _ = O.render(1, world["itris"][: f["ni"]], world["ipos"], world["iattr"], f["mvp_i"], f["vp"], col,
                     depth=dep, stats=st)
    assert e1 == 0 and e2 == 0
    return col, dep, st


def test_batched_64_uneven_frames_full_scenes(full_world):
    w = full_world
    frames = S.batch_frames(64, w["ctris"].shape[0], w["itris"].shape[0], w["nt"], seed=0)
    log(f"c3: {sum(f['nc'] + f['ni'] for f in frames):,} triangles over {len(frames)} frames")
    crates = RB.Mesh(w["cpos"], w["ctris"], w["cuv"])
    icos = RB.Mesh(w["ipos"], w["itris"], w["iattr"])
    tex = RB.Texture(w["tex"])
    fbs, spec = [], []
    for f in frames:
        fb = RB.DeviceFramebuf(f["w"], f["h"], 0, True)
        fbs.append(fb)
        draws = [RB.mesh_draw(crates, rf.shader.TexRepeatPot(w["tex"]), f["mvp_c"], f["vp"], n_tris=f["nc"],
                              texture=tex),
                 RB.mesh_draw(icos, rf.shader.Gouraud(), f["mvp_i"], f["vp"], n_tris=f["ni"])]
        spec.append((fb, rf.Context(clear=True), draws))
    t0 = time.time()
    b = RB.Batch(spec)  # autotuned tile width, span chunk and binsort tiles per wave
    log(f"c3: batch planned in {time.time() - t0:.1f} s: tile_w {b.tile_w}, chunk {b.chunk}")
    for rep in range(2):  # two graph replays
        b.run(sync=True)
        stats = b.wait()
    log(f"c3: replay {b.last_ms():.2f} ms")
    oracle_set = {3, 17, 40, 58}
    for k, f in enumerate(frames):
        gcol, gdep = fbs[k].read()
        pcol, pdep, pst = plain_frame(w, f)
        assert np.array_equal(gcol, pcol), f"frame {k} ({f['w']}x{f['h']}) colour: batched != plain"
        assert np.array_equal(gdep.view(np.uint32), pdep.view(np.uint32)), f"frame {k} depth: batched != plain"
        s = stats[k]
        assert [s.calls, s.prims_i, s.prims_o, s.verts_i, s.verts_o, s.frags_i, s.frags_o] == \
            [pst.calls, pst.prims_i, pst.prims_o, pst.verts_i, pst.verts_o, pst.frags_i, pst.frags_o], f"frame {k}"
        if k in oracle_set:
            ocol, odep, ost = oracle_frame(w, f)
            assert np.array_equal(gcol, ocol), f"frame {k} colour != oracle"
            assert np.array_equal(gdep.view(np.uint32), odep.view(np.uint32)), f"frame {k} depth != oracle"
            assert [s.prims_o, s.verts_o, s.frags_i, s.frags_o] == [int(x) for x in ost[[2, 4, 5, 6]]]
            log(f"c3: frame {k} equals the oracle ({s.frags_o:,} fragments written)")
    log("c3: all 64 frames equal the plain path")
    b.free()
    for fb in fbs:
        fb.free()
    crates.free()
    icos.free()
    tex.free()


def rf_vp(w, h):
    from retrofire_amd import mat as M
    return M.viewport(0, h, w, 0)


def served_draws(meshes, dmeshes, dtex, tex_px, spec):
    """One request's draw: mesh `mi` of the 16 from its pose at w x h."""
    _, w, h, mi, pose = spec
    if meshes[mi][0] == "crates":
        return [RB.mesh_draw(dmeshes[mi], rf.shader.TexRepeatPot(tex_px), S.crates_mvp(w, h, pose), rf_vp(w, h),
                             texture=dtex)]
    return [RB.mesh_draw(dmeshes[mi], rf.shader.Gouraud(), S.spheres_mvp(w, h, pose), rf_vp(w, h))]


def test_served_poisson_stream_16_meshes_under_budget():
    meshes, tex_px = S.serve_meshes()
    dmeshes = [RB.Mesh(p, t, a) for (_, p, a, t) in meshes]
    dtex = RB.Texture(tex_px)
    stream = S.serve_stream(3000.0, 320, seed=9)
    budget = 256 << 20  # the largest request (2048x2048, 640k triangles) needs about 120 MB
    srv = RB.Server(budget_bytes=budget, max_batch=16, n_streams=3)
    results = {}
    try:
        t0 = time.perf_counter()
        i = 0
        deadline = time.perf_counter() + 120.0
        while len(results) < len(stream):
            assert time.perf_counter() < deadline, f"results missing: {len(stream) - len(results)}"
            now = time.perf_counter() - t0
            while i < len(stream) and stream[i][0] <= now:
                _, w, h, _, _ = stream[i]
                srv.submit(i, served_draws(meshes, dmeshes, dtex, tex_px, stream[i]), w, h)
                i += 1
            bufs = [np.empty(2048 * 2048 * 4, np.uint8) for _ in range(8)]
            for r, buf in zip(srv.poll(8, timeout_ms=1 if i < len(stream) else 200, buffers=bufs), bufs):
                assert r.id not in results
                w, h = stream[r.id][1], stream[r.id][2]
                results[r.id] = (r, buf[: w * h * 4].reshape(h, w, 4).copy())
    finally:
        srv.close()
    log(f"c4: {len(results)} results")
    assert sorted(results) == list(range(len(stream))), "a request was lost"
    bad = [(k, r.status) for k, (r, _) in results.items() if r.status != 0]
    assert bad == [], f"requests failed: {bad[:10]}"
    # some requests waited for memory (admitted well after they arrived) ...
    waits = np.array([r.start_s - r.submit_s for r, _ in results.values()])
    assert (waits > 2e-3).sum() >= 20, f"only {(waits > 2e-3).sum()} requests waited"
    # ... and the frames on the device at any time fit the budget (targets alone)
    iv = sorted((r.start_s, r.done_s, r.width * r.height * 8) for r, _ in results.values())
    for s0, _, _ in iv:
        live = sum(b for (a, e, b) in iv if a <= s0 < e)
        assert live <= budget, f"{live} bytes of targets in flight over a budget of {budget}"
    log(f"c4: {(waits > 2e-3).sum()} of {len(waits)} requests waited more than 2 ms; max wait {waits.max() * 1e3:.1f} ms")
    # every result equals the request rendered alone through the plain path
    oracle_ids = [k for k in range(len(stream)) if stream[k][1] <= 512][:4]
    for k, (t, w, h, mi, pose) in enumerate(stream):
        kind, pos, attr, tris = meshes[mi]
        col = np.zeros((h, w, 4), np.uint8)
        dep = np.zeros((h, w), np.float32)
        ctx = rf.Context(clear=True)
        if kind == "crates":
            mvp, shd, okind = S.crates_mvp(w, h, pose), rf.shader.TexRepeatPot(tex_px), 3
        else:
            mvp, shd, okind = S.spheres_mvp(w, h, pose), rf.shader.Gouraud(), 1
        rf.render(tris, (pos, attr), shd, mvp, rf_vp(w, h), rf.Framebuf(col, dep), ctx)
        r, got = results[k]
        assert np.array_equal(got, col), f"request {k} ({kind}, {w}x{h}): served != alone"
        assert (r.stats.prims_o, r.stats.frags_i, r.stats.frags_o) == \
            (ctx.stats.prims_o, ctx.stats.frags_i, ctx.stats.frags_o), f"request {k} stats"
        if k in oracle_ids:
            ocol = np.zeros((h, w, 4), np.uint8)
            ocol[...] = (0, 0, 0, 255)
            odep = np.zeros((h, w), np.float32)
            err, _ = O.render(okind, tris, pos, attr, mvp, rf_vp(w, h), ocol, depth=odep,
                              tex=tex_px if okind == 3 else None)
            assert err == 0 and np.array_equal(got, ocol), f"request {k} != oracle"
    log("c4: every served frame equals the plain path; 4 equal the oracle")
    for m in dmeshes:
        m.free()
    dtex.free()
