"""Benchmark of the render path on one MI355X.

Workload (the metric's 1080p textured case): the crates scene -- 1000
textured unit cubes on a 100 x 100 field (one merged mesh, 24,000 vertices,
12,000 triangles, one 256 x 256 RGBA8 texture, SamplerRepeatPot) -- rendered
at 1920x1080 into an RGBA8 colour buffer and an f32 depth buffer, cleared per
frame, from a first-person camera at height 1.7 (near-plane clipping). One
step is one launch of the batched path over 64 such frames, each with its own
camera heading; meshes, texture and targets are resident in HBM before the
timed region.

Printed as one JSON line: frames/s (value), Mfrag/s, the plain per-frame call
rate, serving p50/p99 latency, the roofline of the rasterizer and the CPU
baseline (the oracle, a single-threaded C++ restatement of the reference CPU
path, on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--frames", type=int, default=64)
    p.add_argument("--tile-w", type=int, default=0, help="0 = autotune")
    p.add_argument("--serve-seconds", type=float, default=2.0)
    p.add_argument("--serve-rate", type=float, default=1000.0)
    p.add_argument("--serve-streams", type=int, default=6, help="concurrent launches of the serving loop")
    p.add_argument("--serve-sweep-seconds", type=float, default=5.0,
                   help="seconds per offered rate of the 16-mesh serving sweep (0 = no sweep)")
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--c3-steps", type=int, default=10, help="timed launches of the 64-frame mixed batch")
    p.add_argument("--no-c3", action="store_true")
    p.add_argument("--c2-steps", type=int, default=10, help="timed launches of 8 icospheres1080 frames")
    p.add_argument("--no-c2", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--serve-only", action="store_true", help="print only the serving measurement")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="also run one device-to-device copy of this many MiB (scale reference for HBM counters)")
    return p.parse_args()


def calib_copy(mib):
    """One device-to-device hipMemcpy of `mib` MiB: a copy kernel whose HBM
    traffic is known (read mib, write mib), the reference that turns the
    FETCH_SIZE / WRITE_SIZE counters of a profiler run into bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError(f"hipMemcpy failed ({rc})")


def measured_traffic(tile_w, tile_h, chunk):
    """HBM bytes per k_raster launch from the committed counter summary, if it
    was taken for this tile shape and span chunk (profiles/k_raster_traffic.json)."""
    try:
        with open(os.path.join(ROOT, "profiles", "k_raster_traffic.json")) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    key = (t.get("tile_w"), t.get("tile_h"), t.get("chunk"))
    return t.get("bytes_per_launch") if key == (tile_w, tile_h, chunk) else None


def camera_frames(n, w, h, seed=12345):
    """n view-projection matrices: first person at (0, 1.7, 0), headings from
    a seeded generator."""
    from retrofire_amd import mat as M
    from retrofire_amd import scenes as S
    rng = S.Xorshift64(S.DEFAULT_SEED + 7 + seed)
    aspect = np.float32(w) / np.float32(h)
    proj = M.perspective(M.fov_diagonal_focal_ratio(90.0, aspect), aspect, 0.1, 1000.0)
    mvps = []
    for i in range(n):
        yaw = float(rng.uniform_f32(0.0, 1.0)) if i else 0.0
        mvps.append(M.then(S.first_person(0.0, 1.7, 0.0, yaw), proj))
    return mvps, M.viewport(0, h, w, 0)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import retrofire_amd as rf
    from retrofire_amd import batch as RB
    from retrofire_amd import scenes as S

    W, H = 1920, 1080
    pos, uv, tris, tex = S.crates_world(1000)
    dev = rf.default_device()
    mesh = RB.Mesh(pos, tris, uv)
    texr = RB.Texture(tex)
    shd = rf.shader.TexRepeatPot(tex)
    if a.serve_only:
        # the serving measurement alone (for an API trace of the serving loop)
        print(json.dumps({"serving": serve_run(rf, RB, mesh, texr, shd, a)}), flush=True)
        return
    mvps, vp = camera_frames(a.frames, W, H)
    fbs = [RB.DeviceFramebuf(W, H, rf.PIX_RGBA8, True) for _ in range(a.frames)]
    frames = [(fbs[i], rf.Context(clear=True), [RB.mesh_draw(mesh, shd, mvps[i], vp, texture=texr)])
              for i in range(a.frames)]
    batch = RB.Batch(frames, tile_w=a.tile_w)

    for _ in range(a.warmup):
        batch.run(sync=True)
    stats = batch.wait()
    frags_i = sum(s.frags_i for s in stats)
    frags_o = sum(s.frags_o for s in stats)
    batch.run(sync=True)  # make sure the device is idle and synchronized
    t0 = time.perf_counter()
    for _ in range(a.steps):
        batch.run(sync=False)
    batch.wait()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / a.steps
    fps = a.frames * a.steps / elapsed
    mfrag = frags_i * a.steps / elapsed / 1e6

    # per-kernel durations (HIP events on the batch stream, eager launches)
    prof = batch.profile(iters=max(3, min(10, a.steps)))
    names = ["vertex", "setup", "scan", "bin", "binsort", "raster"]
    raster_ms = prof[5]
    bytes_per_frame_raster = W * H * (4 + 4) + tex.nbytes  # colour + depth written once, texture once
    alg_bytes = bytes_per_frame_raster * a.frames
    achieved = alg_bytes / (raster_ms * 1e-3) / 1e9
    total_prof_ms = sum(prof)
    alg_frame_bytes = W * H * 8 + pos.nbytes + uv.nbytes + tris.nbytes + tex.nbytes

    out = {
        "metric": "Mfrag/s + frames/s @1080p textured; p50/p99 latency, batched vs plain",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded crates scene, resident in HBM)",
        "config": {"workload": "crates1080: 1000 textured crates (24k verts, 12k tris, 256x256 RGBA8, "
                               "SamplerRepeatPot), 1920x1080 RGBA8 + f32 depth, clear per frame",
                   "frames_per_step": a.frames, "tile_w": batch.tile_w, "tile_h": batch.tile_h,
                   "chunk": batch.chunk,
                   "parallelism": "single GPU"},
        "mfrag_per_s": round(mfrag, 1),
        "frags_i_per_frame": frags_i // a.frames,
        "frags_o_per_frame": frags_o // a.frames,
        "stage_ms": {n: round(x, 4) for n, x in zip(names, prof)},
        "roofline": {"bound": "hbm", "kernel": "k_raster", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(batch.tile_w, batch.tile_h, batch.chunk),
                     "traffic_source": "committed rocprofv3 FETCH_SIZE + WRITE_SIZE measurement of this kernel "
                                       "and workload (profiles/k_raster_traffic.json), not measured in this run",
                     "alg_bytes_per_launch": alg_bytes, "launch_ms": round(raster_ms, 4)},
        "pipeline_alg_GBs": round(alg_frame_bytes * a.frames / (total_prof_ms * 1e-3) / 1e9, 1),
        "autotune_ms": [list(t) for t in batch.tune_results()],
    }

    # plain drop-in call: host buffers, upload + run + readback per frame
    if not a.no_plain:
        col = np.zeros((H, W, 4), np.uint8)
        dep = np.zeros((H, W), np.float32)
        ctx = rf.Context(clear=True)
        lat = []
        for i in range(8):
            t = time.perf_counter()
            rf.render(tris, (pos, uv), shd, mvps[i % a.frames], vp, rf.Framebuf(col, dep), ctx)
            lat.append(time.perf_counter() - t)
        lat = sorted(lat[2:])
        out["plain"] = {"p50_ms": round(1e3 * lat[len(lat) // 2], 3), "frames_per_s": round(1.0 / lat[len(lat) // 2], 1),
                        "note": "one synchronous rf_render per frame incl. host->device upload and readback"}
        out["batched_over_plain"] = round(fps * lat[len(lat) // 2], 1)

    # serving: open-loop Poisson arrivals over resident meshes
    if not a.no_serve and a.serve_seconds > 0:
        out["serving"] = serve_run(rf, RB, mesh, texr, shd, a)

    # the batched workload at its stated size (64 uneven frames of the full
    # crates and icosphere scenes)
    if not a.no_c2:
        out["batched_c2"] = c2_run(rf, RB, a, a.c2_steps)
    if not a.no_c3:
        out["batched_c3"] = c3_run(rf, RB, a, a.c3_steps)

    # CPU baseline: the oracle on the same workloads, one pinned core, bounded sample
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        out["gpu_over_cpu"] = round(fps / out["cpu_baseline"]["value"], 1)

    batch.free()
    for fb in fbs:
        fb.free()
    if a.calib_copy_mib > 0:
        calib_copy(a.calib_copy_mib)
    print(json.dumps(out), flush=True)


def crates_request(RB, mesh, shd, texr, w, h, yaw):
    from retrofire_amd import scenes as S
    from retrofire_amd import mat as M
    return [RB.mesh_draw(mesh, shd, S.crates_mvp(w, h, yaw), M.viewport(0, h, w, 0), texture=texr)]


def serve_stream(RB, reqs, seconds, warm, budget=4 << 30, alone=None, streams=6):
    """Open-loop replay of (time, w, h, draws) requests into a fresh server;
    latency = completion - submission, wait = admission - submission. `warm`
    (w, h, draws) requests run first, untimed, so the pinned readback pool
    and the learned scratch ratios reach their steady state. `alone`, if
    given, holds each request's latency when served on its own (ms): the
    slowdown latency / alone is then reported too."""
    clk = {"open_ns": time.clock_gettime_ns(time.CLOCK_MONOTONIC)}
    srv = RB.Server(budget_bytes=budget, max_batch=32, n_streams=streams)
    n = len(reqs)
    for k, (w, h, draws) in enumerate(warm):
        srv.submit(10 ** 9 + k, draws, w, h)
    got = 0
    t_warm = time.perf_counter() + 60.0
    while got < len(warm) and time.perf_counter() < t_warm:
        got += len(srv.poll(64, timeout_ms=1000))
    t0 = time.perf_counter()
    clk["stream_ns"] = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
    done = []
    i = 0
    deadline = t0 + seconds + 60.0
    while len(done) < n and time.perf_counter() < deadline:
        now = time.perf_counter() - t0
        while i < n and reqs[i][0] <= now:
            _, w, h, draws = reqs[i]
            srv.submit(i, draws, w, h)
            i += 1
        done += srv.poll(64, timeout_ms=0 if i < n else 5)
        if i < n:
            nxt = reqs[i][0] - (time.perf_counter() - t0)
            if nxt > 0.0002:
                time.sleep(min(nxt, 0.001))
    t1 = time.perf_counter()
    clk["end_ns"] = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
    srv.close()
    if len(done) < n:
        return {"requests": n, "completed": len(done), "error": "serving run did not finish"}
    lat = np.array([(r.done_s - r.submit_s) for r in done]) * 1e3
    wait = np.array([(r.start_s - r.submit_s) for r in done]) * 1e3
    svc = lat - wait
    ok = sum(1 for r in done if r.status == 0)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)  # noqa: E731
    # the requests differ in size (five resolutions): latency by resolution
    # separates the spread of the work from queueing
    res = np.array([r.width * 65536 + r.height for r in done])
    by_res = {}
    for key in sorted(set(res.tolist()), key=lambda k: (k >> 16) * (k & 0xFFFF)):
        m = res == key
        by_res[f"{key >> 16}x{key & 0xFFFF}"] = {"n": int(m.sum()), "p50_ms": pct(lat[m], 50), "p99_ms": pct(lat[m], 99)}
    out = {"requests": n, "ok": ok, "frames_per_s": round(n / (t1 - t0), 1),
           "p50_ms": pct(lat, 50), "p90_ms": pct(lat, 90), "p99_ms": pct(lat, 99), "max_ms": round(float(lat.max()), 3),
           "wait_p50_ms": pct(wait, 50), "wait_p99_ms": pct(wait, 99),
           "service_p50_ms": pct(svc, 50), "service_p99_ms": pct(svc, 99), "by_resolution": by_res,
           "max_p99_over_p50_within_resolution": round(max(v["p99_ms"] / v["p50_ms"] for v in by_res.values()), 2),
           "clock_monotonic_ns": clk}
    if alone is not None:
        slow = np.array([lat[k] / alone[r.id] for k, r in enumerate(done)])
        al = np.array(alone, dtype=np.float64)
        # the requests' own spread: the p99 / p50 latency ratio an ideal
        # server (every request served alone, at once) would show
        out.update({"slowdown_p50": pct(slow, 50), "slowdown_p99": pct(slow, 99),
                    "alone_p50_ms": pct(al, 50), "alone_p99_ms": pct(al, 99),
                    "alone_p99_over_p50": round(float(np.percentile(al, 99) / np.percentile(al, 50)), 2)})
    return out


def serve_alone(RB, classes, budget=4 << 30, reps=3):
    """Latency of each request class served on its own (median of `reps`
    requests, each submitted after the previous one completed): a dict
    class -> ms, from (class, w, h, draws) tuples."""
    srv = RB.Server(budget_bytes=budget, max_batch=32, n_streams=3)
    out = {}
    for k, (cls, w, h, draws) in enumerate(classes):
        lat = []
        for rep in range(reps + 1):
            srv.submit(k * (reps + 1) + rep, draws, w, h)
            got = []
            t_end = time.perf_counter() + 30.0
            while not got and time.perf_counter() < t_end:
                got = srv.poll(1, timeout_ms=100)
            if got and rep > 0:  # the first is a warm-up
                lat.append((got[0].done_s - got[0].submit_s) * 1e3)
        out[cls] = float(np.median(lat)) if lat else float("nan")
    srv.close()
    return out


def serve_run(rf, RB, mesh, texr, shd, a):
    """Serving: open-loop Poisson arrivals. Headline at --serve-rate over the
    crates1080 mesh at the five resolutions; then a sweep of offered rates
    over the 16 resident meshes of the served workload (8 crate fields of
    125..1000 crates, 8 icosphere clouds of 250..2000 spheres), each request
    one mesh at one resolution from a random pose, beside the plain
    sequential rate of the same requests (one synchronous rf_render each)."""
    from retrofire_amd import mat as M
    from retrofire_amd import scenes as S
    rng = S.Xorshift64(S.DEFAULT_SEED + 4)
    n = max(1, int(a.serve_rate * a.serve_seconds))
    reqs, t = [], 0.0
    for i in range(n):
        t += -math.log(max(1e-7, 1.0 - float(rng.uniform_f32(0.0, 1.0)))) / a.serve_rate
        w, h = S.RESOLUTIONS[rng.uniform_usize(0, 5)]
        reqs.append((t, w, h, crates_request(RB, mesh, shd, texr, w, h, float(rng.uniform_f32(0.0, 1.0)))))
    out = serve_stream(RB, reqs, a.serve_seconds, [(w, h, reqs[0][3]) for (w, h) in S.RESOLUTIONS],
                       streams=a.serve_streams)
    out["offered_rate"] = a.serve_rate
    out["resolutions"] = "uniform over 256^2, 512^2, 1280x720, 1920x1080, 2048^2"
    if a.serve_sweep_seconds <= 0:
        return out
    meshes, tex_px = S.serve_meshes()
    dmeshes = [RB.Mesh(p, t_, at) for (_, p, at, t_) in meshes]
    gour = rf.shader.Gouraud()

    def draws_of(w, h, mi, pose):
        if meshes[mi][0] == "crates":
            return crates_request(RB, dmeshes[mi], shd, texr, w, h, pose)
        return [RB.mesh_draw(dmeshes[mi], gour, S.spheres_mvp(w, h, pose), M.viewport(0, h, w, 0))]

    warm = [(w, h, draws_of(w, h, mi, 0.0)) for (w, h) in S.RESOLUTIONS for mi in (7, 15)]
    # each (resolution, mesh) class alone, at one pose: the reference for the
    # slowdown of the streamed requests
    alone_cls = serve_alone(RB, [((w, h, mi), w, h, draws_of(w, h, mi, 0.5)) for (w, h) in S.RESOLUTIONS
                                 for mi in range(len(meshes))])
    sweep = []
    for rate in (200.0, 500.0, 1000.0, 2000.0):
        spec = S.serve_stream(rate, max(1, int(rate * a.serve_sweep_seconds)), seed=5)
        reqs = [(t_, w, h, draws_of(w, h, mi, pose)) for (t_, w, h, mi, pose) in spec]
        r = serve_stream(RB, reqs, a.serve_sweep_seconds, warm,
                         alone=[alone_cls[(w, h, mi)] for (_, w, h, mi, _) in spec], streams=a.serve_streams)
        r["offered_rate"] = rate
        sweep.append(r)
    # the plain sequential GPU call on the same kind of requests
    spec = S.serve_stream(1000.0, 200, seed=5)
    # one packed colour and depth array per resolution, as a caller would hold
    tgt = {(w, h): (np.zeros((h, w, 4), np.uint8), np.zeros((h, w), np.float32)) for (w, h) in S.RESOLUTIONS}
    crates_shd = rf.shader.TexRepeatPot(tex_px)
    t0 = time.perf_counter()
    for (_, w, h, mi, pose) in spec:
        kind, p, at, tr = meshes[mi]
        c, d = tgt[(w, h)]
        if kind == "crates":
            rf.render(tr, (p, at), crates_shd, S.crates_mvp(w, h, pose), M.viewport(0, h, w, 0), rf.Framebuf(c, d),
                      rf.Context(clear=True))
        else:
            rf.render(tr, (p, at), gour, S.spheres_mvp(w, h, pose), M.viewport(0, h, w, 0), rf.Framebuf(c, d),
                      rf.Context(clear=True))
    plain_fps = len(spec) / (time.perf_counter() - t0)
    for _, me in enumerate(dmeshes):
        me.free()
    out["sweep_16_meshes"] = sweep
    out["sweep_plain_sequential_frames_per_s"] = round(plain_fps, 1)
    out["sweep_seconds_per_rate"] = a.serve_sweep_seconds
    out["sweep_alone_ms_range"] = [round(min(alone_cls.values()), 3), round(max(alone_cls.values()), 3)]
    return out


def c3_run(rf, RB, a, W_steps):
    """The batched workload at its stated size: 64 frames of uneven
    resolution, each a seeded mix of the full crates and icosphere scenes
    (10-100 % of each) from random poses, in one HIP-graph launch."""
    from retrofire_amd import scenes as S
    cpos, cuv, ctris, tex = S.crates_world(1000)
    ipos, iattr, itris, _, nt = S.icospheres_world(10000)
    frames = S.batch_frames(64, ctris.shape[0], itris.shape[0], nt, seed=0)
    crates = RB.Mesh(cpos, ctris, cuv)
    icos = RB.Mesh(ipos, itris, iattr)
    texr = RB.Texture(tex)
    fbs, spec = [], []
    for f in frames:
        fb = RB.DeviceFramebuf(f["w"], f["h"], 0, True)
        fbs.append(fb)
        spec.append((fb, rf.Context(clear=True), [
            RB.mesh_draw(crates, rf.shader.TexRepeatPot(tex), f["mvp_c"], f["vp"], n_tris=f["nc"], texture=texr),
            RB.mesh_draw(icos, rf.shader.Gouraud(), f["mvp_i"], f["vp"], n_tris=f["ni"])]))
    t0 = time.perf_counter()
    b = RB.Batch(spec)
    plan_s = time.perf_counter() - t0
    b.run(sync=True)
    stats = b.wait()
    t0 = time.perf_counter()
    for _ in range(W_steps):
        b.run(sync=False)
    b.wait()
    el = time.perf_counter() - t0
    prof = b.profile(iters=3)
    alg = sum(f["w"] * f["h"] * 8 + tex.nbytes for f in frames)
    out = {"frames_per_step": 64, "steps": W_steps, "ms_per_step": round(1e3 * el / W_steps, 3),
           "frames_per_s": round(64 * W_steps / el, 1),
           "mfrag_per_s": round(sum(s.frags_i for s in stats) * W_steps / el / 1e6, 1),
           "triangles_per_step": int(sum(f["nc"] + f["ni"] for f in frames)),
           "pixels_per_step": int(sum(f["w"] * f["h"] for f in frames)),
           "frags_o_per_step": int(sum(s.frags_o for s in stats)),
           "tile_w": b.tile_w, "tile_h": b.tile_h, "chunk": b.chunk, "plan_s": round(plan_s, 2),
           "autotune_ms": [list(t) for t in b.tune_results()],
           "stage_ms": {n: round(x, 4) for n, x in zip(["vertex", "setup", "scan", "bin", "binsort", "raster"], prof)},
           "raster_roofline": {"alg_bytes_per_launch": alg, "launch_ms": round(prof[5], 4),
                               "achieved_GBs": round(alg / (prof[5] * 1e-3) / 1e9, 1),
                               "frac": round(alg / (prof[5] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    b.free()
    for fb in fbs:
        fb.free()
    crates.free()
    icos.free()
    texr.free()
    return out


def c2_run(rf, RB, a, W_steps, n_frames=8):
    """icospheres1080 (10,000 Gouraud icospheres, 3.2M triangles per frame)
    batched: n_frames copies of the frame in one HIP-graph launch, tile shape
    and chunk autotuned on it."""
    from retrofire_amd import scenes as S
    sc = S.icospheres1080()
    mesh = RB.Mesh(sc.pos, sc.tris, sc.attr)
    fbs = [RB.DeviceFramebuf(sc.width, sc.height, 0, True) for _ in range(n_frames)]
    b = RB.Batch([(fb, rf.Context(clear=True), [RB.mesh_draw(mesh, rf.shader.Gouraud(), sc.mvp, sc.viewport)]) for fb in fbs])
    b.run(sync=True)
    stats = b.wait()
    t0 = time.perf_counter()
    for _ in range(W_steps):
        b.run(sync=False)
    b.wait()
    el = time.perf_counter() - t0
    prof = b.profile(iters=3)
    alg = n_frames * sc.width * sc.height * 8
    out = {"frames_per_step": n_frames, "steps": W_steps, "ms_per_step": round(1e3 * el / W_steps, 3),
           "frames_per_s": round(n_frames * W_steps / el, 1),
           "mfrag_per_s": round(sum(s.frags_i for s in stats) * W_steps / el / 1e6, 1),
           "tile_w": b.tile_w, "tile_h": b.tile_h, "chunk": b.chunk,
           "autotune_ms": [list(t) for t in b.tune_results()],
           "stage_ms": {n: round(x, 4) for n, x in zip(["vertex", "setup", "scan", "bin", "binsort", "raster"], prof)},
           "raster_roofline": {"alg_bytes_per_launch": alg, "launch_ms": round(prof[5], 4),
                               "achieved_GBs": round(alg / (prof[5] * 1e-3) / 1e9, 1),
                               "frac": round(alg / (prof[5] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    b.free()
    for fb in fbs:
        fb.free()
    mesh.free()
    return out


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds):
    """The oracle (a single-threaded C++ restatement of the reference CPU
    path, -O2, fp-contract off) timed per configuration on one host core:
    the process is pinned to one core while it runs; each frame is
    Frame::clear plus one render() call; median over the sample."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from retrofire_amd import scenes as S
    L = O.lib()
    old_aff = os.sched_getaffinity(0)
    core = min(old_aff)
    os.sched_setaffinity(0, {core})
    budget = {"sphere640": 0.15, "crates1080": 0.35, "icospheres1080": 0.5}
    cfgs = {}
    try:
        for name, key in (("sphere640", "c0"), ("crates1080", "c1"), ("icospheres1080", "c2")):
            sc = S.SCENES[name]()
            col = np.zeros((sc.height, sc.width, 4), np.uint8)
            dep = np.zeros((sc.height, sc.width), np.float32)
            c = O.OracleCtx(1, 1, 1, 1)
            clear = np.array([0, 0, 0, 255], np.uint8)
            flat = np.array([255, 255, 255, 255], np.uint8)
            tr = np.ascontiguousarray(sc.tris.reshape(-1), np.uint32)
            ps = np.ascontiguousarray(sc.pos.reshape(-1), np.float32)
            at = np.ascontiguousarray(sc.attr.reshape(-1), np.float32)
            m = np.ascontiguousarray(sc.mvp.reshape(-1), np.float32)
            vpf = np.ascontiguousarray(sc.viewport.reshape(-1), np.float32)
            tx = np.ascontiguousarray(sc.tex if sc.tex is not None else np.zeros((1, 1, 4), np.uint8))
            times, frags = [], 0
            t_end = time.perf_counter() + seconds * budget[name]
            while (time.perf_counter() < t_end and len(times) < 1000) or len(times) < 3:
                st = np.zeros(8, np.uint64)
                per = (C.c_double * 1)()
                s = L.oracle_time_frames(sc.kind, O._p(tr, C.c_uint32), sc.tris.shape[0], O._p(ps, C.c_float),
                                         O._p(at, C.c_float), sc.pos.shape[0], O._p(m, C.c_float),
                                         O._p(vpf, C.c_float), flat.ctypes.data_as(C.POINTER(C.c_uint8)),
                                         tx.ctypes.data_as(C.POINTER(C.c_uint8)), tx.shape[1], tx.shape[0],
                                         tx.shape[1], col.ctypes.data_as(C.POINTER(C.c_uint8)), sc.width, sc.height,
                                         0, O._p(dep, C.c_float), C.byref(c),
                                         clear.ctypes.data_as(C.POINTER(C.c_uint8)), float("inf"), 1, per,
                                         O._p(st, C.c_uint64))
                assert s >= 0
                times.append(s)
                frags = int(st[5])
            med = float(np.median(times))
            cfgs[key] = {"workload": name, "frames": len(times), "median_ms": round(med * 1e3, 3),
                         "frames_per_s": round(1.0 / med, 3), "mfrag_per_s": round(frags / med / 1e6, 3),
                         "frags_i_per_frame": frags}
    finally:
        os.sched_setaffinity(0, old_aff)
    return {"value": cfgs["c1"]["frames_per_s"], "unit": "frames/s", "cores": 1, "kind": "port",
            "host_cpu": host_cpu_model(), "host_cores": os.cpu_count(), "pinned_core": core,
            "sample": "per configuration, Frame::clear + render() frames on one pinned host core, median frame "
                      "time: sphere640, crates1080 (the value), icospheres1080; the C++ restatement of render() "
                      "(-O2, fp-contract off), not the Rust crate",
            "configs": cfgs}


if __name__ == "__main__":
    main()
