// Batched path: many frames of uneven resolution, each an ordered list of
// draws over resident meshes, planned once and replayed as one HIP graph.
//
// Planning runs the pipeline eagerly to learn the exact clip-fan and bin-run
// counts (the batch's inputs are fixed, so replays need exactly as much), sizes
// the scratch, picks the tile width by timing the candidates when asked to,
// and captures the launch sequence on the batch's own stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

namespace rf {

struct Batch {
    Device* dev = nullptr;
    hipStream_t stream = nullptr;
    std::vector<rf_frame> frames;
    std::vector<std::vector<rf_mesh_draw>> draws;
    std::vector<DrawDesc> h_draws;
    std::vector<FrameDesc> h_frames;
    std::vector<uint64_t> frame_prims_i, frame_verts_i, frame_calls, frame_objs;
    DevBuf d_desc;
    Scratch sc;
    BatchDesc B{};
    uint32_t tile_w = 64, tile_h = DEFAULT_TILE_H;  // rasterizer tile shape
    uint32_t chunk = 8;  // rasterizer span chunk (6 or 8)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    FrameStats* h_stats = nullptr;  // pinned
    uint32_t* h_misc = nullptr;     // pinned
    bool pending = false;
    std::vector<rf_tune_entry> tune;  // the autotuner's timings (tile width, height, chunk, ms)
    // depth sort: the triangle draws of frames whose context asks for it
    SortPass sp;
    std::vector<int32_t> sort_of;     // per draw (in batch order): its draw in sp, or -1
};

// The device descriptor of mesh draw `md` of frame `fi` (draw order
// positions vert_base / tri_base are the caller's).
static DrawDesc draw_desc(const rf_mesh_draw& md, uint32_t fi) {
    const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
    const Tex* t = reinterpret_cast<const Tex*>(md.tex);
    DrawDesc dd;
    std::memset(&dd, 0, sizeof(dd));
    const uint64_t first = md.first_tri;
    const uint64_t nt = md.n_tris ? md.n_tris : m->n_tris - first;
    dd.pos = (const float*)m->pos.p;
    dd.attr = (const float*)m->attr.p;
    dd.tris = (const uint32_t*)m->tris.p + (m->prim == PRIM_EDGE ? 2 : 3) * first;
    dd.prim = m->prim;
    dd.tex = t ? t->texels() : nullptr;
    dd.n_verts = (uint32_t)m->n_verts;
    dd.n_tris = (uint32_t)nt;
    dd.kind = md.kind;
    dd.frame = fi;
    dd.tex_w = t ? t->w : 0;
    dd.tex_h = t ? t->h : 0;
    dd.tex_stride = t ? t->stride : 0;
    dd.flat_rgba = pack_pixel(PIX_RGBA8, md.flat_color[0], md.flat_color[1], md.flat_color[2], md.flat_color[3]);
    dd.attr_stride = m->n_attr;
    std::memcpy(dd.mvp, md.mvp, 64);
    std::memcpy(dd.vp, md.viewport, 64);
    dd.cull = md.cull;
    std::memcpy(dd.bbox, m->bbox, sizeof(dd.bbox));
    return dd;
}

static rf_status build_descs(Batch& b, uint32_t tile_w, uint32_t tile_h, uint32_t chunk) {
    Device& d = *b.dev;
    b.h_draws.clear();
    b.h_frames.clear();
    b.frame_prims_i.assign(b.frames.size(), 0);
    b.frame_verts_i.assign(b.frames.size(), 0);
    b.frame_calls.assign(b.frames.size(), 0);
    b.frame_objs.assign(b.frames.size(), 0);
    uint32_t any_cull = 0;
    uint64_t vbase = 0, tbase = 0;
    uint32_t tile_base = 0;
    for (size_t fi = 0; fi < b.frames.size(); ++fi) {
        const rf_frame& f = b.frames[fi];
        const Framebuf* fb = reinterpret_cast<const Framebuf*>(f.target);
        FrameDesc fd;
        std::memset(&fd, 0, sizeof(fd));
        fd.color = (uint8_t*)fb->color.p;
        fd.depth = fb->has_depth ? (float*)fb->depth.p : nullptr;
        fd.color_stride = fb->stride;
        fd.depth_stride = fb->w;
        fd.width = fb->w;
        fd.height = fb->h;
        fd.fmt = fb->fmt;
        fill_frame_ctx(fd, f.ctx);
        set_grid(fd, tile_w, tile_h, tile_base);
        b.h_frames.push_back(fd);
        for (const rf_mesh_draw& md : b.draws[fi]) {
            const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
            DrawDesc dd = draw_desc(md, (uint32_t)fi);
            const uint64_t nt = dd.n_tris;
            b.frame_prims_i[fi] += nt;
            b.frame_verts_i[fi] += m->n_verts;
            const int32_t si = b.sort_of[b.h_draws.size()];
            if (si >= 0) {
                // a depth-sorted render() call: its sorted clip fans, 3 vertex
                // slots each, filled by the sort pass's gather
                const uint32_t nf = b.sp.cnt[si];
                dd.pre = 1;
                dd.sort = (uint32_t)f.ctx.depth_sort;
                dd.in_tris = dd.n_tris;
                dd.in_verts = dd.n_verts;
                dd.n_tris = nf;
                dd.n_verts = 3 * nf;
                dd.tris = sortpass_idx(b.sp, b.sp.off[si]);
                b.sp.main_of[si] = (uint32_t)b.h_draws.size();
            }
            dd.vert_base = vbase;
            dd.tri_base = tbase;
            any_cull |= md.cull;
            vbase += dd.n_verts;
            tbase += dd.n_tris;
            b.frame_calls[fi] += 1;
            b.frame_objs[fi] += md.cull;
            b.h_draws.push_back(dd);
        }
    }
    if (tbase >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    if (!b.sp.draws.empty() && sortpass_map(b.sp, b.stream)) return RF_E_DEVICE;
    const size_t db = b.h_draws.size() * sizeof(DrawDesc), fb = b.h_frames.size() * sizeof(FrameDesc);
    const size_t foff = (db + 255) & ~size_t(255);
    if (b.d_desc.ensure(foff + fb + 256, &d.bytes)) return RF_E_DEVICE;
    if (db && hipMemcpy(b.d_desc.p, b.h_draws.data(), db, hipMemcpyHostToDevice)) return RF_E_DEVICE;
    if (hipMemcpy((uint8_t*)b.d_desc.p + foff, b.h_frames.data(), fb, hipMemcpyHostToDevice)) return RF_E_DEVICE;
    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)b.d_desc.p;
    B.frames = (const FrameDesc*)((uint8_t*)b.d_desc.p + foff);
    B.n_draws = (uint32_t)b.h_draws.size();
    B.n_frames = (uint32_t)b.h_frames.size();
    B.n_verts = vbase;
    B.n_tris = tbase;
    B.n_groups = (uint32_t)((tbase + WAVE_SIZE - 1) / WAVE_SIZE);
    B.n_tiles = tile_base;
    B.any_cull = any_cull;
    B.tile_w = tile_w;
    B.tile_h = tile_h;
    B.chunk = chunk;
    B.extra_cap = std::max<uint32_t>(b.B.extra_cap, 1024u);
    B.runs_cap = std::max<uint32_t>(b.B.runs_cap, 4 * B.n_groups + B.n_tiles + 1024);
    b.B = B;
    b.tile_w = tile_w;
    b.tile_h = tile_h;
    b.chunk = chunk;
    return RF_OK;
}

// The depth sort pass (if the batch has one): P's fans, their sorts and the
// gather into the batch's current scratch.
static hipError_t sort_pass(Batch& b, const BatchDesc& M) {
    if (b.sp.draws.empty()) return hipSuccess;
    hipError_t e = sortpass_fans(b.sp, b.stream);
    return e ? e : sortpass_sort_gather(b.sp, M, b.stream);
}

// Eager run that sizes the scratch exactly for this batch (with headroom).
static rf_status size_batch(Batch& b) {
    std::vector<FrameStats> fs;
    uint32_t err = 0;
    rf_status st = run_sized(*b.dev, b.sc, b.B, b.stream, [&b](const BatchDesc& M) { return sort_pass(b, M); },
                             fs, &err, nullptr, &b.dev->bytes);
    if (st) return st;
    if ((st = map_err(err))) return st;
    return RF_OK;
}

static float time_eager(Batch& b, int reps) {
    if (hipEventRecord(b.ev0, b.stream)) return -1.0f;
    for (int i = 0; i < reps; ++i)
        if (sort_pass(b, b.B) || launch_pipeline(b.B, b.stream, raster_grid(*b.dev, b.B))) return -1.0f;
    if (hipEventRecord(b.ev1, b.stream) || hipEventSynchronize(b.ev1)) return -1.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b.ev0, b.ev1)) return -1.0f;
    return ms / reps;
}

static void destroy_graph(Batch& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

static rf_status capture(Batch& b) {
    destroy_graph(b);
    if (hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal)) return RF_E_DEVICE;
    hipError_t le = sort_pass(b, b.B);
    if (!le) le = launch_pipeline(b.B, b.stream, raster_grid(*b.dev, b.B));
    const hipError_t ee = hipStreamEndCapture(b.stream, &b.graph);
    if (le || ee || !b.graph) return RF_E_DEVICE;
    if (hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0)) return RF_E_DEVICE;
    return RF_OK;
}

static void batch_release(Batch* b) {
    if (!b) return;
    Device& d = *b->dev;
    (void)hipSetDevice(d.ordinal);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graph(*b);
    b->sc.release(&d.bytes);
    sortpass_release(b->sp, &d.bytes);
    b->d_desc.release(&d.bytes);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->h_stats) (void)hipHostFree(b->h_stats);
    if (b->h_misc) (void)hipHostFree(b->h_misc);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace rf

using namespace rf;

extern "C" {

rf_status rf_batch_create(rf_device* dev, const rf_frame* frames, uint32_t n_frames, uint32_t tile_w,
                          rf_batch** out) {
    if (!dev || !out || !frames || n_frames == 0) return RF_E_ARG;
    if (tile_w != 0 && tile_w != 16 && tile_w != 32 && tile_w != 64) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    for (uint32_t i = 0; i < n_frames; ++i) {
        const rf_frame& f = frames[i];
        if (!f.target || (f.n_draws && !f.draws)) return RF_E_ARG;
        rf_status st = check_ctx(&f.ctx);
        if (st) return st;
        for (uint32_t k = 0; k < f.n_draws; ++k) {
            const rf_mesh_draw& md = f.draws[k];
            if (!md.mesh || md.kind > KIND_LAST || md.cull > 1) return RF_E_ARG;
            const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
            const Tex* t = reinterpret_cast<const Tex*>(md.tex);
            if ((int)m->n_attr < kind_attrs(md.kind)) return RF_E_ARG;
            if (md.first_tri > m->n_tris || md.first_tri + md.n_tris > m->n_tris) return RF_E_ARG;
            if ((st = check_tex(md.kind, t ? t->texels() : nullptr, t ? t->w : 0, t ? t->h : 0, t ? t->stride : 0)))
                return st;
        }
    }
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    Batch* b = new (std::nothrow) Batch();
    if (!b) return RF_E_DEVICE;
    b->dev = &d;
    b->frames.assign(frames, frames + n_frames);
    b->draws.resize(n_frames);
    for (uint32_t i = 0; i < n_frames; ++i) {
        b->draws[i].assign(frames[i].draws, frames[i].draws + frames[i].n_draws);
        b->frames[i].draws = nullptr;
    }
    rf_status st = RF_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) || hipEventCreate(&b->ev0) ||
        hipEventCreate(&b->ev1) || hipHostMalloc((void**)&b->h_stats, n_frames * sizeof(FrameStats), 0) ||
        hipHostMalloc((void**)&b->h_misc, 64, 0)) {
        batch_release(b);
        return RF_E_DEVICE;
    }
    std::memset(b->h_misc, 0, 64);
    // Depth sort (render.rs:180-183): the triangle draws of frames that ask
    // for it go through the sort pass; its planning run gives their fan counts.
    for (uint32_t i = 0; i < n_frames; ++i)
        for (const rf_mesh_draw& md : b->draws[i]) {
            const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
            if (frames[i].ctx.depth_sort && m->prim == PRIM_TRI) {
                b->sort_of.push_back((int32_t)b->sp.draws.size());
                DrawDesc dd = draw_desc(md, i);
                dd.sort = (uint32_t)frames[i].ctx.depth_sort;
                b->sp.draws.push_back(dd);
            } else {
                b->sort_of.push_back(-1);
            }
        }
    if (!b->sp.draws.empty()) {
        b->sp.frames_n = n_frames;
        b->sp.main_of.assign(b->sp.draws.size(), 0);
        uint32_t e = 0;
        st = sortpass_plan(d, b->sp, b->stream, &d.bytes, &e);
        if (!st) st = map_err(e);
        if (st) {
            batch_release(b);
            return st;
        }
    }
    // Tile shape and span chunk: the tile width forced (with the device's
    // tile height and chunk), or the fastest (width, height, chunk) of the
    // 3 x 3 x 2 candidates, each timed on this batch.
    struct Cand {
        uint32_t w, h, c;
    };
    std::vector<Cand> cands;
    if (tile_w)
        cands.push_back({tile_w, d.plain_tile_h, d.plain_chunk});
    else
        for (uint32_t w : {16u, 32u, 64u})
            for (uint32_t h : {16u, 32u, 64u})
                for (uint32_t c : {6u, 8u}) cands.push_back({w, h, c});
    float best = 1e30f;
    Cand bc = cands[0];
    b->tune.clear();
    for (const Cand& c : cands) {
        if ((st = build_descs(*b, c.w, c.h, c.c)) || (st = size_batch(*b))) {
            batch_release(b);
            return st;
        }
        if (cands.size() > 1) {
            const float ms = time_eager(*b, 3);
            if (ms < 0) {
                batch_release(b);
                return RF_E_DEVICE;
            }
            b->tune.push_back({c.w, c.h, c.c, ms});
            if (ms < best) {
                best = ms;
                bc = c;
            }
        }
    }
    if ((bc.w != b->tile_w || bc.h != b->tile_h || bc.c != b->chunk) &&
        ((st = build_descs(*b, bc.w, bc.h, bc.c)) || (st = size_batch(*b)))) {
        batch_release(b);
        return st;
    }
    // Then the tiles per binsort wave: 8 (the default) or 1 (dense scenes'
    // long run lists spread over more waves).
    if (cands.size() > 1) {
        b->B.sort_tpw = 1;
        const float ms1 = time_eager(*b, 3);
        b->B.sort_tpw = 0;
        const float ms8 = time_eager(*b, 3);
        if (ms1 < 0 || ms8 < 0) {
            batch_release(b);
            return RF_E_DEVICE;
        }
        if (ms1 < ms8) b->B.sort_tpw = 1;
    }
    // headroom, then the final scratch and the graph
    b->B.extra_cap += b->B.extra_cap / 8 + 64;
    b->B.runs_cap += b->B.runs_cap / 8 + 64;
    if (b->sc.prepare(b->B, &d.bytes) || (st = capture(*b))) {
        batch_release(b);
        return st ? st : RF_E_DEVICE;
    }
    *out = reinterpret_cast<rf_batch*>(b);
    return RF_OK;
}

rf_status rf_batch_run(rf_batch* bh, int sync) {
    if (!bh) return RF_E_ARG;
    Batch& b = *reinterpret_cast<Batch*>(bh);
    if (hipSetDevice(b.dev->ordinal)) return RF_E_DEVICE;
    if (hipEventRecord(b.ev0, b.stream) || hipGraphLaunch(b.exec, b.stream) || hipEventRecord(b.ev1, b.stream))
        return RF_E_DEVICE;
    if (hipMemcpyAsync(b.h_stats, b.B.stats, b.B.n_frames * sizeof(FrameStats), hipMemcpyDeviceToHost,
                       b.stream) ||
        hipMemcpyAsync(b.h_misc, misc_words(b.B), 16, hipMemcpyDeviceToHost, b.stream))
        return RF_E_DEVICE;
    // the sort pass's error bits (its own counters)
    if (!b.sp.draws.empty() &&
        hipMemcpyAsync(b.h_misc + 8, misc_words(b.sp.P) + 2, 4, hipMemcpyDeviceToHost, b.stream))
        return RF_E_DEVICE;
    b.pending = true;
    if (sync) return rf_batch_wait(bh, nullptr);
    return RF_OK;
}

rf_status rf_batch_wait(rf_batch* bh, rf_stats* stats) {
    if (!bh) return RF_E_ARG;
    Batch& b = *reinterpret_cast<Batch*>(bh);
    if (hipStreamSynchronize(b.stream)) return RF_E_DEVICE;
    b.pending = false;
    rf_status st = map_err(b.h_misc[2] | b.h_misc[8]);
    if (st) return st;
    if (stats) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, b.ev0, b.ev1);
        for (uint32_t i = 0; i < b.B.n_frames; ++i) {
            rf_stats& s = stats[i];
            const FrameStats& h = b.h_stats[i];
            s.calls = b.frame_calls[i] - h.hidden_draws;
            s.prims_i = b.frame_prims_i[i] - h.hidden_prims;
            s.verts_i = b.frame_verts_i[i] - h.hidden_verts;
            s.objs_i = b.frame_objs[i];
            s.objs_o = b.frame_objs[i] - h.hidden_draws;
            s.prims_o = b.h_stats[i].prims_o;
            s.verts_o = b.h_stats[i].verts_o;
            s.frags_i = b.h_stats[i].frags_i;
            s.frags_o = b.h_stats[i].frags_o;
            s.seconds = ms * 1e-3;
        }
    }
    return RF_OK;
}

rf_status rf_batch_free(rf_batch* bh) {
    if (!bh) return RF_E_ARG;
    batch_release(reinterpret_cast<Batch*>(bh));
    return RF_OK;
}

uint32_t rf_batch_tile_width(const rf_batch* bh) {
    return bh ? reinterpret_cast<const Batch*>(bh)->tile_w : 0;
}

uint32_t rf_batch_raster_chunk(const rf_batch* bh) {
    return bh ? reinterpret_cast<const Batch*>(bh)->chunk : 0;
}

uint32_t rf_batch_tile_height(const rf_batch* bh) {
    return bh ? reinterpret_cast<const Batch*>(bh)->tile_h : 0;
}

uint32_t rf_batch_tune_results(const rf_batch* bh, rf_tune_entry* out, uint32_t max) {
    if (!bh) return 0;
    const Batch& b = *reinterpret_cast<const Batch*>(bh);
    const uint32_t n = (uint32_t)std::min<size_t>(b.tune.size(), max);
    for (uint32_t i = 0; i < n && out; ++i) out[i] = b.tune[i];
    return (uint32_t)b.tune.size();
}

rf_status rf_batch_last_ms(const rf_batch* bh, float* total_ms, float* raster_ms) {
    if (!bh) return RF_E_ARG;
    const Batch& b = *reinterpret_cast<const Batch*>(bh);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b.ev0, b.ev1)) return RF_E_DEVICE;
    if (total_ms) *total_ms = ms;
    if (raster_ms) *raster_ms = -1.0f;  // not separable inside the graph; see rf_batch_profile
    return RF_OK;
}

rf_status rf_batch_profile(rf_batch* bh, int iters, float* ms) {
    if (!bh || !ms || iters <= 0) return RF_E_ARG;
    Batch& b = *reinterpret_cast<Batch*>(bh);
    if (hipSetDevice(b.dev->ordinal)) return RF_E_DEVICE;
    hipEvent_t ev[7];
    for (auto& e : ev)
        if (hipEventCreate(&e)) return RF_E_DEVICE;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    rf_status st = RF_OK;
    for (int it = 0; it < iters && !st; ++it) {
        if (launch_pipeline_timed(b.B, b.stream, raster_grid(*b.dev, b.B), ev) ||
            hipEventSynchronize(ev[6])) {
            st = RF_E_DEVICE;
            break;
        }
        for (int k = 0; k < 6; ++k) {
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            acc[k] += t;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 6; ++k) ms[k] = (float)(acc[k] / iters);
    return st;
}

}  // extern "C"
