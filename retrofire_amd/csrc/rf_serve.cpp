// Serving: a request loop over resident meshes and textures.
//
// Requests (one frame each: a resolution, a context and an ordered list of
// draws over resident meshes) are admitted in arrival order while their
// device memory fits the budget; whatever is admitted when a stream is free
// goes out as one pipeline launch (continuous batching). Up to `n_streams`
// batches are on the device at once, so frames that arrive while others are
// rendering overlap them on the GPU. Each stream has a waiter thread that
// blocks on an event recorded behind its batch; results carry submit/start/done
// times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

namespace rf {

namespace {

struct Req {
    rf_request r;
    std::vector<rf_mesh_draw> draws;
    double submit_s = 0;
    uint64_t bytes = 0;       // budget charge
    uint64_t n_tris = 0, n_verts = 0;
};

struct Done {
    rf_result res;
    uint8_t* pixels = nullptr;  // pinned, width * height * bpp
    size_t pix_bytes = 0;
};

using FbKey = std::tuple<uint32_t, uint32_t, uint32_t, int>;

struct Lane {
    hipStream_t stream = nullptr;
    Scratch sc;
    DevBuf desc;
    uint8_t* h_desc = nullptr;  // pinned staging for descriptors
    size_t h_desc_cap = 0;
    FrameStats* h_stats = nullptr;
    uint32_t* h_misc = nullptr;
    uint32_t stats_cap = 0;
    uint32_t extra_cap = 4096, runs_cap = 1 << 16;
    bool busy = false;
    // the batch in flight
    std::vector<Req> reqs;
    std::vector<Framebuf*> fbs;
    std::vector<uint8_t*> pix;
    std::vector<size_t> pix_bytes;
    BatchDesc B{};
    double start_s = 0;
    std::atomic<int> finished{0};
    double done_s = 0;
    int attempts = 0;
    // completion: the worker records `ev` behind the batch and sets `launched`;
    // the lane's waiter thread blocks in hipEventSynchronize (holding no lock),
    // then sets `finished`
    hipEvent_t ev = nullptr;
    bool launched = false;
    std::condition_variable cv_launch;
    std::thread waiter;
};

}  // namespace

struct Serve {
    Device* dev = nullptr;
    uint64_t bytes = 0;  // device memory held by the serving loop (worker thread only)
    uint64_t budget = 0, used = 0;
    uint32_t max_batch = 16;
    std::vector<Lane*> lanes;
    std::thread worker;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req> pending;
    std::deque<Done> done;
    bool stop = false;
    bool stop_waiters = false;
    std::mutex pm;  // guards the pools; never held across a HIP call
    std::map<FbKey, std::vector<Framebuf*>> fb_pool;
    std::map<size_t, std::vector<uint8_t*>> pin_pool;
};

static void waiter_main(Serve* sp, Lane* l) {
    Serve& s = *sp;
    (void)hipSetDevice(s.dev->ordinal);
    std::unique_lock<std::mutex> lk(s.m);
    for (;;) {
        l->cv_launch.wait(lk, [&] { return l->launched || s.stop_waiters; });
        if (!l->launched) break;
        l->launched = false;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(l->ev);
        const double t = now_s();
        lk.lock();
        if (e != hipSuccess) l->h_misc[8] = 1;
        l->done_s = t;
        l->finished.store(1);
        s.cv_work.notify_all();
    }
}

static Framebuf* fb_get(Serve& s, uint32_t w, uint32_t h, uint32_t fmt, int depth) {
    {
        std::lock_guard<std::mutex> g(s.pm);
        auto& v = s.fb_pool[FbKey(w, h, fmt, depth)];
        if (!v.empty()) {
            Framebuf* f = v.back();
            v.pop_back();
            return f;
        }
    }
    Framebuf* f = new Framebuf();
    f->w = w;
    f->h = h;
    f->fmt = fmt;
    f->stride = w * bytes_per_pixel(fmt);
    f->has_depth = depth != 0;
    uint64_t bytes = 0;
    if (f->color.ensure((size_t)h * f->stride, &bytes) || (depth && f->depth.ensure((size_t)w * h * 4, &bytes))) {
        f->color.release(&bytes);
        f->depth.release(&bytes);
        delete f;
        return nullptr;
    }
    return f;
}
static void fb_put(Serve& s, Framebuf* f) {
    if (!f) return;
    std::lock_guard<std::mutex> g(s.pm);
    s.fb_pool[FbKey(f->w, f->h, f->fmt, f->has_depth)].push_back(f);
}

static uint8_t* pin_get(Serve& s, size_t bytes) {
    {
        std::lock_guard<std::mutex> g(s.pm);
        auto& v = s.pin_pool[bytes];
        if (!v.empty()) {
            uint8_t* p = v.back();
            v.pop_back();
            return p;
        }
    }
    uint8_t* p = nullptr;
    if (hipHostMalloc((void**)&p, bytes, 0)) return nullptr;
    return p;
}
static void pin_put(Serve& s, uint8_t* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(s.pm);
    s.pin_pool[bytes].push_back(p);
}

// Device bytes a request holds while in flight: its target plus the pipeline
// scratch its geometry needs.
static uint64_t request_bytes(const Req& q) {
    const uint64_t bpp = bytes_per_pixel(q.r.fmt);
    const uint64_t px = (uint64_t)q.r.width * q.r.height;
    const uint64_t target = px * (bpp + (q.r.with_depth ? 4 : 0));
    const uint64_t geom = q.n_verts * 56 + q.n_tris * (sizeof(ScrTri) + 12) + (q.n_tris / 64 + 1) * 4 * sizeof(Run);
    return target + geom;
}

static void finish_lane(Serve& s, Lane& l, rf_status st) {
    // called with s.m held
    std::vector<FrameStats> fs(l.reqs.size());
    for (size_t i = 0; i < l.reqs.size(); ++i) {
        Req& q = l.reqs[i];
        Done d;
        std::memset(&d.res, 0, sizeof(d.res));
        d.res.id = q.r.id;
        d.res.status = st;
        d.res.width = q.r.width;
        d.res.height = q.r.height;
        d.res.fmt = q.r.fmt;
        d.res.submit_s = q.submit_s;
        d.res.start_s = l.start_s;
        d.res.done_s = l.done_s;
        uint64_t objs = 0;
        for (const rf_mesh_draw& md : q.draws) objs += md.cull;
        d.res.stats.calls = q.draws.size();
        d.res.stats.prims_i = q.n_tris;
        d.res.stats.verts_i = q.n_verts;
        d.res.stats.objs_i = d.res.stats.objs_o = objs;
        if (st == RF_OK) {
            const FrameStats& h = l.h_stats[i];
            d.res.stats.calls -= h.hidden_draws;
            d.res.stats.prims_i -= h.hidden_prims;
            d.res.stats.verts_i -= h.hidden_verts;
            d.res.stats.objs_o -= h.hidden_draws;
            d.res.stats.prims_o = l.h_stats[i].prims_o;
            d.res.stats.verts_o = l.h_stats[i].verts_o;
            d.res.stats.frags_i = l.h_stats[i].frags_i;
            d.res.stats.frags_o = l.h_stats[i].frags_o;
        }
        d.res.stats.seconds = l.done_s - l.start_s;
        d.pixels = l.pix[i];
        d.pix_bytes = l.pix_bytes[i];
        s.done.push_back(d);
        s.used -= q.bytes;
        fb_put(s, l.fbs[i]);
    }
    l.reqs.clear();
    l.fbs.clear();
    l.pix.clear();
    l.pix_bytes.clear();
    l.busy = false;
    s.cv_done.notify_all();
}

// Builds descriptors for the lane's requests and enqueues the pipeline, the
// readbacks and the completion event. Called by the worker without s.m.
static rf_status launch_lane(Serve& s, Lane& l) {
    Device& d = *s.dev;
    std::vector<DrawDesc> dds;
    std::vector<FrameDesc> fds;
    uint64_t vbase = 0, tbase = 0;
    uint32_t tile_base = 0;
#ifndef RF_SERVE_CHUNK
// 6-pixel span chunks: served launches mix meshes of small triangles, where
// they shorten the latency, with larger ones, where they cost nothing
// measurable
#define RF_SERVE_CHUNK 6
#endif
#ifndef RF_SERVE_TILE_W
#define RF_SERVE_TILE_W 32
#endif
    const uint32_t tile_w = RF_SERVE_TILE_W;  // rasterizer tile width of served launches
    uint32_t any_cull = 0;
    for (size_t i = 0; i < l.reqs.size(); ++i) {
        Req& q = l.reqs[i];
        Framebuf* fb = l.fbs[i];
        FrameDesc fd;
        std::memset(&fd, 0, sizeof(fd));
        fd.color = (uint8_t*)fb->color.p;
        fd.depth = fb->has_depth ? (float*)fb->depth.p : nullptr;
        fd.color_stride = fb->stride;
        fd.depth_stride = fb->w;
        fd.width = fb->w;
        fd.height = fb->h;
        fd.fmt = fb->fmt;
        rf_ctx c = q.r.ctx;
        c.clear_color = 1;  // a served frame starts from Frame::clear
        c.clear_depth = 1;
        fill_frame_ctx(fd, c);
        set_grid(fd, tile_w, tile_base);
        fds.push_back(fd);
        for (const rf_mesh_draw& md : q.draws) {
            const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
            const Tex* t = reinterpret_cast<const Tex*>(md.tex);
            DrawDesc dd;
            std::memset(&dd, 0, sizeof(dd));
            const uint64_t nt = md.n_tris ? md.n_tris : m->n_tris - md.first_tri;
            dd.pos = (const float*)m->pos.p;
            dd.attr = (const float*)m->attr.p;
            dd.tris = (const uint32_t*)m->tris.p + (m->prim == PRIM_EDGE ? 2 : 3) * md.first_tri;
            dd.prim = m->prim;
            dd.tex = t ? t->texels() : nullptr;
            dd.n_verts = (uint32_t)m->n_verts;
            dd.n_tris = (uint32_t)nt;
            dd.kind = md.kind;
            dd.frame = (uint32_t)i;
            dd.tex_w = t ? t->w : 0;
            dd.tex_h = t ? t->h : 0;
            dd.tex_stride = t ? t->stride : 0;
            dd.flat_rgba = pack_pixel(PIX_RGBA8, md.flat_color[0], md.flat_color[1], md.flat_color[2],
                                      md.flat_color[3]);
            dd.attr_stride = m->n_attr;
            dd.vert_base = vbase;
            dd.tri_base = tbase;
            std::memcpy(dd.mvp, md.mvp, 64);
            std::memcpy(dd.vp, md.viewport, 64);
            dd.cull = md.cull;
            std::memcpy(dd.bbox, m->bbox, sizeof(dd.bbox));
            any_cull |= md.cull;
            vbase += m->n_verts;
            tbase += nt;
            dds.push_back(dd);
        }
    }
    const size_t db = dds.size() * sizeof(DrawDesc), fb = fds.size() * sizeof(FrameDesc);
    const size_t foff = (db + 255) & ~size_t(255);
    const size_t total = foff + fb + 256;
    if (l.desc.ensure(total, &s.bytes)) return RF_E_DEVICE;
    if (total > l.h_desc_cap) {
        if (l.h_desc) (void)hipHostFree(l.h_desc);
        l.h_desc = nullptr;
        if (hipHostMalloc((void**)&l.h_desc, total * 2, 0)) return RF_E_DEVICE;
        l.h_desc_cap = total * 2;
    }
    if ((uint32_t)fds.size() > l.stats_cap) {
        if (l.h_stats) (void)hipHostFree(l.h_stats);
        l.h_stats = nullptr;
        if (hipHostMalloc((void**)&l.h_stats, fds.size() * 2 * sizeof(FrameStats), 0)) return RF_E_DEVICE;
        l.stats_cap = (uint32_t)fds.size() * 2;
    }
    if (db) std::memcpy(l.h_desc, dds.data(), db);
    std::memcpy(l.h_desc + foff, fds.data(), fb);
    if (hipMemcpyAsync(l.desc.p, l.h_desc, total, hipMemcpyHostToDevice, l.stream)) return RF_E_DEVICE;
    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)l.desc.p;
    B.frames = (const FrameDesc*)((uint8_t*)l.desc.p + foff);
    B.n_draws = (uint32_t)dds.size();
    B.n_frames = (uint32_t)fds.size();
    B.n_verts = vbase;
    B.n_tris = tbase;
    B.n_groups = (uint32_t)((tbase + WAVE_SIZE - 1) / WAVE_SIZE);
    B.any_cull = any_cull;
    B.n_tiles = tile_base;
    B.tile_w = tile_w;
    B.tile_h = TILE_ROWS;
    B.chunk = RF_SERVE_CHUNK;  // span chunk of served launches
    B.sort_tpw = 1;            // few frames a launch: one tile per binsort wave
    B.extra_cap = l.extra_cap;
    B.runs_cap = std::max<uint32_t>(l.runs_cap, 8 * B.n_groups + B.n_tiles + 4096);
    if (l.sc.prepare(B, &s.bytes)) return RF_E_DEVICE;
    l.B = B;
    l.start_s = now_s();
    if (launch_pipeline(B, l.stream, raster_grid(d, B.n_tiles))) return RF_E_DEVICE;
    if (hipMemcpyAsync(l.h_stats, l.sc.stats.p, fds.size() * sizeof(FrameStats), hipMemcpyDeviceToHost,
                       l.stream) ||
        hipMemcpyAsync(l.h_misc, l.sc.misc.p, 12, hipMemcpyDeviceToHost, l.stream))
        return RF_E_DEVICE;
    if (hipMemcpyAsync(l.h_misc + 4, (uint32_t*)l.sc.run_off.p + B.n_tiles, 4, hipMemcpyDeviceToHost, l.stream))
        return RF_E_DEVICE;
    for (size_t i = 0; i < l.reqs.size(); ++i)
        if (hipMemcpyAsync(l.pix[i], l.fbs[i]->color.p, l.pix_bytes[i], hipMemcpyDeviceToHost, l.stream))
            return RF_E_DEVICE;
    l.h_misc[8] = 0;
    if (hipEventRecord(l.ev, l.stream)) return RF_E_DEVICE;
    return RF_OK;
}

// The worker never holds `m` across a HIP call: a stream's host function may
// run on the calling thread when the stream is already idle, and it takes `m`.
static void worker_main(Serve* sp) {
    Serve& s = *sp;
    (void)hipSetDevice(s.dev->ordinal);
    std::unique_lock<std::mutex> lk(s.m);
    for (;;) {
        bool progressed = false;
        std::vector<Lane*> relaunch, launch;
        // 1. retire finished lanes
        for (Lane* l : s.lanes) {
            if (!l->busy || !l->finished.load()) continue;
            progressed = true;
            if (l->h_misc[8]) {
                finish_lane(s, *l, RF_E_DEVICE);
                continue;
            }
            const uint32_t err = l->h_misc[2];
            if ((err & (ERRB_EXTRA | ERRB_RUNS)) && l->attempts < 3) {
                // grow the lane's capacities to what the batch needed and rerun it
                if (err & ERRB_EXTRA) l->extra_cap = l->h_misc[0] + l->h_misc[0] / 4 + 1024;
                if (err & ERRB_RUNS) l->runs_cap = l->h_misc[4] + l->h_misc[4] / 4 + 1024;
                ++l->attempts;
                l->finished.store(0);
                relaunch.push_back(l);
                continue;
            }
            finish_lane(s, *l, map_err(err));
        }
        // 2. admit pending requests onto free lanes
        for (Lane* l : s.lanes) {
            if (l->busy || s.pending.empty()) continue;
            while (!s.pending.empty() && l->reqs.size() < s.max_batch) {
                Req& q = s.pending.front();
                if (q.bytes > s.budget) {
                    // can never fit: fail it
                    Done d;
                    std::memset(&d.res, 0, sizeof(d.res));
                    d.res.id = q.r.id;
                    d.res.status = RF_E_BUDGET;
                    d.res.submit_s = q.submit_s;
                    d.res.start_s = d.res.done_s = now_s();
                    s.done.push_back(d);
                    s.pending.pop_front();
                    s.cv_done.notify_all();
                    continue;
                }
                if (s.used + q.bytes > s.budget) break;  // wait for memory to free up
                s.used += q.bytes;
                l->reqs.push_back(std::move(q));
                s.pending.pop_front();
            }
            if (l->reqs.empty()) continue;
            l->busy = true;
            l->attempts = 0;
            l->finished.store(0);
            progressed = true;
            launch.push_back(l);
        }
        // 3. allocate targets and enqueue, without the lock
        if (!launch.empty() || !relaunch.empty()) {
            lk.unlock();
            std::vector<std::pair<Lane*, rf_status>> failed;
            for (Lane* l : launch) {
                rf_status st = RF_OK;
                for (Req& q : l->reqs) {
                    const size_t pb = (size_t)q.r.width * q.r.height * bytes_per_pixel(q.r.fmt);
                    Framebuf* fb = fb_get(s, q.r.width, q.r.height, q.r.fmt, q.r.with_depth);
                    uint8_t* pin = pin_get(s, pb);
                    l->fbs.push_back(fb);
                    l->pix.push_back(pin);
                    l->pix_bytes.push_back(pb);
                    if (!fb || !pin) st = RF_E_DEVICE;
                }
                if (st == RF_OK) st = launch_lane(s, *l);
                if (st != RF_OK) failed.emplace_back(l, st);
            }
            for (Lane* l : relaunch) {
                const rf_status st = launch_lane(s, *l);
                if (st != RF_OK) failed.emplace_back(l, st);
            }
            lk.lock();
            for (auto& f : failed) {
                f.first->done_s = now_s();
                finish_lane(s, *f.first, f.second);
            }
            auto is_failed = [&](Lane* l) {
                for (auto& f : failed)
                    if (f.first == l) return true;
                return false;
            };
            for (Lane* l : launch)
                if (!is_failed(l)) {
                    l->launched = true;
                    l->cv_launch.notify_all();
                }
            for (Lane* l : relaunch)
                if (!is_failed(l)) {
                    l->launched = true;
                    l->cv_launch.notify_all();
                }
        }
        if (s.stop) {
            bool any = !s.pending.empty();
            for (Lane* l : s.lanes) any = any || l->busy;
            if (!any) break;
        }
        if (!progressed) s.cv_work.wait(lk);
    }
}

}  // namespace rf

using namespace rf;

extern "C" {

rf_status rf_serve_open(rf_device* dev, uint64_t budget_bytes, uint32_t max_batch, uint32_t n_streams,
                        rf_serve** out) {
    if (!dev || !out || budget_bytes == 0 || max_batch == 0 || n_streams == 0 || n_streams > 8) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    Serve* s = new (std::nothrow) Serve();
    if (!s) return RF_E_DEVICE;
    s->dev = &d;
    s->budget = budget_bytes;
    s->max_batch = max_batch;
    for (uint32_t i = 0; i < n_streams; ++i) {
        Lane* l = new Lane();
        if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) ||
            hipEventCreateWithFlags(&l->ev, hipEventBlockingSync | hipEventDisableTiming) ||
            hipHostMalloc((void**)&l->h_misc, 64, 0)) {
            if (l->ev) (void)hipEventDestroy(l->ev);
            if (l->stream) (void)hipStreamDestroy(l->stream);
            delete l;
            rf_serve_close(reinterpret_cast<rf_serve*>(s));
            return RF_E_DEVICE;
        }
        s->lanes.push_back(l);
    }
    for (Lane* l : s->lanes) l->waiter = std::thread(waiter_main, s, l);
    s->worker = std::thread(worker_main, s);
    *out = reinterpret_cast<rf_serve*>(s);
    return RF_OK;
}

rf_status rf_serve_submit(rf_serve* sh, const rf_request* req) {
    if (!sh || !req || req->width == 0 || req->height == 0 || req->fmt > PIX_LAST || (req->n_draws && !req->draws) ||
        req->width > MAX_DIM || req->height > MAX_DIM)
        return RF_E_ARG;
    Serve& s = *reinterpret_cast<Serve*>(sh);
    rf_status st = check_ctx(&req->ctx);
    if (!st && req->ctx.depth_sort) st = RF_E_UNSUPPORTED;  // depth sort: rf_render only
    if (st) return st;
    Req q;
    q.r = *req;
    q.draws.assign(req->draws, req->draws + req->n_draws);
    q.r.draws = nullptr;
    for (const rf_mesh_draw& md : q.draws) {
        if (!md.mesh || md.kind > KIND_LAST || md.cull > 1) return RF_E_ARG;
        const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
        const Tex* t = reinterpret_cast<const Tex*>(md.tex);
        if ((int)m->n_attr < kind_attrs(md.kind)) return RF_E_ARG;
        if (md.first_tri > m->n_tris || md.first_tri + md.n_tris > m->n_tris) return RF_E_ARG;
        if ((st = check_tex(md.kind, t ? t->texels() : nullptr, t ? t->w : 0, t ? t->h : 0, t ? t->stride : 0)))
            return st;
        q.n_tris += md.n_tris ? md.n_tris : m->n_tris - md.first_tri;
        q.n_verts += m->n_verts;
    }
    q.submit_s = now_s();
    q.bytes = request_bytes(q);
    {
        std::lock_guard<std::mutex> g(s.m);
        if (s.stop) return RF_E_ARG;
        s.pending.push_back(std::move(q));
    }
    s.cv_work.notify_all();
    return RF_OK;
}

rf_status rf_serve_poll(rf_serve* sh, rf_result* out, void** color, uint32_t max, uint32_t* n_out,
                        uint32_t timeout_ms) {
    if (!sh || !out || !n_out) return RF_E_ARG;
    Serve& s = *reinterpret_cast<Serve*>(sh);
    std::unique_lock<std::mutex> lk(s.m);
    if (s.done.empty() && timeout_ms)
        s.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s.done.empty(); });
    uint32_t n = 0;
    while (n < max && !s.done.empty()) {
        Done d = s.done.front();
        s.done.pop_front();
        out[n] = d.res;
        if (d.pixels) {
            if (color && color[n]) std::memcpy(color[n], d.pixels, d.pix_bytes);
            pin_put(s, d.pixels, d.pix_bytes);
        }
        ++n;
    }
    *n_out = n;
    return RF_OK;
}

rf_status rf_serve_close(rf_serve* sh) {
    if (!sh) return RF_E_ARG;
    Serve* s = reinterpret_cast<Serve*>(sh);
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop_waiters = true;
    }
    for (Lane* l : s->lanes) {
        l->cv_launch.notify_all();
        if (l->waiter.joinable()) l->waiter.join();
    }
    Device& d = *s->dev;
    (void)hipSetDevice(d.ordinal);
    for (Lane* l : s->lanes) {
        if (l->stream) (void)hipStreamSynchronize(l->stream);
        l->sc.release(&s->bytes);
        l->desc.release(&s->bytes);
        if (l->h_desc) (void)hipHostFree(l->h_desc);
        if (l->h_stats) (void)hipHostFree(l->h_stats);
        if (l->h_misc) (void)hipHostFree(l->h_misc);
        if (l->ev) (void)hipEventDestroy(l->ev);
        if (l->stream) (void)hipStreamDestroy(l->stream);
        delete l;
    }
    for (auto& kv : s->fb_pool)
        for (Framebuf* f : kv.second) {
            f->color.release(nullptr);
            f->depth.release(nullptr);
            delete f;
        }
    for (auto& d2 : s->done)
        if (d2.pixels) (void)hipHostFree(d2.pixels);
    for (auto& kv : s->pin_pool)
        for (uint8_t* p : kv.second) (void)hipHostFree(p);
    delete s;
    return RF_OK;
}

}  // extern "C"
