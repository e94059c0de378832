// Serving: a request loop over resident meshes and textures.
//
// Requests (one frame each: a resolution, a context and an ordered list of
// draws over resident meshes) are admitted oldest first while their device
// memory fits the budget; what is admitted when a stream is free goes out as
// one pipeline launch (continuous batching). A launch groups requests of
// similar cost only: every frame of a launch completes when the launch does,
// so a small frame batched with large ones would wait for them. Up to `n_streams`
// launches are on the device at once, so frames that arrive while others are
// rendering overlap them on the GPU. Each stream has a waiter thread that
// blocks on an event recorded behind its launch; results carry
// submit/start/done times.
//
// Memory: the budget is one device allocation made when the loop opens (the
// arena). A launch takes one block of it for everything it touches on the
// device -- its draw and frame descriptors, the pipeline scratch and the
// targets of its frames -- and gives it back once its results are read back.
// So the budget bounds the loop's device memory, and the steady-state loop
// makes no device allocation (no hipMalloc or hipFree, which would
// synchronise the device under the other streams' work). Read-back buffers
// come from one pinned host block allocated at open as well (a quarter of the
// budget, 64 MiB to 1 GiB), reserved when a request is admitted and returned
// when its result is polled; only a frame larger than that whole block gets
// a pinned buffer of its own. The descriptor staging of each stream is sized
// at open for max_batch frames of up to 8 draws each. The scratch of a
// launch is sized from its counts and from clip-fan and bin-run ratios
// learned from earlier launches; a launch that needs more is rerun with a
// larger block. A depth-sorted request (render.rs:180-183) is launched alone,
// through its stream's sort pass and grow-only scratch (launch_lane_sorted).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

#ifndef RF_SERVE_CHUNK
// 6-pixel span chunks: served launches mix meshes of small triangles, where
// they shorten the latency, with larger ones, where they cost nothing
// measurable
#define RF_SERVE_CHUNK 6
#endif
#ifndef RF_SERVE_TILE_W
#define RF_SERVE_TILE_W 32
#endif

namespace rf {

namespace {

constexpr uint32_t SERVE_TILE_W = RF_SERVE_TILE_W;  // rasterizer tile width of served launches
// Estimated GPU cost of a request in pixel units (about 10 ps each): a
// triangle's vertex, setup, binning and edge work costs about as much as 28
// pixels' fill and write-back (crates1080 and icospheres1080 batches), and a
// launch has a fixed part (its kernel sequence) worth about 0.1 ms.
constexpr double COST_TRI = 28.0;
constexpr double COST_LAUNCH = 1.0e7;
// A launch's estimated cost stays within this factor of its cheapest
// request's cost alone (the slowdown batching may cause that request).
#ifndef RF_SERVE_SLOWDOWN
#define RF_SERVE_SLOWDOWN 3.0
#endif
constexpr double SERVE_SLOWDOWN = RF_SERVE_SLOWDOWN;
constexpr size_t SERVE_SCAN = 256;  // pending requests looked at per launch
constexpr uint64_t ALIGN = 256;
inline uint64_t align_up(uint64_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// First-fit allocator over a byte range; free ranges coalesce. Not locked:
// the device arena is used by the worker thread only, the pinned read-back
// block under Serve::pm.
struct Arena {
    uint8_t* base = nullptr;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> bytes
    void init(uint8_t* b, uint64_t n) {
        base = b;
        size = n;
        used = 0;
        free_.clear();
        if (n) free_[0] = n;
    }
    uint64_t largest() const {
        uint64_t m = 0;
        for (const auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    bool alloc(uint64_t bytes, uint64_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    rf_request r;
    std::vector<rf_mesh_draw> draws;
    double submit_s = 0;
    uint64_t n_tris = 0, n_verts = 0;
    bool sorted = false;  // depth_sort with a triangle draw: launched alone, through the lane's sort pass
};

struct Done {
    rf_result res;
    uint8_t* pixels = nullptr;  // pinned, width * height * bpp
    size_t pix_bytes = 0;
};
constexpr uint32_t STAGE_DRAWS = 8;  // draws per frame the staging is sized for at open

// Counts of a launch: what its block must hold.
struct Plan {
    uint64_t n_verts = 0, n_tris = 0, n_draws = 0, n_frames = 0, n_tiles = 0, targets = 0;
    bool sorted = false;  // a depth-sorted launch: its scratch is the lane's own
    void add(const Req& q) {
        sorted = sorted || q.sorted;
        n_verts += q.n_verts;
        n_tris += q.n_tris;
        n_draws += q.draws.size();
        n_frames += 1;
        n_tiles += (uint64_t)((q.r.width + SERVE_TILE_W - 1) / SERVE_TILE_W) *
                   ((q.r.height + DEFAULT_TILE_H - 1) / DEFAULT_TILE_H);
        const uint64_t px = (uint64_t)q.r.width * q.r.height;
        targets += align_up(px * bytes_per_pixel(q.r.fmt)) + (q.r.with_depth ? align_up(px * 4) : 0);
    }
};

struct Lane {
    hipStream_t stream = nullptr;
    uint8_t* h_desc = nullptr;  // pinned staging for descriptors
    size_t h_desc_cap = 0;
    FrameStats* h_stats = nullptr;
    uint32_t* h_misc = nullptr;
    uint32_t stats_cap = 0;
    bool busy = false;
    bool wait_mem = false;   // rerun needs a larger block than is free now
    // the launch in flight
    std::vector<Req> reqs;
    std::vector<uint8_t*> pix;
    std::vector<size_t> pix_bytes;
    uint64_t blk_off = 0, blk_bytes = 0;  // arena block (blk_bytes 0: none)
    uint32_t extra_cap = 0, runs_cap = 0;
    BatchDesc B{};
    // depth-sorted launches (render.rs:180-183): the sort pass and the
    // pipeline scratch, grow-only device buffers of this lane
    SortPass sp;
    Scratch sort_sc;
    double start_s = 0;
    std::atomic<int> finished{0};
    double done_s = 0;
    int attempts = 0;
    // completion: the worker records `ev` behind the launch and sets `launched`;
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
    void* arena_mem = nullptr;
    Arena arena;             // worker thread only
    uint32_t max_batch = 16;
    // learned capacity ratios (worker thread only): clip fans beyond the
    // first per triangle, bin runs per 64-triangle group
    double extra_ratio = 1.0 / 16, runs_ratio = 8.0;
    std::vector<Lane*> lanes;
    std::thread worker;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req> pending;
    std::deque<Done> done;
    bool stop = false;
    bool stop_waiters = false;
    std::mutex pm;  // guards `pin`; never held across a HIP call (taken after `m`)
    void* pin_mem = nullptr;
    Arena pin;      // read-back buffers in the pinned block
    uint64_t sort_bytes = 0;  // device memory of the lanes' depth-sort buffers (worker thread only)
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

// Read-back buffer of `bytes` from the pinned block (nullptr if it does not
// fit now). Called with `m` held.
static uint8_t* pin_take(Serve& s, size_t bytes) {
    std::lock_guard<std::mutex> g(s.pm);
    uint64_t off = 0;
    if (!s.pin.alloc(align_up(bytes), off)) return nullptr;
    return s.pin.base + off;
}
static bool pin_in_block(const Serve& s, const uint8_t* p) {
    return p >= s.pin.base && p < s.pin.base + s.pin.size;
}
// Returns a read-back buffer: to the pinned block, or (a frame larger than
// the block) to the host.
static void pin_put(Serve& s, uint8_t* p, size_t bytes) {
    if (!p) return;
    if (pin_in_block(s, p)) {
        std::lock_guard<std::mutex> g(s.pm);
        s.pin.release((uint64_t)(p - s.pin.base), align_up(bytes));
    } else {
        (void)hipHostFree(p);
    }
}

static uint32_t est_extra(const Serve& s, const Plan& p) {
    const double e = std::ceil((double)p.n_tris * s.extra_ratio) + 1024.0;
    return (uint32_t)std::min<double>(e, (double)MAX_LAUNCH_PRIMS);
}
static uint32_t est_runs(const Serve& s, const Plan& p) {
    const double groups = std::ceil((double)p.n_tris / WAVE_SIZE);
    const double r = std::ceil(groups * s.runs_ratio) + (double)p.n_tiles + 4096.0;
    return (uint32_t)std::min<double>(r, 4.0e9);
}

// Bytes of the block of a launch with counts `p` and capacities (extra, runs):
// descriptors, scratch and targets.
static uint64_t block_bytes(const Plan& p, uint32_t extra, uint32_t runs) {
    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.n_verts = p.n_verts;
    B.n_tris = p.n_tris;
    B.n_draws = (uint32_t)p.n_draws;
    B.n_frames = (uint32_t)p.n_frames;
    B.n_tiles = (uint32_t)std::min<uint64_t>(p.n_tiles, 0xFFFFFFFFull);
    B.extra_cap = extra;
    B.runs_cap = runs;
    const uint64_t desc = align_up(p.n_draws * sizeof(DrawDesc)) + align_up(p.n_frames * sizeof(FrameDesc));
    return desc + (p.sorted ? 0 : scratch_bytes(B)) + p.targets;
}

static Plan plan_of(const std::vector<Req>& reqs) {
    Plan p;
    for (const Req& q : reqs) p.add(q);
    return p;
}

static void finish_lane(Serve& s, Lane& l, rf_status st) {
    // called with s.m held, on the worker thread
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
            d.res.stats.prims_o = h.prims_o;
            d.res.stats.verts_o = h.verts_o;
            d.res.stats.frags_i = h.frags_i;
            d.res.stats.frags_o = h.frags_o;
        }
        d.res.stats.seconds = l.done_s - l.start_s;
        if (i < l.pix.size()) {
            d.pixels = l.pix[i];
            d.pix_bytes = l.pix_bytes[i];
        }
        s.done.push_back(d);
    }
    if (l.blk_bytes) s.arena.release(l.blk_off, l.blk_bytes);
    l.blk_bytes = 0;
    l.reqs.clear();
    l.pix.clear();
    l.pix_bytes.clear();
    l.busy = false;
    l.wait_mem = false;
    s.cv_done.notify_all();
}

// Fails one request that can never run.
static void reject(Serve& s, const Req& q, rf_status st) {
    Done d;
    std::memset(&d.res, 0, sizeof(d.res));
    d.res.id = q.r.id;
    d.res.status = st;
    d.res.width = q.r.width;
    d.res.height = q.r.height;
    d.res.fmt = q.r.fmt;
    d.res.submit_s = q.submit_s;
    d.res.start_s = d.res.done_s = now_s();
    s.done.push_back(d);
    s.cv_done.notify_all();
}

// Sizes the lane's block for its requests at the current capacities and
// takes it from the arena. Returns false (nothing taken) if no free range is
// large enough now.
static bool take_block(Serve& s, Lane& l) {
    const Plan p = plan_of(l.reqs);
    const uint64_t need = block_bytes(p, l.extra_cap, l.runs_cap);
    uint64_t off = 0;
    if (!s.arena.alloc(need, off)) return false;
    l.blk_off = off;
    l.blk_bytes = need;
    return true;
}

// A served frame's descriptor, its targets taken from `tgt` onwards (which
// advances past them).
static FrameDesc frame_desc(const Req& q, uint8_t*& tgt, uint32_t& tile_base) {
    const uint64_t px = (uint64_t)q.r.width * q.r.height;
    FrameDesc f;
    std::memset(&f, 0, sizeof(f));
    f.color = tgt;
    tgt += align_up(px * bytes_per_pixel(q.r.fmt));
    if (q.r.with_depth) {
        f.depth = (float*)tgt;
        tgt += align_up(px * 4);
    }
    f.color_stride = (uint64_t)q.r.width * bytes_per_pixel(q.r.fmt);
    f.depth_stride = q.r.width;
    f.width = q.r.width;
    f.height = q.r.height;
    f.fmt = q.r.fmt;
    rf_ctx c = q.r.ctx;
    c.clear_color = 1;  // a served frame starts from Frame::clear
    c.clear_depth = 1;
    fill_frame_ctx(f, c);
    set_grid(f, SERVE_TILE_W, DEFAULT_TILE_H, tile_base);
    return f;
}

// A draw's descriptor (vert_base and tri_base left at 0).
static DrawDesc draw_desc(const rf_mesh_draw& md, uint32_t fi) {
    const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
    const Tex* t = reinterpret_cast<const Tex*>(md.tex);
    DrawDesc dd;
    std::memset(&dd, 0, sizeof(dd));
    dd.pos = (const float*)m->pos.p;
    dd.attr = (const float*)m->attr.p;
    dd.tris = (const uint32_t*)m->tris.p + (m->prim == PRIM_EDGE ? 2 : 3) * md.first_tri;
    dd.prim = m->prim;
    dd.tex = t ? t->texels() : nullptr;
    dd.n_verts = (uint32_t)m->n_verts;
    dd.n_tris = (uint32_t)(md.n_tris ? md.n_tris : m->n_tris - md.first_tri);
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

// Grows the lane's pinned descriptor staging and stats to hold `desc_bytes`
// and `frames` (host memory; sized at open, grown only past that).
static rf_status stage_room(Lane& l, size_t desc_bytes, uint32_t frames) {
    if (desc_bytes > l.h_desc_cap) {
        if (l.h_desc) (void)hipHostFree(l.h_desc);
        l.h_desc = nullptr;
        l.h_desc_cap = 0;
        if (hipHostMalloc((void**)&l.h_desc, desc_bytes * 2, 0)) return RF_E_DEVICE;
        l.h_desc_cap = desc_bytes * 2;
    }
    if (frames > l.stats_cap) {
        if (l.h_stats) (void)hipHostFree(l.h_stats);
        l.h_stats = nullptr;
        l.stats_cap = 0;
        if (hipHostMalloc((void**)&l.h_stats, (size_t)frames * 2 * sizeof(FrameStats), 0)) return RF_E_DEVICE;
        l.stats_cap = frames * 2;
    }
    return RF_OK;
}

// A depth-sorted request (alone in its launch; render.rs:180-183, 209-219):
// its triangle draws go through the lane's sort pass, like rf_render's, and
// the pipeline runs on the lane's own grow-only scratch, synchronously (the
// sort pass learns the fan counts first); the block holds the descriptors
// and the targets. Then the read-back and the completion event, as for any
// launch.
static rf_status launch_lane_sorted(Serve& s, Lane& l) {
    Device& d = *s.dev;
    const Req& q = l.reqs[0];
    uint8_t* blk = (uint8_t*)s.arena.base + l.blk_off;
    const uint64_t db = align_up(q.draws.size() * sizeof(DrawDesc)), fb = align_up(sizeof(FrameDesc));
    uint8_t* tgt = blk + db + fb;
    uint32_t tile_base = 0;
    const FrameDesc f = frame_desc(q, tgt, tile_base);
    if (rf_status st = stage_room(l, db + fb, 1)) return st;
    std::memcpy(l.h_desc + db, &f, sizeof(f));
    l.start_s = now_s();
    if (hipMemcpyAsync(blk + db, l.h_desc + db, sizeof(f), hipMemcpyHostToDevice, l.stream)) return RF_E_DEVICE;
    // the sort pass over the request's triangle draws
    SortPass& S = l.sp;
    S.draws.clear();
    std::vector<int32_t> sort_of;
    for (const rf_mesh_draw& md : q.draws) {
        const Mesh* m = reinterpret_cast<const Mesh*>(md.mesh);
        if (m->prim == PRIM_TRI) {
            sort_of.push_back((int32_t)S.draws.size());
            DrawDesc dd = draw_desc(md, 0);
            dd.sort = (uint32_t)q.r.ctx.depth_sort;
            S.draws.push_back(dd);
        } else {
            sort_of.push_back(-1);
        }
    }
    S.frames = (const FrameDesc*)(blk + db);
    S.frames_n = 1;
    S.main_of.assign(S.draws.size(), 0);
    uint32_t e = 0;
    rf_status st = sortpass_plan(d, S, l.stream, &s.sort_bytes, &e);
    if (!st) st = map_err(e);
    if (st) return st;
    // the main draws: the sorted fans of each triangle draw as its vertex
    // slots and index triples
    std::vector<DrawDesc> dds;
    uint64_t vbase = 0, tbase = 0;
    uint32_t any_cull = 0;
    for (size_t k = 0; k < q.draws.size(); ++k) {
        DrawDesc dd = draw_desc(q.draws[k], 0);
        if (const int32_t si = sort_of[k]; si >= 0) {
            const uint32_t nf = S.cnt[si];
            dd.pre = 1;
            dd.sort = (uint32_t)q.r.ctx.depth_sort;
            dd.in_tris = dd.n_tris;
            dd.in_verts = dd.n_verts;
            dd.n_tris = nf;
            dd.n_verts = 3 * nf;
            dd.tris = sortpass_idx(S, S.off[si]);
            S.main_of[si] = (uint32_t)k;
        }
        dd.vert_base = vbase;
        dd.tri_base = tbase;
        any_cull |= dd.cull;
        vbase += dd.n_verts;
        tbase += dd.n_tris;
        dds.push_back(dd);
    }
    if (tbase >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    if (sortpass_map(S, l.stream)) return RF_E_DEVICE;
    std::memcpy(l.h_desc, dds.data(), dds.size() * sizeof(DrawDesc));
    if (hipMemcpyAsync(blk, l.h_desc, dds.size() * sizeof(DrawDesc), hipMemcpyHostToDevice, l.stream))
        return RF_E_DEVICE;
    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)blk;
    B.frames = (const FrameDesc*)(blk + db);
    B.n_draws = (uint32_t)dds.size();
    B.n_frames = 1;
    B.n_verts = vbase;
    B.n_tris = tbase;
    B.n_groups = (uint32_t)((tbase + WAVE_SIZE - 1) / WAVE_SIZE);
    B.n_tiles = tile_base;
    B.tile_w = SERVE_TILE_W;
    B.tile_h = DEFAULT_TILE_H;
    B.chunk = RF_SERVE_CHUNK;
    B.sort_tpw = 1;
    B.any_cull = any_cull;
    B.extra_cap = l.extra_cap;
    B.runs_cap = std::max<uint32_t>(4 * B.n_groups + B.n_tiles + 1024, l.runs_cap);
    std::vector<FrameStats> fs;
    uint32_t err = 0;
    st = run_sized(d, l.sort_sc, B, l.stream,
                   [&S, &l](const BatchDesc& M) -> hipError_t {
                       hipError_t r = sortpass_fans(S, l.stream);
                       return r ? r : sortpass_sort_gather(S, M, l.stream);
                   },
                   fs, &err, nullptr, &s.sort_bytes);
    if (!st) st = map_err(err);
    if (st) return st;
    l.B = B;
    l.h_stats[0] = fs[0];
    for (int k = 0; k < 8; ++k) l.h_misc[k] = 0;  // errors and capacities: settled by run_sized
    if (hipMemcpyAsync(l.pix[0], f.color, l.pix_bytes[0], hipMemcpyDeviceToHost, l.stream)) return RF_E_DEVICE;
    l.h_misc[8] = 0;
    if (hipEventRecord(l.ev, l.stream)) return RF_E_DEVICE;
    return RF_OK;
}

// Lays the lane's block out (descriptors, scratch, targets), uploads the
// descriptors and enqueues the pipeline, the readbacks and the completion
// event. Called by the worker without s.m.
static rf_status launch_lane(Serve& s, Lane& l) {
    if (l.reqs.size() == 1 && l.reqs[0].sorted) return launch_lane_sorted(s, l);
    Device& d = *s.dev;
    uint8_t* blk = (uint8_t*)s.arena.base + l.blk_off;
    const Plan p = plan_of(l.reqs);
    const uint64_t db = align_up(p.n_draws * sizeof(DrawDesc)), fb = align_up(p.n_frames * sizeof(FrameDesc));
    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)blk;
    B.frames = (const FrameDesc*)(blk + db);
    B.n_draws = (uint32_t)p.n_draws;
    B.n_frames = (uint32_t)p.n_frames;
    B.n_verts = p.n_verts;
    B.n_tris = p.n_tris;
    B.n_groups = (uint32_t)((p.n_tris + WAVE_SIZE - 1) / WAVE_SIZE);
    B.n_tiles = (uint32_t)p.n_tiles;
    B.tile_w = SERVE_TILE_W;
    B.tile_h = DEFAULT_TILE_H;
    B.chunk = RF_SERVE_CHUNK;  // span chunk of served launches
    B.sort_tpw = 1;            // few frames a launch: one tile per binsort wave
    B.extra_cap = l.extra_cap;
    B.runs_cap = l.runs_cap;
    const uint64_t sb = scratch_bytes(B);
    scratch_carve(B, blk + db + fb);
    uint8_t* tgt = blk + db + fb + sb;

    std::vector<DrawDesc> dds;
    std::vector<FrameDesc> fds;
    dds.reserve(p.n_draws);
    fds.reserve(p.n_frames);
    uint64_t vbase = 0, tbase = 0;
    uint32_t tile_base = 0, any_cull = 0;
    for (size_t i = 0; i < l.reqs.size(); ++i) {
        Req& q = l.reqs[i];
        fds.push_back(frame_desc(q, tgt, tile_base));
        for (const rf_mesh_draw& md : q.draws) {
            DrawDesc dd = draw_desc(md, (uint32_t)i);
            dd.vert_base = vbase;
            dd.tri_base = tbase;
            any_cull |= md.cull;
            vbase += dd.n_verts;
            tbase += dd.n_tris;
            dds.push_back(dd);
        }
    }
    B.any_cull = any_cull;
    const size_t total = db + fb;
    if (rf_status st = stage_room(l, total, (uint32_t)fds.size())) return st;
    std::memcpy(l.h_desc, dds.data(), dds.size() * sizeof(DrawDesc));
    std::memcpy(l.h_desc + db, fds.data(), fds.size() * sizeof(FrameDesc));
    l.B = B;
    l.start_s = now_s();
    // Nothing is queued before this point; from here on a failure leaves
    // work on the stream, which the caller drains before the block is reused.
    if (hipMemcpyAsync(blk, l.h_desc, total, hipMemcpyHostToDevice, l.stream)) return RF_E_DEVICE;
    if (launch_pipeline(B, l.stream, raster_grid(d, B))) return RF_E_DEVICE;
    if (hipMemcpyAsync(l.h_stats, B.stats, fds.size() * sizeof(FrameStats), hipMemcpyDeviceToHost, l.stream) ||
        hipMemcpyAsync(l.h_misc, misc_words(B), 12, hipMemcpyDeviceToHost, l.stream) ||
        hipMemcpyAsync(l.h_misc + 4, B.run_off + B.n_tiles, 4, hipMemcpyDeviceToHost, l.stream))
        return RF_E_DEVICE;
    for (size_t i = 0; i < l.reqs.size(); ++i)
        if (hipMemcpyAsync(l.pix[i], fds[i].color, l.pix_bytes[i], hipMemcpyDeviceToHost, l.stream))
            return RF_E_DEVICE;
    l.h_misc[8] = 0;
    if (hipEventRecord(l.ev, l.stream)) return RF_E_DEVICE;
    return RF_OK;
}

// The worker never holds `m` across a HIP call that may block on the device.
static void worker_main(Serve* sp) {
    Serve& s = *sp;
    (void)hipSetDevice(s.dev->ordinal);
    std::unique_lock<std::mutex> lk(s.m);
    for (;;) {
        bool progressed = false;
        std::vector<Lane*> launch;
        // 1. retire finished lanes; grow and rerun a launch that ran out of
        //    clip-fan or bin-run capacity
        for (Lane* l : s.lanes) {
            if (!l->busy || l->wait_mem || !l->finished.load()) continue;
            progressed = true;
            l->finished.store(0);
            if (l->h_misc[8]) {
                finish_lane(s, *l, RF_E_DEVICE);
                continue;
            }
            const uint32_t err = l->h_misc[2];
            if ((err & (ERRB_EXTRA | ERRB_RUNS)) && l->attempts < 3) {
                const Plan p = plan_of(l->reqs);
                if (err & ERRB_EXTRA) {
                    const uint32_t need = l->h_misc[0];
                    s.extra_ratio = std::max(s.extra_ratio, 1.25 * need / std::max<double>(1.0, (double)p.n_tris));
                    l->extra_cap = std::max(l->extra_cap, grow_extra_cap(need));
                }
                if (err & ERRB_RUNS) {
                    const uint32_t need = l->h_misc[4];
                    const double groups = std::ceil((double)p.n_tris / WAVE_SIZE);
                    s.runs_ratio = std::max(s.runs_ratio, 1.25 * need / std::max(1.0, groups));
                    l->runs_cap = std::max(l->runs_cap, grow_runs_cap(need));
                }
                ++l->attempts;
                s.arena.release(l->blk_off, l->blk_bytes);
                l->blk_bytes = 0;
                if (block_bytes(p, l->extra_cap, l->runs_cap) > s.arena.size) {
                    l->done_s = now_s();
                    finish_lane(s, *l, RF_E_BUDGET);
                    continue;
                }
                l->wait_mem = true;  // takes its larger block below, or waits for one
                continue;
            }
            finish_lane(s, *l, map_err(err));
        }
        // 2. reruns waiting for a larger block
        for (Lane* l : s.lanes) {
            if (!l->wait_mem) continue;
            if (take_block(s, *l)) {
                l->wait_mem = false;
                progressed = true;
                launch.push_back(l);
            } else {
                bool others = false;
                for (Lane* o : s.lanes) others = others || (o != l && o->busy && !o->wait_mem);
                if (!others) {  // the arena is as free as it gets and still too fragmented or small
                    l->done_s = now_s();
                    finish_lane(s, *l, RF_E_BUDGET);
                    progressed = true;
                }
            }
        }
        // 3. admit pending requests onto free lanes, the oldest first, while
        //    the launch's block fits the largest free range of the arena (not
        //    while a rerun waits for memory: it goes first). Behind the oldest
        //    one, a request joins only if the launch's estimated cost stays
        //    within SERVE_SLOWDOWN times its cheapest member's cost alone;
        //    the others stay pending, in order, for the next launch.
        bool rerun_waits = false;
        for (Lane* l : s.lanes) rerun_waits = rerun_waits || l->wait_mem;
        for (Lane* l : s.lanes) {
            if (l->busy || s.pending.empty() || rerun_waits) continue;
            const uint64_t room = s.arena.largest();
            Plan p;
            double cost = COST_LAUNCH, cheapest = 0.0;
            std::vector<size_t> take;
            size_t i = 0;
            while (i < s.pending.size() && i < SERVE_SCAN && take.size() < s.max_batch) {
                Req& q = s.pending[i];
                Plan alone;
                alone.add(q);
                if (block_bytes(alone, est_extra(s, alone), est_runs(s, alone)) > s.arena.size) {
                    reject(s, q, RF_E_BUDGET);  // can never fit
                    s.pending.erase(s.pending.begin() + (ptrdiff_t)i);
                    progressed = true;
                    continue;
                }
                const double c = (double)q.r.width * q.r.height + COST_TRI * (double)q.n_tris;
                const double cheap = take.empty() ? c : std::min(cheapest, c);
                if (!take.empty() && (q.sorted || cost + c > SERVE_SLOWDOWN * (COST_LAUNCH + cheap))) {
                    ++i;  // another cost class, or depth-sorted: a later launch
                    continue;
                }
                Plan p2 = p;
                p2.add(q);
                if (p2.n_tris >= MAX_LAUNCH_PRIMS) break;
                if (block_bytes(p2, est_extra(s, p2), est_runs(s, p2)) > room) break;  // wait for memory
                // its read-back buffer: reserved now, or (larger than the
                // whole pinned block) allocated at launch
                const size_t pb = (size_t)q.r.width * q.r.height * bytes_per_pixel(q.r.fmt);
                uint8_t* pin = nullptr;
                if (align_up(pb) <= s.pin.size && !(pin = pin_take(s, pb))) break;  // wait for polls
                l->pix.push_back(pin);
                l->pix_bytes.push_back(pb);
                p = p2;
                cost += c;
                cheapest = cheap;
                take.push_back(i);
                ++i;
                if (q.sorted) break;  // a depth-sorted request goes alone
            }
            for (size_t k : take) l->reqs.push_back(std::move(s.pending[k]));
            for (auto it = take.rbegin(); it != take.rend(); ++it) s.pending.erase(s.pending.begin() + (ptrdiff_t)*it);
            if (l->reqs.empty()) continue;
            l->extra_cap = est_extra(s, p);
            l->runs_cap = est_runs(s, p);
            if (!take_block(s, *l)) {  // cannot happen: the block fits `room`
                for (auto it = l->reqs.rbegin(); it != l->reqs.rend(); ++it) s.pending.push_front(std::move(*it));
                l->reqs.clear();
                for (size_t k = 0; k < l->pix.size(); ++k) pin_put(s, l->pix[k], l->pix_bytes[k]);
                l->pix.clear();
                l->pix_bytes.clear();
                continue;
            }
            l->busy = true;
            l->attempts = 0;
            l->finished.store(0);
            progressed = true;
            launch.push_back(l);
        }
        // 4. readback buffers and enqueue, without the lock
        if (!launch.empty()) {
            lk.unlock();
            std::vector<std::pair<Lane*, rf_status>> failed;
            for (Lane* l : launch) {
                rf_status st = RF_OK;
                // frames larger than the pinned block: a buffer of their own
                for (size_t k = 0; k < l->pix.size(); ++k)
                    if (!l->pix[k] && hipHostMalloc((void**)&l->pix[k], l->pix_bytes[k], 0)) {
                        l->pix[k] = nullptr;
                        st = RF_E_DEVICE;
                    }
                if (st == RF_OK) st = launch_lane(s, *l);
                if (st != RF_OK) {
                    // work may be queued already: drain it before the block
                    // and the readback buffers are reused
                    (void)hipStreamSynchronize(l->stream);
                    failed.emplace_back(l, st);
                }
            }
            lk.lock();
            for (auto& f : failed) {
                f.first->done_s = now_s();
                finish_lane(s, *f.first, f.second);
            }
            for (Lane* l : launch) {
                bool bad = false;
                for (auto& f : failed) bad = bad || f.first == l;
                if (!bad) {
                    l->launched = true;
                    l->cv_launch.notify_all();
                }
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
    s->max_batch = max_batch;
    budget_bytes &= ~(ALIGN - 1);
    if (!budget_bytes || hipMalloc(&s->arena_mem, budget_bytes)) {
        delete s;
        return RF_E_DEVICE;
    }
    s->arena.init((uint8_t*)s->arena_mem, budget_bytes);
    const uint64_t pin_bytes = std::min<uint64_t>(1ull << 30, std::max<uint64_t>(64ull << 20, budget_bytes / 4));
    if (hipHostMalloc(&s->pin_mem, pin_bytes, 0)) {
        s->pin_mem = nullptr;
        rf_serve_close(reinterpret_cast<rf_serve*>(s));
        return RF_E_DEVICE;
    }
    s->pin.init((uint8_t*)s->pin_mem, pin_bytes);
    const size_t stage = 2 * (align_up((size_t)max_batch * STAGE_DRAWS * sizeof(DrawDesc)) +
                              align_up((size_t)max_batch * sizeof(FrameDesc)));
    for (uint32_t i = 0; i < n_streams; ++i) {
        Lane* l = new Lane();
        s->lanes.push_back(l);
        if (hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) ||
            hipEventCreateWithFlags(&l->ev, hipEventBlockingSync | hipEventDisableTiming) ||
            hipHostMalloc((void**)&l->h_misc, 64, 0) || hipHostMalloc((void**)&l->h_desc, stage, 0) ||
            hipHostMalloc((void**)&l->h_stats, (size_t)max_batch * 2 * sizeof(FrameStats), 0)) {
            rf_serve_close(reinterpret_cast<rf_serve*>(s));
            return RF_E_DEVICE;
        }
        l->h_desc_cap = stage;
        l->stats_cap = max_batch * 2;
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
        q.sorted = q.sorted || (req->ctx.depth_sort && m->prim == PRIM_TRI);
    }
    if (q.n_tris >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    q.submit_s = now_s();
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
    lk.unlock();
    if (n) s.cv_work.notify_all();  // admissions may wait for read-back space
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
        sortpass_release(l->sp, &s->sort_bytes);
        l->sort_sc.release(&s->sort_bytes);
        if (l->h_desc) (void)hipHostFree(l->h_desc);
        if (l->h_stats) (void)hipHostFree(l->h_stats);
        if (l->h_misc) (void)hipHostFree(l->h_misc);
        if (l->ev) (void)hipEventDestroy(l->ev);
        if (l->stream) (void)hipStreamDestroy(l->stream);
        delete l;
    }
    if (s->arena_mem) (void)hipFree(s->arena_mem);
    for (auto& d2 : s->done)
        if (d2.pixels && !pin_in_block(*s, d2.pixels)) (void)hipHostFree(d2.pixels);
    if (s->pin_mem) (void)hipHostFree(s->pin_mem);
    delete s;
    return RF_OK;
}

}  // extern "C"
