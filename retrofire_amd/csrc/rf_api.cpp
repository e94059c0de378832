// Host runtime behind include/retrofire_amd.h: device handles, the
// synchronous drop-in `rf_render`, resident meshes/textures/framebuffers, and
// batches captured into HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

namespace rf {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

hipError_t DevBuf::ensure(size_t bytes, uint64_t* counter) {
    if (bytes <= cap) return hipSuccess;
    size_t want = std::max(bytes, cap + cap / 2);  // grow by half at least
    if (p) {
        (void)hipFree(p);
        if (counter) *counter -= cap;
        p = nullptr;
        cap = 0;
    }
    want = (want + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return e;
    }
    cap = want;
    if (counter) *counter += cap;
    return hipSuccess;
}
void DevBuf::release(uint64_t* counter) {
    if (p) {
        (void)hipFree(p);
        if (counter) *counter -= cap;
    }
    p = nullptr;
    cap = 0;
}

uint32_t bytes_per_pixel(uint32_t fmt) { return pixfmt_bytes(fmt); }

uint32_t pack_clear(uint32_t fmt, const uint8_t* c) { return pack_pixel(fmt, c[0], c[1], c[2], c[3]); }

void fill_frame_ctx(FrameDesc& f, const rf_ctx& c) {
    f.depth_test = c.depth_test;
    f.face_cull = c.face_cull;
    f.color_write = c.color_write;
    f.depth_write = c.depth_write;
    f.clear_color = c.clear_color;
    f.clear_depth = c.clear_depth;
    f.clear_pixel = pack_clear(f.fmt, c.color_clear);
    f.clear_depth_value = 1.0f / c.depth_clear;  // depth_clear.recip() (front/src/lib.rs:111-119)
}

rf_status check_ctx(const rf_ctx* c) {
    if (!c) return RF_E_ARG;
    if (c->depth_sort < 0 || c->depth_sort > 2) return RF_E_ARG;
    if (c->depth_test < 0 || c->depth_test > 3 || c->face_cull < 0 || c->face_cull > 2) return RF_E_ARG;
    return RF_OK;
}

rf_status check_tex(uint32_t kind, const void* tex, uint32_t w, uint32_t h, uint32_t stride) {
    if (kind == KIND_FLAT || kind == KIND_GOURAUD) return RF_OK;
    if (kind > KIND_LAST) return RF_E_UNSUPPORTED;
    if (!tex || w == 0 || h == 0 || stride < w) return RF_E_ARG;
    if ((uint64_t)stride * h > 0xFFFFFFFFull) return RF_E_ARG;  // texel indices are 32-bit
    if (kind == KIND_TEX_REPEAT_POT && ((w & (w - 1)) || (h & (h - 1)))) return RF_E_ARG;
    return RF_OK;
}

rf_status map_err(uint32_t e) {
    if (e & ERRB_INDEX) return RF_E_INDEX;
    if (e & (ERRB_BOUNDS | ERRB_TEXEL)) return RF_E_BOUNDS;
    if (e & ERRB_NEGY) return RF_E_VIEWPORT;
    if (e & ERRB_NONFINITE) return RF_E_NONFINITE;
    if (e & (ERRB_MARGIN | ERRB_EXTRA | ERRB_RUNS | ERRB_CLIP)) return RF_E_INTERNAL;
    return RF_OK;
}

// Frame tile grid for a tile width.
void set_grid(FrameDesc& f, uint32_t tile_w, uint32_t& tile_base) {
    f.ntx = (f.width + tile_w - 1) / tile_w;
    f.nty = (f.height + TILE_ROWS - 1) / TILE_ROWS;
    f.tile_base = tile_base;
    tile_base += f.ntx * f.nty;
}

// Scratch sized for a batch; fills the scratch pointers of `B`.
hipError_t Scratch::prepare(BatchDesc& B, uint64_t* counter) {
    hipError_t e;
    const uint64_t nv = std::max<uint64_t>(B.n_verts, 1), nt = std::max<uint64_t>(B.n_tris, 1);
    if ((e = cv.ensure(nv * sizeof(float4), counter))) return e;
    if ((e = ca.ensure(nv * sizeof(float4), counter))) return e;
    if ((e = sv.ensure(nv * 6 * sizeof(float), counter))) return e;
    if ((e = hdr.ensure(nt * sizeof(uint32_t), counter))) return e;
    if ((e = scr.ensure(nt * sizeof(ScrTri), counter))) return e;
    if ((e = extra.ensure(std::max<uint64_t>(B.extra_cap, 1) * sizeof(ScrTri), counter))) return e;
    if ((e = rect.ensure(nt * 2 * sizeof(uint32_t), counter))) return e;
    if ((e = group_draw.ensure((nt / 64 + 1) * sizeof(uint32_t), counter))) return e;
    if ((e = vblock_draw.ensure((nv / 256 + 1) * sizeof(uint32_t), counter))) return e;
    const uint64_t ntl = std::max<uint64_t>(B.n_tiles, 1);
    if ((e = run_count.ensure(ntl * sizeof(uint32_t), counter))) return e;
    if ((e = run_off.ensure((ntl + 1) * sizeof(uint32_t), counter))) return e;
    if ((e = run_cursor.ensure(ntl * sizeof(uint32_t), counter))) return e;
    if ((e = scan_blk.ensure(2 * (ntl / 1024 + 2) * sizeof(uint32_t), counter))) return e;
    if ((e = tile_list.ensure(ntl * sizeof(uint32_t), counter))) return e;
    if ((e = tile_frame.ensure(ntl * sizeof(uint32_t), counter))) return e;
    if ((e = runs.ensure(std::max<uint64_t>(B.runs_cap, 1) * sizeof(Run), counter))) return e;
    if ((e = misc.ensure(256, counter))) return e;
    if ((e = stats.ensure(std::max<uint64_t>(B.n_frames, 1) * sizeof(FrameStats), counter))) return e;
    if ((e = draw_vis.ensure(std::max<uint64_t>(B.n_draws, 1) * sizeof(uint32_t), counter))) return e;
    if ((e = prims_part.ensure(std::max<uint64_t>(B.n_frames, 1) * PRIM_PARTS * PRIM_PART_STRIDE * 8, counter)))
        return e;
    B.cv = (float4*)cv.p;
    B.ca = (float4*)ca.p;
    B.sv = (float*)sv.p;
    B.hdr = (uint32_t*)hdr.p;
    B.scr = (ScrTri*)scr.p;
    B.extra = (ScrTri*)extra.p;
    B.rect = (uint32_t*)rect.p;
    B.group_draw = (uint32_t*)group_draw.p;
    B.vblock_draw = (uint32_t*)vblock_draw.p;
    B.run_count = (uint32_t*)run_count.p;
    B.run_off = (uint32_t*)run_off.p;
    B.run_cursor = (uint32_t*)run_cursor.p;
    B.scan_blk = (uint32_t*)scan_blk.p;
    B.tile_frame = (uint32_t*)tile_frame.p;
    B.tile_list = (uint32_t*)tile_list.p;
    B.busy_count = (uint32_t*)misc.p + 8;
    B.runs = (Run*)runs.p;
    B.extra_count = (uint32_t*)misc.p;
    B.tile_counter = (uint32_t*)misc.p + 16;  // TILE_PARTS + 1 words (busy partitions, empty tiles)
    B.err = (uint32_t*)misc.p + 2;
    B.stats = (FrameStats*)stats.p;
    B.draw_vis = (uint32_t*)draw_vis.p;
    B.prims_part = (unsigned long long*)prims_part.p;
    return hipSuccess;
}

void Scratch::release(uint64_t* counter) {
    for (DevBuf* b : {&cv, &ca, &sv, &hdr, &scr, &extra, &rect, &run_count, &run_off, &run_cursor,
                      &scan_blk, &tile_frame, &tile_list, &group_draw, &vblock_draw, &runs, &misc, &stats, &draw_vis, &prims_part})
        b->release(counter);
}

uint32_t raster_grid(const Device& d, uint32_t n_tiles) {
    const uint32_t g = (uint32_t)d.num_cu * RASTER_WG_PER_CU;
    return std::max(1u, std::min(n_tiles, g));
}

// Runs the pipeline on `s` and waits; grows the clip-fan and run capacities
// and runs again when they were exceeded. `restore` re-initialises anything
// the aborted run may have written (the targets) before a rerun.
rf_status run_sized(Device& d, Scratch& sc, BatchDesc& B, hipStream_t s, const std::vector<FrameStats*>&,
                    std::vector<FrameStats>& stats_out, uint32_t* err_out,
                    const std::function<hipError_t()>& restore) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (sc.prepare(B, &d.bytes) != hipSuccess) return RF_E_DEVICE;
        if (launch_pipeline(B, s, raster_grid(d, B.n_tiles)) != hipSuccess) return RF_E_DEVICE;
        uint32_t misc[3];
        uint32_t total_runs = 0;
        if (hipMemcpyAsync(misc, sc.misc.p, sizeof(misc), hipMemcpyDeviceToHost, s) != hipSuccess)
            return RF_E_DEVICE;
        if (hipMemcpyAsync(&total_runs, (uint32_t*)sc.run_off.p + B.n_tiles, 4, hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return RF_E_DEVICE;
        stats_out.resize(B.n_frames);
        if (B.n_frames &&
            hipMemcpyAsync(stats_out.data(), sc.stats.p, B.n_frames * sizeof(FrameStats),
                           hipMemcpyDeviceToHost, s) != hipSuccess)
            return RF_E_DEVICE;
        if (hipStreamSynchronize(s) != hipSuccess) return RF_E_DEVICE;
        const uint32_t err = misc[2];
        bool again = false;
        if (err & ERRB_EXTRA) {
            B.extra_cap = misc[0] + misc[0] / 4 + 1024;
            again = true;
        }
        if (err & ERRB_RUNS) {
            B.runs_cap = total_runs + total_runs / 4 + 1024;
            again = true;
        }
        if (!again) {
            *err_out = err;
            return RF_OK;
        }
        if (restore && restore() != hipSuccess) return RF_E_DEVICE;
    }
    return RF_E_INTERNAL;
}

}  // namespace rf

using namespace rf;

extern "C" {

const char* rf_status_str(rf_status s) {
    switch (s) {
        case RF_OK: return "ok";
        case RF_E_INDEX: return "vertex index out of range";
        case RF_E_BOUNDS: return "span outside the render target";
        case RF_E_ARG: return "invalid argument";
        case RF_E_DEVICE: return "HIP device error";
        case RF_E_UNSUPPORTED: return "unsupported shader, target or context (use the CPU path)";
        case RF_E_VIEWPORT: return "to_screen maps geometry above the target's first row";
        case RF_E_NONFINITE: return "non-finite screen-space vertex";
        case RF_E_INTERNAL: return "internal capacity check failed";
        case RF_E_BUDGET: return "request exceeds the serving memory budget";
    }
    return "unknown status";
}

double rf_now_s(void) { return now_s(); }

rf_status rf_device_open(int ordinal, rf_device** out) {
    if (!out) return RF_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || ordinal < 0 || ordinal >= n) return RF_E_DEVICE;
    if (hipSetDevice(ordinal) != hipSuccess) return RF_E_DEVICE;
    Device* d = new (std::nothrow) Device();
    if (!d) return RF_E_DEVICE;
    d->ordinal = ordinal;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) {
        delete d;
        return RF_E_DEVICE;
    }
    d->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
        delete d;
        return RF_E_DEVICE;
    }
    *out = reinterpret_cast<rf_device*>(d);
    return RF_OK;
}

rf_status rf_device_close(rf_device* dev) {
    if (!dev) return RF_E_ARG;
    Device* d = reinterpret_cast<Device*>(dev);
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    d->plain.release(&d->bytes);
    for (DevBuf* b : {&d->up_pos, &d->up_attr, &d->up_tris, &d->up_tex, &d->up_color, &d->up_depth,
                      &d->up_desc, &d->ds_fans, &d->ds_keys, &d->ds_vals, &d->ds_tmp, &d->ds_idx})
        b->release(&d->bytes);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return RF_OK;
}

uint64_t rf_device_bytes(const rf_device* dev) {
    return dev ? reinterpret_cast<const Device*>(dev)->bytes : 0;
}

uint32_t rf_tile_height(void) { return TILE_ROWS; }

rf_status rf_device_set_tile_width(rf_device* dev, uint32_t tile_w) {
    if (!dev || (tile_w != 0 && tile_w != 16 && tile_w != 32 && tile_w != 64)) return RF_E_ARG;
    reinterpret_cast<Device*>(dev)->plain_tile_w = tile_w ? tile_w : 32;
    return RF_OK;
}

rf_status rf_device_set_raster_chunk(rf_device* dev, uint32_t chunk) {
    if (!dev || (chunk != 0 && chunk != 6 && chunk != 8)) return RF_E_ARG;
    reinterpret_cast<Device*>(dev)->plain_chunk = chunk ? chunk : 8;
    return RF_OK;
}

rf_status rf_render(rf_device* dev, const rf_draw* draw, rf_target* target, const rf_ctx* ctx,
                    rf_stats* stats) {
    const double t0 = now_s();
    if (!dev || !draw || !target) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    rf_status st = check_ctx(ctx);
    if (st) return st;
    if (draw->kind > KIND_LAST) return RF_E_UNSUPPORTED;
    if (draw->prim > PRIM_EDGE) return RF_E_ARG;
    const uint32_t ipp = draw->prim == PRIM_EDGE ? 2u : 3u;  // indices per primitive
    if ((st = check_tex(draw->kind, draw->tex, draw->tex_w, draw->tex_h, draw->tex_stride))) return st;
    if (!target->color || target->width == 0 || target->height == 0 || target->fmt > PIX_LAST ||
        target->width > MAX_DIM || target->height > MAX_DIM)
        return RF_E_ARG;
    const uint32_t bpp = bytes_per_pixel(target->fmt);
    if (target->color_stride < target->width * bpp) return RF_E_ARG;
    if (target->depth && target->depth_stride < target->width) return RF_E_ARG;
    if ((draw->n_tris && (!draw->tris || !draw->pos)) || (draw->n_verts && !draw->pos)) return RF_E_ARG;
    if (draw->n_tris > 0xFFFFFFFFull || draw->n_verts > 0xFFFFFFFFull) return RF_E_ARG;
    const int na = kind_attrs(draw->kind);
    if (na && draw->n_verts && !draw->attr) return RF_E_ARG;
    if (hipSetDevice(d.ordinal) != hipSuccess) return RF_E_DEVICE;
    hipStream_t s = d.stream;

    // uploads
    const size_t pos_b = draw->n_verts * 12, attr_b = draw->n_verts * na * 4, tri_b = draw->n_tris * ipp * 4;
    const bool textured = draw->kind >= KIND_TEX_CLAMP;
    const size_t tex_b = textured ? (size_t)draw->tex_h * draw->tex_stride * 4 : 0;
    const size_t col_b = (size_t)target->height * target->color_stride;
    const size_t dep_b = target->depth ? (size_t)target->height * target->depth_stride * 4 : 0;
    if (d.up_pos.ensure(std::max<size_t>(pos_b, 4), &d.bytes) || d.up_attr.ensure(std::max<size_t>(attr_b, 4), &d.bytes) ||
        d.up_tris.ensure(std::max<size_t>(tri_b, 4), &d.bytes) || d.up_tex.ensure(std::max<size_t>(tex_b, 4), &d.bytes) ||
        d.up_color.ensure(col_b, &d.bytes) || d.up_depth.ensure(std::max<size_t>(dep_b, 4), &d.bytes) ||
        d.up_desc.ensure(sizeof(DrawDesc) + sizeof(FrameDesc) + 256, &d.bytes))
        return RF_E_DEVICE;
    if (pos_b && hipMemcpyAsync(d.up_pos.p, draw->pos, pos_b, hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
    if (attr_b && hipMemcpyAsync(d.up_attr.p, draw->attr, attr_b, hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
    if (tri_b && hipMemcpyAsync(d.up_tris.p, draw->tris, tri_b, hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
    if (tex_b && hipMemcpyAsync(d.up_tex.p, draw->tex, tex_b, hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
    const bool load_color = !ctx->clear_color, load_depth = target->depth && !ctx->clear_depth;
    auto upload_target = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (load_color) e = hipMemcpyAsync(d.up_color.p, target->color, col_b, hipMemcpyHostToDevice, s);
        if (!e && load_depth) e = hipMemcpyAsync(d.up_depth.p, target->depth, dep_b, hipMemcpyHostToDevice, s);
        return e;
    };
    if (upload_target()) return RF_E_DEVICE;

    DrawDesc dd;
    std::memset(&dd, 0, sizeof(dd));
    dd.prim = draw->prim;
    dd.pos = (const float*)d.up_pos.p;
    dd.attr = (const float*)d.up_attr.p;
    dd.tris = (const uint32_t*)d.up_tris.p;
    dd.tex = textured ? (const uint8_t*)d.up_tex.p : nullptr;
    dd.n_verts = (uint32_t)draw->n_verts;
    dd.n_tris = (uint32_t)draw->n_tris;
    dd.kind = draw->kind;
    dd.frame = 0;
    dd.tex_w = draw->tex_w;
    dd.tex_h = draw->tex_h;
    dd.tex_stride = draw->tex_stride;
    dd.flat_rgba = pack_pixel(PIX_RGBA8, draw->flat_color[0], draw->flat_color[1], draw->flat_color[2],
                              draw->flat_color[3]);
    dd.attr_stride = (uint32_t)na;
    dd.vert_base = 0;
    dd.tri_base = 0;
    std::memcpy(dd.mvp, draw->mvp, 64);
    std::memcpy(dd.vp, draw->viewport, 64);

    FrameDesc fd;
    std::memset(&fd, 0, sizeof(fd));
    fd.color = (uint8_t*)d.up_color.p;
    fd.depth = target->depth ? (float*)d.up_depth.p : nullptr;
    fd.color_stride = target->color_stride;
    fd.depth_stride = target->depth_stride;
    fd.width = target->width;
    fd.height = target->height;
    fd.fmt = target->fmt;
    fill_frame_ctx(fd, *ctx);
    const uint32_t tw = d.plain_tile_w;
    uint32_t tile_base = 0;
    set_grid(fd, tw, tile_base);

    uint8_t* desc = (uint8_t*)d.up_desc.p;
    if (hipMemcpyAsync(desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync(desc + 256 * ((sizeof(dd) + 255) / 256), &fd, sizeof(fd), hipMemcpyHostToDevice, s))
        return RF_E_DEVICE;

    BatchDesc B;
    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)desc;
    B.frames = (const FrameDesc*)(desc + 256 * ((sizeof(dd) + 255) / 256));
    B.n_draws = 1;
    B.n_frames = 1;
    B.n_verts = draw->n_verts;
    B.n_tris = draw->n_tris;
    B.n_groups = (uint32_t)((draw->n_tris + WAVE_SIZE - 1) / WAVE_SIZE);
    B.n_tiles = tile_base;
    B.tile_w = tw;
    B.chunk = d.plain_chunk;
    B.sort_tpw = 1;  // one frame: one tile per binsort wave costs little and spreads dense tiles
    B.tile_h = TILE_ROWS;
    B.extra_cap = std::max<uint32_t>(d.plain_extra_cap, 1024u);
    B.runs_cap = std::max<uint32_t>(d.plain_runs_cap, 4 * B.n_groups + B.n_tiles + 1024);

    // Depth sort (render.rs:180-183): the clip fans are sorted by depth on the
    // device and drawn in that order as unclipped triangles. Edges keep their
    // order: Render::depth is +inf for every edge (render.rs:67-70), and
    // sort_unstable_by leaves a slice of equal keys as it is.
    if (ctx->depth_sort && draw->prim == PRIM_TRI && draw->n_tris) {
        if (draw->n_tris >= (1ull << 29)) return RF_E_ARG;  // triangle << 3 | fan must fit 32 bits
        uint32_t cap = std::max<uint32_t>(d.ds_cap, (uint32_t)std::min<uint64_t>(2 * draw->n_tris + 1024, 1u << 30));
        uint32_t n_f = 0;
        for (int attempt = 0;; ++attempt) {
            if (attempt == 4) return RF_E_INTERNAL;
            if (d.plain.prepare(B, &d.bytes) || d.ds_fans.ensure((size_t)cap * FAN_FLOATS * 4, &d.bytes) ||
                d.ds_keys.ensure((size_t)cap * 16, &d.bytes) || d.ds_vals.ensure((size_t)cap * 8, &d.bytes))
                return RF_E_DEVICE;
            BatchDesc A = B;
            A.depth_sort = (uint32_t)ctx->depth_sort;
            A.fan_cap = cap;
            A.fans = (float*)d.ds_fans.p;
            A.fan_keys = (unsigned long long*)d.ds_keys.p;
            A.fan_vals = (uint32_t*)d.ds_vals.p;
            A.fan_count = (uint32_t*)d.plain.misc.p + 32;
            uint32_t misc[33];
            if (launch_fans(A, s) ||
                hipMemcpyAsync(misc, d.plain.misc.p, sizeof(misc), hipMemcpyDeviceToHost, s) ||
                hipStreamSynchronize(s))
                return RF_E_DEVICE;
            const uint32_t e = misc[2];
            if (e & ERRB_INDEX) return RF_E_INDEX;
            if (e & ERRB_CLIP) return RF_E_INTERNAL;
            if (e & ERRB_EXTRA) {
                cap = misc[32] + misc[32] / 4 + 1024;
                continue;
            }
            n_f = misc[32];
            break;
        }
        d.ds_cap = cap;
        unsigned long long* keys = (unsigned long long*)d.ds_keys.p;
        uint32_t* vals = (uint32_t*)d.ds_vals.p;
        size_t tb = 0;
        if (sort_pairs_u64(nullptr, tb, keys, keys + cap, vals, vals + cap, n_f, s) ||
            d.ds_tmp.ensure(std::max<size_t>(tb, 4), &d.bytes) ||
            sort_pairs_u64(d.ds_tmp.p, tb, keys, keys + cap, vals, vals + cap, n_f, s) ||
            d.ds_idx.ensure(std::max<size_t>((size_t)n_f * 12, 4), &d.bytes))
            return RF_E_DEVICE;
        // the sorted fans as vertex slots and index triples of one draw
        dd.pre = 1;
        dd.n_verts = 3 * n_f;
        dd.n_tris = n_f;
        dd.tris = (const uint32_t*)d.ds_idx.p;
        if (hipMemcpyAsync(desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
        B.n_verts = 3ull * n_f;
        B.n_tris = n_f;
        B.n_groups = (uint32_t)((n_f + WAVE_SIZE - 1) / WAVE_SIZE);
        B.runs_cap = std::max<uint32_t>(B.runs_cap, 4 * B.n_groups + B.n_tiles + 1024);
        B.fans = (float*)d.ds_fans.p;  // the gather reads the fans of the first launch
        if (d.plain.prepare(B, &d.bytes) ||
            launch_fan_gather(B, vals + cap, n_f, (uint32_t*)d.ds_idx.p, s))
            return RF_E_DEVICE;
        B.fans = nullptr;
    }

    std::vector<FrameStats> fs;
    uint32_t err = 0;
    st = run_sized(d, d.plain, B, s, {}, fs, &err, upload_target);
    if (st) return st;
    d.plain_extra_cap = std::max(d.plain_extra_cap, B.extra_cap);
    d.plain_runs_cap = std::max(d.plain_runs_cap, B.runs_cap);
    if ((st = map_err(err))) return st;

    if (hipMemcpyAsync(target->color, d.up_color.p, col_b, hipMemcpyDeviceToHost, s)) return RF_E_DEVICE;
    if (target->depth && hipMemcpyAsync(target->depth, d.up_depth.p, dep_b, hipMemcpyDeviceToHost, s))
        return RF_E_DEVICE;
    if (hipStreamSynchronize(s)) return RF_E_DEVICE;
    if (stats) {
        stats->calls += 1;
        stats->prims_i += draw->n_tris;
        stats->verts_i += draw->n_verts;
        stats->prims_o += fs[0].prims_o;
        stats->verts_o += fs[0].verts_o;
        stats->frags_i += fs[0].frags_i;
        stats->frags_o += fs[0].frags_o;
        stats->seconds += now_s() - t0;
    }
    return RF_OK;
}

// ---- residency -----------------------------------------------------------

static rf_status mesh_create(rf_device* dev, const float* pos, const float* attr, uint32_t n_attr,
                             uint64_t n_verts, const uint32_t* tris, uint64_t n_tris, uint32_t prim,
                             rf_mesh** out) {
    const uint64_t ipp = prim == PRIM_EDGE ? 2 : 3;  // indices per primitive
    if (!dev || !out || (n_verts && !pos) || (n_tris && !tris) || (n_attr && n_verts && !attr) || n_attr > 3)
        return RF_E_ARG;
    if (n_verts > 0xFFFFFFFFull || n_tris > 0xFFFFFFFFull) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    Mesh* m = new (std::nothrow) Mesh();
    if (!m) return RF_E_DEVICE;
    m->n_verts = n_verts;
    m->n_tris = n_tris;
    m->n_attr = n_attr;
    m->prim = prim;
    const size_t pb = std::max<size_t>(n_verts * 12, 4), ab = std::max<size_t>(n_verts * n_attr * 4, 4),
                 tb = std::max<size_t>(n_tris * ipp * 4, 4);
    if (m->pos.ensure(pb, &d.bytes) || m->attr.ensure(ab, &d.bytes) || m->tris.ensure(tb, &d.bytes)) {
        m->pos.release(&d.bytes);
        m->attr.release(&d.bytes);
        m->tris.release(&d.bytes);
        delete m;
        return RF_E_DEVICE;
    }
    hipError_t e = hipSuccess;
    if (n_verts) e = hipMemcpy(m->pos.p, pos, n_verts * 12, hipMemcpyHostToDevice);
    if (!e && n_verts && n_attr) e = hipMemcpy(m->attr.p, attr, n_verts * n_attr * 4, hipMemcpyHostToDevice);
    if (!e && n_tris) e = hipMemcpy(m->tris.p, tris, n_tris * ipp * 4, hipMemcpyHostToDevice);
    if (e) {
        rf_mesh_free(dev, reinterpret_cast<rf_mesh*>(m));
        return RF_E_DEVICE;
    }
    // BBox::of (scene.rs:37-45, 110-114): from (+inf, -inf), extended by each
    // position with f32::min / f32::max, which skip a NaN operand like fminf / fmaxf
    for (int k = 0; k < 3; ++k) {
        m->bbox[k] = INFINITY;
        m->bbox[3 + k] = -INFINITY;
    }
    for (uint64_t i = 0; i < n_verts; ++i)
        for (int k = 0; k < 3; ++k) {
            m->bbox[k] = std::fmin(m->bbox[k], pos[3 * i + k]);
            m->bbox[3 + k] = std::fmax(m->bbox[3 + k], pos[3 * i + k]);
        }
    *out = reinterpret_cast<rf_mesh*>(m);
    return RF_OK;
}

rf_status rf_mesh_create(rf_device* dev, const float* pos, const float* attr, uint32_t n_attr,
                         uint64_t n_verts, const uint32_t* tris, uint64_t n_tris, rf_mesh** out) {
    return mesh_create(dev, pos, attr, n_attr, n_verts, tris, n_tris, PRIM_TRI, out);
}

rf_status rf_mesh_create_edges(rf_device* dev, const float* pos, const float* attr, uint32_t n_attr,
                               uint64_t n_verts, const uint32_t* edges, uint64_t n_edges, rf_mesh** out) {
    return mesh_create(dev, pos, attr, n_attr, n_verts, edges, n_edges, PRIM_EDGE, out);
}

rf_status rf_mesh_free(rf_device* dev, rf_mesh* mesh) {
    if (!dev || !mesh) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    Mesh* m = reinterpret_cast<Mesh*>(mesh);
    m->pos.release(&d.bytes);
    m->attr.release(&d.bytes);
    m->tris.release(&d.bytes);
    delete m;
    return RF_OK;
}

rf_status rf_texture_create(rf_device* dev, const uint8_t* rgba, uint32_t w, uint32_t h, uint32_t stride,
                            rf_texture** out) {
    if (!dev || !out || !rgba || !w || !h || stride < w) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    Tex* t = new (std::nothrow) Tex();
    if (!t) return RF_E_DEVICE;
    t->w = w;
    t->h = h;
    t->stride = stride;
    if (t->data.ensure((size_t)h * stride * 4, &d.bytes) ||
        hipMemcpy(t->data.p, rgba, (size_t)h * stride * 4, hipMemcpyHostToDevice)) {
        t->data.release(&d.bytes);
        delete t;
        return RF_E_DEVICE;
    }
    *out = reinterpret_cast<rf_texture*>(t);
    return RF_OK;
}

rf_status rf_texture_view(rf_device* dev, const rf_texture* tex, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                          rf_texture** out) {
    if (!dev || !tex || !out || !w || !h) return RF_E_ARG;
    const Tex* p = reinterpret_cast<const Tex*>(tex);
    if ((uint64_t)x + w > p->w || (uint64_t)y + h > p->h) return RF_E_ARG;  // Slice2 bounds (buf.rs:459-480)
    Tex* t = new (std::nothrow) Tex();
    if (!t) return RF_E_DEVICE;
    t->parent = p;
    t->offset = ((uint64_t)y * p->stride + x) * 4;
    t->w = w;
    t->h = h;
    t->stride = p->stride;
    *out = reinterpret_cast<rf_texture*>(t);
    return RF_OK;
}

rf_status rf_texture_free(rf_device* dev, rf_texture* tex) {
    if (!dev || !tex) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    Tex* t = reinterpret_cast<Tex*>(tex);
    t->data.release(&d.bytes);  // a view owns nothing
    delete t;
    return RF_OK;
}

rf_status rf_framebuf_create(rf_device* dev, uint32_t w, uint32_t h, uint32_t fmt, int with_depth,
                             rf_framebuf** out) {
    if (!dev || !out || !w || !h || fmt > PIX_LAST || w > MAX_DIM || h > MAX_DIM) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    Framebuf* f = new (std::nothrow) Framebuf();
    if (!f) return RF_E_DEVICE;
    f->w = w;
    f->h = h;
    f->fmt = fmt;
    f->stride = w * bytes_per_pixel(fmt);
    if (f->color.ensure((size_t)h * f->stride, &d.bytes) ||
        (with_depth && f->depth.ensure((size_t)h * w * 4, &d.bytes))) {
        f->color.release(&d.bytes);
        f->depth.release(&d.bytes);
        delete f;
        return RF_E_DEVICE;
    }
    f->has_depth = with_depth != 0;
    (void)hipMemset(f->color.p, 0, (size_t)h * f->stride);
    if (with_depth) (void)hipMemset(f->depth.p, 0, (size_t)h * w * 4);
    *out = reinterpret_cast<rf_framebuf*>(f);
    return RF_OK;
}

rf_status rf_framebuf_free(rf_device* dev, rf_framebuf* fb) {
    if (!dev || !fb) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    Framebuf* f = reinterpret_cast<Framebuf*>(fb);
    f->color.release(&d.bytes);
    f->depth.release(&d.bytes);
    delete f;
    return RF_OK;
}

rf_status rf_framebuf_target(const rf_framebuf* fb, rf_target* out) {
    if (!fb || !out) return RF_E_ARG;
    const Framebuf* f = reinterpret_cast<const Framebuf*>(fb);
    out->color = f->color.p;
    out->width = f->w;
    out->height = f->h;
    out->color_stride = f->stride;
    out->fmt = f->fmt;
    out->depth = f->has_depth ? (float*)f->depth.p : nullptr;
    out->depth_stride = f->w;
    return RF_OK;
}

rf_status rf_framebuf_read(rf_device* dev, const rf_framebuf* fb, void* color, float* depth) {
    if (!dev || !fb) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    const Framebuf* f = reinterpret_cast<const Framebuf*>(fb);
    if (hipSetDevice(d.ordinal)) return RF_E_DEVICE;
    if (color && hipMemcpy(color, f->color.p, (size_t)f->h * f->stride, hipMemcpyDeviceToHost)) return RF_E_DEVICE;
    if (depth && f->has_depth && hipMemcpy(depth, f->depth.p, (size_t)f->h * f->w * 4, hipMemcpyDeviceToHost))
        return RF_E_DEVICE;
    return RF_OK;
}

}  // extern "C"
