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
void set_grid(FrameDesc& f, uint32_t tile_w, uint32_t tile_h, uint32_t& tile_base) {
    f.ntx = (f.width + tile_w - 1) / tile_w;
    f.nty = (f.height + tile_h - 1) / tile_h;
    f.tile_base = tile_base;
    tile_base += f.ntx * f.nty;
}

// The scratch arrays of a batch in block order: (bytes, where B keeps the
// pointer). The misc words come first, so a block's first bytes are the
// counters every launch reads back.
template <class Fn>
static void scratch_arrays(BatchDesc& B, Fn&& fn) {
    const uint64_t nv = std::max<uint64_t>(B.n_verts, 1), nt = std::max<uint64_t>(B.n_tris, 1);
    const uint64_t ntl = std::max<uint64_t>(B.n_tiles, 1), nf = std::max<uint64_t>(B.n_frames, 1);
    void* misc = nullptr;
    fn(MISC_BYTES, &misc);
    fn(nv * sizeof(float4), (void**)&B.cv);
    fn(nv * sizeof(float4), (void**)&B.ca);
    fn(nv * 6 * sizeof(float), (void**)&B.sv);
    fn(nt * sizeof(uint32_t), (void**)&B.hdr);
    fn(nt * sizeof(ScrTri), (void**)&B.scr);
    fn(std::max<uint64_t>(B.extra_cap, 1) * sizeof(ScrTri), (void**)&B.extra);
    fn(nt * 2 * sizeof(uint32_t), (void**)&B.rect);
    fn((nt / 64 + 1) * sizeof(uint32_t), (void**)&B.group_draw);
    fn((nv / 256 + 1) * sizeof(uint32_t), (void**)&B.vblock_draw);
    fn(ntl * sizeof(uint32_t), (void**)&B.run_count);
    fn((ntl + 1) * sizeof(uint32_t), (void**)&B.run_off);
    fn(ntl * sizeof(uint32_t), (void**)&B.run_cursor);
    fn(2 * (ntl / 1024 + 2) * sizeof(uint32_t), (void**)&B.scan_blk);
    fn(ntl * sizeof(uint32_t), (void**)&B.tile_list);
    fn(ntl * sizeof(uint32_t), (void**)&B.tile_frame);
    fn(std::max<uint64_t>(B.runs_cap, 1) * sizeof(Run), (void**)&B.runs);
    fn(nf * sizeof(FrameStats), (void**)&B.stats);
    fn(std::max<uint64_t>(B.n_draws, 1) * sizeof(uint32_t), (void**)&B.draw_vis);
    fn(nf * PRIM_PARTS * PRIM_PART_STRIDE * 8, (void**)&B.prims_part);
    uint32_t* m = (uint32_t*)misc;
    B.extra_count = m;
    B.err = m ? m + 2 : nullptr;
    B.busy_count = m ? m + 8 : nullptr;
    B.tile_counter = m ? m + 16 : nullptr;  // TILE_PARTS + 1 words (busy partitions, empty tiles)
}

uint64_t scratch_bytes(const BatchDesc& B0) {
    BatchDesc B = B0;
    uint64_t off = 0;
    scratch_arrays(B, [&](uint64_t bytes, void** p) {
        *p = nullptr;
        off += (bytes + 255) & ~uint64_t(255);
    });
    return off;
}

void scratch_carve(BatchDesc& B, uint8_t* base) {
    uint64_t off = 0;
    scratch_arrays(B, [&](uint64_t bytes, void** p) {
        *p = base + off;
        off += (bytes + 255) & ~uint64_t(255);
    });
}

hipError_t Scratch::prepare(BatchDesc& B, uint64_t* counter) {
    const hipError_t e = mem.ensure(scratch_bytes(B), counter);
    if (e != hipSuccess) return e;
    scratch_carve(B, (uint8_t*)mem.p);
    return hipSuccess;
}

uint32_t grow_extra_cap(uint32_t needed) {
    const uint64_t c = (uint64_t)needed + needed / 4 + 1024;
    return (uint32_t)std::min<uint64_t>(c, MAX_LAUNCH_PRIMS);
}
uint32_t grow_runs_cap(uint32_t needed) {
    const uint64_t c = (uint64_t)needed + needed / 4 + 1024;
    return (uint32_t)std::min<uint64_t>(c, 0xFFFFFF00ull);
}

uint32_t raster_grid(const Device& d, const BatchDesc& B) {
    const uint32_t g = (uint32_t)d.num_cu * raster_wg_per_cu(B.tile_w, B.chunk, B.tile_h);
    return std::max(1u, std::min(B.n_tiles, g));
}

// Runs the pipeline on `s` and waits; grows the clip-fan and run capacities
// and runs again when they were exceeded. `pre` (may be empty) enqueues what
// must precede every run once B's scratch is carved (the depth-sorted fans);
// `restore` re-initialises anything the aborted run may have written (the
// targets) before a rerun.
rf_status run_sized(Device& d, Scratch& sc, BatchDesc& B, hipStream_t s,
                    const std::function<hipError_t(const BatchDesc&)>& pre, std::vector<FrameStats>& stats_out,
                    uint32_t* err_out, const std::function<hipError_t()>& restore, uint64_t* counter) {
    if (B.n_tris >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    for (int attempt = 0; attempt < 4; ++attempt) {
        if (sc.prepare(B, counter) != hipSuccess) return RF_E_DEVICE;
        if (pre && pre(B) != hipSuccess) return RF_E_DEVICE;
        if (launch_pipeline(B, s, raster_grid(d, B)) != hipSuccess) return RF_E_DEVICE;
        uint32_t misc[3];
        uint32_t total_runs = 0;
        if (hipMemcpyAsync(misc, misc_words(B), sizeof(misc), hipMemcpyDeviceToHost, s) != hipSuccess)
            return RF_E_DEVICE;
        if (hipMemcpyAsync(&total_runs, B.run_off + B.n_tiles, 4, hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return RF_E_DEVICE;
        stats_out.resize(B.n_frames);
        if (B.n_frames &&
            hipMemcpyAsync(stats_out.data(), B.stats, B.n_frames * sizeof(FrameStats),
                           hipMemcpyDeviceToHost, s) != hipSuccess)
            return RF_E_DEVICE;
        if (hipStreamSynchronize(s) != hipSuccess) return RF_E_DEVICE;
        const uint32_t err = misc[2];
        bool again = false;
        if (err & ERRB_EXTRA) {
            // more clip fans than a launch can index: refused, not wrapped
            if (B.extra_cap >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
            B.extra_cap = grow_extra_cap(misc[0]);
            again = true;
        }
        if (err & ERRB_RUNS) {
            if (B.runs_cap >= 0xFFFFFF00u) return RF_E_UNSUPPORTED;
            B.runs_cap = grow_runs_cap(total_runs);
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


// ---- the plain path's pieces, shared by rf_render and rf_render_noalloc ----

rf_status plain_check(const rf_draw* draw, const rf_target* target, const rf_ctx* ctx) {
    rf_status st = check_ctx(ctx);
    if (st) return st;
    if (draw->kind > KIND_LAST) return RF_E_UNSUPPORTED;
    if (draw->prim > PRIM_EDGE) return RF_E_ARG;
    if ((st = check_tex(draw->kind, draw->tex, draw->tex_w, draw->tex_h, draw->tex_stride))) return st;
    if (!target->color || target->width == 0 || target->height == 0 || target->fmt > PIX_LAST ||
        target->width > MAX_DIM || target->height > MAX_DIM)
        return RF_E_ARG;
    const uint32_t bpp = bytes_per_pixel(target->fmt);
    if (target->color_stride < target->width * bpp) return RF_E_ARG;
    if (target->depth && target->depth_stride < target->width) return RF_E_ARG;
    if ((draw->n_tris && (!draw->tris || !draw->pos)) || (draw->n_verts && !draw->pos)) return RF_E_ARG;
    if (draw->n_verts > 0xFFFFFFFFull) return RF_E_ARG;
    if (draw->n_tris >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    if (kind_attrs(draw->kind) && draw->n_verts && !draw->attr) return RF_E_ARG;
    return RF_OK;
}

// A Slice2 view spans (h - 1) * stride + w elements from its first
// (buf.rs:570-580): rows of width elements at the view's stride, the gaps
// between them belonging to the parent buffer (an atlas, a pitched SDL
// surface). The device copy keeps the view's stride. Uploads read that span
// as one block (the gaps are the parent's own memory, only read); downloads
// write only the rows, as one block when they are packed (stride = row
// bytes) and by a 2D copy otherwise, so no byte outside the view is written.
PlainLayout plain_layout(const rf_draw* draw, const rf_target* target) {
    PlainLayout L;
    const uint32_t ipp = draw->prim == PRIM_EDGE ? 2u : 3u;  // indices per primitive
    const uint32_t bpp = bytes_per_pixel(target->fmt);
    L.na = kind_attrs(draw->kind);
    L.textured = draw->kind >= KIND_TEX_CLAMP;
    L.pos_b = draw->n_verts * 12;
    L.attr_b = draw->n_verts * L.na * 4;
    L.tri_b = draw->n_tris * ipp * 4;
    L.tex_b = L.textured ? ((size_t)(draw->tex_h - 1) * draw->tex_stride + draw->tex_w) * 4 : 0;
    L.col_row = (size_t)target->width * bpp;
    L.col_pitch = target->color_stride;
    L.col_b = (size_t)(target->height - 1) * L.col_pitch + L.col_row;
    L.dep_row = (size_t)target->width * 4;
    L.dep_pitch_f = target->depth_stride;
    L.dep_b = target->depth ? ((size_t)(target->height - 1) * L.dep_pitch_f * 4 + L.dep_row) : 0;
    return L;
}

hipError_t plain_upload_inputs(const rf_draw* draw, const PlainLayout& L, const PlainBufs& P, hipStream_t s) {
    hipError_t e = hipSuccess;
    if (L.pos_b) e = hipMemcpyAsync(P.pos, draw->pos, L.pos_b, hipMemcpyHostToDevice, s);
    if (!e && L.attr_b) e = hipMemcpyAsync(P.attr, draw->attr, L.attr_b, hipMemcpyHostToDevice, s);
    if (!e && L.tri_b) e = hipMemcpyAsync(P.tris, draw->tris, L.tri_b, hipMemcpyHostToDevice, s);
    if (!e && L.tex_b) e = hipMemcpyAsync(P.tex, draw->tex, L.tex_b, hipMemcpyHostToDevice, s);
    return e;
}

hipError_t plain_upload_target(const rf_target* target, const rf_ctx* ctx, const PlainLayout& L, const PlainBufs& P,
                               hipStream_t s) {
    hipError_t e = hipSuccess;
    if (!ctx->clear_color) e = hipMemcpyAsync(P.color, target->color, L.col_b, hipMemcpyHostToDevice, s);
    if (!e && target->depth && !ctx->clear_depth)
        e = hipMemcpyAsync(P.depth, target->depth, L.dep_b, hipMemcpyHostToDevice, s);
    return e;
}

// Largest view read-back staged through pinned memory (colour and depth).
constexpr size_t STAGE_MAX = 256ull << 20;

hipError_t plain_download_sync(Device& d, const rf_target* target, const PlainLayout& L, const PlainBufs& P,
                               hipStream_t s) {
    const bool col_view = L.col_pitch != L.col_row;
    const bool dep_view = target->depth && L.dep_pitch_f * 4 != L.dep_row;
    const uint32_t h = target->height;
    const size_t cb = col_view ? (size_t)h * L.col_row : 0, db = dep_view ? (size_t)h * L.dep_row : 0;
    // Packed rows go straight into the caller's memory in one copy. A view's
    // rows would be a 2D copy into pageable memory, which the runtime stages
    // row by row (about 10x slower at 2048x2048); into pinned memory it is one
    // DMA transfer, so they come into d.stage and then move on the host,
    // writing only the view's bytes.
    bool staged = (col_view || dep_view) && cb + db <= STAGE_MAX;
    if (staged && cb + db > d.stage_cap) {
        if (d.stage) (void)hipHostFree(d.stage);
        d.stage = nullptr;
        d.stage_cap = 0;
        if (hipHostMalloc((void**)&d.stage, cb + db, 0)) {
            d.stage = nullptr;
            staged = false;
        } else {
            d.stage_cap = cb + db;
        }
    }
    if (!staged) {
        const hipError_t e = plain_download(target, L, P, s);
        return e ? e : hipStreamSynchronize(s);
    }
    hipError_t e = col_view ? hipMemcpy2DAsync(d.stage, L.col_row, P.color, L.col_pitch, L.col_row, h,
                                               hipMemcpyDeviceToHost, s)
                            : hipMemcpyAsync(target->color, P.color, L.col_b, hipMemcpyDeviceToHost, s);
    if (!e && target->depth)
        e = dep_view ? hipMemcpy2DAsync(d.stage + cb, L.dep_row, P.depth, L.dep_pitch_f * 4, L.dep_row, h,
                                        hipMemcpyDeviceToHost, s)
                     : hipMemcpyAsync(target->depth, P.depth, L.dep_b, hipMemcpyDeviceToHost, s);
    if (!e) e = hipStreamSynchronize(s);
    if (e) return e;
    uint8_t* const col = static_cast<uint8_t*>(target->color);
    uint8_t* const dep = reinterpret_cast<uint8_t*>(target->depth);
    for (uint32_t y = 0; y < h; ++y) {
        if (col_view) std::memcpy(col + (size_t)y * L.col_pitch, d.stage + (size_t)y * L.col_row, L.col_row);
        if (dep_view)
            std::memcpy(dep + (size_t)y * L.dep_pitch_f * 4, d.stage + cb + (size_t)y * L.dep_row, L.dep_row);
    }
    return hipSuccess;
}

hipError_t plain_download(const rf_target* target, const PlainLayout& L, const PlainBufs& P, hipStream_t s) {
    // packed rows: one block; a view's rows: a 2D copy (the gaps stay the parent's)
    hipError_t e = L.col_pitch == L.col_row
                       ? hipMemcpyAsync(target->color, P.color, L.col_b, hipMemcpyDeviceToHost, s)
                       : hipMemcpy2DAsync(target->color, L.col_pitch, P.color, L.col_pitch, L.col_row, target->height,
                                          hipMemcpyDeviceToHost, s);
    if (!e && target->depth)
        e = L.dep_pitch_f * 4 == L.dep_row
                ? hipMemcpyAsync(target->depth, P.depth, L.dep_b, hipMemcpyDeviceToHost, s)
                : hipMemcpy2DAsync(target->depth, L.dep_pitch_f * 4, P.depth, L.dep_pitch_f * 4, L.dep_row,
                                   target->height, hipMemcpyDeviceToHost, s);
    return e;
}

void plain_descs(const rf_draw* draw, const rf_target* target, const rf_ctx* ctx, const PlainLayout& L,
                 const PlainBufs& P, uint32_t tile_w, uint32_t tile_h, uint32_t chunk, DrawDesc& dd, FrameDesc& fd,
                 BatchDesc& B) {
    std::memset(&dd, 0, sizeof(dd));
    dd.prim = draw->prim;
    dd.pos = (const float*)P.pos;
    dd.attr = (const float*)P.attr;
    dd.tris = (const uint32_t*)P.tris;
    dd.tex = L.textured ? (const uint8_t*)P.tex : nullptr;
    dd.n_verts = (uint32_t)draw->n_verts;
    dd.n_tris = (uint32_t)draw->n_tris;
    dd.kind = draw->kind;
    dd.frame = 0;
    dd.tex_w = draw->tex_w;
    dd.tex_h = draw->tex_h;
    dd.tex_stride = draw->tex_stride;
    dd.flat_rgba = pack_pixel(PIX_RGBA8, draw->flat_color[0], draw->flat_color[1], draw->flat_color[2],
                              draw->flat_color[3]);
    dd.attr_stride = (uint32_t)L.na;
    dd.vert_base = 0;
    dd.tri_base = 0;
    std::memcpy(dd.mvp, draw->mvp, 64);
    std::memcpy(dd.vp, draw->viewport, 64);

    std::memset(&fd, 0, sizeof(fd));
    fd.color = (uint8_t*)P.color;
    fd.depth = target->depth ? (float*)P.depth : nullptr;
    fd.color_stride = L.col_pitch;
    fd.depth_stride = L.dep_pitch_f;
    fd.width = target->width;
    fd.height = target->height;
    fd.fmt = target->fmt;
    fill_frame_ctx(fd, *ctx);
    uint32_t tile_base = 0;
    set_grid(fd, tile_w, tile_h, tile_base);

    std::memset(&B, 0, sizeof(B));
    B.draws = (const DrawDesc*)P.desc;
    B.frames = (const FrameDesc*)((uint8_t*)P.desc + 256 * ((sizeof(DrawDesc) + 255) / 256));
    B.n_draws = 1;
    B.n_frames = 1;
    B.n_verts = draw->n_verts;
    B.n_tris = draw->n_tris;
    B.n_groups = (uint32_t)((draw->n_tris + WAVE_SIZE - 1) / WAVE_SIZE);
    B.n_tiles = tile_base;
    B.tile_w = tile_w;
    B.chunk = chunk;
    B.sort_tpw = 1;  // one frame: one tile per binsort wave costs little and spreads dense tiles
    B.tile_h = tile_h;
}

void plain_stats(rf_stats* stats, const rf_draw* draw, const FrameStats& fs, double t0) {
    if (!stats) return;
    stats->calls += 1;
    stats->prims_i += draw->n_tris;
    stats->verts_i += draw->n_verts;
    stats->prims_o += fs.prims_o;
    stats->verts_o += fs.verts_o;
    stats->frags_i += fs.frags_i;
    stats->frags_o += fs.frags_o;
    stats->seconds += now_s() - t0;
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
        case RF_E_VIEWPORT: return "to_screen maps geometry more than 2^22 rows above the target";
        case RF_E_NONFINITE: return "infinite screen-space vertex coordinate";
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
    // the device address of rf_render_noalloc's static array on this GPU
    // (this device is current); on failure that entry point reports RF_E_DEVICE
    if (noalloc_init(*d) != hipSuccess) d->noalloc = nullptr;
    *out = reinterpret_cast<rf_device*>(d);
    return RF_OK;
}

rf_status rf_device_close(rf_device* dev) {
    if (!dev) return RF_E_ARG;
    Device* d = reinterpret_cast<Device*>(dev);
    (void)hipSetDevice(d->ordinal);
    (void)hipStreamSynchronize(d->stream);
    d->plain.release(&d->bytes);
    sortpass_release(d->ds, &d->bytes);
    for (DevBuf* b : {&d->up_pos, &d->up_attr, &d->up_tris, &d->up_tex, &d->up_color, &d->up_depth,
                      &d->up_desc})
        b->release(&d->bytes);
    if (d->stage) (void)hipHostFree(d->stage);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return RF_OK;
}

uint64_t rf_device_bytes(const rf_device* dev) {
    return dev ? reinterpret_cast<const Device*>(dev)->bytes : 0;
}

uint32_t rf_tile_height(void) { return DEFAULT_TILE_H; }

rf_status rf_device_set_tile_height(rf_device* dev, uint32_t tile_h) {
    if (!dev || (tile_h != 0 && !valid_tile_h(tile_h))) return RF_E_ARG;
    reinterpret_cast<Device*>(dev)->plain_tile_h = tile_h ? tile_h : DEFAULT_TILE_H;
    return RF_OK;
}

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
    rf_status st = plain_check(draw, target, ctx);
    if (st) return st;
    if (hipSetDevice(d.ordinal) != hipSuccess) return RF_E_DEVICE;
    hipStream_t s = d.stream;
    const PlainLayout L = plain_layout(draw, target);

    // uploads
    if (d.up_pos.ensure(std::max<size_t>(L.pos_b, 4), &d.bytes) ||
        d.up_attr.ensure(std::max<size_t>(L.attr_b, 4), &d.bytes) ||
        d.up_tris.ensure(std::max<size_t>(L.tri_b, 4), &d.bytes) ||
        d.up_tex.ensure(std::max<size_t>(L.tex_b, 4), &d.bytes) || d.up_color.ensure(L.col_b, &d.bytes) ||
        d.up_depth.ensure(std::max<size_t>(L.dep_b, 4), &d.bytes) ||
        d.up_desc.ensure(PLAIN_DESC_BYTES, &d.bytes))
        return RF_E_DEVICE;
    PlainBufs P{d.up_pos.p, d.up_attr.p, d.up_tris.p, d.up_tex.p, d.up_color.p, d.up_depth.p, d.up_desc.p};
    if (plain_upload_inputs(draw, L, P, s)) return RF_E_DEVICE;
    auto upload_target = [&]() -> hipError_t { return plain_upload_target(target, ctx, L, P, s); };
    if (upload_target()) return RF_E_DEVICE;

    DrawDesc dd;
    FrameDesc fd;
    BatchDesc B;
    plain_descs(draw, target, ctx, L, P, d.plain_tile_w, d.plain_tile_h, d.plain_chunk, dd, fd, B);
    uint8_t* desc = (uint8_t*)P.desc;
    if (hipMemcpyAsync(desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync((void*)B.frames, &fd, sizeof(fd), hipMemcpyHostToDevice, s))
        return RF_E_DEVICE;
    B.extra_cap = std::max<uint32_t>(d.plain_extra_cap, 1024u);
    B.runs_cap = std::max<uint32_t>(d.plain_runs_cap, 4 * B.n_groups + B.n_tiles + 1024);

    // Depth sort (render.rs:180-183): the clip fans are sorted by depth on the
    // device and drawn in that order as unclipped triangles. Edges keep their
    // order: Render::depth is +inf for every edge (render.rs:67-70), and
    // sort_unstable_by leaves a slice of equal keys as it is.
    std::function<hipError_t(const BatchDesc&)> pre;
    if (ctx->depth_sort && draw->prim == PRIM_TRI && draw->n_tris) {
        SortPass& S = d.ds;
        S.draws.assign(1, dd);
        S.draws[0].sort = (uint32_t)ctx->depth_sort;
        S.main_of.assign(1, 0);
        S.frames = B.frames;
        S.frames_n = 1;
        uint32_t e = 0;
        if ((st = sortpass_plan(d, S, s, &d.bytes, &e))) return st;
        if (e & ERRB_INDEX) return RF_E_INDEX;
        if (e & ERRB_CLIP) return RF_E_INTERNAL;
        if (sortpass_map(S, s)) return RF_E_DEVICE;
        // the sorted fans as vertex slots and index triples of one draw
        const uint32_t n_f = S.n_f;
        dd.pre = 1;
        dd.in_tris = dd.n_tris;
        dd.in_verts = dd.n_verts;
        dd.n_verts = 3 * n_f;
        dd.n_tris = n_f;
        dd.tris = sortpass_idx(S, 0);
        if (hipMemcpyAsync(desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s)) return RF_E_DEVICE;
        B.n_verts = 3ull * n_f;
        B.n_tris = n_f;
        B.n_groups = (uint32_t)((n_f + WAVE_SIZE - 1) / WAVE_SIZE);
        B.runs_cap = std::max<uint32_t>(B.runs_cap, 4 * B.n_groups + B.n_tiles + 1024);
        // P's fans again and their sort into the (re)carved scratch, before every run
        pre = [&S, s](const BatchDesc& M) -> hipError_t {
            hipError_t r = sortpass_fans(S, s);
            return r ? r : sortpass_sort_gather(S, M, s);
        };
    }

    std::vector<FrameStats> fs;
    uint32_t err = 0;
    st = run_sized(d, d.plain, B, s, pre, fs, &err, upload_target, &d.bytes);
    if (st) return st;
    d.plain_extra_cap = std::max(d.plain_extra_cap, B.extra_cap);
    d.plain_runs_cap = std::max(d.plain_runs_cap, B.runs_cap);
    if ((st = map_err(err))) return st;

    if (plain_download_sync(d, target, L, P, s)) return RF_E_DEVICE;
    plain_stats(stats, draw, fs[0], t0);
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
