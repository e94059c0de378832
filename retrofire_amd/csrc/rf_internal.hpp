// Internal host-side structures of the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_math.hpp"
#include "rf_types.hpp"

namespace rf {

constexpr uint32_t WAVE_SIZE = 64;
constexpr uint32_t MAX_DIM = 16384;         // largest target side (span records pack x in 16 bits)
// Primitives (triangles or edges) of one launch, and clip fans stored beyond
// the first of each triangle: the rasterizer's fan queue packs a primitive
// index as index << 3 | fan and the per-primitive header an extra-fan base as
// base << 3 | count, both in 32 bits. Larger launches are refused
// (RF_E_UNSUPPORTED), never wrapped.
constexpr uint64_t MAX_LAUNCH_PRIMS = 1ull << 29;
// Clip-fan and bin-run capacities grown by a rerun that needs more.
uint32_t grow_extra_cap(uint32_t needed);
uint32_t grow_runs_cap(uint32_t needed);
constexpr uint32_t DEFAULT_TILE_H = 32;  // rasterizer tile height of rf_render and served launches
inline bool valid_tile_h(uint32_t th) { return th == 16 || th == 32 || th == 64; }

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes, uint64_t* counter);
    void release(uint64_t* counter);
};

// Pipeline scratch of a batch: one block of device memory carved into the
// arrays BatchDesc points to (the counters' words first, then per-vertex,
// per-triangle, per-tile and per-frame arrays, each 256-byte aligned).
// scratch_bytes gives the block's size for B's counts and capacities;
// scratch_carve sets B's pointers into a block of at least that size.
constexpr uint64_t MISC_BYTES = 256;  // counters: extra count, error word, busy count, tile counters, fan count
uint64_t scratch_bytes(const BatchDesc& B);
void scratch_carve(BatchDesc& B, uint8_t* base);
inline uint32_t* misc_words(const BatchDesc& B) { return B.extra_count; }

// A grow-only block holding a batch's scratch.
struct Scratch {
    DevBuf mem;
    hipError_t prepare(BatchDesc& B, uint64_t* counter);
    void release(uint64_t* counter) { mem.release(counter); }
};

// Depth sort of some render() calls of a batch (render.rs:180-183,
// 209-219; rf_dsort.cpp). Their draws, over their own meshes, form a batch P
// of their own: P's pass clips every triangle into clip fans with depth keys,
// a radix sort orders the fans by depth (then triangle and fan, so equal
// depths keep submission order), a second stable sort groups them by draw,
// and the gather writes each draw's fans, in that order, as the vertex slots
// and index triples of the corresponding pre draw of the main batch. The
// planning run learns each draw's fan count (fixed for fixed inputs), which
// sizes the main batch's pre draws.
struct SortPass {
    std::vector<DrawDesc> draws;      // P's draws (host copies; vert_base / tri_base in P)
    std::vector<uint32_t> main_of;    // P draw -> its draw in the main batch
    std::vector<uint32_t> cnt, off;   // per P draw: fans, first position in the sorted order
    uint32_t n_f = 0;                 // fans in all
    uint32_t cap = 0;                 // fan capacity of P's pass
    uint32_t sort_bits = 0;           // key bits of the second sort (0: one draw, none)
    size_t tmp_bytes = 0;
    BatchDesc P{};
    Scratch sc;
    DevBuf desc, fans, keys, vals, fdraw, fcnt, key2, tmp, idx, map;
    uint32_t frames_n = 0;            // main frames (P's draws keep their frame indices)
    const FrameDesc* frames = nullptr;
};


struct Device {
    int ordinal = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    uint64_t bytes = 0;  // device memory held
    Scratch plain;
    DevBuf up_pos, up_attr, up_tris, up_tex, up_color, up_depth, up_desc;
    SortPass ds;  // rf_render's depth sort
    uint32_t plain_tile_w = 32;  // rasterizer tile width of rf_render
    uint32_t plain_tile_h = DEFAULT_TILE_H;  // and its tile height
    uint32_t plain_chunk = 8;    // rasterizer span chunk of rf_render and of batches with a forced tile width
    uint32_t plain_extra_cap = 0, plain_runs_cap = 0;
    // rf_render_noalloc's static device array on this GPU (from
    // hipGetSymbolAddress at rf_device_open; null if that failed)
    uint8_t* noalloc = nullptr;
    // pinned host staging of rf_render's read-back into a view (grow-only)
    uint8_t* stage = nullptr;
    size_t stage_cap = 0;
};
// Looks up the device address of rf_render_noalloc's array on the current
// device (rf_noalloc.cpp).
hipError_t noalloc_init(Device& d);

struct Mesh {
    DevBuf pos, attr, tris;
    uint64_t n_verts = 0, n_tris = 0;  // n_tris: primitives (triangles or edges)
    uint32_t n_attr = 0;
    uint32_t prim = 0;                  // PrimType
    float bbox[6] = {0, 0, 0, 0, 0, 0};  // BBox::of over the positions (scene.rs:37-39)
};

struct Tex {
    DevBuf data;                        // owned texels (empty for a view)
    const Tex* parent = nullptr;        // a view's texture (Atlas::get, tex.rs:163-172)
    uint64_t offset = 0;                // a view's first texel in its parent, in bytes
    uint32_t w = 0, h = 0, stride = 0;  // stride in texels
    const uint8_t* texels() const { return parent ? parent->texels() + offset : (const uint8_t*)data.p; }
};

struct Framebuf {
    DevBuf color, depth;
    uint32_t w = 0, h = 0, fmt = 0, stride = 0;
    bool has_depth = false;
};

double now_s();
uint32_t bytes_per_pixel(uint32_t fmt);
void fill_frame_ctx(FrameDesc& f, const rf_ctx& c);
rf_status check_ctx(const rf_ctx* c);
rf_status check_tex(uint32_t kind, const void* tex, uint32_t w, uint32_t h, uint32_t stride);
rf_status map_err(uint32_t e);
void set_grid(FrameDesc& f, uint32_t tile_w, uint32_t tile_h, uint32_t& tile_base);
// Persistent rasterizer workgroups for batch B (resident workgroups per CU of
// its instance, times the CUs, at most one per tile).
uint32_t raster_grid(const Device& d, const BatchDesc& B);
// Resident one-wave rasterizer workgroups per CU of an instance (rf_kernels.hip).
uint32_t raster_wg_per_cu(uint32_t tile_w, uint32_t chunk, uint32_t tile_h);
// Runs B's pipeline on `s` synchronously, growing B's clip-fan and bin-run
// capacities (and the scratch, counted in *counter) until they suffice;
// `pre` runs before each attempt's pipeline, `restore` before each retry.
rf_status run_sized(Device& d, Scratch& sc, BatchDesc& B, hipStream_t s,
                    const std::function<hipError_t(const BatchDesc&)>& pre, std::vector<FrameStats>& stats_out,
                    uint32_t* err_out, const std::function<hipError_t()>& restore, uint64_t* counter);

// The plain path (one render() call with host buffers): argument checks,
// the device layout of its inputs and target, the host<->device copies, the
// descriptors and the stats. rf_render keeps the buffers in grow-only device
// allocations; rf_render_noalloc carves them from a fixed static array.
struct PlainLayout {
    int na = 0;          // attribute floats per vertex
    bool textured = false;
    size_t pos_b = 0, attr_b = 0, tri_b = 0, tex_b = 0;
    size_t col_row = 0, col_pitch = 0, col_b = 0;    // colour: row bytes, pitch (the view's), span bytes
    size_t dep_row = 0, dep_pitch_f = 0, dep_b = 0;  // depth: row bytes, pitch (floats), span bytes
};
struct PlainBufs {
    void *pos, *attr, *tris, *tex, *color, *depth, *desc;
};
constexpr size_t PLAIN_DESC_BYTES = 256 * ((sizeof(DrawDesc) + 255) / 256) + 256 * ((sizeof(FrameDesc) + 255) / 256);
rf_status plain_check(const rf_draw* draw, const rf_target* target, const rf_ctx* ctx);
PlainLayout plain_layout(const rf_draw* draw, const rf_target* target);
hipError_t plain_upload_inputs(const rf_draw* draw, const PlainLayout& L, const PlainBufs& P, hipStream_t s);
hipError_t plain_upload_target(const rf_target* target, const rf_ctx* ctx, const PlainLayout& L, const PlainBufs& P,
                               hipStream_t s);
hipError_t plain_download(const rf_target* target, const PlainLayout& L, const PlainBufs& P, hipStream_t s);
// rf_render's read-back, synchronous: packed rows straight into the target;
// a view's rows by one DMA transfer into d.stage (pinned, up to 256 MiB), then
// on the host into the view (larger views, or no pinned memory: a 2D copy).
hipError_t plain_download_sync(Device& d, const rf_target* target, const PlainLayout& L, const PlainBufs& P,
                               hipStream_t s);
void plain_descs(const rf_draw* draw, const rf_target* target, const rf_ctx* ctx, const PlainLayout& L,
                 const PlainBufs& P, uint32_t tile_w, uint32_t tile_h, uint32_t chunk, DrawDesc& dd, FrameDesc& fd,
                 BatchDesc& B);
void plain_stats(rf_stats* stats, const rf_draw* draw, const FrameStats& fs, double t0);

// Depth sort kernels (rf_kernels.hip, rf_sort.hip): the clip fans of B's
// draws and their keys (B.fan_*); the radix sorts of (key, fan) pairs; the
// draw of each fan in an order; the sorted fans written as vertex slots of
// the main batch M's depth-sorted draws, with their index triples.
hipError_t launch_fans(const BatchDesc& B, hipStream_t s);
hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* keys_in,
                          unsigned long long* keys_out, const uint32_t* vals_in, uint32_t* vals_out, uint32_t n,
                          hipStream_t s);
hipError_t sort_pairs_u32(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t s);
hipError_t launch_fan_draw_keys(const uint32_t* order, const uint32_t* fan_draw, uint32_t n, uint32_t* keys,
                                hipStream_t s);
hipError_t launch_fan_gather(const BatchDesc& P, const BatchDesc& M, const uint32_t* order, const uint32_t* draws_of,
                             const uint32_t* map, uint32_t n_fans, uint32_t* idx, hipStream_t s);

// Uploads P's draws, runs P's pass eagerly (growing the fan capacity as
// needed) and sizes the sorts and the index triples from the fan counts.
// *err_out receives P's error bits.
rf_status sortpass_plan(Device& d, SortPass& S, hipStream_t s, uint64_t* counter, uint32_t* err_out);
// Uploads the (first position, main draw) table once main_of is set.
hipError_t sortpass_map(SortPass& S, hipStream_t s);
// P's pass (reset, culling, vertex shader, clip fans and keys).
hipError_t sortpass_fans(const SortPass& S, hipStream_t s);
// The sorts and the gather into M (whose pre draws and scratch are in place).
hipError_t sortpass_sort_gather(const SortPass& S, const BatchDesc& M, hipStream_t s);
void sortpass_release(SortPass& S, uint64_t* counter);
// Device pointer to index triple `first_fan` of the sorted order.
inline const uint32_t* sortpass_idx(const SortPass& S, uint32_t first_fan) {
    return (const uint32_t*)S.idx.p + 3ull * first_fan;
}

// Enqueues the whole pipeline for batch `B` on stream `s` (rf_kernels.hip).
hipError_t launch_pipeline(const BatchDesc& B, hipStream_t s, uint32_t raster_grid);
// Same, with an event recorded after each stage (for per-kernel timing);
// `ev` holds 8 events.
hipError_t launch_pipeline_timed(const BatchDesc& B, hipStream_t s, uint32_t raster_grid, hipEvent_t* ev);
// The rasterizer's launch alone: rf_kernels.hip compiled with RF_RASTER_TU
// (its own object, scheduled for instruction-level parallelism).
hipError_t launch_raster(const BatchDesc& B, hipStream_t s, uint32_t raster_grid);

}  // namespace rf
