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
constexpr uint32_t TILE_ROWS = TILE_HEIGHT;  // rasterizer tile height
#ifndef RF_RASTER_WG_PER_CU
#define RF_RASTER_WG_PER_CU (TILE_HEIGHT > 32 ? 4 : 8)
#endif
constexpr uint32_t RASTER_WG_PER_CU = RF_RASTER_WG_PER_CU;  // persistent rasterizer workgroups per CU

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes, uint64_t* counter);
    void release(uint64_t* counter);
};

// Per-pipeline scratch buffers.
struct Scratch {
    DevBuf cv, ca, sv, hdr, scr, extra, rect, run_count, run_off, run_cursor, scan_blk, tile_frame, tile_list, group_draw, vblock_draw, runs, misc,
        stats, draw_vis, prims_part;
    hipError_t prepare(BatchDesc& B, uint64_t* counter);
    void release(uint64_t* counter);
};

struct Device {
    int ordinal = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    uint64_t bytes = 0;  // device memory held
    Scratch plain;
    DevBuf up_pos, up_attr, up_tris, up_tex, up_color, up_depth, up_desc;
    DevBuf ds_fans, ds_keys, ds_vals, ds_tmp, ds_idx;  // depth sort: clip fans, keys, order, sort scratch
    uint32_t ds_cap = 0;
    uint32_t plain_tile_w = 32;  // rasterizer tile width of rf_render
    uint32_t plain_chunk = 8;    // rasterizer span chunk of rf_render and of batches with a forced tile width
    uint32_t plain_extra_cap = 0, plain_runs_cap = 0;
};

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
void set_grid(FrameDesc& f, uint32_t tile_w, uint32_t& tile_base);
uint32_t raster_grid(const Device& d, uint32_t n_tiles);
rf_status run_sized(Device& d, Scratch& sc, BatchDesc& B, hipStream_t s, const std::vector<FrameStats*>&,
                    std::vector<FrameStats>& stats_out, uint32_t* err_out,
                    const std::function<hipError_t()>& restore);

// Depth sort (rf_kernels.hip, rf_sort.hip): clip fans and their keys for the
// single draw of `B`; the sort of (key, fan) pairs; the sorted fans written
// as vertex slots 3i..3i+2 with index triples `idx`.
hipError_t launch_fans(const BatchDesc& B, hipStream_t s);
hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* keys_in,
                          unsigned long long* keys_out, const uint32_t* vals_in, uint32_t* vals_out, uint32_t n,
                          hipStream_t s);
hipError_t launch_fan_gather(const BatchDesc& B, const uint32_t* order, uint32_t n_fans, uint32_t* idx,
                             hipStream_t s);

// Enqueues the whole pipeline for batch `B` on stream `s` (rf_kernels.hip).
hipError_t launch_pipeline(const BatchDesc& B, hipStream_t s, uint32_t raster_grid);
// Same, with an event recorded after each stage (for per-kernel timing);
// `ev` holds 8 events.
hipError_t launch_pipeline_timed(const BatchDesc& B, hipStream_t s, uint32_t raster_grid, hipEvent_t* ev);
// The rasterizer's launch alone: rf_kernels.hip compiled with RF_RASTER_TU
// (its own object, scheduled for instruction-level parallelism).
hipError_t launch_raster(const BatchDesc& B, hipStream_t s, uint32_t raster_grid);

}  // namespace rf
