// rf_render_noalloc: the plain drop-in call without device allocation.
//
// Everything rf_render keeps on the device -- the draw's vertices, indices
// and texels, the target's colour and depth, the descriptors and the
// pipeline scratch -- is carved from `scratch`, a fixed array in device
// memory that the library holds on every GPU it runs on until the process
// ends. The call allocates nothing: no hipMalloc, no hipFree, no library
// routine that allocates. A draw whose layout does not fit the array is
// refused with RF_E_BUDGET before any device work.
//
// The array is shared by every call on one GPU and keeps the last call's
// contents; every launch initialises what it reads (counters are zeroed by
// the pipeline's first kernel, the target is uploaded or cleared). Calls that
// overlap on one GPU (several host threads on one rf_device) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

namespace rf {

constexpr uint64_t NOALLOC_BYTES = 64ull << 20;  // the largest layout rf_render_noalloc takes

// In host code `scratch` names a host-side stand-in of the array, not its
// device address: only the address hipGetSymbolAddress returns for the
// current device (Device::noalloc) is used.
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

hipError_t noalloc_init(Device& d) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    d.noalloc = (uint8_t*)p;
    return hipSuccess;
}

namespace {

constexpr uint64_t A256 = 256;
inline uint64_t up256(uint64_t x) { return (x + A256 - 1) & ~(A256 - 1); }

// The array's layout for one call: descriptors, inputs, target, then the
// pipeline scratch of the remaining bytes.
struct Carve {
    uint64_t desc = 0, pos = 0, attr = 0, tris = 0, tex = 0, color = 0, depth = 0, scratch = 0, end = 0;
};

Carve carve(const PlainLayout& L) {
    Carve c;
    uint64_t o = 0;
    c.desc = o;
    o += up256(PLAIN_DESC_BYTES);
    c.pos = o;
    o += up256(L.pos_b);
    c.attr = o;
    o += up256(L.attr_b);
    c.tris = o;
    o += up256(L.tri_b);
    c.tex = o;
    o += up256(L.tex_b);
    c.color = o;
    o += up256(L.col_b);
    c.depth = o;
    o += up256(L.dep_b);
    c.scratch = o;
    c.end = o;
    return c;
}

}  // namespace
}  // namespace rf

using namespace rf;

extern "C" {

rf_status rf_render_noalloc(rf_device* dev, const rf_draw* draw, rf_target* target, const rf_ctx* ctx,
                            rf_stats* stats) {
    const double t0 = now_s();
    if (!dev || !draw || !target) return RF_E_ARG;
    Device& d = *reinterpret_cast<Device*>(dev);
    rf_status st = plain_check(draw, target, ctx);
    if (st) return st;
    if (ctx->depth_sort) return RF_E_UNSUPPORTED;  // its radix sort needs temporary storage: rf_render
    if (!d.noalloc) return RF_E_DEVICE;
    const PlainLayout L = plain_layout(draw, target);
    const Carve C = carve(L);
    if (C.end > NOALLOC_BYTES) return RF_E_BUDGET;
    if (hipSetDevice(d.ordinal) != hipSuccess) return RF_E_DEVICE;
    hipStream_t s = d.stream;
    uint8_t* base = d.noalloc;
    PlainBufs P{base + C.pos, base + C.attr, base + C.tris, base + C.tex, base + C.color, base + C.depth,
                base + C.desc};

    DrawDesc dd;
    FrameDesc fd;
    BatchDesc B;
    plain_descs(draw, target, ctx, L, P, d.plain_tile_w, d.plain_tile_h, d.plain_chunk, dd, fd, B);
    // capacities: the clip fans and bin runs rf_render has needed on this
    // device, at least its starting sizes; a run that needs more reruns with
    // what it needed, while the layout fits
    B.extra_cap = std::max<uint32_t>(d.plain_extra_cap, 1024u);
    B.runs_cap = std::max<uint32_t>(d.plain_runs_cap, 4 * B.n_groups + B.n_tiles + 1024);
    if (C.scratch + scratch_bytes(B) > NOALLOC_BYTES) return RF_E_BUDGET;

    if (plain_upload_inputs(draw, L, P, s) || plain_upload_target(target, ctx, L, P, s) ||
        hipMemcpyAsync(P.desc, &dd, sizeof(dd), hipMemcpyHostToDevice, s) ||
        hipMemcpyAsync((void*)B.frames, &fd, sizeof(fd), hipMemcpyHostToDevice, s))
        return RF_E_DEVICE;
    FrameStats fs;
    uint32_t err = 0;
    for (int attempt = 0;; ++attempt) {
        if (attempt == 4) return RF_E_INTERNAL;
        scratch_carve(B, base + C.scratch);
        uint32_t misc[3] = {0, 0, 0}, total_runs = 0;
        if (launch_pipeline(B, s, raster_grid(d, B)) ||
            hipMemcpyAsync(misc, misc_words(B), sizeof(misc), hipMemcpyDeviceToHost, s) ||
            hipMemcpyAsync(&total_runs, B.run_off + B.n_tiles, 4, hipMemcpyDeviceToHost, s) ||
            hipMemcpyAsync(&fs, B.stats, sizeof(fs), hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s))
            return RF_E_DEVICE;
        err = misc[2];
        if (!(err & (ERRB_EXTRA | ERRB_RUNS))) break;
        if (err & ERRB_EXTRA) {
            if (B.extra_cap >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
            B.extra_cap = grow_extra_cap(misc[0]);
        }
        if (err & ERRB_RUNS) B.runs_cap = grow_runs_cap(total_runs);
        if (C.scratch + scratch_bytes(B) > NOALLOC_BYTES) return RF_E_BUDGET;
        // the aborted run may have written the target: start it over
        if (plain_upload_target(target, ctx, L, P, s)) return RF_E_DEVICE;
    }
    if ((st = map_err(err))) return st;
    if (plain_download(target, L, P, s) || hipStreamSynchronize(s)) return RF_E_DEVICE;
    plain_stats(stats, draw, fs, t0);
    return RF_OK;
}

}  // extern "C"
