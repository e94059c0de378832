// Depth sort of render() calls (render.rs:180-183, 209-219) on the device,
// for rf_render and for batches: see SortPass in rf_internal.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/retrofire_amd.h"
#include "rf_internal.hpp"

namespace rf {

static void point_fans(SortPass& S) {
    BatchDesc& P = S.P;
    P.fan_cap = S.cap;
    P.fans = (float*)S.fans.p;
    P.fan_keys = (unsigned long long*)S.keys.p;
    P.fan_vals = (uint32_t*)S.vals.p;
    P.fan_draw = (uint32_t*)S.fdraw.p;
    P.fan_cnt = (uint32_t*)S.fcnt.p;
    P.fan_count = misc_words(P) + 32;
}

rf_status sortpass_plan(Device& d, SortPass& S, hipStream_t s, uint64_t* counter, uint32_t* err_out) {
    *err_out = 0;
    const uint32_t nd = (uint32_t)S.draws.size();
    uint64_t nv = 0, nt = 0;
    uint32_t any_cull = 0;
    for (DrawDesc& dd : S.draws) {
        dd.vert_base = nv;
        dd.tri_base = nt;
        dd.pre = 0;
        nv += dd.n_verts;
        nt += dd.n_tris;
        any_cull |= dd.cull;
    }
    if (nt >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
    const size_t db = std::max<size_t>(nd, 1) * sizeof(DrawDesc);
    if (S.desc.ensure(db, counter) || S.fcnt.ensure(std::max<size_t>(nd, 1) * 4, counter) ||
        (nd && hipMemcpyAsync(S.desc.p, S.draws.data(), nd * sizeof(DrawDesc), hipMemcpyHostToDevice, s)))
        return RF_E_DEVICE;
    BatchDesc& P = S.P;
    std::memset(&P, 0, sizeof(P));
    P.draws = (const DrawDesc*)S.desc.p;
    P.frames = S.frames;
    P.n_draws = nd;
    P.n_frames = S.frames_n;
    P.n_verts = nv;
    P.n_tris = nt;
    P.n_groups = (uint32_t)((nt + WAVE_SIZE - 1) / WAVE_SIZE);
    P.any_cull = any_cull;
    P.tile_w = 32;
    P.tile_h = DEFAULT_TILE_H;
    P.chunk = 8;
    uint32_t cap = std::max<uint32_t>(S.cap, (uint32_t)std::min<uint64_t>(2 * nt + 1024, MAX_LAUNCH_PRIMS));
    std::vector<uint32_t> misc(33);
    for (int attempt = 0;; ++attempt) {
        if (attempt == 4) return RF_E_INTERNAL;
        if (S.sc.prepare(P, counter) || S.fans.ensure((size_t)cap * FAN_FLOATS * 4, counter) ||
            S.keys.ensure((size_t)cap * 16, counter) || S.vals.ensure((size_t)cap * 8, counter) ||
            S.fdraw.ensure((size_t)cap * 4, counter))
            return RF_E_DEVICE;
        S.cap = cap;
        point_fans(S);
        S.cnt.assign(nd, 0);
        if (launch_fans(P, s) ||
            hipMemcpyAsync(misc.data(), misc_words(P), 33 * 4, hipMemcpyDeviceToHost, s) ||
            (nd && hipMemcpyAsync(S.cnt.data(), S.fcnt.p, nd * 4, hipMemcpyDeviceToHost, s)) ||
            hipStreamSynchronize(s))
            return RF_E_DEVICE;
        const uint32_t e = misc[2];
        if (e & ERRB_EXTRA) {
            if (cap >= MAX_LAUNCH_PRIMS) return RF_E_UNSUPPORTED;
            cap = grow_extra_cap(misc[32]);
            continue;
        }
        *err_out = e;
        S.n_f = misc[32];
        break;
    }
    S.off.assign(nd, 0);
    uint32_t acc = 0;
    for (uint32_t i = 0; i < nd; ++i) {
        S.off[i] = acc;
        acc += S.cnt[i];
    }
    S.sort_bits = 0;
    while (nd > 1 && (1u << S.sort_bits) < nd) ++S.sort_bits;
    // sort scratch: the larger of the two sorts' needs
    size_t t64 = 0, t32 = 0;
    unsigned long long* k = (unsigned long long*)S.keys.p;
    uint32_t* v = (uint32_t*)S.vals.p;
    if (sort_pairs_u64(nullptr, t64, k, k + S.cap, v, v + S.cap, S.n_f, s)) return RF_E_DEVICE;
    if (S.sort_bits &&
        sort_pairs_u32(nullptr, t32, v, v, v, v, S.n_f, S.sort_bits, s))
        return RF_E_DEVICE;
    S.tmp_bytes = std::max(t64, t32);
    if (S.tmp.ensure(std::max<size_t>(S.tmp_bytes, 4), counter) ||
        S.key2.ensure(std::max<size_t>((size_t)S.n_f * 8, 4), counter) ||
        S.idx.ensure(std::max<size_t>((size_t)S.n_f * 12, 4), counter) ||
        S.map.ensure(std::max<size_t>(nd, 1) * 8, counter))
        return RF_E_DEVICE;
    return RF_OK;
}

hipError_t sortpass_map(SortPass& S, hipStream_t s) {
    std::vector<uint32_t> m(2 * S.draws.size());
    for (size_t i = 0; i < S.draws.size(); ++i) {
        m[2 * i] = S.off[i];
        m[2 * i + 1] = S.main_of[i];
    }
    if (m.empty()) return hipSuccess;
    hipError_t e = hipMemcpyAsync(S.map.p, m.data(), m.size() * 4, hipMemcpyHostToDevice, s);
    if (!e) e = hipStreamSynchronize(s);  // `m` is a local
    return e;
}

hipError_t sortpass_fans(const SortPass& S, hipStream_t s) { return launch_fans(S.P, s); }

hipError_t sortpass_sort_gather(const SortPass& S, const BatchDesc& M, hipStream_t s) {
    if (S.n_f == 0) return hipSuccess;
    unsigned long long* k = (unsigned long long*)S.keys.p;
    uint32_t* v = (uint32_t*)S.vals.p;
    uint32_t* k2 = (uint32_t*)S.key2.p;
    size_t tb = S.tmp_bytes;
    hipError_t e = sort_pairs_u64(S.tmp.p, tb, k, k + S.cap, v, v + S.cap, S.n_f, s);
    const uint32_t* order = v + S.cap;
    const uint32_t* draws_of = nullptr;
    if (!e && S.sort_bits) {
        // stable by draw: each draw's fans contiguous, in depth order
        e = launch_fan_draw_keys(v + S.cap, (const uint32_t*)S.fdraw.p, S.n_f, k2, s);
        tb = S.tmp_bytes;
        if (!e) e = sort_pairs_u32(S.tmp.p, tb, k2, k2 + S.n_f, v + S.cap, v, S.n_f, S.sort_bits, s);
        order = v;
        draws_of = k2 + S.n_f;
    }
    if (!e) e = launch_fan_gather(S.P, M, order, draws_of, (const uint32_t*)S.map.p, S.n_f, (uint32_t*)S.idx.p, s);
    return e;
}

void sortpass_release(SortPass& S, uint64_t* counter) {
    S.sc.release(counter);
    for (DevBuf* b : {&S.desc, &S.fans, &S.keys, &S.vals, &S.fdraw, &S.fcnt, &S.key2, &S.tmp, &S.idx, &S.map})
        b->release(counter);
}

}  // namespace rf
