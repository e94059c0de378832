// HIP kernels of the render path for gfx950 (CDNA4, wave64).
//
//   k_vertex   one thread per vertex slot: the built-in vertex shader
//              (ProjMat3::apply, mat.rs:965-969), the outcode (clip.rs:303-309)
//              and the screen-space vertex (prim.rs:62-88) used when the
//              triangle needs no clipping.
//   k_setup    one wave per group of 64 consecutive triangles: primitive
//              assembly (prim.rs:17-19), frustum clipping (clip.rs:350-400),
//              to_screen, face culling (render.rs:185-196), the tile rectangle
//              of each triangle, and the per-tile count of bin runs.
//   k_scan     exclusive scan of the run counts.
//   k_bin      one wave per group again: writes one run (group, 64-bit mask)
//              per distinct tile the group touches.
//   k_binsort  one workgroup per tile: sorts the tile's runs by group, which
//              restores primitive submission order inside the tile.
//   k_raster   one wave per tile (persistent): the tile's depth and colour
//              sources live in LDS; a lane walks a triangle's edges, then
//              lanes take its row spans in submission order and apply
//              pixels under ordered (row, 8-column) cell claims (tri_fill/scan
//              raster.rs:185-302, ScanlineIter::next raster.rs:84-113,
//              rasterize_fb target.rs:163-198); the tile is written to HBM
//              once.
//
// The file is compiled twice: as the pipeline (every kernel but k_raster, and
// the launchers), and with RF_RASTER_TU defined as k_raster alone, under the
// max-ilp scheduler. That scheduling shortens k_raster's dependent chains
// (1.145 -> 1.131 ms per 64 crates1080 frames) and would lengthen k_setup's
// (0.092 -> 0.096 ms), so it applies to k_raster only.
//
// Exactness: the serial recurrences of the reference (edge stepping per row,
// varying stepping per pixel) are reproduced as serial sums; no closed forms.
#include <hip/hip_runtime.h>

#include "rf_ff.hpp"
#include "rf_math.hpp"
#include "rf_types.hpp"
#include "rf_internal.hpp"

namespace rf {

constexpr int WAVE = 64;
constexpr int BIN_MARGIN = 2;  // pixels around a triangle's bounding box, besides the drift bound
// Rows above the target (y < 0) that a triangle may fold onto row 0; more is
// RF_E_VIEWPORT (the row counter y0 + k, a float in the reference, stays
// exact below 2^22).
constexpr int32_t FOLD_ROWS_MAX = 1 << 22;
// Tuning constants (overridable at build time for sweeps).
#ifndef RF_FF_MIN
#define RF_FF_MIN 12
#endif
constexpr int FF_MIN = RF_FF_MIN;  // step counts above this use the exact fast-forward

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Cross-lane steps through DPP (row shifts, mirrors and row broadcasts: VALU
// operations, no LDS round trip). Every caller runs with all 64 lanes active.
template <int CTRL, int ROW_MASK = 0xF, int BANK_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, BANK_MASK, false);
}
constexpr int DPP_QUAD_1032 = 0xB1, DPP_QUAD_2301 = 0x4E, DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141,
              DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR3 = 0x113, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp_u32<DPP_QUAD_1032>(v, v));
    v = min(v, dpp_u32<DPP_QUAD_2301>(v, v));
    v = min(v, dpp_u32<DPP_ROW_HALF_MIRROR>(v, v));
    v = min(v, dpp_u32<DPP_ROW_MIRROR>(v, v));
    v = min(v, dpp_u32<DPP_ROW_BCAST15, 0xA>(v, v));
    v = min(v, dpp_u32<DPP_ROW_BCAST31, 0xC>(v, v));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
// Inclusive scan with operator op over the lanes; `id` is op's identity.
template <class Op>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t id, Op op) {
    uint32_t x = op(v, dpp_u32<DPP_ROW_SHR1>(id, v));
    x = op(x, dpp_u32<DPP_ROW_SHR2>(id, v));
    x = op(x, dpp_u32<DPP_ROW_SHR3>(id, v));
    x = op(x, dpp_u32<DPP_ROW_SHR4, 0xF, 0xE>(id, x));
    x = op(x, dpp_u32<DPP_ROW_SHR8, 0xF, 0xC>(id, x));
    x = op(x, dpp_u32<DPP_ROW_BCAST15, 0xA>(id, x));
    x = op(x, dpp_u32<DPP_ROW_BCAST31, 0xC>(id, x));
    return x;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
// Exclusive prefix sum over the wave.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
    const uint32_t x = wave_incl_scan(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
    *total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    return x - v;
}

// Rust's saturating float-to-integer casts (`as i32`, `as u32`, `as u8`) in
// one instruction each: V_CVT_I32_F32 and V_CVT_U32_F32 truncate toward
// zero, saturate out-of-range values (infinities included) and convert NaN to
// 0, which is exactly the `as` semantics; a C cast would leave those cases
// undefined.
__device__ __forceinline__ int32_t cvt_i32(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ uint32_t cvt_u32(float x) {
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ uint32_t cvt_u8(float x) {
    const uint32_t r = cvt_u32(x);
    return r < 255u ? r : 255u;
}

__device__ __forceinline__ uint32_t find_draw_tri(const DrawDesc* d, uint32_t n, uint64_t g) {
    uint32_t lo = 0, hi = n;  // last draw with tri_base <= g
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (d[mid].tri_base <= g) lo = mid; else hi = mid;
    }
    return lo;
}
// The same searches from a hint: the last draw with base <= a key at or
// before this one (draw bases never decrease); usually one load.
__device__ __forceinline__ uint32_t draw_of_tri(const DrawDesc* d, uint32_t n, uint32_t hint, uint64_t g) {
    uint32_t k = hint;
    while (k + 1 < n && d[k + 1].tri_base <= g) ++k;
    return k;
}
__device__ __forceinline__ uint32_t draw_of_vert(const DrawDesc* d, uint32_t n, uint32_t hint, uint64_t s) {
    uint32_t k = hint;
    while (k + 1 < n && d[k + 1].vert_base <= s) ++k;
    return k;
}
__device__ __forceinline__ uint32_t find_draw_vert(const DrawDesc* d, uint32_t n, uint64_t s) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (d[mid].vert_base <= s) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// k_vertex
// ---------------------------------------------------------------------------
#ifndef RF_RASTER_TU
__global__ __launch_bounds__(256) void k_vertex(BatchDesc B) {
    const uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= B.n_verts) return;
    const uint32_t di = draw_of_vert(B.draws, B.n_draws, B.vblock_draw[slot / 256], slot);
    if (B.any_cull && !B.draw_vis[di]) return;  // a draw the bounding-box test hid: nothing reads these slots
    const DrawDesc* D = &B.draws[di];
    if (D->pre) return;  // filled by k_fan_gather
    const uint64_t i = slot - D->vert_base;
    const float px = D->pos[3 * i + 0], py = D->pos[3 * i + 1], pz = D->pos[3 * i + 2];
    // vertex(mvp.apply(&v.pos), v.attrib): ProjMat3::apply, to_hom w = 1
    const float* m = D->mvp;
    const float x = dot4(m + 0, px, py, pz, 1.0f);
    const float y = dot4(m + 4, px, py, pz, 1.0f);
    const float z = dot4(m + 8, px, py, pz, 1.0f);
    const float w = dot4(m + 12, px, py, pz, 1.0f);
    const int na = kind_attrs(D->kind);
    const uint64_t as = D->attr_stride;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    if (na >= 2) {
        a0 = D->attr[as * i + 0];
        a1 = D->attr[as * i + 1];
    }
    if (na == 3) a2 = D->attr[as * i + 2];
    const uint32_t oc = outcode(x, y, z, w);
    B.cv[slot] = make_float4(x, y, z, w);
    B.ca[slot] = make_float4(a0, a1, a2, u2f(oc));
    // to_screen (prim.rs:62-88): (x/w, y/w, 1/w), viewport, attribute z_div
    const float nx = x / w, ny = y / w, nz = 1.0f / w;
    const float* v = D->vp;
    float* s = B.sv + 6 * slot;
    s[0] = dot4(v + 0, nx, ny, nz, 1.0f);
    s[1] = dot4(v + 4, nx, ny, nz, 1.0f);
    s[2] = dot4(v + 8, nx, ny, nz, 1.0f);
    const bool persp = kind_persp(D->kind);
    s[3] = persp ? a0 / w : a0;
    s[4] = persp ? a1 / w : a1;
    s[5] = persp ? a2 / w : a2;
}
#endif  // RF_RASTER_TU

// ---------------------------------------------------------------------------
// k_setup helpers: clipping in registers
// ---------------------------------------------------------------------------
struct CV {
    float p[4];
    float a[3];
    uint32_t oc;
};

__device__ __forceinline__ void cv_make(CV& o, const float* p, const float* a) {
    for (int k = 0; k < 4; ++k) o.p[k] = p[k];
    for (int k = 0; k < 3; ++k) o.a[k] = a[k];
    o.oc = outcode(p[0], p[1], p[2], p[3]);
}

// Screen vertex of a clip vertex (to_screen, prim.rs:62-88).
__device__ __forceinline__ void to_screen(const CV& c, const float* vp, bool persp, float* s) {
    const float w = c.p[3];
    const float nx = c.p[0] / w, ny = c.p[1] / w, nz = 1.0f / w;
    s[0] = dot4(vp + 0, nx, ny, nz, 1.0f);
    s[1] = dot4(vp + 4, nx, ny, nz, 1.0f);
    s[2] = dot4(vp + 8, nx, ny, nz, 1.0f);
    s[3] = persp ? c.a[0] / w : c.a[0];
    s[4] = persp ? c.a[1] / w : c.a[1];
    s[5] = persp ? c.a[2] / w : c.a[2];
}

// Tri::winding() == Cw (geom/prim.rs:288-295).
__device__ __forceinline__ bool backface(const float* s0, const float* s1, const float* s2) {
    const float ux = s1[0] - s0[0], uy = s1[1] - s0[1];
    const float vx = s2[0] - s0[0], vy = s2[1] - s0[1];
    float z = 0.0f;
    z = z + (-uy) * vx;
    z = z + ux * vy;
    return z < 0.0f;
}
__device__ __forceinline__ bool culled(int32_t fc, bool bf) {
    return (fc == FC_BACK && bf) || (fc == FC_FRONT && !bf);
}

// Sutherland-Hodgman over the six planes (clip.rs:167-193, 283-301).
// `poly` holds n vertices on entry; returns the output count (0..CLIP_MAX), or
// -1 if a plane would produce more than CLIP_MAX vertices (a convex polygon
// gains at most one vertex per plane, 3 + 6 = 9; more only arises from a
// numerically non-convex polygon and is reported instead of stored).
constexpr int CLIP_MAX = 9;
__device__ int clip_polygon(CV* poly, int n, CV* tmp) {
    CV* in = poly;
    CV* out = tmp;
    for (int p = 0; p < 6; ++p) {
        const uint32_t bit = 1u << p;
        int m = 0;
        for (int k = 0; k < n; ++k) {
            const CV& v0 = in[k];
            const CV& v1 = in[k + 1 < n ? k + 1 : 0];
            if ((bit & v0.oc) == 0) {
                if (m == CLIP_MAX) return -1;
                out[m++] = v0;
            }
            const float d0 = plane_dist(p, v0.p[0], v0.p[1], v0.p[2], v0.p[3]);
            const float d1 = plane_dist(p, v1.p[0], v1.p[1], v1.p[2], v1.p[3]);
            if (d0 * d1 < 0.0f) {
                if (m == CLIP_MAX) return -1;
                const float t = -d0 / (d1 - d0);
                float q[4], a[3];
                for (int c = 0; c < 4; ++c) q[c] = v0.p[c] + (v1.p[c] - v0.p[c]) * t;
                for (int c = 0; c < 3; ++c) a[c] = v0.a[c] + (v1.a[c] - v0.a[c]) * t;
                cv_make(out[m++], q, a);
            }
        }
        n = m;
        if (n == 0) break;
        CV* s = in; in = out; out = s;
    }
    if (n > 0 && in != poly)
        for (int k = 0; k < n; ++k) poly[k] = in[k];
    return n;
}

__device__ __forceinline__ void write_scr(ScrTri* dst, const float* s0, const float* s1,
                                          const float* s2, uint32_t draw, uint32_t kind) {
    float* d = &dst->v[0][0];
    for (int c = 0; c < 6; ++c) d[c] = s0[c];
    for (int c = 0; c < 6; ++c) d[6 + c] = s1[c];
    for (int c = 0; c < 6; ++c) d[12 + c] = s2[c];
    dst->draw = draw;
    dst->kind = kind;
}

// Screen-space extent of a primitive's vertices. fminf / fmaxf skip NaN
// operands, so the extent is that of the vertices' numbers; `nan` records a
// NaN x or y, `inf` an infinite one.
struct RectAcc {
    float xmin, xmax, ymin, ymax;
    bool nan, inf;
    __device__ void init() {
        xmin = ymin = __builtin_inff();
        xmax = ymax = -__builtin_inff();
        nan = inf = false;
    }
    __device__ void add(const float* s) {
        nan = nan || s[0] != s[0] || s[1] != s[1];
        inf = inf || __builtin_isinf(s[0]) || __builtin_isinf(s[1]);
        xmin = fminf(xmin, s[0]);
        xmax = fmaxf(xmax, s[0]);
        ymin = fminf(ymin, s[1]);
        ymax = fmaxf(ymax, s[1]);
    }
};

// Tile rectangle of a triangle's pixels. Rows are exact: tri_fill covers rows
// floor(top_y + 0.5) .. floor(bot_y + 0.5) (scan, raster.rs:286-300). Columns
// are the vertex x range widened by a margin that bounds how far the serially
// stepped edges drift from the exact edges (the rasterizer checks every span
// stays inside):
//   - each row step x + dl rounds by at most half an ulp of the values it
//     passes, so n steps drift by at most n * ulp(X) / 2, X bounding |x|;
//   - dl itself, (x1 - x0) * (1 / (y1 - y0)), is off by a few ulps relative,
//     which over the edge's rows adds at most |x1 - x0| * 2^-22.
// margin_px = BIN_MARGIN + round(D), with D = rows * ulp(X) +
// (xmax - xmin) * 2^-20 (twice and four times those bounds), tolerates a
// drift of D + 1 pixel beyond the rounding of the span ends (round_up_to_half)
// that BIN_MARGIN alone covers. Ordinary content keeps the base margin
// (1080p: X < 2048, a few hundred rows, D < 0.5); a 16k-row edge at
// x = 16,000 gets 18 pixels.
__device__ __forceinline__ int32_t bin_margin(const RectAcc& r, float rows) {
    const float xmin = __builtin_isfinite(r.xmin) ? r.xmin : 0.0f, xmax = __builtin_isfinite(r.xmax) ? r.xmax : 0.0f;
    const float X = fmaxf(fabsf(xmin), fabsf(xmax)) + 1.0f;
    const float ulp = u2f((f2u(X) & 0x7F800000u) - (23u << 23));  // ulp of X (X >= 1)
    const float D = rows * ulp + (xmax - xmin) * 0x1p-20f;
    return BIN_MARGIN + (int32_t)floorf(D + 0.5f);
}
//
// A triangle with a NaN vertex (the screen position of a w = 0 clip vertex,
// 0/0) draws no pixel, but its rows whose left edge is NaN are spans
// `0..x1` (`x0 as usize` of NaN is 0, raster.rs:104-107) that count in
// frags.i: its columns start at 0.
__device__ __forceinline__ bool tile_rect(const RectAcc& r, const FrameDesc& F, uint32_t tw, uint32_t th_log2,
                                          uint32_t* out) {
    const float r0 = floorf(r.ymin + 0.5f), r1 = floorf(r.ymax + 0.5f);
    if (!(r1 > r0)) return false;
    int64_t row0 = (int64_t)r0, row1 = (int64_t)r1 - 1;
    const float xmin = r.nan ? 0.0f : r.xmin, xmax = r.xmax >= xmin ? r.xmax : xmin;
    const int32_t margin = bin_margin(r, r1 - r0);
    int64_t c0 = (int64_t)floorf(xmin) - margin, c1 = (int64_t)floorf(xmax) + margin;
    // rows above the target fold onto row 0 (`self.y as usize`, raster.rs:106)
    if (row0 < 0) row0 = 0;
    if (row1 < 0) row1 = 0;
    if (c0 < 0) c0 = 0;
    if (row1 >= (int64_t)F.height) row1 = (int64_t)F.height - 1;
    if (c1 >= (int64_t)F.width) c1 = (int64_t)F.width - 1;
    if (row1 < row0 || c1 < c0) return false;
    const uint32_t tx0 = (uint32_t)(c0 / tw), tx1 = (uint32_t)(c1 / tw);
    const uint32_t ty0 = (uint32_t)row0 >> th_log2, ty1 = (uint32_t)row1 >> th_log2;
    out[0] = tx0 | (ty0 << 16);
    out[1] = tx1 | (ty1 << 16);
    return true;
}

// Visit, over the whole wave, each distinct tile in the union of the lanes'
// rectangles, in increasing tile order. For every tile, `fn(tile, mask)` runs
// on the lowest lane whose rectangle contains it; `mask` holds those lanes.
// Each lane walks its own rectangle in row-major (= increasing id) order; the
// wave minimum of the lanes' next tiles is processed next, so a lane whose
// rectangle contains that tile always has it as its next one.
template <class Fn>
__device__ __forceinline__ void for_each_tile(bool valid, uint32_t r0, uint32_t r1, uint32_t base,
                                              uint32_t ntx, Fn&& fn) {
    uint32_t tx0 = r0 & 0xFFFF, ty0 = r0 >> 16, tx1 = r1 & 0xFFFF, ty1 = r1 >> 16;
    uint32_t cx = tx0, cy = ty0;
    while (true) {
        const uint32_t next = valid ? base + cy * ntx + cx : 0xFFFFFFFFu;
        const uint32_t t = wave_min_u32(next);
        if (t == 0xFFFFFFFFu) break;
        const bool mine = valid && next == t;
        const unsigned long long m = __ballot(mine);
        if (mine) {
            if (lane_id() == (uint32_t)(__ffsll((long long)m) - 1)) fn(t, m);
            if (++cx > tx1) {
                cx = tx0;
                if (++cy > ty1) valid = false;
            }
        }
    }
}

// Tile rectangle of an exact pixel rectangle (rows r0..r1, columns c0..c1).
__device__ __forceinline__ void tile_rect_px(uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1, uint32_t tw,
                                             uint32_t th_log2, uint32_t* out) {
    out[0] = (c0 / tw) | ((r0 >> th_log2) << 16);
    out[1] = (c1 / tw) | ((r1 >> th_log2) << 16);
}

// One Edge<usize> primitive (prim.rs:41-60): assembly, edge clipping
// (clip.rs:311-348), to_screen, face culling (an edge is never a backface,
// render.rs:72-75, so only FaceCull::Front drops it, ctx.rs:95-101), and the
// set-up of raster::line (raster.rs:122-177). Returns the primitives kept (0
// or 1). The ScrTri holds in v[0] the line's first end after the swaps (its z
// and varyings are every fragment's) and in v[1] its pixel walk: first
// column (wide) or row (tall), pixel count, the stepped coordinate's start
// and step, and the wide flag. The binning rectangle is the exact pixel
// rectangle; a pixel outside the target is the reference's out-of-bounds
// panic (target.rs:173-174).
__device__ uint32_t edge_setup(const BatchDesc& B, const DrawDesc* D, const FrameDesc& F, uint32_t di, uint64_t g,
                               uint64_t t, uint32_t* rect, bool& has_rect) {
    const uint32_t i0 = D->tris[2 * t], i1 = D->tris[2 * t + 1];
    if (i0 >= D->n_verts || i1 >= D->n_verts) {
        atomicOr(B.err, ERRB_INDEX);
        return 0;
    }
    const uint64_t s0 = D->vert_base + i0, s1 = D->vert_base + i1;
    const uint32_t o0 = f2u(B.ca[s0].w), o1 = f2u(B.ca[s1].w);
    if (o0 & o1) return 0;  // both ends outside one plane
    const bool persp = kind_persp(D->kind);
    float a[6], b[6];
    if ((o0 | o1) == 0) {
        for (int c = 0; c < 6; ++c) {
            a[c] = B.sv[6 * s0 + c];
            b[c] = B.sv[6 * s1 + c];
        }
    } else {
        CV ea, eb;
        {
            const float4 p0 = B.cv[s0], q0 = B.ca[s0], p1 = B.cv[s1], q1 = B.ca[s1];
            const float pa[4] = {p0.x, p0.y, p0.z, p0.w}, aa[3] = {q0.x, q0.y, q0.z};
            const float pb[4] = {p1.x, p1.y, p1.z, p1.w}, ab[3] = {q1.x, q1.y, q1.z};
            cv_make(ea, pa, aa);
            cv_make(eb, pb, ab);
        }
        for (int p = 0; p < 6; ++p) {
            const uint32_t bit = 1u << p;
            const bool a_in = (bit & ea.oc) == 0, b_in = (bit & eb.oc) == 0;
            if (!a_in && !b_in) return 0;
            // ClipPlane::intersect (clip.rs:121-145)
            const float d0 = plane_dist(p, ea.p[0], ea.p[1], ea.p[2], ea.p[3]);
            const float d1 = plane_dist(p, eb.p[0], eb.p[1], eb.p[2], eb.p[3]);
            if (d0 * d1 < 0.0f) {
                const float tt = -d0 / (d1 - d0);
                float q[4], qa[3];
                for (int c = 0; c < 4; ++c) q[c] = ea.p[c] + (eb.p[c] - ea.p[c]) * tt;
                for (int c = 0; c < 3; ++c) qa[c] = ea.a[c] + (eb.a[c] - ea.a[c]) * tt;
                CV v;
                cv_make(v, q, qa);
                if (a_in) eb = v;
                else if (b_in) ea = v;
            }
        }
        to_screen(ea, D->vp, persp, a);
        to_screen(eb, D->vp, persp, b);
    }
    if (F.face_cull == FC_FRONT) return 0;
    if (!(__builtin_isfinite(a[0]) && __builtin_isfinite(a[1]) && __builtin_isfinite(b[0]) &&
          __builtin_isfinite(b[1]))) {
        atomicOr(B.err, ERRB_NONFINITE);
        return 0;
    }
    // raster::line: top end first; a wide line then runs left to right
    float v0[6], v1[6];
    const bool sw = a[1] > b[1];
    for (int c = 0; c < 6; ++c) {
        v0[c] = sw ? b[c] : a[c];
        v1[c] = sw ? a[c] : b[c];
    }
    const float dx = v1[0] - v0[0], dy = v1[1] - v0[1];
    const bool wide = fabsf(dx) > dy;
    if (wide && dx < 0.0f)
        for (int c = 0; c < 6; ++c) {
            const float tmp = v0[c];
            v0[c] = v1[c];
            v1[c] = tmp;
        }
    // wide: columns (x0 as usize)..(x1 as usize), y stepped by dy/dx from its
    // value at the rounded x0; tall: rows likewise, x stepped by dx/dy
    const int ax = wide ? 0 : 1;  // the axis walked pixel by pixel
    const float r0 = round_up_to_half(v0[ax]), r1 = round_up_to_half(v1[ax]);
    const float step = wide ? dy / dx : dx / dy;
    const float base = v0[1 - ax] + step * (r0 - v0[ax]);
    const uint64_t S64 = as_usize(r0), E64 = as_usize(r1);
    const uint64_t n64 = E64 > S64 ? E64 - S64 : 0;
    float* d = &B.scr[g].v[0][0];
    for (int c = 0; c < 6; ++c) d[c] = v0[c];
    B.scr[g].draw = di;
    B.scr[g].kind = D->kind | KIND_LINE;
    if (n64 == 0) {
        d[6] = u2f(0);
        d[7] = u2f(0);
        return 1;  // counted, draws nothing
    }
    const uint64_t lim_walk = wide ? F.width : F.height, lim_step = wide ? F.height : F.width;
    if (E64 > lim_walk) {
        atomicOr(B.err, ERRB_BOUNDS);
        return 1;
    }
    const uint32_t S = (uint32_t)S64, n = (uint32_t)n64;
    const float last = n > 1 ? ff_add(base, step, n - 1) : base;
    // the stepped coordinate is monotone along the walk, so its pixels lie
    // between those of the first and last steps
    const uint64_t e0 = as_usize(base), e1 = as_usize(last);
    const uint64_t lo = e0 < e1 ? e0 : e1, hi = e0 < e1 ? e1 : e0;
    if (hi >= lim_step) {
        atomicOr(B.err, ERRB_BOUNDS);
        return 1;
    }
    d[6] = u2f(S);
    d[7] = u2f(n);
    d[8] = base;
    d[9] = step;
    d[10] = u2f(wide ? 1u : 0u);
    d[11] = 0.0f;
    const uint32_t th_log2 = (uint32_t)__builtin_ctz(B.tile_h);
    if (wide) tile_rect_px((uint32_t)lo, (uint32_t)hi, S, S + n - 1, B.tile_w, th_log2, rect);
    else tile_rect_px(S, S + n - 1, (uint32_t)lo, (uint32_t)hi, B.tile_w, th_log2, rect);
    has_rect = true;
    return 1;
}

// ---------------------------------------------------------------------------
// k_setup
// ---------------------------------------------------------------------------
#ifndef RF_RASTER_TU
// Eight waves per SIMD (at most 64 VGPRs): 0.092 -> 0.089 ms on crates1080.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_setup(BatchDesc B) {
    const uint32_t lane = lane_id();
    const uint32_t group = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (group >= B.n_groups) return;  // whole wave exits together
    const uint64_t g = (uint64_t)group * WAVE + lane;
    bool valid = g < B.n_tris;
    uint32_t nf = 0, extra_base = 0;
    uint32_t rect[2] = {0, 0};
    bool has_rect = false;
    uint32_t frame = 0xFFFFFFFFu;
    if (valid) {
        const uint32_t di = draw_of_tri(B.draws, B.n_draws, B.group_draw[group], g);
        const DrawDesc* D = &B.draws[di];
        const FrameDesc& F = B.frames[D->frame];
        frame = D->frame;
        const uint64_t t = g - D->tri_base;
        if (B.any_cull && !B.draw_vis[di]) {
            // the caller skips a hidden object before render() (crates.rs:118-122)
        } else if (D->prim == PRIM_EDGE) {
            nf = edge_setup(B, D, F, di, g, t, rect, has_rect);
        } else if (D->tris[3 * t] >= D->n_verts || D->tris[3 * t + 1] >= D->n_verts ||
                   D->tris[3 * t + 2] >= D->n_verts) {
            atomicOr(B.err, ERRB_INDEX);
        } else {
            const uint32_t i0 = D->tris[3 * t], i1 = D->tris[3 * t + 1], i2 = D->tris[3 * t + 2];
            const uint64_t s0 = D->vert_base + i0, s1 = D->vert_base + i1, s2 = D->vert_base + i2;
            const uint32_t o0 = f2u(B.ca[s0].w), o1 = f2u(B.ca[s1].w), o2 = f2u(B.ca[s2].w);
            const uint32_t all_out = o0 & o1 & o2, any_out = o0 | o1 | o2;
            RectAcc ra;
            ra.init();
            if (all_out != 0) {
                // Hidden
            } else if (any_out == 0) {
                // Visible: screen vertices were computed per vertex
                float v0[6], v1[6], v2[6];
                for (int c = 0; c < 6; ++c) {
                    v0[c] = B.sv[6 * s0 + c];
                    v1[c] = B.sv[6 * s1 + c];
                    v2[c] = B.sv[6 * s2 + c];
                }
                if (!culled(F.face_cull, backface(v0, v1, v2))) {
                    write_scr(&B.scr[g], v0, v1, v2, di, D->kind);
                    ra.add(v0);
                    ra.add(v1);
                    ra.add(v2);
                    nf = 1;
                }
            } else {
                // Clipped
                CV poly[CLIP_MAX], tmp[CLIP_MAX];
                const uint64_t ss[3] = {s0, s1, s2};
                for (int k = 0; k < 3; ++k) {
                    const float4 p = B.cv[ss[k]];
                    const float4 a = B.ca[ss[k]];
                    const float pp[4] = {p.x, p.y, p.z, p.w};
                    const float aa[3] = {a.x, a.y, a.z};
                    cv_make(poly[k], pp, aa);
                }
                const int n = clip_polygon(poly, 3, tmp);
                if (n < 0) atomicOr(B.err, ERRB_CLIP);
                if (n >= 3) {
                    const bool persp = kind_persp(D->kind);
                    // screen vertices in registers, two passes over the fan
                    // (recomputed, same bits), rather than a per-lane array
                    // in scratch memory
                    float s0v[6], sa[6], sb[6];
                    to_screen(poly[0], D->vp, persp, s0v);
                    // fan [p, e0, e1] over rest.windows(2) (clip.rs:373-396)
                    uint32_t keep = 0, cnt = 0;
                    to_screen(poly[1], D->vp, persp, sa);
                    for (int k = 1; k + 1 < n; ++k) {
                        to_screen(poly[k + 1], D->vp, persp, sb);
                        if (!culled(F.face_cull, backface(s0v, sa, sb))) {
                            keep |= 1u << k;
                            ++cnt;
                        }
                        for (int c = 0; c < 6; ++c) sa[c] = sb[c];
                    }
                    if (cnt > 1) {
                        extra_base = atomicAdd(B.extra_count, cnt - 1);
                        if (extra_base + cnt - 1 > B.extra_cap) {
                            atomicOr(B.err, ERRB_EXTRA);
                            cnt = 0;
                        }
                    }
                    uint32_t j = 0;
                    to_screen(poly[1], D->vp, persp, sa);
                    for (int k = 1; k + 1 < n && j < cnt; ++k) {
                        to_screen(poly[k + 1], D->vp, persp, sb);
                        if (keep & (1u << k)) {
                            ScrTri* dst = (j == 0) ? &B.scr[g] : &B.extra[extra_base + j - 1];
                            write_scr(dst, s0v, sa, sb, di, D->kind);
                            ra.add(s0v);
                            ra.add(sa);
                            ra.add(sb);
                            ++j;
                        }
                        for (int c = 0; c < 6; ++c) sa[c] = sb[c];
                    }
                    nf = cnt;
                }
            }
            if (nf > 0) {
                if (ra.inf) {
                    // an infinite screen coordinate: the reference walks up
                    // to 2^32 scanlines for it; reported instead
                    atomicOr(B.err, ERRB_NONFINITE);
                    nf = 0;
                } else if (ra.ymin < -(float)FOLD_ROWS_MAX) {
                    // rows above the target fold onto row 0 through the
                    // saturating `self.y as usize` (raster.rs:106), each a
                    // span of its own; beyond this many the float row
                    // counter of the reference loses its halves
                    atomicOr(B.err, ERRB_NEGY);
                    nf = 0;
                } else {
                    // rows at or below the last one panic in the reference
                    // (target.rs:173-174); so does a span starting past the right edge
                    const float rr0 = floorf(ra.ymin + 0.5f), rr1 = floorf(ra.ymax + 0.5f);
                    if (rr1 > rr0 && (rr1 > (float)F.height || floorf(ra.xmin) > (float)F.width + 1.0f))
                        atomicOr(B.err, ERRB_BOUNDS);
                    has_rect = tile_rect(ra, F, B.tile_w, (uint32_t)__builtin_ctz(B.tile_h), rect);
                }
            }
        }
    }
    if (valid) {
        B.hdr[g] = nf | (extra_base << 3);
        B.rect[2 * g] = has_rect ? rect[0] : 0xFFFFFFFFu;
        B.rect[2 * g + 1] = rect[1];
    }
    // prims.o (render.rs:195; verts.o is 3 per primitive, derived in
    // k_scan_top): one atomic per wave when the wave's triangles all belong to
    // one frame, else one per lane, into the group's partial counter
    {
        const uint32_t fmin = wave_min_u32(nf > 0 ? frame : 0xFFFFFFFFu);
        const uint32_t fmax = ~wave_min_u32(nf > 0 ? ~frame : 0xFFFFFFFFu);
        const uint32_t part = (group % PRIM_PARTS) * PRIM_PART_STRIDE;
        if (fmin == fmax) {
            const unsigned long long tot = wave_sum_u64(nf);
            if (lane == 0 && tot) atomicAdd(&B.prims_part[(uint64_t)fmin * PRIM_PARTS * PRIM_PART_STRIDE + part], tot);
        } else if (nf > 0) {
            atomicAdd(&B.prims_part[(uint64_t)frame * PRIM_PARTS * PRIM_PART_STRIDE + part], (unsigned long long)nf);
        }
    }
    // per-tile run counts
    const FrameDesc* F = has_rect ? &B.frames[frame] : nullptr;
    const uint32_t base = has_rect ? F->tile_base : 0, ntx = has_rect ? F->ntx : 0;
    for_each_tile(has_rect, rect[0], rect[1], base, ntx,
                  [&](uint32_t tile, unsigned long long) { atomicAdd(&B.run_count[tile], 1u); });
}
#endif  // RF_RASTER_TU

__device__ __forceinline__ uint32_t find_frame_tile(const FrameDesc* f, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (f[mid].tile_base <= t) lo = mid; else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// k_scan_*: exclusive scan of the run counts over all tiles, in three steps
// (blocks of 1024 tiles; the block totals; their offsets added back)
// ---------------------------------------------------------------------------
// Two scans side by side: run counts (offsets of the tiles' run lists) and
// "has runs" flags: B.tile_list lists the tiles with runs first, then the
// others (which the rasterizer only clears).
#ifndef RF_RASTER_TU
__global__ __launch_bounds__(1024) void k_scan_local(BatchDesc B) {
    __shared__ uint32_t wave_tot[2][16];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / WAVE;
    const uint32_t i = blockIdx.x * 1024 + tid;
    const uint32_t v = i < B.n_tiles ? B.run_count[i] : 0u;
    const uint32_t f = v != 0 ? 1u : 0u;
    uint32_t tot, ftot;
    const uint32_t ex = wave_excl_scan(v, &tot);
    const uint32_t fex = wave_excl_scan(f, &ftot);
    if (lane == 0) {
        wave_tot[0][w] = tot;
        wave_tot[1][w] = ftot;
    }
    __syncthreads();
    uint32_t off = 0, foff = 0;
    for (uint32_t k = 0; k < w; ++k) {
        off += wave_tot[0][k];
        foff += wave_tot[1][k];
    }
    if (i < B.n_tiles) {
        B.run_off[i] = off + ex;
        B.run_cursor[i] = foff + fex;  // block-local list position until k_scan_add
        B.tile_frame[i] = find_frame_tile(B.frames, B.n_frames, i);
    }
    if (tid == 1023) {
        B.scan_blk[2 * blockIdx.x] = off + ex + v;
        B.scan_blk[2 * blockIdx.x + 1] = foff + fex + f;
    }
}

__global__ __launch_bounds__(1024) void k_scan_top(BatchDesc B) {
    __shared__ uint32_t wave_tot[2][16];
    __shared__ uint32_t carry[2];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / WAVE;
    const uint32_t nb = (B.n_tiles + 1023) / 1024;
    if (tid < 2) carry[tid] = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t v = i < nb ? B.scan_blk[2 * i] : 0u, f = i < nb ? B.scan_blk[2 * i + 1] : 0u;
        uint32_t tot, ftot;
        const uint32_t ex = wave_excl_scan(v, &tot);
        const uint32_t fex = wave_excl_scan(f, &ftot);
        if (lane == 0) {
            wave_tot[0][w] = tot;
            wave_tot[1][w] = ftot;
        }
        __syncthreads();
        uint32_t off = carry[0], foff = carry[1];
        for (uint32_t k = 0; k < w; ++k) {
            off += wave_tot[0][k];
            foff += wave_tot[1][k];
        }
        if (i < nb) {
            B.scan_blk[2 * i] = off + ex;
            B.scan_blk[2 * i + 1] = foff + fex;
        }
        __syncthreads();
        if (tid < 2)
            for (uint32_t k = 0; k < 16; ++k) carry[tid] += wave_tot[tid][k];
        __syncthreads();
    }
    // prims.o of every frame from its partial counters; verts.o is 3 per
    // primitive (render.rs:196)
    for (uint32_t f = tid; f < B.n_frames; f += blockDim.x) {
        unsigned long long p = 0;
        for (uint32_t k = 0; k < PRIM_PARTS; ++k) p += B.prims_part[((uint64_t)f * PRIM_PARTS + k) * PRIM_PART_STRIDE];
        B.stats[f].prims_o = p;
        B.stats[f].verts_o = 3ull * p;
    }
    if (tid == 0) {
        B.run_off[B.n_tiles] = carry[0];
        *B.busy_count = carry[1];
        if (carry[0] > B.runs_cap) atomicOr(B.err, ERRB_RUNS);
        for (uint32_t k = 0; k <= TILE_PARTS; ++k) B.tile_counter[k] = 0;
    }
}

__global__ __launch_bounds__(1024) void k_scan_add(BatchDesc B) {
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    if (i >= B.n_tiles) return;
    if (blockIdx.x > 0) B.run_off[i] += B.scan_blk[2 * blockIdx.x];
    const uint32_t busy_before = B.run_cursor[i] + B.scan_blk[2 * blockIdx.x + 1];
    B.tile_list[B.run_count[i] != 0 ? busy_before : *B.busy_count + (i - busy_before)] = i;
    B.run_cursor[i] = 0;
}
#endif  // RF_RASTER_TU

#ifndef RF_SORT_LDS
#define RF_SORT_LDS 1024
#endif
#ifndef RF_SORT_TPW
#define RF_SORT_TPW 8
#endif
constexpr int SORT_LDS = RF_SORT_LDS;  // bins up to this many runs sort in LDS
constexpr int SORT_TPW = RF_SORT_TPW;  // default tiles per k_binsort wave (their sorts run one after another)
static_assert(SORT_TPW >= 1 && SORT_TPW <= 64, "one lane per tile");
// Tiles per k_binsort wave of this batch (1..64; 0: the default). One tile a
// wave spreads dense scenes' long run lists over many waves; 8 a wave costs
// fewer waves where lists are short.
__host__ __device__ __forceinline__ uint32_t sort_tpw(const BatchDesc& B) {
    return B.sort_tpw >= 1 && B.sort_tpw <= 64 ? B.sort_tpw : (uint32_t)SORT_TPW;
}

// Sorts the runs held one per lane (lanes >= n hold none) by group: a bitonic
// network over the 64 lanes; lane i ends with the i-th run. Groups are
// distinct within a tile.
__device__ __forceinline__ void wave_sort_runs(uint32_t n, uint32_t& group, unsigned long long& mask,
                                               uint32_t& multi) {
    const uint32_t lane = lane_id();
    const bool in = lane < n;
    const uint32_t g = in ? group : 0xFFFFFFFFu;
    const uint32_t nextg = __shfl_down(g, 1, WAVE);
    if (!__ballot(in && lane + 1 < n && g > nextg)) return;  // already in order
    unsigned long long kk = ((unsigned long long)g << 6) | lane;
    for (uint32_t ksz = 2; ksz <= (uint32_t)WAVE; ksz <<= 1) {
        for (uint32_t js = ksz >> 1; js > 0; js >>= 1) {
            const unsigned long long other = __shfl_xor(kk, js, WAVE);
            const bool take_min = ((lane & js) == 0) == ((lane & ksz) == 0);
            kk = take_min ? (other < kk ? other : kk) : (other > kk ? other : kk);
        }
    }
    mask = __shfl(mask, (uint32_t)(kk & 63u), WAVE);
    multi = __shfl(multi, (uint32_t)(kk & 63u), WAVE);
    group = (uint32_t)(kk >> 6);
}

// Orders the runs of each tile with more than 64 of them by group (k_bin
// appends them in arrival order; the rasterizer sorts lists of up to 64 runs
// itself, in registers). One wave per sort_tpw(B) tiles, whose long lists it
// sorts one after another: bins of up to SORT_LDS runs in LDS, larger ones in
// place in global memory.
#ifndef RF_RASTER_TU
__global__ __launch_bounds__(64) void k_binsort(BatchDesc B) {
    __shared__ unsigned long long key[SORT_LDS];
    __shared__ unsigned long long msk[SORT_LDS];
    const uint32_t lane = lane_id();
    const uint32_t tpw = sort_tpw(B);
    const uint32_t t0 = blockIdx.x * tpw;
    const uint32_t t = t0 + lane;
    uint32_t off = 0, n = 0;
    if (lane < tpw && t < B.n_tiles) {
        off = B.run_off[t];
        n = B.run_off[t + 1] - off;
        if (off >= B.runs_cap) n = 0;
        else if (off + n > B.runs_cap) n = B.runs_cap - off;
    }
    unsigned long long todo = __ballot(n > (uint32_t)WAVE);
    while (todo) {
        const uint32_t j = (uint32_t)(__ffsll((long long)todo) - 1);
        todo &= todo - 1;
        const uint32_t o = __shfl(off, j, WAVE), nn = __shfl(n, j, WAVE);
        Run* runs = B.runs + o;
        if (nn <= (uint32_t)SORT_LDS) {
            bool unsorted = false;
            for (uint32_t i = lane; i + 1 < nn; i += WAVE) unsorted |= runs[i].group > runs[i + 1].group;
            if (!__ballot(unsorted)) continue;
            uint32_t p2 = 1;
            while (p2 < nn) p2 <<= 1;
            for (uint32_t i = lane; i < nn; i += WAVE) {
                // groups are distinct in a tile, so the flag and index below
                // the group never decide the order
                key[i] = ((unsigned long long)runs[i].group << 32) | ((unsigned long long)runs[i].multi << 31) | i;
                msk[i] = runs[i].mask;
            }
            __syncthreads();
            // all-ascending bitonic network: the first step of each merge
            // compares i with i ^ (k - 1), later steps i with i ^ j; indices
            // >= nn act as +inf keys
            for (uint32_t ksz = 2; ksz <= p2; ksz <<= 1) {
                for (uint32_t js = ksz >> 1; js > 0; js >>= 1) {
                    const uint32_t x = (js == (ksz >> 1)) ? (ksz - 1) : js;
                    for (uint32_t i = lane; i < nn; i += WAVE) {
                        const uint32_t l = i ^ x;
                        if (l > i && l < nn) {
                            const unsigned long long a = key[i], c = key[l];
                            if (a > c) {
                                key[i] = c;
                                key[l] = a;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for (uint32_t i = lane; i < nn; i += WAVE) {
                Run w;
                w.group = (uint32_t)(key[i] >> 32);
                w.multi = (uint32_t)(key[i] >> 31) & 1u;
                w.mask = msk[(uint32_t)key[i] & 0x7FFFFFFFu];
                runs[i] = w;
            }
            __syncthreads();
        } else {
            // very large bin: the same network in place on the run records
            uint32_t p2 = 1;
            while (p2 < nn) p2 <<= 1;
            for (uint32_t ksz = 2; ksz <= p2; ksz <<= 1) {
                for (uint32_t js = ksz >> 1; js > 0; js >>= 1) {
                    const uint32_t x = (js == (ksz >> 1)) ? (ksz - 1) : js;
                    for (uint32_t i = lane; i < nn; i += WAVE) {
                        const uint32_t l = i ^ x;
                        if (l > i && l < nn) {
                            const Run a = runs[i], c = runs[l];
                            if (a.group > c.group) {
                                runs[i] = c;
                                runs[l] = a;
                            }
                        }
                    }
                    __threadfence_block();
                    __syncthreads();
                }
            }
        }
    }
}
#endif  // RF_RASTER_TU

// ---------------------------------------------------------------------------
// k_raster
// ---------------------------------------------------------------------------
struct Trap {
    float y0i;             // first row (floor(y0 + 0.5))
    uint32_t n;            // row count
    float lx, lz, la[3];   // left state at the first row
    float dlx, dlz, dla[3];
    float rx, drx;         // right x at the first row, step
    float dvz, dva[3];     // d(varyings)/dx
};
struct Slot {
    Trap t[2];
    int32_t ylo, yhi;      // rows covered (absolute)
};

// scan (raster.rs:248-302) for one trapezoid; y components are not needed by
// any output and are left out.
__device__ __forceinline__ void scan_setup(Trap& T, float y0, float y1, const float* l0,
                                           const float* l1, const float* r0, const float* r1,
                                           int na) {
    // l*/r* layout: x, y, z, a0, a1, a2
    const float recip_dy = 1.0f / (y1 - y0);
    float dl[6], dr[6];
    for (int c = 0; c < 6; ++c) {
        dl[c] = (l1[c] - l0[c]) * recip_dy;
        dr[c] = (r1[c] - r0[c]) * recip_dy;
    }
    float l0s[6], r0s[6];
    for (int c = 0; c < 6; ++c) {
        l0s[c] = l0[c] + dl[c];
        r0s[c] = r0[c] + dr[c];
    }
    const float dx = r0s[0] - l0s[0];
    const float rdx = 1.0f / dx;
    T.dvz = (r0s[2] - l0s[2]) * rdx;
    for (int c = 0; c < 3; ++c) T.dva[c] = (c < na) ? (r0s[3 + c] - l0s[3 + c]) * rdx : 0.0f;
    const float y0r = round_up_to_half(y0);
    const float y1r = round_up_to_half(y1);
    const float tweak = y0r - y0;
    // l0.lerp(&l0.step(&dl_dy), y_tweak) == l0 + ((l0 + dl) - l0) * tweak
    T.lx = l0[0] + (l0s[0] - l0[0]) * tweak;
    T.lz = l0[2] + (l0s[2] - l0[2]) * tweak;
    for (int c = 0; c < 3; ++c) T.la[c] = (c < na) ? l0[3 + c] + (l0s[3 + c] - l0[3 + c]) * tweak : 0.0f;
    T.dlx = dl[0];
    T.dlz = dl[2];
    for (int c = 0; c < 3; ++c) T.dla[c] = (c < na) ? dl[3 + c] : 0.0f;
    T.rx = r0[0] + dr[0] * tweak;
    T.drx = dr[0];
    T.n = as_u32(y1r - y0r);
    T.y0i = y0r - 0.5f;
}

// tri_fill (raster.rs:185-224) setup for one screen triangle (vertices as
// x, y, z, a0, a1, a2; sorted in registers).
__device__ __forceinline__ void swap6(float* a, float* b) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const float t = a[k];
        a[k] = b[k];
        b[k] = t;
    }
}
// The total_cmp order key of a screen y. A NaN y (0/0, the screen position
// of a w = 0 clip vertex) is keyed as x86-64's default NaN, which has its
// sign bit set and so sorts before -inf: that is the NaN the reference's
// arithmetic produces there (the GPU's own NaN has no defined sign).
__device__ __forceinline__ int32_t y_key(float y) { return total_key(y == y ? y : u2f(0xFFC00000u)); }

__device__ __forceinline__ void tri_setup(Slot& S, float (&va)[6], float (&vb)[6], float (&vc)[6], int na) {
    // stable sort by total_cmp of y (slice::sort_by is stable)
    if (y_key(vb[1]) < y_key(va[1])) swap6(va, vb);
    if (y_key(vc[1]) < y_key(vb[1])) swap6(vb, vc);
    if (y_key(vb[1]) < y_key(va[1])) swap6(va, vb);
    const float* top = va;
    const float* mid0 = vb;
    const float* bot = vc;
    const float top_y = top[1], mid_y = mid0[1], bot_y = bot[1];
    const float tt = (mid_y - top_y) / (bot_y - top_y);
    float mid1[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) mid1[k] = top[k] + (bot[k] - top[k]) * tt;
    const bool l_is_mid0 = mid0[0] < mid1[0];
    float left[6], right[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        left[k] = l_is_mid0 ? mid0[k] : mid1[k];
        right[k] = l_is_mid0 ? mid1[k] : mid0[k];
    }
    scan_setup(S.t[0], top_y, mid_y, top, left, top, right, na);
    scan_setup(S.t[1], mid_y, bot_y, left, bot, right, bot, na);
    int32_t lo = 0x7FFFFFFF, hi = -0x7FFFFFFF;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (S.t[h].n == 0) continue;
        const int32_t y0 = (int32_t)S.t[h].y0i;
        const int32_t y1 = y0 + (int32_t)S.t[h].n;
        lo = y0 < lo ? y0 : lo;
        hi = y1 > hi ? y1 : hi;
    }
    S.ylo = lo;
    S.yhi = hi;
}

// 16-byte framebuffer stores. The rasterizer writes every target pixel once
// and never reads it back in the launch, so with RF_NT_STORE the stores are
// streaming (non-temporal): a gigabyte of colour and depth per launch then
// does not displace from L2 the triangles, runs and texels that other tiles
// still read.
#ifndef RF_NT_STORE
#define RF_NT_STORE 1
#endif
typedef uint32_t rf_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st16(void* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
#if RF_NT_STORE
    __builtin_nontemporal_store(rf_u32x4{a, b, c, d}, (rf_u32x4*)p);
#else
    *(uint4*)p = make_uint4(a, b, c, d);
#endif
}
__device__ __forceinline__ void st16f(float* p, float a, float b, float c, float d) {
    st16(p, __float_as_uint(a), __float_as_uint(b), __float_as_uint(c), __float_as_uint(d));
}

__device__ __forceinline__ void store_pixel(const FrameDesc& F, uint32_t x, uint32_t y, uint32_t v) {
    uint8_t* p = F.color + (uint64_t)y * F.color_stride;
    if (F.fmt == PIX_RGB8) {
        p += 3ull * x;
        p[0] = (uint8_t)v;
        p[1] = (uint8_t)(v >> 8);
        p[2] = (uint8_t)(v >> 16);
        return;
    }
    if (F.fmt >= PIX_RGBA4444) {
        *(uint16_t*)(p + 2ull * x) = (uint16_t)v;
        return;
    }
    *(uint32_t*)(p + 4ull * x) = v;
}


#ifndef RF_RUN_MULTI
#define RF_RUN_MULTI 1  // runs flag groups with several fans per triangle
#endif
#ifndef RF_CLEAR_ROWS
#define RF_CLEAR_ROWS 1  // clear runs of empty tiles row segment by row segment
#endif
#ifndef RF_EARLY_CLEAR
#define RF_EARLY_CLEAR 0  // workgroups that clear the tiles without runs first
#endif
#ifndef RF_VEC_INIT
#define RF_VEC_INIT 1  // tile init of a cleared frame with 16-byte LDS fills
#endif
#ifndef RF_CLEAR_IN_RASTER
#define RF_CLEAR_IN_RASTER 1
#endif
#ifndef RF_EMPTY_GRAB
#define RF_EMPTY_GRAB 32
#endif
constexpr int EMPTY_GRAB = RF_EMPTY_GRAB;  // tiles without runs a wave takes at once

// The fused clear (Frame::clear) of one tile that nothing draws into: 16-byte
// stores where whole rows of 4-byte pixels are aligned.
template <int TW, int TILE_H>
__device__ __forceinline__ void clear_tile(const FrameDesc& F, uint32_t tile, uint32_t lane) {
    const bool has_depth = F.depth != nullptr;
    if (!F.clear_color && !(has_depth && F.clear_depth)) return;
    const uint32_t lt = tile - F.tile_base;
    const uint32_t x0 = (lt % F.ntx) * TW, y0 = (lt / F.ntx) * TILE_H;
    const uint32_t ncols = min((uint32_t)TW, F.width - x0), nrows = min((uint32_t)TILE_H, F.height - y0);
    const bool vec = ncols == (uint32_t)TW && pixfmt_bytes(F.fmt) == 4u && ((uintptr_t)F.color & 15u) == 0 &&
                     (F.color_stride & 15u) == 0 &&
                     (!has_depth || (((uintptr_t)F.depth & 15u) == 0 && (F.depth_stride & 3u) == 0));
    if (vec) {
        constexpr uint32_t QR = TW / 4;  // quads per row
        const uint32_t cp = F.clear_pixel;
        const float cz = F.clear_depth_value;
        for (uint32_t q = lane; q < nrows * QR; q += WAVE) {
            const uint32_t r = q / QR, c = (q - r * QR) * 4;
            if (F.clear_color)
                st16(F.color + (uint64_t)(y0 + r) * F.color_stride + 4ull * (x0 + c), cp, cp, cp, cp);
            if (has_depth && F.clear_depth) st16f(F.depth + (uint64_t)(y0 + r) * F.depth_stride + x0 + c, cz, cz, cz, cz);
        }
    } else {
        for (uint32_t i = lane; i < ncols * nrows; i += WAVE) {
            const uint32_t r = i / ncols, c = i - r * ncols;
            if (F.clear_color) store_pixel(F, x0 + c, y0 + r, F.clear_pixel);
            if (has_depth && F.clear_depth)
                F.depth[(uint64_t)(y0 + r) * F.depth_stride + x0 + c] = F.clear_depth_value;
        }
    }
}

// Clears tiles without runs (listed after the others in B.tile_list): wave
// `w` takes EMPTY_GRAB list entries.
__device__ __forceinline__ void clear_empty_tiles(const BatchDesc& B, uint32_t w, uint32_t lane) {
    const uint32_t n_busy = *B.busy_count, n_empty = B.n_tiles - n_busy;
    const uint32_t e0 = w * EMPTY_GRAB;
    if (e0 >= n_empty) return;
    const uint32_t ne = min((uint32_t)EMPTY_GRAB, n_empty - e0);
    uint32_t my_tile = 0, my_fi = 0;
    if (lane < ne) {
        my_tile = B.tile_list[n_busy + e0 + lane];
        my_fi = B.tile_frame[my_tile];
    }
    uint32_t cur_fi = 0xFFFFFFFFu;
    FrameDesc F;
    for (uint32_t j = 0; j < ne; ++j) {
        const uint32_t tile = (uint32_t)__builtin_amdgcn_readlane((int)my_tile, (int)j);
        const uint32_t fi = (uint32_t)__builtin_amdgcn_readlane((int)my_fi, (int)j);
        if (fi != cur_fi) {
            F = B.frames[fi];
            cur_fi = fi;
        }
        if (B.tile_h == 16) {
            if (B.tile_w == 16) clear_tile<16, 16>(F, tile, lane);
            else if (B.tile_w == 32) clear_tile<32, 16>(F, tile, lane);
            else clear_tile<64, 16>(F, tile, lane);
        } else if (B.tile_h == 64) {
            if (B.tile_w == 16) clear_tile<16, 64>(F, tile, lane);
            else if (B.tile_w == 32) clear_tile<32, 64>(F, tile, lane);
            else clear_tile<64, 64>(F, tile, lane);
        } else {
            if (B.tile_w == 16) clear_tile<16, 32>(F, tile, lane);
            else if (B.tile_w == 32) clear_tile<32, 32>(F, tile, lane);
            else clear_tile<64, 32>(F, tile, lane);
        }
    }
}

// Blocks past the groups' clear the tiles without runs: binning waits on
// atomics and loads while the clears only store, so they share the GPU well.
#ifndef RF_RASTER_TU
__global__ __launch_bounds__(256) void k_bin(BatchDesc B) {
    const uint32_t lane = lane_id();
    const uint32_t bin_blocks = (B.n_groups + 3) / 4;
    if (blockIdx.x >= bin_blocks) {
        if (!B.clear_in_raster) clear_empty_tiles(B, (blockIdx.x - bin_blocks) * 4 + threadIdx.x / WAVE, lane);
        return;
    }
    const uint32_t group = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (group >= B.n_groups) return;
    const uint64_t g = (uint64_t)group * WAVE + lane;
    bool valid = false;
    uint32_t r0 = 0, r1 = 0, base = 0, ntx = 0;
    if (g < B.n_tris) {
        r0 = B.rect[2 * g];
        r1 = B.rect[2 * g + 1];
        valid = r0 != 0xFFFFFFFFu;
        if (valid) {
            const uint32_t di = draw_of_tri(B.draws, B.n_draws, B.group_draw[group], g);
            const FrameDesc& F = B.frames[B.draws[di].frame];
            base = F.tile_base;
            ntx = F.ntx;
        }
    }
    const uint32_t cap = B.runs_cap;
    // whether a triangle of the group has more than one clip fan
    const uint32_t multi = __ballot(valid && (B.hdr[g] & 7u) > 1u) ? 1u : 0u;
    // The tiles of the group in increasing order (as in for_each_tile); lane k
    // collects the k-th (tile, mask), and every 64 of them take their run slots
    // with one round of atomics.
    const uint32_t tx0 = r0 & 0xFFFF, tx1 = r1 & 0xFFFF, ty1 = r1 >> 16;
    uint32_t cx = tx0, cy = r0 >> 16;
    uint32_t my_tile = 0, k = 0;
    unsigned long long my_mask = 0;
    auto flush = [&]() {
        if (lane < k) {
            const uint32_t pos = B.run_off[my_tile] + atomicAdd(&B.run_cursor[my_tile], 1u);
            if (pos < cap) {
                Run r;
                r.group = group;
                r.multi = multi;
                r.mask = my_mask;
                B.runs[pos] = r;
            }
        }
        k = 0;
    };
    while (true) {
        const uint32_t next = valid ? base + cy * ntx + cx : 0xFFFFFFFFu;
        const uint32_t t = wave_min_u32(next);
        if (t == 0xFFFFFFFFu) break;
        const bool mine = valid && next == t;
        const unsigned long long m = __ballot(mine);
        if (lane == k) {
            my_tile = t;
            my_mask = m;
        }
        if (++k == (uint32_t)WAVE) flush();
        if (mine && ++cx > tx1) {
            cx = tx0;
            if (++cy > ty1) valid = false;
        }
    }
    flush();
}
#endif  // RF_RASTER_TU
// Per-triangle data of a batch slot (64 B).
struct SlotP {
    float dv[2][4];        // d(z, a0, a1, a2)/dx of the upper and lower trapezoid
    uint32_t kind_tag;     // shader kind | (draw - the tile's draw base) << 16
    uint32_t p0, p1, p2;   // flat: packed colour; textured: width, height, row stride (texels)
    uint32_t rec;          // the triangle's first edge record in the batch
    int32_t ysplit;        // first row of the lower trapezoid (absolute)
    uint32_t rxb;          // pixel columns of the binned tiles, first | end << 16 (margin check)
    int32_t ra;            // row of that record, relative to the tile's first row (negative:
                           // a row above the target, folded onto row 0)
};
// Edge record of one (triangle, tile row): the left and right edge state
// after the serial row steps of ScanlineIter::next (raster.rs:84-113).
struct EdgeRec {
    float lx, rx, lz, l0, l1, l2;
};

// Edge records per batch of a tile th rows tall (a batch's first triangle
// has at most th rows in a tile, unless rows above the target fold onto row 0).
__host__ __device__ constexpr int rec_cap(int th) { return th > 32 ? 384 : 128; }
constexpr int FQ_CAP = 512;          // fan queue (ring); one run adds at most 64 * 7 fans
#ifndef RF_CELL_W
#define RF_CELL_W 8
#endif
constexpr int CELL_W = RF_CELL_W;    // pixel columns per ordering cell
#ifndef RF_TILE_GRAB
#define RF_TILE_GRAB 1
#endif
constexpr int TILE_GRAB = RF_TILE_GRAB;  // tiles a wave takes from the shared counter at once
#ifndef RF_TILE_TAIL
#define RF_TILE_TAIL 512
#endif
constexpr int TILE_TAIL = RF_TILE_TAIL;  // the last tiles of a partition are taken one at a time
#ifndef RF_KMAX
#define RF_KMAX 600
#endif
#ifndef RF_JSER
#define RF_JSER 256
#endif
constexpr int KMAX = RF_KMAX;  // row steps taken serially; more use the exact fast-forward
constexpr int JSER = RF_JSER;  // pixel steps taken serially; more use the exact fast-forward
#ifndef RF_REFILL_MIN
#define RF_REFILL_MIN 16
#endif
constexpr int REFILL_MIN = RF_REFILL_MIN;  // free lanes that trigger a refill while others still work

// Tile rasterizer. One wave owns one tile (TILE_H rows x TW columns) at a
// time; the tile's depth and each pixel's colour source live in LDS.
// Triangles (clip fans) are consumed in submission order in batches of up to
// 64:
//   setup  lane = triangle: tri_fill/scan setup (raster.rs:185-302);
//   edges  lane = triangle: walks its rows inside the tile with the serial
//          edge recurrence (ScanlineIter::next, raster.rs:84-113), entering
//          the tile through the exact fast-forward when it starts far above,
//          and writes one span record per row, in triangle-major order; a
//          record's index is its order key;
//   spans  lanes take records in key order and keep one until it is done:
//          the varyings are stepped serially pixel by pixel from the span's
//          start (Scanline::fragments, raster.rs:61-68; the exact
//          fast-forward enters a tile far from the start). In rounds, every
//          record claims the (row, 8-column) cells its remaining pixels touch
//          with an LDS atomic minimum of its key, and a record whose next 8
//          pixels lie in cells it holds applies them: depth test, depth write
//          and colour source (rasterize_fb, target.rs:163-198). A cell held by
//          a record blocks every later one, so each pixel sees its fragments
//          in submission order; the record with the smallest key always
//          proceeds.
// A passing fragment that writes colour leaves in the pixel's record what its
// colour is made of: the colour itself for the flat and Gouraud shaders, the
// draw (relative to the tile's first draw) and the sampled texel's index for
// the textured ones (the samplers' coordinate arithmetic runs here; only the
// texel fetch waits). The colour a pixel ends with is that of its last
// passing fragment, so the fetches run once per written pixel at the end of
// the tile, all pixels in parallel, and colour and depth go straight to the
// target.
constexpr uint16_t TAG_NONE = 0xFFFFu;   // no fragment wrote colour
constexpr uint16_t TAG_COLOR = 0xFFFEu;  // payload is the colour; else tag = draw - base, payload = texel
constexpr uint32_t LINE_TAG = 0x8000u;   // SlotP::kind_tag bit of a line primitive

__device__ __forceinline__ uint32_t texel_pixel(const BatchDesc& B, uint32_t fmt, uint32_t draw, uint32_t texel) {
    const uint32_t t = ((const uint32_t*)B.draws[draw].tex)[texel];
    return pack_pixel(fmt, t & 0xFF, (t >> 8) & 0xFF, (t >> 16) & 0xFF, t >> 24);
}

// n serial additions of d to each of k values, serially for small n and
// through the exact fast-forward otherwise (only the first nlive values).
template <int K>
__device__ __forceinline__ void advance(float (&v)[K], const float (&d)[K], uint32_t n, uint32_t nser, int nlive) {
    if (n <= nser) {
        uint32_t q = 0;
        for (; q + 4 <= n; q += 4) {  // four steps per trip: fewer loop branches
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int i = 0; i < K; ++i) v[i] = v[i] + d[i];
        }
        for (; q < n; ++q) {
#pragma unroll
            for (int i = 0; i < K; ++i) v[i] = v[i] + d[i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (i < nlive) v[i] = ff_add(v[i], d[i], n);
    }
}

__device__ __forceinline__ int32_t wave_incl_max_i32(int32_t v) {
    // v >= -1: order-preserving shift to unsigned
    const uint32_t x = wave_incl_scan((uint32_t)(v + 1), 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; });
    return (int32_t)x - 1;
}

// The depth test of Context::depth_test (ctx.rs:86-89) without branches.
__device__ __forceinline__ bool depth_pass_u(int32_t dt, float nz, float cur) {
    return (dt == DT_NONE) | ((dt == DT_LESS) & (cur < nz)) | ((dt == DT_EQUAL) & (cur == nz)) |
           ((dt == DT_GREATER) & (cur > nz));
}

#ifdef RF_RASTER_TU
// TW: tile width (16, 32, 64); CH: pixels per lane applied together in a span
// round (6 or 8: short spans waste fewer lanes' slots and registers at 6,
// long spans take fewer rounds at 8).
template <int TW, int CH, int TH>
__global__ __launch_bounds__(64, 2) void k_raster(BatchDesc B) {
    constexpr int TILE_H = TH;  // tile rows
    constexpr int REC_CAP = rec_cap(TH);
    constexpr int LD = TILE_H + 1;  // padded column stride in LDS (bank spread)
    constexpr int NPX = TW * LD;
    constexpr int CW = CELL_W < TW ? CELL_W : TW;  // ordering cell width in this tile
    constexpr int NCG = TW / CW;
    // depth, then the payloads, in one array: a chunk's depth reads run up to
    // CH - 1 columns past the last one, into the payloads (values never
    // used), so that they are unconditional and pair up (ds_read2)
    __shared__ __align__(16) float s_px[2 * NPX];
    float* const s_dep = s_px;
    uint32_t* const s_pay = reinterpret_cast<uint32_t*>(s_px + NPX);  // the pixel's colour source: payload
    static_assert((CH - 1) * LD + TILE_H <= NPX, "depth over-reads stay in the array");
    __shared__ __align__(16) uint16_t s_tag[NPX];   //   and tag
    static_assert((NPX * 4) % 16 == 0 && (NPX * 2) % 16 == 0, "16-byte LDS fills");
    __shared__ SlotP s_slot[WAVE];
    __shared__ EdgeRec s_rec[REC_CAP];
    __shared__ uint32_t s_fq[FQ_CAP];  // g << 3 | fan
    __shared__ int8_t s_own[REC_CAP];  // slot whose records start here, or -1
    __shared__ __align__(16) uint32_t s_cell[TILE_H * NCG];

    const uint32_t lane = lane_id();
    for (uint32_t i = lane; i < (uint32_t)(TILE_H * NCG); i += WAVE) s_cell[i] = 0xFFFFFFFFu;
    // Tiles come from TILE_PARTS contiguous partitions with a counter each.
    // Workgroups are dispatched round-robin over the XCDs, so a workgroup
    // starts on partition blockIdx.x % TILE_PARTS: neighbouring tiles (which
    // share triangles) meet in one XCD's L2. A wave takes TILE_GRAB tiles per
    // atomic and moves on to the next partition when its own is done.
    const uint32_t n_busy = *B.busy_count;  // tiles with runs, listed in B.tile_list
    // The tiles without runs, EMPTY_GRAB per atomic: their fused clear is the
    // only write. The first EARLY_CLEAR workgroups start with them, so these
    // stores overlap the busy tiles' work; every wave takes what is left at
    // its end.
    const uint32_t n_empty = B.clear_in_raster ? B.n_tiles - n_busy : 0u;
    auto clear_empties = [&]() {
        uint32_t ecur_fi = 0xFFFFFFFFu;
        FrameDesc E;
        for (;;) {
            uint32_t e0 = 0;
            if (lane == 0) e0 = atomicAdd(&B.tile_counter[TILE_PARTS], (uint32_t)EMPTY_GRAB);
            e0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)e0);
            if (e0 >= n_empty) break;
            const uint32_t ne = min((uint32_t)EMPTY_GRAB, n_empty - e0);
            uint32_t my_tile = 0, my_fi = 0;
            if (lane < ne) {
                my_tile = B.tile_list[n_busy + e0 + lane];
                my_fi = B.tile_frame[my_tile];
            }
            if (RF_CLEAR_ROWS) {
                // ne consecutive full tiles of one tile row of one frame (the
                // common case: the list is in tile order): clear the row
                // segment they make, each store instruction 1 KB of one pixel row
                const uint32_t t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)my_tile);
                const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)my_fi);
                const bool run = __ballot(lane < ne && (my_tile != t0 + lane || my_fi != f0)) == 0ull;
                if (run) {
                    if (f0 != ecur_fi) {
                        E = B.frames[f0];
                        ecur_fi = f0;
                    }
                    const uint32_t lt = t0 - E.tile_base, tx = lt % E.ntx, ty = lt / E.ntx;
                    const uint32_t x0 = tx * TW, y0 = ty * TILE_H;
                    const bool has_depth = E.depth != nullptr;
                    const bool ok = tx + ne <= E.ntx && x0 + ne * TW <= E.width && y0 + TILE_H <= E.height &&
                                    pixfmt_bytes(E.fmt) == 4u && ((uintptr_t)E.color & 15u) == 0 &&
                                    (E.color_stride & 15u) == 0 &&
                                    (!has_depth || (((uintptr_t)E.depth & 15u) == 0 && (E.depth_stride & 3u) == 0));
                    if (ok) {
                        if (E.clear_color || (has_depth && E.clear_depth)) {
                            const uint32_t qpr = ne * (TW / 4);  // 16-byte quads per pixel row
                            const uint32_t cp = E.clear_pixel;
                            const float cz = E.clear_depth_value;
                            for (uint32_t q = lane; q < TILE_H * qpr; q += WAVE) {
                                const uint32_t r = q / qpr, c = (q - r * qpr) * 4;
                                if (E.clear_color)
                                    st16(E.color + (uint64_t)(y0 + r) * E.color_stride + 4ull * (x0 + c), cp, cp, cp,
                                         cp);
                                if (has_depth && E.clear_depth)
                                    st16f(E.depth + (uint64_t)(y0 + r) * E.depth_stride + x0 + c, cz, cz, cz, cz);
                            }
                        }
                        continue;
                    }
                }
            }
            for (uint32_t j = 0; j < ne; ++j) {
                const uint32_t tile = (uint32_t)__builtin_amdgcn_readlane((int)my_tile, (int)j);
                const uint32_t fi = (uint32_t)__builtin_amdgcn_readlane((int)my_fi, (int)j);
                if (fi != ecur_fi) {
                    E = B.frames[fi];
                    ecur_fi = fi;
                }
                clear_tile<TW, TH>(E, tile, lane);
            }
        }
    };
    if (blockIdx.x < (uint32_t)RF_EARLY_CLEAR) clear_empties();
    uint32_t part = blockIdx.x % TILE_PARTS, parts_left = TILE_PARTS;
    uint32_t tcur = 0, tend = 0;
    bool tail_near = false;
    for (;;) {
        while (tcur >= tend && parts_left) {
            const uint32_t pb = (uint32_t)(((uint64_t)n_busy * part) / TILE_PARTS);
            const uint32_t pe = (uint32_t)(((uint64_t)n_busy * (part + 1)) / TILE_PARTS);
            // near a partition's end single tiles keep the tail short
            const uint32_t grab = tail_near ? 1u : (uint32_t)TILE_GRAB;
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(&B.tile_counter[part], grab);
            t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)t0);
            if (t0 < pe - pb) {
                tcur = pb + t0;
                tend = min(tcur + grab, pe);
                tail_near = t0 + (uint32_t)TILE_TAIL >= pe - pb;
            } else {
                part = (part + 1) % TILE_PARTS;
                --parts_left;
                tail_near = true;  // other partitions are in their tails too
            }
        }
        if (tcur >= tend) break;
        const uint32_t tile = B.tile_list[tcur++];
        const uint32_t fi = B.tile_frame[tile];
        const FrameDesc F = B.frames[fi];
        const uint32_t lt = tile - F.tile_base;
        const uint32_t tx = lt % F.ntx, ty = lt / F.ntx;
        const uint32_t x0 = tx * TW, y0 = ty * TILE_H;
        const uint32_t off = B.run_off[tile];
        uint32_t nruns = B.run_off[tile + 1] - off;
        if (off >= B.runs_cap) nruns = 0;
        else if (off + nruns > B.runs_cap) nruns = B.runs_cap - off;
        const bool has_depth = F.depth != nullptr;
        const bool dwrite = has_depth && F.depth_write;
        const int32_t dtest = has_depth ? F.depth_test : (int32_t)DT_NONE;
        const uint32_t ncols = min((uint32_t)TW, F.width - x0);
        const uint32_t nrows = min((uint32_t)TILE_H, F.height - y0);
        const uint32_t npx = ncols * nrows;

        // 16-byte stores: whole tile row of 4-byte pixels, 16-byte aligned rows
        const bool vec = ncols == (uint32_t)TW && pixfmt_bytes(F.fmt) == 4u && ((uintptr_t)F.color & 15u) == 0 &&
                         (F.color_stride & 15u) == 0 && (!has_depth || (((uintptr_t)F.depth & 15u) == 0 &&
                                                                         (F.depth_stride & 3u) == 0));
        if (nruns == 0) {
            clear_tile<TW, TH>(F, tile, lane);
            continue;
        }

        // tile init (WAVE / TILE_H lanes per row, each a share of the columns):
        // depth from the fused clear or the target; no fragment yet
        if (RF_VEC_INIT && (F.clear_depth || !has_depth)) {
            // constant fills of the whole arrays (padding and any part beyond
            // the target included, never written back), 16 bytes a store
            const float cd = F.clear_depth_value;
            if (has_depth)
                for (uint32_t i = lane; i < (uint32_t)NPX / 4; i += WAVE)
                    reinterpret_cast<float4*>(s_dep)[i] = make_float4(cd, cd, cd, cd);
            for (uint32_t i = lane; i < (uint32_t)NPX / 8; i += WAVE)
                reinterpret_cast<uint4*>(s_tag)[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        } else {
            constexpr uint32_t LPR = WAVE / TILE_H;  // lanes per row
            const uint32_t r = lane % TILE_H, part = lane / TILE_H;
            if (r < nrows) {
                const float* drow = has_depth ? F.depth + (uint64_t)(y0 + r) * F.depth_stride + x0 : nullptr;
                for (uint32_t c = part; c < ncols; c += LPR) {
                    if (has_depth) s_dep[c * LD + r] = F.clear_depth ? F.clear_depth_value : drow[c];
                    s_tag[c * LD + r] = TAG_NONE;
                }
            }
        }
        __syncthreads();

        unsigned long long fi_acc = 0, fo_acc = 0;
        uint32_t errs = 0;
        uint32_t fq_head = 0, fq_tail = 0;  // ring indices (wave-uniform)
        uint32_t head_skip = 0;             // rows of the head triangle already taken (wave-uniform)
        uint32_t dbase = 0;                 // draw of the tile's first triangle
        bool dbase_set = false;

        // the tile's runs in group order: lists of up to 64 are loaded one per
        // lane and sorted here; longer ones were sorted by k_binsort
        uint32_t lrun_g = 0, lrun_x = 0;
        unsigned long long lrun_m = 0;
        if (nruns <= (uint32_t)WAVE) {
            if (lane < nruns) {
                const Run r = B.runs[off + lane];
                lrun_g = r.group;
                lrun_m = r.mask;
                lrun_x = r.multi;
            }
            wave_sort_runs(nruns, lrun_g, lrun_m, lrun_x);
        }
        // Batches of up to 64 fans follow one another without a pause: once a
        // batch has handed out its last record, the next one is set up while
        // the items in flight finish (an item keeps in registers all it needs
        // of its slot and record). Order keys run on across the batches.
        uint32_t ri = 0;          // next run to append
        uint32_t next = 0;        // next record of the current batch to hand out
        uint32_t nrec = 0;        // records of the current batch
        uint32_t key_base = 0;    // order key of the current batch's first record
        int32_t carry = -1;       // slot of the last record handed out
        bool have = false;
        uint32_t key = 0, row = 0, ikind = 0, ip0 = 0, ip1 = 0, ip2 = 0, itag = 0;
        float dz = 0, d0 = 0, d1 = 0, d2 = 0, z = 0, c0 = 0, c1 = 0, c2 = 0;
        int32_t xs = 0, jc = 0, j1 = 0;
        for (;;) {
            if (next >= nrec) {
                // append runs until a full batch of fans is queued (or none are left)
                while (ri < nruns && fq_tail - fq_head < (uint32_t)WAVE) {
                    Run run;
                    if (nruns <= (uint32_t)WAVE) {
                        run.group = __shfl(lrun_g, ri, WAVE);
                        run.mask = __shfl(lrun_m, ri, WAVE);
                        run.multi = __shfl(lrun_x, ri, WAVE);
                    } else {
                        run = B.runs[off + ri];
                    }
                    const uint64_t g = (uint64_t)run.group * WAVE + lane;
                    const bool has = (run.mask >> lane) & 1ull;
                    // a triangle in the mask has at least one fan; only groups
                    // with clipped triangles need their headers
                    const uint32_t nf = has ? (RF_RUN_MULTI && !run.multi ? 1u : (B.hdr[g] & 7u)) : 0u;
                    uint32_t tot;
                    const uint32_t pre = wave_excl_scan(nf, &tot);
                    for (uint32_t j = 0; j < nf; ++j) s_fq[(fq_tail + pre + j) % FQ_CAP] = ((uint32_t)g << 3) | j;
                    fq_tail += tot;
                    ++ri;
                    __syncthreads();
                }
                const uint32_t avail = fq_tail - fq_head;
                if (avail == 0) {
                    if (__ballot(have) == 0ull) break;  // the tile is done
                    goto rounds;
                }
                const uint32_t nb = avail < (uint32_t)WAVE ? avail : (uint32_t)WAVE;
                // ---- setup: lane = triangle
                const bool act = lane < nb;
                Slot S;
                uint32_t rows_in = 0, di = 0, kind = 0;
                int32_t ra = 0, rb = 0;
                uint32_t rx0b = 0, rx1b = 0;
                int na = 0;
                // a line (KIND_LINE): one record per pixel of its walk that lies in
                // the tile's columns (wide) or rows (tall), from step lka on
                bool is_line = false, lwide = false;
                uint32_t lS = 0, lka = 0;
                float lbase = 0.0f, lstep = 0.0f, lz = 0.0f, la0 = 0.0f, la1 = 0.0f, la2 = 0.0f;
                if (act) {
                    const uint32_t e = s_fq[(fq_head + lane) % FQ_CAP];
                    const uint64_t gg = e >> 3;
                    const uint32_t j = e & 7u;
                    const ScrTri* st = (j == 0) ? &B.scr[gg] : &B.extra[(B.hdr[gg] >> 3) + j - 1];
                    float va[6], vb[6], vc[6];
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        va[c] = st->v[0][c];
                        vb[c] = st->v[1][c];
                        vc[c] = st->v[2][c];
                    }
                    di = st->draw;
                    kind = st->kind & 0xFFu;
                    is_line = (st->kind & KIND_LINE) != 0u;
                    na = kind_attrs(kind);
                    if (is_line) {
                        lS = f2u(vb[0]);
                        const uint32_t ln = f2u(vb[1]);
                        lbase = vb[2];
                        lstep = vb[3];
                        lwide = f2u(vb[4]) != 0u;
                        lz = va[2];
                        la0 = va[3];
                        la1 = va[4];
                        la2 = va[5];
                        const uint32_t lo = lwide ? x0 : y0, hi = lo + (lwide ? ncols : nrows);
                        lka = lo > lS ? lo - lS : 0u;
                        const uint32_t kend = hi > lS ? min(ln, hi - lS) : 0u;
                        rows_in = kend > lka ? kend - lka : 0u;
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            S.t[h].n = 0;
                            S.t[h].dvz = 0.0f;
                            for (int c = 0; c < 3; ++c) S.t[h].dva[c] = 0.0f;
                        }
                    } else {
                        tri_setup(S, va, vb, vc, na);
                        const uint32_t rr0 = B.rect[2 * gg], rr1 = B.rect[2 * gg + 1];
                        rx0b = (rr0 & 0xFFFF) * TW;
                        rx1b = ((rr1 & 0xFFFF) + 1) * TW;
                        // rows of the triangle in this tile; in the first tile
                        // row, the rows above the target too: `self.y as
                        // usize` folds each onto row 0 as a span of its own
                        // (raster.rs:106), in row order before row 0's own
                        ra = S.ylo - (int32_t)y0;
                        rb = S.yhi - (int32_t)y0;
                        if (y0 > 0 || ra > 0) ra = ra < 0 ? 0 : ra;
                        rb = rb > (int32_t)nrows ? (int32_t)nrows : rb;
                        if (lane == 0) ra += (int32_t)head_skip;  // rows an earlier batch took
                        rows_in = rb > ra ? (uint32_t)(rb - ra) : 0u;
                    }
                }
                // A triangle with more rows here than a batch holds (only
                // with folded rows: otherwise at most TILE_H) goes in parts
                // of REC_CAP rows, as the sole triangle of its batches.
                const bool split = lane == 0 && rows_in > (uint32_t)REC_CAP;
                if (split) rows_in = REC_CAP;
                const bool head_split = __ballot(split) != 0ull;
                uint32_t rec_tot;
                const uint32_t rec_base = wave_excl_scan(rows_in, &rec_tot);
                // take the leading triangles whose records fit (lane 0 always
                // fits: it has at most REC_CAP rows)
                const bool fits = rec_base + rows_in <= (uint32_t)REC_CAP;
                const unsigned long long nf_mask = ~__ballot(act && fits && !head_split);
                const uint32_t nuse = head_split ? 1u
                                    : nf_mask == 0ull ? (uint32_t)WAVE : (uint32_t)(__ffsll((long long)nf_mask) - 1);
                const bool use = lane < nuse;
                const uint32_t nrec_new = __shfl(rec_base + rows_in, nuse - 1, WAVE);
                // draw tags are relative to the tile's first draw (draws only grow
                // along the triangle order); resolve pending texels before the
                // range would overflow
                const uint32_t dfirst = __shfl(di, 0, WAVE), dlast = __shfl(di, nuse - 1, WAVE);
                const bool rebase = dbase_set && dlast - dbase >= (uint32_t)TAG_COLOR;
                if (rebase && __ballot(have) != 0ull) {
                    // items in flight still tag with the old base: let them finish first
                    goto rounds;
                }
                if (!dbase_set) {
                    dbase = dfirst;
                    dbase_set = true;
                } else if (rebase) {
                    for (uint32_t i = lane; i < npx; i += WAVE) {
                        const uint32_t r = i / ncols, c = i - r * ncols;
                        const uint32_t li = c * LD + r;
                        const uint32_t tg = s_tag[li];
                        if (tg < TAG_COLOR) {
                            s_pay[li] = texel_pixel(B, F.fmt, dbase + tg, s_pay[li]);
                            s_tag[li] = TAG_COLOR;
                        }
                    }
                    dbase = dfirst;
                }
                for (uint32_t i = lane; i < nrec_new; i += WAVE) s_own[i] = -1;
                __syncthreads();
                if (use) {
                    SlotP& P = s_slot[lane];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        P.dv[h][0] = S.t[h].dvz;
                        for (int c = 0; c < 3; ++c) P.dv[h][1 + c] = S.t[h].dva[c];
                    }
                    P.kind_tag = kind | (is_line ? LINE_TAG : 0u) | ((di - dbase) << 16);
                    const DrawDesc& D = B.draws[di];
                    if (kind == KIND_FLAT) {
                        const uint32_t fc = D.flat_rgba;
                        P.p0 = pack_pixel(F.fmt, fc & 0xFF, (fc >> 8) & 0xFF, (fc >> 16) & 0xFF, fc >> 24);
                    } else {
                        P.p0 = D.tex_w;
                    }
                    P.p1 = D.tex_h;
                    P.p2 = D.tex_stride;
                    P.rec = rec_base;
                    P.ra = ra;
                    P.ysplit = S.t[0].n ? (int32_t)S.t[0].y0i + (int32_t)S.t[0].n : -0x7FFFFFFF;
                    P.rxb = rx0b | (rx1b << 16);
                    if (rows_in) s_own[rec_base] = (int8_t)lane;
                }
                // ---- edges: lane = triangle, walk its rows in the tile
                if (use && rows_in && is_line) {
                    // the line's pixels: the walked coordinate is the pixel
                    // index, the stepped one goes on serially (raster.rs:138-175)
                    float lv[1] = {lbase};
                    const float ld[1] = {lstep};
                    advance<1>(lv, ld, lka, KMAX, 1);
                    EdgeRec* out = &s_rec[rec_base];
                    for (uint32_t k = 0; k < rows_in; ++k) {
                        const float pix = (float)(lS + lka + k);
                        out[k] = lwide ? EdgeRec{pix, lv[0], lz, la0, la1, la2} : EdgeRec{lv[0], pix, lz, la0, la1, la2};
                        lv[0] = lv[0] + lstep;
                    }
                } else if (use && rows_in) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const Trap& T = S.t[h];
                        if (T.n == 0) continue;
                        const int32_t ty0 = (int32_t)T.y0i;  // absolute first row
                        const int32_t rend = ra + (int32_t)rows_in;  // rows recorded in this batch
                        int32_t a = ty0 > (int32_t)y0 + ra ? ty0 : (int32_t)y0 + ra;
                        int32_t b = ty0 + (int32_t)T.n;
                        if (b > (int32_t)y0 + rend) b = (int32_t)y0 + rend;
                        if (a >= b) continue;
                        float ev[6] = {T.lx, T.rx, T.lz, T.la[0], T.la[1], T.la[2]};
                        const float ed[6] = {T.dlx, T.drx, T.dlz, T.dla[0], T.dla[1], T.dla[2]};
                        advance<6>(ev, ed, (uint32_t)(a - ty0), KMAX, 3 + na);
                        EdgeRec* out = &s_rec[rec_base + (uint32_t)(a - (int32_t)y0 - ra)];
                        int32_t y = a;
                        for (; y + 2 <= b; y += 2, out += 2) {  // two rows per trip
                            out[0] = EdgeRec{ev[0], ev[1], ev[2], ev[3], ev[4], ev[5]};
#pragma unroll
                            for (int i = 0; i < 6; ++i) ev[i] = ev[i] + ed[i];
                            out[1] = EdgeRec{ev[0], ev[1], ev[2], ev[3], ev[4], ev[5]};
#pragma unroll
                            for (int i = 0; i < 6; ++i) ev[i] = ev[i] + ed[i];
                        }
                        if (y < b) *out = EdgeRec{ev[0], ev[1], ev[2], ev[3], ev[4], ev[5]};
                    }
                }
                __syncthreads();

                key_base += nrec;
                nrec = nrec_new;
                next = 0;
                carry = -1;
                if (head_split) {
                    head_skip += REC_CAP;  // the same triangle heads the next batch
                } else {
                    fq_head += nuse;
                    head_skip = 0;
                }
            }
        rounds:
            const unsigned long long need = __ballot(!have);
            if (next < nrec && (uint32_t)__popcll(need) >= (uint32_t)REFILL_MIN) {
                uint32_t cnt;
                const uint32_t rank = wave_excl_scan(have ? 0u : 1u, &cnt);
                const uint32_t k = next + rank;
                const bool valid = !have && k < nrec;
                int32_t f = wave_incl_max_i32(valid ? (int32_t)s_own[k] : -1);
                f = f > carry ? f : carry;
                f = f < 0 ? 0 : f;  // every valid record has a slot at or before it
                carry = __shfl(f, WAVE - 1, WAVE);  // the scan is non-decreasing
                next = next + cnt < nrec ? next + cnt : nrec;
                if (valid) {
                    const EdgeRec R = s_rec[k];
                    const SlotP& P = s_slot[f];
                    if (P.kind_tag & LINE_TAG) {
                        // a line pixel (Scanline x..x + 1 with the first end's
                        // single varying value); only the tile holding it draws it
                        const uint32_t px = cvt_u32(R.lx), py = cvt_u32(R.rx);
                        if (px >= x0 && px < x0 + ncols && py >= y0 && py < y0 + nrows) {
                            fi_acc += 1;
                            ikind = P.kind_tag & 0xFFu;
                            dz = d0 = d1 = d2 = 0.0f;
                            z = R.lz;
                            c0 = R.l0;
                            c1 = R.l1;
                            c2 = R.l2;
                            have = true;
                            key = key_base + k;
                            row = py - y0;
                            xs = (int32_t)px;
                            jc = 0;
                            j1 = 1;
                            ip0 = P.p0;
                            ip1 = P.p1;
                            ip2 = P.p2;
                            itag = P.kind_tag >> 16;
                        }
                    } else {
                    const int32_t r = P.ra + (int32_t)(k - P.rec);
                    const int32_t y = (int32_t)y0 + r;  // absolute row; negative: folded onto row 0
                    const uint32_t ty = y > 0 ? (uint32_t)y : 0u;
                    const int h = y < P.ysplit ? 0 : 1;
                    ikind = P.kind_tag & 0xFFu;
                    dz = P.dv[h][0];
                    d0 = P.dv[h][1];
                    d1 = P.dv[h][2];
                    d2 = P.dv[h][3];
                    // the span of this row (ScanlineIter::next, raster.rs:84-113)
                    const float rx0 = round_up_to_half(R.lx), rx1 = round_up_to_half(R.rx);
                    // v0.lerp(&v0.step(&dv_dx), x0 - v0.x)
                    const float tw_ = rx0 - R.lx;
                    float vv[4] = {R.lz + ((R.lz + dz) - R.lz) * tw_, R.l0 + ((R.l0 + d0) - R.l0) * tw_,
                                   R.l1 + ((R.l1 + d1) - R.l1) * tw_, R.l2 + ((R.l2 + d2) - R.l2) * tw_};
                    // 32-bit saturating casts: any value the 64-bit `as usize`
                    // keeps apart from these is beyond the target and fails
                    // the bounds check the same way
                    const uint32_t cnt_f = cvt_u32(rx1 - rx0);
                    const uint32_t xs32 = cvt_u32(rx0);
                    uint32_t xe = cvt_u32(rx1);
                    if (xe < xs32) xe = xs32;
                    int32_t jj0 = 0, jj1 = 0;
                    if (xe > F.width || ty >= F.height) {
                        errs |= ERRB_BOUNDS;
                    } else {
                        if (xe > xs32 && (xs32 < (P.rxb & 0xFFFFu) || xe > (P.rxb >> 16))) errs |= ERRB_MARGIN;
                        // frags.i: the part of [xs, xe) inside this tile
                        const uint32_t pa = xs32 > x0 ? xs32 : x0;
                        const uint32_t pb = xe < x0 + TW ? xe : x0 + TW;
                        if (pb > pa) fi_acc += pb - pa;
                        const uint32_t nfrag = cnt_f < xe - xs32 ? cnt_f : xe - xs32;
                        xs = (int32_t)xs32;
                        jj0 = (int32_t)x0 - xs > 0 ? (int32_t)x0 - xs : 0;
                        jj1 = (int32_t)x0 + TW - xs;
                        if (jj1 > (int32_t)nfrag) jj1 = (int32_t)nfrag;
                    }
                    if (jj0 < jj1) {
                        const float dd[4] = {dz, d0, d1, d2};
                        // the fragment iterator stepped jj0 times (Scanline::fragments)
                        advance<4>(vv, dd, (uint32_t)jj0, JSER, 1 + kind_attrs(ikind));
                        z = vv[0];
                        c0 = vv[1];
                        c1 = vv[2];
                        c2 = vv[3];
                        have = true;
                        key = key_base + k;
                        row = ty - y0;
                        jc = jj0;
                        j1 = jj1;
                        ip0 = P.p0;
                        ip1 = P.p1;
                        ip2 = P.p2;
                        itag = P.kind_tag >> 16;
                    }
                    }
                }
            }
            const unsigned long long busy = __ballot(have);
            if (busy == 0ull) continue;
            // claim every cell of the remaining pixels
            int32_t cl0 = 0, cl1 = -1, cw1 = -1;
            // a lane without a record reads and frees row 0's cells: harmless
            const uint32_t crow = have ? row : 0u;
            if (have) {
                cl0 = (xs + jc - (int32_t)x0) / CW;
                cl1 = (xs + j1 - 1 - (int32_t)x0) / CW;
                const int32_t je = jc + CH < j1 ? jc + CH : j1;
                cw1 = (xs + je - 1 - (int32_t)x0) / CW;
#pragma unroll
                for (int32_t c = 0; c < NCG; ++c)
                    if (c >= cl0 && c <= cl1) atomicMin(&s_cell[row * NCG + c], key);
            }
            // the row's cells in one (vector) read, after every lane's claims
            uint32_t cells[NCG];
            if constexpr (NCG % 4 == 0) {
#pragma unroll
                for (int32_t c = 0; c < NCG; c += 4) {
                    const uint4 q = *reinterpret_cast<const uint4*>(&s_cell[crow * NCG + c]);
                    cells[c] = q.x;
                    cells[c + 1] = q.y;
                    cells[c + 2] = q.z;
                    cells[c + 3] = q.w;
                }
            } else {
#pragma unroll
                for (int32_t c = 0; c < NCG; ++c) cells[c] = s_cell[crow * NCG + c];
            }
            bool win = have;
#pragma unroll
            for (int32_t c = 0; c < NCG; ++c)
                if (c >= cl0 && c <= cw1) win = win && cells[c] == key;
            if (win) {
                const int32_t rem = j1 - jc;
                const uint32_t li = (uint32_t)(xs + jc - (int32_t)x0) * LD + row;
                float zz[CH], v0[CH], v1[CH], v2[CH], cur[CH];
                bool ps[CH];
#pragma unroll
                for (int p = 0; p < CH; ++p) {
                    zz[p] = z;
                    v0[p] = c0;
                    v1[p] = c1;
                    v2[p] = c2;
                    z = z + dz;
                    c0 = c0 + d0;
                    c1 = c1 + d1;
                    c2 = c2 + d2;
                }
                // the pixels of one chunk are distinct: read all depths first
                if (has_depth) {
#pragma unroll
                    for (int p = 0; p < CH; ++p) cur[p] = s_px[li + p * LD];  // p >= rem: unused
                } else {
#pragma unroll
                    for (int p = 0; p < CH; ++p) cur[p] = 0.0f;
                }
                uint32_t npass = 0;
#pragma unroll
                for (int p = 0; p < CH; ++p) {
                    ps[p] = (p < rem) & depth_pass_u(dtest, zz[p], cur[p]);
                    npass += ps[p] ? 1u : 0u;
                }
                if (ikind == KIND_TEX_ONCE) {
                    // the shader runs for every passing fragment, colour write or not
                    const float tw = (float)ip0, th = (float)ip1;
#pragma unroll
                    for (int p = 0; p < CH; ++p)
                        if (ps[p] && (cvt_u32(tw * (v0[p] / zz[p])) >= ip0 || cvt_u32(th * (v1[p] / zz[p])) >= ip1))
                            errs |= ERRB_TEXEL;
                }
                if (F.color_write) {
                    fo_acc += npass;
                    // colour sources of the chunk's pixels (the shader kind is
                    // decided once per chunk)
                    uint32_t pay[CH];
                    uint32_t tg;
                    if (ikind == KIND_FLAT) {
                        tg = TAG_COLOR;
#pragma unroll
                        for (int p = 0; p < CH; ++p) pay[p] = ip0;
                    } else if (ikind == KIND_GOURAUD) {  // Color3f::to_color4 (color.rs:253-263)
                        tg = TAG_COLOR;
#pragma unroll
                        for (int p = 0; p < CH; ++p)
                            pay[p] = pack_pixel(F.fmt, cvt_u8(256.0f * v0[p]), cvt_u8(256.0f * v1[p]),
                                                cvt_u8(256.0f * v2[p]), 0xFF);
                    } else {
                        tg = itag;
                        const float tw = (float)ip0, th = (float)ip1;
                        if (ikind == KIND_TEX_CLAMP) {  // SamplerClamp (tex.rs:280-303)
                            const float mu = tw - 1.0f, mv = th - 1.0f;
#pragma unroll
                            for (int p = 0; p < CH; ++p) {
                                // Scanline::fragments z_div (raster.rs:61-68, vec.rs:666-676)
                                float su = (v0[p] / zz[p]) * tw, sv = (v1[p] / zz[p]) * th;
                                su = su < 0.0f ? 0.0f : su;
                                su = su > mu ? mu : su;
                                sv = sv < 0.0f ? 0.0f : sv;
                                sv = sv > mv ? mv : sv;
                                pay[p] = cvt_u32(floorf(sv)) * ip2 + cvt_u32(floorf(su));
                            }
                        } else if (ikind == KIND_TEX_REPEAT_POT) {  // SamplerRepeatPot (tex.rs:240-266)
#pragma unroll
                            for (int p = 0; p < CH; ++p) {
                                const uint32_t ui = (uint32_t)cvt_i32(floorf(tw * (v0[p] / zz[p]))) & (ip0 - 1);
                                const uint32_t vi = (uint32_t)cvt_i32(floorf(th * (v1[p] / zz[p]))) & (ip1 - 1);
                                pay[p] = vi * ip2 + ui;
                            }
                        } else {  // SamplerOnce (tex.rs:343-357): `as u32`, no wrap
#pragma unroll
                            for (int p = 0; p < CH; ++p) {
                                const uint32_t ui = cvt_u32(tw * (v0[p] / zz[p]));
                                const uint32_t vi = cvt_u32(th * (v1[p] / zz[p]));
                                // outside the texture the reference panics (checked
                                // below); never fetch there
                                pay[p] = (ui < ip0 && vi < ip1) ? vi * ip2 + ui : 0u;
                            }
                        }
                    }
#pragma unroll
                    for (int p = 0; p < CH; ++p) {
                        if (ps[p]) {
                            s_tag[li + p * LD] = (uint16_t)tg;
                            s_pay[li + p * LD] = pay[p];
                        }
                    }
                }
                if (dwrite) {
#pragma unroll
                    for (int p = 0; p < CH; ++p)
                        if (ps[p]) s_dep[li + p * LD] = zz[p];
                }
                jc += CH;
                if (jc >= j1) have = false;
            }
            // release the claims: every lane frees its whole row (one vector
            // write). Every claimed cell lies in a claimant's row, so all
            // cells are free again, as each round needs them at its start.
#pragma unroll
            for (int32_t c = 0; c < NCG; ++c) s_cell[crow * NCG + c] = 0xFFFFFFFFu;
                }
        __syncthreads();

        // write-back: fetch the texels, colour and depth to the target
        if (vec && (F.clear_color || F.color_write)) {
            // four pixels per lane; every load of a quad is issued before use
            constexpr uint32_t QR = TW / 4;
            for (uint32_t q = lane; q < nrows * QR; q += WAVE) {
                const uint32_t r = q / QR, c = (q - r * QR) * 4;
                uint8_t* crow = F.color + (uint64_t)(y0 + r) * F.color_stride + 4ull * (x0 + c);
                uint32_t tg[4], py[4], px[4];
                const uint32_t* tp[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    tg[k] = s_tag[(c + k) * LD + r];
                    py[k] = s_pay[(c + k) * LD + r];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    tp[k] = tg[k] < TAG_COLOR ? (const uint32_t*)B.draws[dbase + tg[k]].tex : nullptr;
                uint4 old = make_uint4(0, 0, 0, 0);
                const bool keep = !F.clear_color && (tg[0] == TAG_NONE || tg[1] == TAG_NONE || tg[2] == TAG_NONE ||
                                                     tg[3] == TAG_NONE);
                if (keep) old = *(const uint4*)crow;
                const uint32_t oldk[4] = {old.x, old.y, old.z, old.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (tg[k] == TAG_COLOR) {
                        px[k] = py[k];
                    } else if (tg[k] != TAG_NONE) {
                        const uint32_t t = tp[k][py[k]];
                        px[k] = pack_pixel(F.fmt, t & 0xFF, (t >> 8) & 0xFF, (t >> 16) & 0xFF, t >> 24);
                    } else {
                        px[k] = F.clear_color ? F.clear_pixel : oldk[k];
                    }
                }
                st16(crow, px[0], px[1], px[2], px[3]);
                if (has_depth)
                    st16f(F.depth + (uint64_t)(y0 + r) * F.depth_stride + x0 + c, s_dep[c * LD + r],
                          s_dep[(c + 1) * LD + r], s_dep[(c + 2) * LD + r], s_dep[(c + 3) * LD + r]);
            }
        } else {
            for (uint32_t i = lane; i < npx; i += WAVE) {
                const uint32_t r = i / ncols, c = i - r * ncols;
                const uint32_t li = c * LD + r;
                const uint32_t tg = s_tag[li];
                if (tg == TAG_COLOR) {
                    store_pixel(F, x0 + c, y0 + r, s_pay[li]);
                } else if (tg != TAG_NONE) {
                    store_pixel(F, x0 + c, y0 + r, texel_pixel(B, F.fmt, dbase + tg, s_pay[li]));
                } else if (F.clear_color) {
                    store_pixel(F, x0 + c, y0 + r, F.clear_pixel);
                }
                if (has_depth) F.depth[(uint64_t)(y0 + r) * F.depth_stride + x0 + c] = s_dep[li];
            }
        }
        const unsigned long long fi_s = wave_sum_u64(fi_acc), fo_s = wave_sum_u64(fo_acc);
        if (lane == 0) {
            if (fi_s) atomicAdd(&B.stats[fi].frags_i, fi_s);
            if (fo_s) atomicAdd(&B.stats[fi].frags_o, fo_s);
        }
        uint32_t e = errs;
        for (int o = 32; o > 0; o >>= 1) e |= __shfl_xor(e, o, WAVE);
        if (lane == 0 && e) atomicOr(B.err, e);
        __syncthreads();
    }

    // Then the tiles without runs (waves that finish the busy tiles early
    // fill the tail with them).
    clear_empties();
}

// The rasterizer's instances: tile width {16, 32, 64} x span chunk {6, 8}
// x tile height {16, 32, 64}.
using RasterFn = void (*)(BatchDesc);
#define RF_K3(TW, CH) {k_raster<TW, CH, 16>, k_raster<TW, CH, 32>, k_raster<TW, CH, 64>}
static const RasterFn RASTER_FNS[3][2][3] = {
    {RF_K3(16, 6), RF_K3(16, 8)}, {RF_K3(32, 6), RF_K3(32, 8)}, {RF_K3(64, 6), RF_K3(64, 8)}};
#undef RF_K3
static int raster_index(uint32_t tw, uint32_t ch, uint32_t th, int* i, int* j, int* k) {
    *i = tw == 16 ? 0 : (tw == 32 ? 1 : 2);
    *j = ch == 6 ? 0 : 1;
    *k = th == 16 ? 0 : (th == 64 ? 2 : 1);
    return 0;
}

hipError_t launch_raster(const BatchDesc& B, hipStream_t s, uint32_t raster_grid) {
    int i, j, k;
    raster_index(B.tile_w, B.chunk, B.tile_h, &i, &j, &k);
    hipLaunchKernelGGL(RASTER_FNS[i][j][k], dim3(raster_grid), dim3(64), 0, s, B);
    return hipGetLastError();
}

uint32_t raster_wg_per_cu(uint32_t tile_w, uint32_t chunk, uint32_t tile_h) {
    // resident one-wave workgroups per CU of each instance (registers and
    // LDS decide), asked of the runtime once
    static int occ[3][2][3] = {};
    int i, j, k;
    raster_index(tile_w, chunk, tile_h, &i, &j, &k);
    int& n = occ[i][j][k];
    if (n == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, RASTER_FNS[i][j][k], 64, 0) != hipSuccess || v <= 0)
            v = 4;
        n = v;
    }
    return (uint32_t)n;
}
#endif  // RF_RASTER_TU

#ifndef RF_RASTER_TU
// ---------------------------------------------------------------------------
// Depth sort (render.rs:180-183, 209-219), before to_screen and face culling:
//   k_fans        one thread per triangle: assembly and clipping as k_setup
//                 does, each resulting clip-space triangle (a clip fan, in
//                 the order clip_tris emits them) stored with its key:
//                 Render::depth, (a.z + b.z + c.z) / 3 (prim.rs:21-23), as a
//                 total_cmp order key, then the triangle and fan index, so
//                 equal depths keep submission order;
//   radix sort of (key, fan) pairs; with several sorted draws, a second,
//                 stable sort by draw, so each draw's fans are contiguous and
//                 in order (each render() call sorts its own primitives);
//   k_fan_gather  fan i of the sorted order becomes vertex slots
//                 3i'..3i'+2 of its draw, i' its rank in the draw (clip
//                 position, attributes, outcode 0, screen vertex): the main
//                 pipeline then draws them as unclipped triangles.
// The sorted draws form a batch of their own (their original meshes); the
// gather writes into the main batch, whose descriptors of these draws hold
// the fan counts the planning run found.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void put_fan(const BatchDesc& B, const DrawDesc* D, uint32_t di, const CV& a,
                                        const CV& b, const CV& c, uint64_t g, uint32_t j) {
    const uint32_t f = atomicAdd(B.fan_count, 1u);
    if (B.fan_cnt) atomicAdd(&B.fan_cnt[di], 1u);
    if (f >= B.fan_cap) {
        atomicOr(B.err, ERRB_EXTRA);
        return;
    }
    const float z = ((a.p[2] + b.p[2]) + c.p[2]) / 3.0f;
    uint32_t k = (uint32_t)total_key(z) ^ 0x80000000u;  // ascending total order as unsigned
    if (D->sort == 2) k = ~k;                           // back to front
    B.fan_keys[f] = ((unsigned long long)k << 32) | (g << 3) | j;
    B.fan_vals[f] = f;
    if (B.fan_draw) B.fan_draw[f] = di;
    float* o = B.fans + (uint64_t)f * FAN_FLOATS;
    const CV* v[3] = {&a, &b, &c};
    for (int q = 0; q < 3; ++q) {
        for (int e = 0; e < 4; ++e) o[7 * q + e] = v[q]->p[e];
        for (int e = 0; e < 3; ++e) o[7 * q + 4 + e] = v[q]->a[e];
    }
}

__global__ __launch_bounds__(256) void k_fans(BatchDesc B) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= B.n_tris) return;
    const uint32_t di = draw_of_tri(B.draws, B.n_draws, B.group_draw[g / WAVE], g);
    if (B.any_cull && !B.draw_vis[di]) return;  // a draw the bounding-box test hid
    const DrawDesc* D = &B.draws[di];
    const uint64_t t = g - D->tri_base;
    const uint32_t i0 = D->tris[3 * t], i1 = D->tris[3 * t + 1], i2 = D->tris[3 * t + 2];
    if (i0 >= D->n_verts || i1 >= D->n_verts || i2 >= D->n_verts) {
        atomicOr(B.err, ERRB_INDEX);
        return;
    }
    CV poly[CLIP_MAX], tmp[CLIP_MAX];
    const uint64_t ix[3] = {D->vert_base + i0, D->vert_base + i1, D->vert_base + i2};
    for (int k = 0; k < 3; ++k) {
        const float4 p = B.cv[ix[k]], a = B.ca[ix[k]];
        const float pp[4] = {p.x, p.y, p.z, p.w}, aa[3] = {a.x, a.y, a.z};
        cv_make(poly[k], pp, aa);
    }
    const uint32_t all_out = poly[0].oc & poly[1].oc & poly[2].oc, any_out = poly[0].oc | poly[1].oc | poly[2].oc;
    if (all_out != 0) return;  // Hidden
    if (any_out == 0) {        // Visible
        put_fan(B, D, di, poly[0], poly[1], poly[2], g, 0);
        return;
    }
    const int n = clip_polygon(poly, 3, tmp);
    if (n < 0) atomicOr(B.err, ERRB_CLIP);
    for (int k = 1; k + 1 < n; ++k) put_fan(B, D, di, poly[0], poly[k], poly[k + 1], g, (uint32_t)(k - 1));
}

// The second sort's keys: the draw of each fan in the first sort's order.
__global__ __launch_bounds__(256) void k_fan_draw_keys(const uint32_t* order, const uint32_t* fan_draw, uint32_t n,
                                                       uint32_t* keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = fan_draw[order[i]];
}

// Fan i of the final order (its draw d = draws_of[i], 0 if null) becomes
// vertex slots of main draw map[2d + 1] at rank i - map[2d] (the draw's first
// position) and index triples idx[3i..3i+2] (relative to that draw).
__global__ __launch_bounds__(256) void k_fan_gather(BatchDesc P, BatchDesc M, const uint32_t* order,
                                                    const uint32_t* draws_of, const uint32_t* map, uint32_t n_fans,
                                                    uint32_t* idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_fans) return;
    const uint32_t d = draws_of ? draws_of[i] : 0u;
    const uint32_t rank = i - map[2 * d];
    const DrawDesc* D = &M.draws[map[2 * d + 1]];
    const float* src = P.fans + (uint64_t)order[i] * FAN_FLOATS;
    const bool persp = kind_persp(D->kind);
    for (int q = 0; q < 3; ++q) {
        const uint64_t slot = D->vert_base + 3ull * rank + q;
        CV c;
        for (int e = 0; e < 4; ++e) c.p[e] = src[7 * q + e];
        for (int e = 0; e < 3; ++e) c.a[e] = src[7 * q + 4 + e];
        M.cv[slot] = make_float4(c.p[0], c.p[1], c.p[2], c.p[3]);
        M.ca[slot] = make_float4(c.a[0], c.a[1], c.a[2], u2f(0u));  // inside: clipped already
        to_screen(c, D->vp, persp, M.sv + 6 * slot);
        idx[3ull * i + q] = 3u * rank + (uint32_t)q;
    }
}

// ---------------------------------------------------------------------------
// k_cull: one thread per draw. A draw that asks for it is skipped when its
// mesh's bounding box is Hidden: the 8 corners (BBox::verts, scene.rs:59-66)
// through the draw's ProjMat3 (scene.rs:81-87) and view_frustum::status
// (clip.rs:245-267) — all corners outside one same plane. What the skipped
// render() call would have counted is added to the frame's hidden totals.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cull(BatchDesc B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n_draws) return;
    const DrawDesc* D = &B.draws[i];
    uint32_t vis = 1;
    if (D->cull) {
        const float* m = D->mvp;
        uint32_t all_out = ~0u;
        for (int c = 0; c < 8; ++c) {
            const float px = D->bbox[(c & 4) ? 3 : 0], py = D->bbox[(c & 2) ? 4 : 1], pz = D->bbox[(c & 1) ? 5 : 2];
            const float x = dot4(m + 0, px, py, pz, 1.0f);
            const float y = dot4(m + 4, px, py, pz, 1.0f);
            const float z = dot4(m + 8, px, py, pz, 1.0f);
            const float w = dot4(m + 12, px, py, pz, 1.0f);
            all_out &= outcode(x, y, z, w);
        }
        if (all_out != 0) {
            vis = 0;
            FrameStats& st = B.stats[D->frame];
            atomicAdd(&st.hidden_draws, 1ull);
            // a depth-sorted draw holds its fans: its render() call's own counts
            atomicAdd(&st.hidden_prims, (unsigned long long)(D->pre ? D->in_tris : D->n_tris));
            atomicAdd(&st.hidden_verts, (unsigned long long)(D->pre ? D->in_verts : D->n_verts));
        }
    }
    B.draw_vis[i] = vis;
}

// Zeroes per-batch counters (a kernel, so it can live inside a captured graph).
__global__ void k_reset(BatchDesc B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B.n_tiles) B.run_count[i] = 0;
    // the draw of each group's first triangle and of each block's first vertex
    if (i < B.n_groups) B.group_draw[i] = find_draw_tri(B.draws, B.n_draws, (uint64_t)i * WAVE);
    if ((uint64_t)i * 256 < B.n_verts) B.vblock_draw[i] = find_draw_vert(B.draws, B.n_draws, (uint64_t)i * 256);
    if (i < B.n_frames * PRIM_PARTS) B.prims_part[(uint64_t)i * PRIM_PART_STRIDE] = 0;
    if (i < B.n_frames) {
        B.stats[i].prims_o = 0;
        B.stats[i].verts_o = 0;
        B.stats[i].frags_i = 0;
        B.stats[i].frags_o = 0;
        B.stats[i].hidden_draws = 0;
        B.stats[i].hidden_prims = 0;
        B.stats[i].hidden_verts = 0;
    }
    if (i == 0) {
        *B.extra_count = 0;
        *B.err = 0;
        if (B.fan_count) *B.fan_count = 0;
        for (uint32_t k = 0; k < TILE_PARTS; ++k) B.tile_counter[k] = 0;
    }
    if (B.fan_cnt && i < B.n_draws) B.fan_cnt[i] = 0;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static hipError_t launch_impl(const BatchDesc& B0, hipStream_t s, uint32_t raster_grid, hipEvent_t* ev) {
    BatchDesc B = B0;
    // the rasterizer's waves clear the tiles without runs at their end (faster,
    // measured, than k_bin's extra blocks doing it beside the binning)
    B.clear_in_raster = RF_CLEAR_IN_RASTER;
    uint32_t reset_n = B.n_tiles > B.n_frames * PRIM_PARTS ? B.n_tiles : B.n_frames * PRIM_PARTS;
    reset_n = reset_n > B.n_groups ? reset_n : B.n_groups;
    const uint32_t n_vblocks = (uint32_t)((B.n_verts + 255) / 256);
    reset_n = reset_n > n_vblocks ? reset_n : n_vblocks;
    auto mark = [&](int i) {
        if (ev) (void)hipEventRecord(ev[i], s);
    };
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t r) {
        if (r != hipSuccess && e == hipSuccess) e = r;
    };
    mark(0);
    hipLaunchKernelGGL(k_reset, dim3((reset_n + 255) / 256 + 1), dim3(256), 0, s, B);
    if (B.any_cull) hipLaunchKernelGGL(k_cull, dim3((B.n_draws + 255) / 256), dim3(256), 0, s, B);
    if (B.n_verts)
        hipLaunchKernelGGL(k_vertex, dim3((uint32_t)((B.n_verts + 255) / 256)), dim3(256), 0, s, B);
    mark(1);
    if (B.n_groups) hipLaunchKernelGGL(k_setup, dim3((B.n_groups + 3) / 4), dim3(256), 0, s, B);
    mark(2);
    const uint32_t nb = (B.n_tiles + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_local, dim3(nb ? nb : 1), dim3(1024), 0, s, B);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, B);
    hipLaunchKernelGGL(k_scan_add, dim3(nb ? nb : 1), dim3(1024), 0, s, B);
    mark(3);
    {
        const uint32_t bin_blocks = (B.n_groups + 3) / 4;
        const uint32_t clear_blocks = B.clear_in_raster ? 0u : (B.n_tiles + 4 * EMPTY_GRAB - 1) / (4 * EMPTY_GRAB);
        if (bin_blocks + clear_blocks)
            hipLaunchKernelGGL(k_bin, dim3(bin_blocks + clear_blocks), dim3(256), 0, s, B);
    }
    mark(4);
    hipLaunchKernelGGL(k_binsort, dim3((B.n_tiles + sort_tpw(B) - 1) / sort_tpw(B)), dim3(WAVE), 0, s, B);
    mark(5);
    chk(hipGetLastError());
    chk(launch_raster(B, s, raster_grid));
    mark(6);
    chk(hipGetLastError());
    return e;
}
hipError_t launch_fans(const BatchDesc& B, hipStream_t s) {
    if (B.n_tris && (!B.fans || !B.fan_keys || !B.fan_vals || !B.fan_count || !B.fan_cap)) return hipErrorInvalidValue;
    uint32_t reset_n = B.n_tiles > B.n_frames * PRIM_PARTS ? B.n_tiles : B.n_frames * PRIM_PARTS;
    reset_n = reset_n > B.n_groups ? reset_n : B.n_groups;
    reset_n = reset_n > B.n_draws ? reset_n : B.n_draws;
    const uint32_t n_vblocks = (uint32_t)((B.n_verts + 255) / 256);
    reset_n = reset_n > n_vblocks ? reset_n : n_vblocks;
    hipLaunchKernelGGL(k_reset, dim3((reset_n + 255) / 256 + 1), dim3(256), 0, s, B);
    if (B.any_cull) hipLaunchKernelGGL(k_cull, dim3((B.n_draws + 255) / 256), dim3(256), 0, s, B);
    if (B.n_verts) hipLaunchKernelGGL(k_vertex, dim3(n_vblocks), dim3(256), 0, s, B);
    if (B.n_tris) hipLaunchKernelGGL(k_fans, dim3((uint32_t)((B.n_tris + 255) / 256)), dim3(256), 0, s, B);
    return hipGetLastError();
}
hipError_t launch_fan_draw_keys(const uint32_t* order, const uint32_t* fan_draw, uint32_t n, uint32_t* keys,
                                hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fan_draw_keys, dim3((n + 255) / 256), dim3(256), 0, s, order, fan_draw, n, keys);
    return hipGetLastError();
}
hipError_t launch_fan_gather(const BatchDesc& P, const BatchDesc& M, const uint32_t* order, const uint32_t* draws_of,
                             const uint32_t* map, uint32_t n_fans, uint32_t* idx, hipStream_t s) {
    if (n_fans && (!P.fans || !order || !map || !idx || !M.cv || !M.ca || !M.sv)) return hipErrorInvalidValue;
    if (n_fans)
        hipLaunchKernelGGL(k_fan_gather, dim3((n_fans + 255) / 256), dim3(256), 0, s, P, M, order, draws_of, map,
                           n_fans, idx);
    return hipGetLastError();
}
hipError_t launch_pipeline(const BatchDesc& B, hipStream_t s, uint32_t raster_grid) {
    return launch_impl(B, s, raster_grid, nullptr);
}
hipError_t launch_pipeline_timed(const BatchDesc& B, hipStream_t s, uint32_t raster_grid, hipEvent_t* ev) {
    return launch_impl(B, s, raster_grid, ev);
}
#endif  // RF_RASTER_TU

}  // namespace rf
