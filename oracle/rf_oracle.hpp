// Oracle: single-threaded CPU restatement of retrofire's triangle rendering
// path, `retrofire_core::render::render` (core/src/render.rs:134-207).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
// against; it is never linked into the product library. Only tests/, the
// smoke check and the CPU-baseline leg of bench.py may load it.
//
// Every function follows the cited reference code operation for operation,
// in the same order, so that with -ffp-contract=off and no fast-math the
// float results are those the Rust code computes (Rust never contracts a*b+c
// into an FMA and its float ops are IEEE-754 binary32 round-to-nearest-even).
//
// Pinning (what checks this restatement independently of the HIP path):
//   - core/tests/rendering.rs:19-60 golden image textured_quad.ppm (byte-exact)
//   - core/examples/hello_tri.rs:47-53 centre pixel, and core/triangle.ppm
//     coverage (51,200 px) for the `fp` variant
//   - raster.rs:326-368 shared-edge grid, raster.rs:371-401 gradient grid,
//     raster.rs:404-436 fragment KATs
//   - clip.rs:425-664 exact clip outputs, clip.rs:667-719 exhaustive histogram
//   - tex.rs:381-418 sampler KATs (SamplerOnce included), vary.rs:164-191
//     stepping KATs
//   - pixfmt.rs:143-213 pixel format KATs (u32, byte and u16 forms)
//   - scene.rs:144-153 BBox extend; closed forms and brute force for
//     BBox::visibility
//   - clip.rs:455-486 edge KATs through the edge-list clipper; brute force
//     for edge clipping and raster::line geometry; the first-end varying rule
//   - depth sort: back to front without depth test equals the Less test
//     (screen-parallel layers); the order of equal depths is parity unpinned
//     (the reference leaves it to Rust's sort_unstable_by)
//   - rand.rs doctests (Xorshift64, Uniform<f32>)
//   - brute-force and closed-form checks in tests/test_oracle_kats.py
// Parity of bit patterns on arbitrary triangles beyond these pins rests on
// following the cited lines (the Rust crate cannot be built here).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace oracle {

// ---------------------------------------------------------------------------
// Rust `as` casts from f32: truncate toward zero, saturate, NaN -> 0.
// ---------------------------------------------------------------------------
inline uint32_t as_u32(float x) {
    if (!(x > 0.0f)) return 0;  // NaN, negatives, zeros
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
inline uint64_t as_usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)x;
}
inline int32_t as_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
inline uint8_t as_u8(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint8_t)x;
}

// `f32::total_cmp` ordering key (core::f32::total_cmp).
inline int32_t total_key(float x) {
    int32_t b;
    std::memcpy(&b, &x, 4);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

// raster.rs:305-307, with std floor (math/float.rs:139).
inline float round_up_to_half(float x) { return std::floor(x + 0.5f) + 0.5f; }

// ---------------------------------------------------------------------------
// Varyings. `Vy<N>` is the tuple `(ScreenPt, V)` of raster.rs:58 when used in
// the raster stage, or `(ProjVec3, V)` in the clip stage: every Vary/Lerp op
// on it is componentwise (vary.rs:106-133, space.rs:193-216, math.rs:159-210,
// vec.rs:624-664, point.rs:259-277, color.rs:593-626).
// ---------------------------------------------------------------------------
template <int N>
struct Attr {
    float c[N > 0 ? N : 1] = {};
};

// a + d  (Affine::add / Vary::step)
template <int N>
inline Attr<N> a_add(const Attr<N>& a, const Attr<N>& d) {
    Attr<N> r;
    for (int i = 0; i < N; ++i) r.c[i] = a.c[i] + d.c[i];
    return r;
}
// (b - a) * t (Affine::sub then Linear::mul)
template <int N>
inline Attr<N> a_dvdt(const Attr<N>& a, const Attr<N>& b, float t) {
    Attr<N> r;
    for (int i = 0; i < N; ++i) r.c[i] = (b.c[i] - a.c[i]) * t;
    return r;
}
// a + (b - a) * t (Lerp blanket impl, math.rs:196-198)
template <int N>
inline Attr<N> a_lerp(const Attr<N>& a, const Attr<N>& b, float t) {
    Attr<N> r;
    for (int i = 0; i < N; ++i) r.c[i] = a.c[i] + (b.c[i] - a.c[i]) * t;
    return r;
}

// (ScreenPt, V) varyings of the raster stage.
template <int N>
struct Vy {
    float x, y, z;
    Attr<N> v;
};
template <int N>
inline Vy<N> vy_step(const Vy<N>& a, const Vy<N>& d) {
    return Vy<N>{a.x + d.x, a.y + d.y, a.z + d.z, a_add<N>(a.v, d.v)};
}
template <int N>
inline Vy<N> vy_dvdt(const Vy<N>& a, const Vy<N>& b, float r) {
    return Vy<N>{(b.x - a.x) * r, (b.y - a.y) * r, (b.z - a.z) * r,
                 a_dvdt<N>(a.v, b.v, r)};
}
template <int N>
inline Vy<N> vy_lerp(const Vy<N>& a, const Vy<N>& b, float t) {
    return Vy<N>{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t,
                 a.z + (b.z - a.z) * t, a_lerp<N>(a.v, b.v, t)};
}

// `dot` is a left fold starting from +0.0 (vec.rs:231-238).
inline float dot4(const float* a, const float* b) {
    float acc = 0.0f;
    acc = acc + a[0] * b[0];
    acc = acc + a[1] * b[1];
    acc = acc + a[2] * b[2];
    acc = acc + a[3] * b[3];
    return acc;
}

// ---------------------------------------------------------------------------
// Clipping (clip.rs)
// ---------------------------------------------------------------------------
template <int N>
struct ClipVert {
    float pos[4];
    Attr<N> attrib;
    uint8_t outcode;
};

// view_frustum::PLANES (clip.rs:215-222); ClipPlane::new stores [x,y,z,-off].
static const float PLANES[6][4] = {
    {0.0f, 0.0f, -1.0f, -1.0f},  // near   0x01
    {0.0f, 0.0f, 1.0f, -1.0f},   // far    0x02
    {-1.0f, 0.0f, 0.0f, -1.0f},  // left   0x04
    {1.0f, 0.0f, 0.0f, -1.0f},   // right  0x08
    {0.0f, -1.0f, 0.0f, -1.0f},  // bottom 0x10
    {0.0f, 1.0f, 0.0f, -1.0f},   // top    0x20
};

// ClipPlane::signed_dist (clip.rs:101-103)
inline float signed_dist(int p, const float* pos) { return dot4(PLANES[p], pos); }

// view_frustum::outcode (clip.rs:240-242), ClipPlane::outcode (clip.rs:109-111)
inline uint8_t outcode(const float* pos) {
    uint8_t code = 0;
    for (int p = 0; p < 6; ++p) code += (uint8_t)((signed_dist(p, pos) > 0.0f) ? (1u << p) : 0u);
    return code;
}

// ClipVert::new (clip.rs:303-309)
template <int N>
inline ClipVert<N> clip_vert(const float* pos, const Attr<N>& a) {
    ClipVert<N> v;
    for (int i = 0; i < 4; ++i) v.pos[i] = pos[i];
    v.attrib = a;
    v.outcode = outcode(pos);
    return v;
}

enum Status { VISIBLE, CLIPPED, HIDDEN };

// view_frustum::status (clip.rs:245-267)
template <int N>
inline Status status(const ClipVert<N>* vs, int n) {
    uint8_t all_out = 0xFF, any_out = 0;
    for (int i = 0; i < n; ++i) {
        all_out &= vs[i].outcode;
        any_out |= vs[i].outcode;
    }
    if (all_out != 0) return HIDDEN;
    if (any_out == 0) return VISIBLE;
    return CLIPPED;
}

// ClipPlane::intersect (clip.rs:121-145). Returns true and writes `out` if the
// edge crosses the plane.
template <int N>
inline bool intersect(int p, const ClipVert<N>& v0, const ClipVert<N>& v1, ClipVert<N>& out) {
    float d0 = signed_dist(p, v0.pos);
    float d1 = signed_dist(p, v1.pos);
    if (!(d0 * d1 < 0.0f)) return false;
    float t = -d0 / (d1 - d0);
    float pos[4];
    for (int i = 0; i < 4; ++i) pos[i] = v0.pos[i] + (v1.pos[i] - v0.pos[i]) * t;
    out = clip_vert<N>(pos, a_lerp<N>(v0.attrib, v1.attrib, t));
    return true;
}

// ClipPlane::clip_simple_polygon (clip.rs:167-193)
template <int N>
inline void clip_plane_polygon(int p, const std::vector<ClipVert<N>>& in,
                               std::vector<ClipVert<N>>& out) {
    if (in.empty()) return;
    const size_t n = in.size();
    const uint8_t bit = (uint8_t)(1u << p);
    for (size_t k = 0; k < n; ++k) {
        const ClipVert<N>& v0 = in[k];
        const ClipVert<N>& v1 = in[(k + 1) % n];  // chain(&verts_in[..1])
        if ((bit & v0.outcode) == 0) out.push_back(v0);  // is_inside
        ClipVert<N> v;
        if (intersect<N>(p, v0, v1, v)) out.push_back(v);
    }
}

// clip_simple_polygon over a plane list (clip.rs:283-301); `planes` holds
// indices into PLANES. view_frustum::clip passes all six in order.
static const int ALL_PLANES[6] = {0, 1, 2, 3, 4, 5};
template <int N>
inline void clip_polygon(const int* planes, int np, std::vector<ClipVert<N>>& vin,
                         std::vector<ClipVert<N>>& vout) {
    for (int i = 0; i < np; ++i) {
        clip_plane_polygon<N>(planes[i], vin, vout);
        vin.clear();
        if (vout.empty()) break;
        if (i < np - 1) std::swap(vin, vout);
    }
}

template <int N>
struct TriC {
    ClipVert<N> v[3];
};

// Edge<ClipVert<A>> (geom Edge, prim.rs:41-60 Render::Clip).
template <int N>
struct EdgeC {
    ClipVert<N> v[2];
};

// `Clip for [Edge<ClipVert<A>>]` (clip.rs:311-348): an edge with both ends
// outside one plane is dropped, one with neither end outside any plane is
// kept; otherwise it is cut plane by plane in PLANES order, the end outside
// each plane being replaced by the intersection, and dropped if both ends
// are outside some plane on the way (the outcodes consulted are those the
// ends carry, recomputed for each intersection vertex).
// `planes` lists PLANES indices (all six, in order, by default).
template <int N>
inline void clip_edges(const std::vector<EdgeC<N>>& edges, std::vector<EdgeC<N>>& out,
                       const int* planes = nullptr, int np = 6) {
    static const int ALL[6] = {0, 1, 2, 3, 4, 5};
    if (!planes) planes = ALL;
    for (const EdgeC<N>& e : edges) {
        const bool both_outside = (e.v[0].outcode & e.v[1].outcode) != 0;
        const bool neither_outside = (e.v[0].outcode | e.v[1].outcode) == 0;
        ClipVert<N> a = e.v[0], b = e.v[1];
        if (both_outside) continue;
        if (neither_outside) {
            out.push_back(EdgeC<N>{{a, b}});
            continue;
        }
        bool dropped = false;
        for (int i = 0; i < np; ++i) {
            const int p = planes[i];
            const uint8_t bit = (uint8_t)(1u << p);
            const bool a_in = (bit & a.outcode) == 0;
            const bool b_in = (bit & b.outcode) == 0;
            if (!a_in && !b_in) {
                dropped = true;
                break;
            }
            ClipVert<N> v;
            if (intersect<N>(p, a, b, v)) {
                if (a_in) b = v;
                else if (b_in) a = v;
            }
        }
        if (!dropped) out.push_back(EdgeC<N>{{a, b}});
    }
}

// `Clip for [Tri<ClipVert<A>>]` (clip.rs:350-400)
template <int N>
inline void clip_tris(const std::vector<TriC<N>>& tris, std::vector<TriC<N>>& out,
                      const int* planes = ALL_PLANES, int np = 6) {
    std::vector<ClipVert<N>> vin, vout;
    vin.reserve(10);
    vout.reserve(10);
    for (const TriC<N>& t : tris) {
        Status s = status<N>(t.v, 3);
        if (s == VISIBLE) {
            out.push_back(t);
            continue;
        }
        if (s == HIDDEN) continue;
        vin.assign(t.v, t.v + 3);
        clip_polygon<N>(planes, np, vin, vout);
        if (vout.size() >= 1) {
            // out.extend(rest.windows(2).map(|e| Tri([p, e[0], e[1]])))
            for (size_t k = 1; k + 1 < vout.size(); ++k) {
                TriC<N> f;
                f.v[0] = vout[0];
                f.v[1] = vout[k];
                f.v[2] = vout[k + 1];
                out.push_back(f);
            }
        }
        vout.clear();
    }
}

// ---------------------------------------------------------------------------
// Screen space (prim.rs:62-88) and winding (geom/prim.rs:288-295)
// ---------------------------------------------------------------------------
template <int N>
struct ScreenVert {
    float pos[3];
    Attr<N> attrib;
};

// `Mat4<Src,Dst>::apply` for points (mat.rs:945-949): rows 0..2 dotted with
// the homogeneous point [x, y, z, 1].
inline void apply_pt(const float* m, const float* p3, float* out3) {
    const float h[4] = {p3[0], p3[1], p3[2], 1.0f};
    out3[0] = dot4(m + 0, h);
    out3[1] = dot4(m + 4, h);
    out3[2] = dot4(m + 8, h);
}
// `ProjMat3::apply` (mat.rs:965-969): all four rows.
inline void apply_proj(const float* m, const float* p3, float* out4) {
    const float h[4] = {p3[0], p3[1], p3[2], 1.0f};
    out4[0] = dot4(m + 0, h);
    out4[1] = dot4(m + 4, h);
    out4[2] = dot4(m + 8, h);
    out4[3] = dot4(m + 12, h);
}

// ---------------------------------------------------------------------------
// Whole-object culling by the caller (scene.rs, crates.rs:114-133)
// ---------------------------------------------------------------------------
// BBox::of (scene.rs:37-39): BBox::default() = (+inf, -inf) (scene.rs:107-112)
// extended by every position (scene.rs:42-46, 115-119) with f32::min / f32::max,
// which return the other operand when one is NaN (as fminf / fmaxf do).
inline void bbox_of(const float* pos, uint64_t n, float* bb6) {
    for (int k = 0; k < 3; ++k) {
        bb6[k] = INFINITY;
        bb6[3 + k] = -INFINITY;
    }
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            bb6[k] = std::fmin(bb6[k], pos[3 * i + k]);
            bb6[3 + k] = std::fmax(bb6[3 + k], pos[3 * i + k]);
        }
}

// BBox::visibility (scene.rs:81-87): the 8 corners of BBox::verts
// (scene.rs:59-66) through ProjMat3::apply, as ClipVerts, then
// view_frustum::status.
inline Status bbox_visibility(const float* bb6, const float* mvp) {
    const float x[2] = {bb6[0], bb6[3]}, y[2] = {bb6[1], bb6[4]}, z[2] = {bb6[2], bb6[5]};
    ClipVert<0> vs[8];
    int n = 0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                const float p[3] = {x[a], y[b], z[c]};
                float clip[4];
                apply_proj(mvp, p, clip);
                vs[n++] = clip_vert<0>(clip, Attr<0>{});
            }
    return status<0>(vs, 8);
}

// to_screen (prim.rs:62-88): pt3(x, y, 1).z_div(w) is (x/w, y/w, 1/w); the
// attribute is z_div'ed by w componentwise when the varying type divides
// (vec.rs:666-676 for Vec2/TexCoord, vary.rs:135-140 for f32), and left as is
// for colors (color.rs:628, default ZDiv body vary.rs:12-14).
template <int N, bool PERSP>
inline ScreenVert<N> to_screen(const ClipVert<N>& v, const float* vp) {
    const float x = v.pos[0], y = v.pos[1], w = v.pos[3];
    const float p[3] = {x / w, y / w, 1.0f / w};
    ScreenVert<N> s;
    apply_pt(vp, p, s.pos);
    for (int i = 0; i < N; ++i) s.attrib.c[i] = PERSP ? v.attrib.c[i] / w : v.attrib.c[i];
    return s;
}

// Tri::winding == Cw (geom/prim.rs:288-295; perp vec.rs:443-445; perp_dot
// vec.rs:480-482): z = dot(perp(u), v) with perp(u) = (-u.y, u.x).
template <int N>
inline bool is_backface(const ScreenVert<N>* t) {
    const float ux = t[1].pos[0] - t[0].pos[0];
    const float uy = t[1].pos[1] - t[0].pos[1];
    const float vx = t[2].pos[0] - t[0].pos[0];
    const float vy = t[2].pos[1] - t[0].pos[1];
    float z = 0.0f;
    z = z + (-uy) * vx;
    z = z + ux * vy;
    return z < 0.0f;
}

// ---------------------------------------------------------------------------
// Scan conversion (raster.rs)
// ---------------------------------------------------------------------------
template <int N>
struct Scanline {
    uint64_t y;
    uint64_t x0, x1;  // xs = x0..x1 (may be x1 < x0)
    Vy<N> v0;         // first value of the `vs` iterator
    Vy<N> dv_dx;      // step of the `vs` iterator
    uint32_t count;   // max items of the `vs` iterator
};

// ScanlineIter (raster.rs:45-51, next 84-113)
template <int N>
struct ScanlineIter {
    float y;
    Vy<N> left, dl;  // left: Iter over Varyings with step dl
    float right, dr; // right: Iter over f32 with step dr
    Vy<N> dv_dx;
    uint32_t n;

    bool next(Scanline<N>& sl) {
        if (n == 0) return false;
        Vy<N> v0 = left;
        left = vy_step<N>(left, dl);
        float x1 = right;
        right = right + dr;
        const float rx0 = round_up_to_half(v0.x);
        const float rx1 = round_up_to_half(x1);
        v0 = vy_lerp<N>(v0, vy_step<N>(v0, dv_dx), rx0 - v0.x);
        sl.v0 = v0;
        sl.dv_dx = dv_dx;
        sl.count = as_u32(rx1 - rx0);
        sl.y = as_usize(y);
        sl.x0 = as_usize(rx0);
        sl.x1 = as_usize(rx1);
        y += 1.0f;
        n -= 1;
        return true;
    }
};

// scan (raster.rs:248-302)
template <int N>
inline ScanlineIter<N> scan(float y0, float y1, const Vy<N>& l0, const Vy<N>& l1,
                            const Vy<N>& r0, const Vy<N>& r1) {
    const float recip_dy = 1.0f / (y1 - y0);
    const Vy<N> dl_dy = vy_dvdt<N>(l0, l1, recip_dy);
    const Vy<N> dr_dy = vy_dvdt<N>(r0, r1, recip_dy);
    Vy<N> dv_dx;
    {
        const Vy<N> l0s = vy_step<N>(l0, dl_dy), r0s = vy_step<N>(r0, dr_dy);
        const float dx = r0s.x - l0s.x;
        dv_dx = vy_dvdt<N>(l0s, r0s, 1.0f / dx);
    }
    const float y0r = round_up_to_half(y0);
    const float y1r = round_up_to_half(y1);
    const float y_tweak = y0r - y0;
    ScanlineIter<N> it;
    it.left = vy_lerp<N>(l0, vy_step<N>(l0, dl_dy), y_tweak);
    it.right = r0.x + dr_dy.x * y_tweak;
    it.y = y0r;
    it.dl = dl_dy;
    it.dr = dr_dy.x;
    it.dv_dx = dv_dx;
    it.n = as_u32(y1r - y0r);
    return it;
}

// tri_fill (raster.rs:185-224). `f` receives each Scanline.
template <int N, class F>
inline void tri_fill(const ScreenVert<N>* in, F&& f) {
    ScreenVert<N> v[3] = {in[0], in[1], in[2]};
    // stable sort by y using total_cmp (slice::sort_by is stable)
    std::stable_sort(v, v + 3, [](const ScreenVert<N>& a, const ScreenVert<N>& b) {
        return total_key(a.pos[1]) < total_key(b.pos[1]);
    });
    const Vy<N> top{v[0].pos[0], v[0].pos[1], v[0].pos[2], v[0].attrib};
    const Vy<N> mid0{v[1].pos[0], v[1].pos[1], v[1].pos[2], v[1].attrib};
    const Vy<N> bot{v[2].pos[0], v[2].pos[1], v[2].pos[2], v[2].attrib};
    const float top_y = v[0].pos[1], mid_y = v[1].pos[1], bot_y = v[2].pos[1];
    const Vy<N> mid1 = vy_lerp<N>(top, bot, (mid_y - top_y) / (bot_y - top_y));
    const bool l_is_mid0 = mid0.x < mid1.x;
    const Vy<N>& left = l_is_mid0 ? mid0 : mid1;
    const Vy<N>& right = l_is_mid0 ? mid1 : mid0;
    Scanline<N> sl;
    {
        ScanlineIter<N> it = scan<N>(top_y, mid_y, top, left, top, right);
        while (it.next(sl)) f(sl);
    }
    {
        ScanlineIter<N> it = scan<N>(mid_y, bot_y, left, bot, right, bot);
        while (it.next(sl)) f(sl);
    }
}

// ---------------------------------------------------------------------------
// Context, stats, targets (ctx.rs, stats.rs, target.rs, pixfmt.rs)
// ---------------------------------------------------------------------------
enum DepthTest { DT_NONE = 0, DT_LESS = 1, DT_EQUAL = 2, DT_GREATER = 3 };
enum FaceCull { FC_NONE = 0, FC_BACK = 1, FC_FRONT = 2 };

struct Ctx {
    int depth_test = DT_LESS;  // ctx.rs:122
    int face_cull = FC_BACK;   // ctx.rs:119
    bool color_write = true;
    bool depth_write = true;
    int depth_sort = 0;        // 0 None, 1 FrontToBack, 2 BackToFront (ctx.rs:39, 68-72)
};

// Context::depth_test (ctx.rs:86-89): `curr.partial_cmp(&new) == depth_test`.
inline bool depth_pass(int dt, float new_z, float curr) {
    switch (dt) {
        case DT_NONE: return true;
        case DT_LESS: return curr < new_z;
        case DT_EQUAL: return curr == new_z;
        case DT_GREATER: return curr > new_z;
    }
    return false;
}
// Context::face_cull (ctx.rs:95-101)
inline bool face_culled(int fc, bool backface) {
    if (fc == FC_BACK) return backface;
    if (fc == FC_FRONT) return !backface;
    return false;
}

struct Stats {
    uint64_t calls = 0, prims_i = 0, prims_o = 0, verts_i = 0, verts_o = 0, frags_i = 0, frags_o = 0;
};

// Pixel formats of the colour buffer.
enum PixFmt {
    PF_RGBA8 = 0,      // [u8;4] r,g,b,a: Buf2<Color4>, or Rgba8888 (pixfmt.rs:93-97)
    PF_RGB8 = 1,       // [u8;3] r,g,b: Buf2<Color3> via Color4::to_rgb (target.rs:127-136)
    PF_U32_RGBA = 2,   // u32 0xRRGGBBAA: Rgba8888 into u32 (pixfmt.rs:76-84)
    PF_U32_XRGB = 3,   // u32 0x00RRGGBB: Xrgb8888 (pixfmt.rs:86-92)
    PF_U32_ARGB = 4,   // u32 0xAARRGGBB: Argb8888 into u32; in memory the Bgra8888 bytes (pixfmt.rs:99-104)
    PF_U32_BGRA = 5,   // u32 0xBBGGRRAA: Bgra8888 into u32; in memory the Argb8888 bytes (pixfmt.rs:105-110)
    PF_RGBA4444 = 6,   // u16 r,g,b,a nibbles from the top: Rgba4444 (pixfmt.rs:122-133)
    PF_RGB565 = 7,     // u16 5,6,5 bits r,g,b: Rgb565 of Color4::to_rgb (pixfmt.rs:58-65, 134-138)
};

inline uint64_t pixfmt_bytes(int fmt) { return fmt == PF_RGB8 ? 3 : (fmt >= PF_RGBA4444 ? 2 : 4); }

struct Target {
    uint8_t* color = nullptr;   // row-major, `color_stride` bytes per row
    uint64_t width = 0, height = 0, color_stride = 0;
    int fmt = PF_RGBA8;
    float* depth = nullptr;     // null => colour-only target (no z-buffer)
    uint64_t depth_stride = 0;  // floats per row
};

// rgba as four bytes (Color4)
struct Rgba {
    uint8_t r, g, b, a;
};

inline void store_pixel(const Target& t, uint64_t x, uint64_t y, Rgba c) {
    switch (t.fmt) {
        case PF_RGBA8: {
            uint8_t* p = t.color + y * t.color_stride + x * 4;
            p[0] = c.r; p[1] = c.g; p[2] = c.b; p[3] = c.a;
            break;
        }
        case PF_RGB8: {
            uint8_t* p = t.color + y * t.color_stride + x * 3;
            p[0] = c.r; p[1] = c.g; p[2] = c.b;
            break;
        }
        case PF_U32_RGBA: {
            uint32_t v = ((uint32_t)c.r << 24) | ((uint32_t)c.g << 16) | ((uint32_t)c.b << 8) | c.a;
            std::memcpy(t.color + y * t.color_stride + x * 4, &v, 4);
            break;
        }
        case PF_U32_XRGB: {
            uint32_t v = ((uint32_t)c.r << 16) | ((uint32_t)c.g << 8) | c.b;
            std::memcpy(t.color + y * t.color_stride + x * 4, &v, 4);
            break;
        }
        case PF_U32_ARGB: {  // u32::from_be_bytes([a, r, g, b])
            uint32_t v = ((uint32_t)c.a << 24) | ((uint32_t)c.r << 16) | ((uint32_t)c.g << 8) | c.b;
            std::memcpy(t.color + y * t.color_stride + x * 4, &v, 4);
            break;
        }
        case PF_U32_BGRA: {  // u32::from_be_bytes([b, g, r, a])
            uint32_t v = ((uint32_t)c.b << 24) | ((uint32_t)c.g << 16) | ((uint32_t)c.r << 8) | c.a;
            std::memcpy(t.color + y * t.color_stride + x * 4, &v, 4);
            break;
        }
        case PF_RGBA4444: {  // (r >> 4) << 12 | (g >> 4) << 8 | (b >> 4) << 4 | a >> 4
            uint16_t v = (uint16_t)(((c.r >> 4) << 12) | ((c.g >> 4) << 8) | ((c.b >> 4) << 4) | (c.a >> 4));
            std::memcpy(t.color + y * t.color_stride + x * 2, &v, 2);
            break;
        }
        case PF_RGB565: {  // (r >> 3 & 0x1F) << 11 | (g >> 2 & 0x3F) << 5 | (b >> 3 & 0x1F)
            uint16_t v = (uint16_t)((((c.r >> 3) & 0x1F) << 11) | (((c.g >> 2) & 0x3F) << 5) | ((c.b >> 3) & 0x1F));
            std::memcpy(t.color + y * t.color_stride + x * 2, &v, 2);
            break;
        }
    }
}

// raster::line (raster.rs:122-177): a one-pixel-thick line. The ends are
// ordered top to bottom; a line wider than tall is then drawn left to right,
// one pixel per column x in (round_up_to_half(x0) as usize)..(round_up_to_half(x1)
// as usize) on row `y as usize`, y stepped serially by dy/dx from the value
// adjusted to the rounded x0 (dx and dy are taken before the left-right swap);
// otherwise one pixel per row, x stepped serially by dx/dy. Every pixel's
// scanline is `x..x + 1` with the single varying value (v0.pos, v0.attrib) of
// the first end (`vary_to(vs, 1)` emits only its start), so a line has one
// depth and one colour.
template <int N, class F>
inline void line(ScreenVert<N> v0, ScreenVert<N> v1, F&& scan_fn) {
    if (v0.pos[1] > v1.pos[1]) std::swap(v0, v1);
    const float dx = v1.pos[0] - v0.pos[0];
    const float dy = v1.pos[1] - v0.pos[1];
    auto scan = [&](uint64_t y, uint64_t x, const ScreenVert<N>& v) {
        Scanline<N> sl;
        sl.y = y;
        sl.x0 = x;
        sl.x1 = x + 1;
        sl.v0 = Vy<N>{v.pos[0], v.pos[1], v.pos[2], v.attrib};
        sl.dv_dx = vy_dvdt<N>(sl.v0, sl.v0, 1.0f);
        sl.count = 1;
        scan_fn(sl);
    };
    if (std::fabs(dx) > dy) {
        if (dx < 0.0f) std::swap(v0, v1);
        const float x0 = round_up_to_half(v0.pos[0]);
        const float x1 = round_up_to_half(v1.pos[0]);
        const float dy_dx = dy / dx;
        const float y0 = v0.pos[1] + dy_dx * (x0 - v0.pos[0]);
        float y = y0;
        for (uint64_t x = as_usize(x0), xe = as_usize(x1); x < xe; ++x) {
            scan(as_usize(y), x, v0);
            y = y + dy_dx;
        }
    } else {
        const float y0 = round_up_to_half(v0.pos[1]);
        const float y1 = round_up_to_half(v1.pos[1]);
        const float dx_dy = dx / dy;
        const float x0 = v0.pos[0] + dx_dy * (y0 - v0.pos[1]);
        float x = x0;
        for (uint64_t y = as_usize(y0), ye = as_usize(y1); y < ye; ++y) {
            scan(y, as_usize(x), v0);
            x = x + dx_dy;
        }
    }
}

enum Err { OK = 0, E_INDEX = -1, E_BOUNDS = -2, E_ARG = -3 };

// Framebuf::rasterize -> rasterize_fb (target.rs:163-198), or the colour-only
// `rasterize` (target.rs:138-161) when the target has no depth buffer.
// `fs(z, var, out)` returns false to discard.
template <int N, bool PERSP, class Fs>
inline int rasterize_span(const Target& t, const Scanline<N>& sl, const Fs& fs, const Ctx& ctx,
                          Stats& st) {
    const uint64_t x0 = sl.x0;
    const uint64_t x1 = std::max(sl.x1, x0);
    // `&mut buf[sl.y][x0..x1]` panics when out of bounds
    if (sl.y >= t.height || x1 > t.width) return E_BOUNDS;
    st.frags_i += x1 - x0;
    const uint64_t nfrag = std::min<uint64_t>(sl.count, x1 - x0);  // zip stops at the shorter
    Vy<N> val = sl.v0;
    for (uint64_t j = 0; j < nfrag; ++j) {
        // Scanline::fragments (raster.rs:61-68): z_div by the interpolated z
        const float z = val.z;
        Attr<N> var = val.v;
        if (PERSP)
            for (int i = 0; i < N; ++i) var.c[i] = var.c[i] / z;
        const uint64_t x = x0 + j;
        if (t.depth) {
            float* zp = t.depth + sl.y * t.depth_stride + x;
            Rgba c;
            if (depth_pass(ctx.depth_test, z, *zp) && fs(z, var, c)) {
                if (ctx.color_write) {
                    st.frags_o += 1;
                    store_pixel(t, x, sl.y, c);
                }
                if (ctx.depth_write) *zp = z;
            }
        } else {
            Rgba c;
            if (fs(z, var, c) && ctx.color_write) {
                st.frags_o += 1;
                store_pixel(t, x, sl.y, c);
            }
        }
        val = vy_step<N>(val, sl.dv_dx);
    }
    return OK;
}

// render() (render.rs:134-207) for triangle meshes with the built-in vertex
// shader `vertex(mvp.apply(&v.pos), v.attrib)` (the form used by every
// built-in shader; ProjMat3::apply mat.rs:965-969).
//   tris:  n_tris index triples, `pos`: n_verts xyz, `attr`: n_verts x N.
//   P = 3: `tris` holds n_tris index triples (Tri<usize>); P = 2: index pairs
//   (Edge<usize>, prim.rs:41-60: edge clipping, no backfaces, raster::line).
template <int N, bool PERSP, class Fs, int P = 3>
inline int render(const uint32_t* tris, uint64_t n_tris, const float* pos, const float* attr,
                  uint64_t n_verts, const float* mvp, const float* to_screen_m, const Fs& fs,
                  const Target& t, const Ctx& ctx, Stats& out_stats) {
    Stats st;
    st.calls = 1;
    st.prims_i = n_tris;
    st.verts_i = n_verts;

    // 1. vertex shader + ClipVert::new (render.rs:157-165)
    std::vector<ClipVert<N>> verts(n_verts);
    for (uint64_t i = 0; i < n_verts; ++i) {
        float clip[4];
        apply_proj(mvp, pos + 3 * i, clip);
        Attr<N> a;
        for (int k = 0; k < N; ++k) a.c[k] = attr[N * i + k];
        verts[i] = clip_vert<N>(clip, a);
    }
    if (P == 2) {
        std::vector<EdgeC<N>> edges(n_tris);
        for (uint64_t i = 0; i < n_tris; ++i)
            for (int k = 0; k < 2; ++k) {
                const uint32_t ix = tris[2 * i + k];
                if (ix >= n_verts) return E_INDEX;
                edges[i].v[k] = verts[ix];
            }
        std::vector<EdgeC<N>> clipped;
        clipped.reserve(n_tris / 2);
        clip_edges<N>(edges, clipped);
        int err = OK;
        for (const EdgeC<N>& c : clipped) {
            const ScreenVert<N> s0 = to_screen<N, PERSP>(c.v[0], to_screen_m);
            const ScreenVert<N> s1 = to_screen<N, PERSP>(c.v[1], to_screen_m);
            // Render::is_backface defaults to false for edges (render.rs:72-75)
            if (face_culled(ctx.face_cull, false)) continue;
            st.prims_o += 1;
            st.verts_o += 3;  // render.rs:193 counts 3 for every primitive
            line<N>(s0, s1, [&](const Scanline<N>& sl) {
                if (err != OK) return;
                err = rasterize_span<N, PERSP>(t, sl, fs, ctx, st);
            });
            if (err != OK) return err;
        }
        out_stats.calls += st.calls;
        out_stats.prims_i += st.prims_i;
        out_stats.prims_o += st.prims_o;
        out_stats.verts_i += st.verts_i;
        out_stats.verts_o += st.verts_o;
        out_stats.frags_i += st.frags_i;
        out_stats.frags_o += st.frags_o;
        return OK;
    }
    // 2. primitive assembly (render.rs:167-172); indexing panics when out of range
    std::vector<TriC<N>> prims(n_tris);
    for (uint64_t i = 0; i < n_tris; ++i) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t ix = tris[3 * i + k];
            if (ix >= n_verts) return E_INDEX;
            prims[i].v[k] = verts[ix];
        }
    }
    // 3. clipping (render.rs:174-177)
    std::vector<TriC<N>> clipped;
    clipped.reserve(n_tris / 2);
    clip_tris<N>(prims, clipped);
    // optional depth sort (render.rs:180-183, 209-219) by Render::depth
    // ((a.z + b.z + c.z) / 3.0 in clip space, prim.rs:21-23) under
    // f32::total_cmp, ascending front to back, descending back to front. The
    // reference's sort_unstable_by leaves the order of equal keys to the Rust
    // toolchain's sort; this restatement keeps them in submission order.
    if (ctx.depth_sort) {
        auto depth = [](const TriC<N>& t) { return ((t.v[0].pos[2] + t.v[1].pos[2]) + t.v[2].pos[2]) / 3.0f; };
        const bool back = ctx.depth_sort == 2;
        std::stable_sort(clipped.begin(), clipped.end(), [&](const TriC<N>& a, const TriC<N>& b) {
            const int32_t ka = total_key(depth(a)), kb = total_key(depth(b));
            return back ? kb < ka : ka < kb;
        });
    }

    // 4. per primitive (render.rs:185-205)
    int err = OK;
    for (const TriC<N>& c : clipped) {
        ScreenVert<N> s[3];
        for (int k = 0; k < 3; ++k) s[k] = to_screen<N, PERSP>(c.v[k], to_screen_m);
        if (face_culled(ctx.face_cull, is_backface<N>(s))) continue;
        st.prims_o += 1;
        st.verts_o += 3;
        tri_fill<N>(s, [&](const Scanline<N>& sl) {
            if (err != OK) return;
            err = rasterize_span<N, PERSP>(t, sl, fs, ctx, st);
        });
        if (err != OK) return err;
    }
    out_stats.calls += st.calls;
    out_stats.prims_i += st.prims_i;
    out_stats.prims_o += st.prims_o;
    out_stats.verts_i += st.verts_i;
    out_stats.verts_o += st.verts_o;
    out_stats.frags_i += st.frags_i;
    out_stats.frags_o += st.frags_o;
    return OK;
}

// ---------------------------------------------------------------------------
// Built-in fragment shaders (shader.rs:68-76 closure form)
// ---------------------------------------------------------------------------
struct Texture {
    const uint8_t* data = nullptr;  // Color4 texels, [u8;4] r,g,b,a
    uint32_t w = 0, h = 0, stride = 0;  // stride in texels
};

// Flat: a constant Color4 for every fragment.
struct FsFlat {
    Rgba color;
    bool operator()(float, const Attr<0>&, Rgba& out) const {
        out = color;
        return true;
    }
};
// Gouraud: Color3f::to_color4 (color.rs:253-263): [(256*c) as u8 ..., 255]
struct FsGouraud {
    bool operator()(float, const Attr<3>& v, Rgba& out) const {
        out.r = as_u8(256.0f * v.c[0]);
        out.g = as_u8(256.0f * v.c[1]);
        out.b = as_u8(256.0f * v.c[2]);
        out.a = 0xFF;
        return true;
    }
};
inline Rgba texel(const Texture& t, uint32_t u, uint32_t v) {
    const uint8_t* p = t.data + ((uint64_t)v * t.stride + u) * 4;
    return Rgba{p[0], p[1], p[2], p[3]};
}
// SamplerRepeatPot::sample (tex.rs:240-266)
struct FsTexRepeatPot {
    Texture tex;
    bool operator()(float, const Attr<2>& uv, Rgba& out) const {
        const float tw = (float)tex.w, th = (float)tex.h;
        const float su = tw * uv.c[0], sv = th * uv.c[1];
        const uint32_t u = (uint32_t)as_i32(std::floor(su)) & (tex.w - 1);
        const uint32_t v = (uint32_t)as_i32(std::floor(sv)) & (tex.h - 1);
        out = texel(tex, u, v);
        return true;
    }
};
// f32::clamp: NaN passes through (core::f32::clamp)
inline float fclamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
// SamplerClamp::sample (tex.rs:280-303)
struct FsTexClamp {
    Texture tex;
    bool operator()(float, const Attr<2>& uv, Rgba& out) const {
        const float tw = (float)tex.w, th = (float)tex.h;
        const float su = uv.c[0] * tw, sv = uv.c[1] * th;
        const uint32_t u = as_u32(std::floor(fclamp(su, 0.0f, tw - 1.0f)));
        const uint32_t v = as_u32(std::floor(fclamp(sv, 0.0f, th - 1.0f)));
        out = texel(tex, u, v);
        return true;
    }
};
// SamplerOnce::sample (tex.rs:327-357): uv scaled by the texture size, then
// `as u32`; indexing outside the texture panics (Slice2 index, buf.rs:731-733
// via to_index_strict buf.rs:444-457), recorded in `oob` (the call fails).
struct FsTexOnce {
    Texture tex;
    bool* oob;
    bool operator()(float, const Attr<2>& uv, Rgba& out) const {
        const float tw = (float)tex.w, th = (float)tex.h;
        const uint32_t u = as_u32(tw * uv.c[0]);
        const uint32_t v = as_u32(th * uv.c[1]);
        if (u >= tex.w || v >= tex.h) {
            *oob = true;
            out = Rgba{0, 0, 0, 0};
            return true;
        }
        out = texel(tex, u, v);
        return true;
    }
};

}  // namespace oracle
