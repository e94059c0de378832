// Oracle C entry points, loaded by the tests (ctypes) and by the CPU-baseline
// leg of bench.py. TEST INFRASTRUCTURE ONLY: the product library never links
// this file.
#include <array>
#include <chrono>
#include <cmath>

#include "rf_oracle.hpp"

using namespace oracle;

extern "C" {

// Shader kinds, same numbering as the product header.
enum { K_FLAT = 0, K_GOURAUD = 1, K_TEX_CLAMP = 2, K_TEX_REPEAT_POT = 3, K_TEX_ONCE = 4 };

struct OracleCtx {
    int32_t depth_test;   // 0 none, 1 less, 2 equal, 3 greater
    int32_t face_cull;    // 0 none, 1 back, 2 front
    int32_t color_write;
    int32_t depth_write;
    int32_t depth_sort;   // 0 none, 1 front to back, 2 back to front
};

// One render() call. Stats are accumulated into stats[0..7]:
// calls, prims_i, prims_o, verts_i, verts_o, frags_i, frags_o.
// Returns 0, or -1 (index out of range), -2 (span outside target), -3 (bad arg).
// `prim` 0: `tris` holds index triples (Tri<usize>); 1: index pairs (Edge<usize>).
int oracle_render_prims(int prim, int kind, const uint32_t* tris, uint64_t n_tris, const float* pos,
                        const float* attr, uint64_t n_verts, const float* mvp, const float* viewport,
                        const uint8_t* flat_rgba, const uint8_t* tex, uint32_t tex_w, uint32_t tex_h,
                        uint32_t tex_stride, uint8_t* color, uint64_t width, uint64_t height,
                        uint64_t color_stride_bytes, int32_t fmt, float* depth,
                        uint64_t depth_stride, const OracleCtx* c, uint64_t* stats) {
    if (prim != 0 && prim != 1) return E_ARG;
    Target t;
    t.color = color;
    t.width = width;
    t.height = height;
    t.color_stride = color_stride_bytes;
    t.fmt = fmt;
    t.depth = depth;
    t.depth_stride = depth_stride;
    Ctx ctx;
    ctx.depth_test = c->depth_test;
    ctx.face_cull = c->face_cull;
    ctx.color_write = c->color_write != 0;
    ctx.depth_write = c->depth_write != 0;
    ctx.depth_sort = c->depth_sort;
    if (ctx.depth_sort < 0 || ctx.depth_sort > 2) return E_ARG;
    Stats st;
    int err = E_ARG;
    switch (kind) {
        case K_FLAT: {
            FsFlat fs{Rgba{flat_rgba[0], flat_rgba[1], flat_rgba[2], flat_rgba[3]}};
            err = prim ? render<0, false, FsFlat, 2>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st)
                       : render<0, false>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st);
            break;
        }
        case K_GOURAUD: {
            FsGouraud fs;
            err = prim ? render<3, false, FsGouraud, 2>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st)
                       : render<3, false>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st);
            break;
        }
        case K_TEX_CLAMP: {
            FsTexClamp fs{Texture{tex, tex_w, tex_h, tex_stride}};
            err = prim ? render<2, true, FsTexClamp, 2>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st)
                       : render<2, true>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st);
            break;
        }
        case K_TEX_REPEAT_POT: {
            // SamplerRepeatPot::new asserts power-of-two sides (tex.rs:228-233)
            if (tex_w == 0 || tex_h == 0 || (tex_w & (tex_w - 1)) || (tex_h & (tex_h - 1)))
                return E_ARG;
            FsTexRepeatPot fs{Texture{tex, tex_w, tex_h, tex_stride}};
            err = prim ? render<2, true, FsTexRepeatPot, 2>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st)
                       : render<2, true>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st);
            break;
        }
        case K_TEX_ONCE: {
            if (tex_w == 0 || tex_h == 0) return E_ARG;
            bool oob = false;
            FsTexOnce fs{Texture{tex, tex_w, tex_h, tex_stride}, &oob};
            err = prim ? render<2, true, FsTexOnce, 2>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st)
                       : render<2, true>(tris, n_tris, pos, attr, n_verts, mvp, viewport, fs, t, ctx, st);
            if (err == OK && oob) return E_BOUNDS;
            break;
        }
        default:
            return E_ARG;
    }
    if (stats) {
        stats[0] += st.calls;
        stats[1] += st.prims_i;
        stats[2] += st.prims_o;
        stats[3] += st.verts_i;
        stats[4] += st.verts_o;
        stats[5] += st.frags_i;
        stats[6] += st.frags_o;
    }
    return err;
}

// One render() call over triangles (oracle_render_prims with prim 0).
int oracle_render(int kind, const uint32_t* tris, uint64_t n_tris, const float* pos,
                  const float* attr, uint64_t n_verts, const float* mvp, const float* viewport,
                  const uint8_t* flat_rgba, const uint8_t* tex, uint32_t tex_w, uint32_t tex_h,
                  uint32_t tex_stride, uint8_t* color, uint64_t width, uint64_t height,
                  uint64_t color_stride_bytes, int32_t fmt, float* depth,
                  uint64_t depth_stride, const OracleCtx* c, uint64_t* stats) {
    return oracle_render_prims(0, kind, tris, n_tris, pos, attr, n_verts, mvp, viewport, flat_rgba, tex, tex_w,
                               tex_h, tex_stride, color, width, height, color_stride_bytes, fmt, depth,
                               depth_stride, c, stats);
}

// The `fp` variant of core/examples/hello_tri.rs:11-19: vertex colours go
// through Color3f::to_linear (powf(c, GAMMA), color.rs:276-285) and fragments
// through to_srgb (powf(c, INV_GAMMA), color.rs:383-391) then to_color4. This
// is a user-closure shader, used only to pin coverage against core/triangle.ppm.
int oracle_render_hello_fp(const uint32_t* tris, uint64_t n_tris, const float* pos,
                           const float* attr, uint64_t n_verts, const float* mvp,
                           const float* viewport, uint8_t* color, uint64_t width,
                           uint64_t height, uint64_t* stats) {
    const float GAMMA = 2.2f;
    const float INV_GAMMA = 1.0f / GAMMA;
    std::vector<float> lin(3 * n_verts);
    for (uint64_t i = 0; i < 3 * n_verts; ++i) lin[i] = std::pow(attr[i], GAMMA);
    struct FsSrgb {
        float ig;
        bool operator()(float, const Attr<3>& v, Rgba& out) const {
            out.r = as_u8(256.0f * std::pow(v.c[0], ig));
            out.g = as_u8(256.0f * std::pow(v.c[1], ig));
            out.b = as_u8(256.0f * std::pow(v.c[2], ig));
            out.a = 0xFF;
            return true;
        }
    } fs{INV_GAMMA};
    Target t;
    t.color = color;
    t.width = width;
    t.height = height;
    t.color_stride = width * 4;
    t.fmt = PF_RGBA8;
    Ctx ctx;
    Stats st;
    int err = render<3, false>(tris, n_tris, pos, lin.data(), n_verts, mvp, viewport, fs, t, ctx, st);
    if (stats) {
        stats[0] += st.calls; stats[1] += st.prims_i; stats[2] += st.prims_o;
        stats[3] += st.verts_i; stats[4] += st.verts_o; stats[5] += st.frags_i;
        stats[6] += st.frags_o;
    }
    return err;
}

// tri_fill over three screen vertices (x, y, z, value) with an f32 varying
// (which z_divs, vary.rs:135-140). Writes per scanline (y, x0, x1, count) into
// `sl` (max `max_sl`) and per fragment (z, value/z) into `frags` (max
// `max_frags` pairs). Returns the number of scanlines, or -1 on overflow.
int64_t oracle_tri_fill_f32(const float* v, int64_t* sl, int64_t max_sl, float* frags,
                            int64_t max_frags, int64_t* n_frags_out) {
    ScreenVert<1> s[3];
    for (int k = 0; k < 3; ++k) {
        s[k].pos[0] = v[4 * k + 0];
        s[k].pos[1] = v[4 * k + 1];
        s[k].pos[2] = v[4 * k + 2];
        s[k].attrib.c[0] = v[4 * k + 3];
    }
    int64_t nsl = 0, nf = 0;
    bool overflow = false;
    tri_fill<1>(s, [&](const Scanline<1>& l) {
        if (nsl >= max_sl) {
            overflow = true;
            return;
        }
        sl[4 * nsl + 0] = (int64_t)l.y;
        sl[4 * nsl + 1] = (int64_t)l.x0;
        sl[4 * nsl + 2] = (int64_t)l.x1;
        sl[4 * nsl + 3] = (int64_t)l.count;
        ++nsl;
        Vy<1> val = l.v0;
        for (uint32_t j = 0; j < l.count; ++j) {
            if (nf >= max_frags) {
                overflow = true;
                return;
            }
            frags[2 * nf + 0] = val.z;
            frags[2 * nf + 1] = val.v.c[0] / val.z;
            ++nf;
            val = vy_step<1>(val, l.dv_dx);
        }
    });
    *n_frags_out = nf;
    return overflow ? -1 : nsl;
}

// Vary::vary over f32 (vary.rs:142-155, space.rs:200-214).
void oracle_vary_f32(float start, float step, uint32_t n, float* out) {
    float v = start;
    for (uint32_t i = 0; i < n; ++i) {
        out[i] = v;
        v = v + step;
    }
}
// Vary::vary_to over f32 (vary.rs:66-76).
void oracle_vary_to_f32(float a, float b, uint32_t n, float* out) {
    const float recip = (n <= 1) ? 1.0f : 1.0f / (float)(n - 1);
    oracle_vary_f32(a, (b - a) * recip, n, out);
}

// Scanline::fragments over `vary_to((p0, v0), (p1, v1), n)` for a varying
// (ScreenPt, f32); writes (x, y, z, v/z) per fragment. Mirrors the KAT at
// raster.rs:404-436.
void oracle_scanline_frags_f32(const float* a, const float* b, uint32_t n, float* out) {
    Vy<1> va{a[0], a[1], a[2], {{a[3]}}};
    Vy<1> vb{b[0], b[1], b[2], {{b[3]}}};
    const float recip = (n <= 1) ? 1.0f : 1.0f / (float)(n - 1);
    const Vy<1> d = vy_dvdt<1>(va, vb, recip);
    Vy<1> val = va;
    for (uint32_t i = 0; i < n; ++i) {
        out[4 * i + 0] = val.x;
        out[4 * i + 1] = val.y;
        out[4 * i + 2] = val.z;
        out[4 * i + 3] = val.v.c[0] / val.z;
        val = vy_step<1>(val, d);
    }
}

// BBox::of over n positions (x, y, z) into bb6 = low xyz, high xyz.
void oracle_bbox_of(const float* pos, uint64_t n, float* bb6) { bbox_of(pos, n, bb6); }
// BBox::visibility: 0 Visible, 1 Clipped, 2 Hidden (clip.rs:65-72).
int oracle_bbox_visibility(const float* bb6, const float* mvp) { return (int)bbox_visibility(bb6, mvp); }

float oracle_signed_dist(int plane, const float* pos) { return signed_dist(plane, pos); }
int oracle_outcode(const float* pos) { return outcode(pos); }

// `[tris].clip(planes, out)` for ClipVert<f32> (clip.rs:350-400) with
// `planes` a list of PLANES indices. `in` holds n_in triangles of 3 x
// (x, y, z, w, attr); `out` receives up to max_out triangles in the same
// layout. Returns the number of output triangles (or -1 on overflow).
int64_t oracle_clip_tris_f32(const float* in, int64_t n_in, const int32_t* planes, int32_t np,
                             float* out, int64_t max_out) {
    std::vector<TriC<1>> tris(n_in), res;
    for (int64_t i = 0; i < n_in; ++i)
        for (int k = 0; k < 3; ++k) {
            const float* p = in + (i * 3 + k) * 5;
            Attr<1> a{{p[4]}};
            tris[i].v[k] = clip_vert<1>(p, a);
        }
    std::vector<int> pl(planes, planes + np);
    clip_tris<1>(tris, res, pl.data(), np);
    if ((int64_t)res.size() > max_out) return -1;
    for (size_t i = 0; i < res.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            float* p = out + (i * 3 + k) * 5;
            for (int c = 0; c < 4; ++c) p[c] = res[i].v[k].pos[c];
            p[4] = res[i].v[k].attrib.c[0];
        }
    return (int64_t)res.size();
}

// `[edges].clip(planes, out)` for ClipVert<f32> (clip.rs:311-348). `in` holds
// n_in edges of 2 x (x, y, z, w, attr); returns the output count (or -1).
int64_t oracle_clip_edges_f32(const float* in, int64_t n_in, const int32_t* planes, int32_t np,
                              float* out, int64_t max_out) {
    std::vector<EdgeC<1>> edges(n_in), res;
    for (int64_t i = 0; i < n_in; ++i)
        for (int k = 0; k < 2; ++k) {
            const float* p = in + (i * 2 + k) * 5;
            Attr<1> a{{p[4]}};
            edges[i].v[k] = clip_vert<1>(p, a);
        }
    std::vector<int> pl(planes, planes + np);
    clip_edges<1>(edges, res, pl.data(), np);
    if ((int64_t)res.size() > max_out) return -1;
    for (size_t i = 0; i < res.size(); ++i)
        for (int k = 0; k < 2; ++k) {
            float* p = out + (i * 2 + k) * 5;
            for (int c = 0; c < 4; ++c) p[c] = res[i].v[k].pos[c];
            p[4] = res[i].v[k].attrib.c[0];
        }
    return (int64_t)res.size();
}

// raster::line (raster.rs:122-177) between screen vertices a and b, each
// (x, y, z, var). Writes up to max_out pixels as (x, y) and each fragment's
// (z, var); returns the pixel count (or -1 if more).
int64_t oracle_line_f32(const float* a, const float* b, int64_t* out_xy, float* out_zv, int64_t max_out) {
    ScreenVert<1> v0, v1;
    for (int c = 0; c < 3; ++c) {
        v0.pos[c] = a[c];
        v1.pos[c] = b[c];
    }
    v0.attrib.c[0] = a[3];
    v1.attrib.c[0] = b[3];
    int64_t n = 0;
    bool over = false;
    line<1>(v0, v1, [&](const Scanline<1>& sl) {
        if (n >= max_out) {
            over = true;
            return;
        }
        out_xy[2 * n] = (int64_t)sl.x0;
        out_xy[2 * n + 1] = (int64_t)sl.y;
        out_zv[2 * n] = sl.v0.z;
        out_zv[2 * n + 1] = sl.v0.v.c[0];
        ++n;
    });
    return over ? -1 : n;
}

// ClipPlane::clip_simple_polygon for one plane over an n-vertex polygon
// (x, y, z, w, attr each). Returns the output vertex count.
int64_t oracle_clip_plane_polygon_f32(int32_t plane, const float* in, int64_t n, float* out,
                                      int64_t max_out) {
    std::vector<ClipVert<1>> vin, vout;
    for (int64_t i = 0; i < n; ++i) {
        Attr<1> a{{in[5 * i + 4]}};
        vin.push_back(clip_vert<1>(in + 5 * i, a));
    }
    clip_plane_polygon<1>(plane, vin, vout);
    if ((int64_t)vout.size() > max_out) return -1;
    for (size_t i = 0; i < vout.size(); ++i) {
        for (int c = 0; c < 4; ++c) out[5 * i + c] = vout[i].pos[c];
        out[5 * i + 4] = vout[i].attrib.c[0];
    }
    return (int64_t)vout.size();
}

// The exhaustive enumeration of clip.rs:667-719: every triangle with vertices
// on the 5x5x5 lattice {-2..2}^3 (w = 1). Writes the histogram of output
// triangle counts (8 bins), and returns the number of output vertices whose
// |v/w| exceeds 1.00001 (clip.rs:725-729), which must be zero.
int64_t oracle_clip_histogram(uint64_t* hist) {
    float xs[5];
    oracle_vary_f32(-2.0f, 1.0f, 5, xs);
    std::vector<std::array<float, 4>> pts;
    for (float x : xs)
        for (float y : xs)
            for (float z : xs) pts.push_back({x, y, z, 1.0f});
    for (int i = 0; i < 8; ++i) hist[i] = 0;
    int64_t oob = 0;
    std::vector<TriC<1>> one(1), res;
    Attr<1> zero{{0.0f}};
    for (auto& a : pts)
        for (auto& b : pts)
            for (auto& c : pts) {
                one[0].v[0] = clip_vert<1>(a.data(), zero);
                one[0].v[1] = clip_vert<1>(b.data(), zero);
                one[0].v[2] = clip_vert<1>(c.data(), zero);
                res.clear();
                clip_tris<1>(one, res);
                if (res.size() < 8) hist[res.size()] += 1;
                for (auto& t : res)
                    for (auto& v : t.v)
                        for (int k = 0; k < 4; ++k)
                            if (!(std::fabs(v.pos[k] / v.pos[3]) <= 1.00001f)) ++oob;
            }
    return oob;
}

// Samplers on a Color4 texture (tex.rs:240-303). kind 2 = clamp, 3 = repeat-pot.
int oracle_sample(int kind, const uint8_t* tex, uint32_t w, uint32_t h, uint32_t stride, float u,
                  float v, uint8_t* out) {
    Texture t{tex, w, h, stride};
    Attr<2> uv{{u, v}};
    Rgba c;
    if (kind == K_TEX_CLAMP)
        FsTexClamp{t}(1.0f, uv, c);
    else if (kind == K_TEX_REPEAT_POT)
        FsTexRepeatPot{t}(1.0f, uv, c);
    else if (kind == K_TEX_ONCE) {
        bool oob = false;
        FsTexOnce{t, &oob}(1.0f, uv, c);
        if (oob) return E_BOUNDS;
    } else
        return E_ARG;
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = c.a;
    return OK;
}

// Wall-clock seconds of `reps` frames, each a Frame::clear (front/src/lib.rs:
// 103-120: colour filled with `clear_rgba`, depth with 1/depth_clear) followed
// by one oracle_render call; both are timed.
double oracle_time_frames(int kind, const uint32_t* tris, uint64_t n_tris, const float* pos,
                          const float* attr, uint64_t n_verts, const float* mvp,
                          const float* viewport, const uint8_t* flat_rgba, const uint8_t* tex,
                          uint32_t tex_w, uint32_t tex_h, uint32_t tex_stride, uint8_t* color,
                          uint64_t width, uint64_t height, int32_t fmt, float* depth,
                          const OracleCtx* c, const uint8_t* clear_rgba, float depth_clear,
                          int32_t reps, double* per_frame, uint64_t* stats) {
    const uint64_t stride = width * pixfmt_bytes(fmt);
    double total = 0.0;
    for (int r = 0; r < reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        // Frame::clear
        for (uint64_t y = 0; y < height; ++y)
            for (uint64_t x = 0; x < width; ++x) {
                Target t;
                t.color = color; t.width = width; t.height = height; t.color_stride = stride; t.fmt = fmt;
                store_pixel(t, x, y, Rgba{clear_rgba[0], clear_rgba[1], clear_rgba[2], clear_rgba[3]});
                if (depth) depth[y * width + x] = 1.0f / depth_clear;
            }
        int err = oracle_render(kind, tris, n_tris, pos, attr, n_verts, mvp, viewport, flat_rgba,
                                tex, tex_w, tex_h, tex_stride, color, width, height, stride, fmt,
                                depth, width, c, stats);
        auto t1 = std::chrono::steady_clock::now();
        if (err) return -1.0;
        double s = std::chrono::duration<double>(t1 - t0).count();
        if (per_frame) per_frame[r] = s;
        total += s;
    }
    return total;
}

}  // extern "C"
