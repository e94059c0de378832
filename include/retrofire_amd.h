/*
 * retrofire_amd: the triangle rendering path of retrofire
 * (`retrofire_core::render::render`, core/src/render.rs:134-207) on an
 * MI355X (gfx950), behind a C ABI.
 *
 * The reference has no C ABI: its interface is the generic Rust function
 *
 *   pub fn render<Prim, Vtx, Var, Uni, Shd>(prims, verts, shader: &Shd,
 *       uniform: Uni, to_screen: Mat4<Ndc, Screen>, target: &mut impl Target,
 *       ctx: &Context)                                   (render.rs:134-146)
 *
 * monomorphised over the shader and target types. The entry points below are
 * what a Rust FFI crate binds to replace that call when the shader is one of
 * the built-in kinds and the target one of the supported buffer layouts; any
 * other shader stays on the reference's own CPU path.
 *
 * Conventions
 *   - Every function returns an rf_status (0 = success, negative = error).
 *   - Matrices are row-major float[16], applied as row . column
 *     (mat.rs:929-970): `mvp` is the uniform of the built-in vertex shader
 *     `vertex(mvp.apply(&v.pos), v.attrib)`, `viewport` is `to_screen`.
 *   - Index triples are uint32 (the reference's Tri<usize>).
 *   - One rf_device per host thread, like the reference's !Sync Context
 *     (ctx.rs:64). A serving handle owns its own worker thread.
 */
#ifndef RETROFIRE_AMD_H
#define RETROFIRE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rf_status;
#define RF_OK 0
#define RF_E_INDEX (-1)       /* vertex index out of range (prim.rs:17-19 panics)   */
#define RF_E_BOUNDS (-2)      /* span outside the target (target.rs:173-174 panics) */
#define RF_E_ARG (-3)         /* invalid argument (e.g. non-power-of-two texture for
                                 SamplerRepeatPot, tex.rs:228-233 panics)            */
#define RF_E_DEVICE (-4)      /* HIP runtime error or out of device memory         */
#define RF_E_UNSUPPORTED (-5) /* not a built-in shader kind / target layout, a
                                 launch of 2^29 primitives or more, or depth_sort
                                 in rf_render_noalloc: the caller takes another
                                 path                                               */
#define RF_E_VIEWPORT (-6)    /* to_screen puts a vertex more than 2^22 rows above
                                 the target (rows above row 0 fold onto it, as
                                 `self.y as usize` does, raster.rs:106, up to that) */
#define RF_E_NONFINITE (-7)   /* a screen-space vertex coordinate is infinite (a NaN
                                 one, from a w = 0 vertex, is drawn as the
                                 reference draws it: empty spans)                 */
#define RF_E_INTERNAL (-8)    /* internal capacity or binning check failed         */
#define RF_E_BUDGET (-9)      /* request larger than the serving memory budget     */

/* Built-in shader kinds (shader.rs:57-126 closure and Shader forms). */
typedef enum rf_shader_kind {
    RF_SHADER_FLAT = 0,            /* |_| color: constant Color4, varying ()            */
    RF_SHADER_GOURAUD = 1,         /* Color3f varying, frag.var.to_color4()             */
    RF_SHADER_TEX_CLAMP = 2,       /* TexCoord varying, SamplerClamp.sample (tex.rs:272) */
    RF_SHADER_TEX_REPEAT_POT = 3,  /* TexCoord varying, SamplerRepeatPot (tex.rs:218)    */
    RF_SHADER_TEX_ONCE = 4         /* TexCoord varying, SamplerOnce (tex.rs:313-357): a
                                    * texel outside the texture is RF_E_BOUNDS, where
                                    * the reference's indexing panics (buf.rs:731-733) */
} rf_shader_kind;

/* Colour buffer element layouts (target.rs:35-136, pixfmt.rs:58-138). The
 * u32 forms are stored little-endian, so RF_PIX_U32_ARGB is also the byte
 * layout of Colorbuf<Buf2<[u8; 4]>, Bgra8888> and RF_PIX_U32_BGRA that of
 * Colorbuf<Buf2<[u8; 4]>, Argb8888>. The u16 forms take 2 bytes a pixel. */
typedef enum rf_pixfmt {
    RF_PIX_RGBA8 = 0,      /* [u8; 4] r,g,b,a: Buf2<Color4> or Colorbuf<_, Rgba8888> */
    RF_PIX_RGB8 = 1,       /* [u8; 3] r,g,b:   Buf2<Color3>                          */
    RF_PIX_U32_RGBA = 2,   /* u32 0xRRGGBBAA:  Colorbuf<Buf2<u32>, Rgba8888>          */
    RF_PIX_U32_XRGB = 3,   /* u32 0x00RRGGBB:  Colorbuf<Buf2<u32>, Xrgb8888>          */
    RF_PIX_U32_ARGB = 4,   /* u32 0xAARRGGBB:  Colorbuf<Buf2<u32>, Argb8888>          */
    RF_PIX_U32_BGRA = 5,   /* u32 0xBBGGRRAA:  Colorbuf<Buf2<u32>, Bgra8888>          */
    RF_PIX_RGBA4444 = 6,   /* u16 rrrrggggbbbbaaaa: Colorbuf<Buf2<u16>, Rgba4444>     */
    RF_PIX_RGB565 = 7      /* u16 rrrrrggggggbbbbb: Colorbuf<Buf2<u16>, Rgb565>       */
} rf_pixfmt;

/* Context (ctx.rs:11-65). depth_test holds the reference's
 * Option<Ordering>: 0 = None, 1 = Less, 2 = Equal, 3 = Greater. face_cull:
 * 0 = None, 1 = Back, 2 = Front. depth_sort (Option<DepthSort>, ctx.rs:68-72):
 * 0 = None, 1 = FrontToBack, 2 = BackToFront; rf_render sorts the clipped
 * triangles by Render::depth under f32::total_cmp (render.rs:209-219), equal
 * depths in submission order (the reference's sort_unstable_by leaves their
 * order to the Rust toolchain); batched and served frames sort each of their
 * triangle draws the same way (a served depth-sorted request is launched on
 * its own). clear_color / clear_depth fuse the frontend's
 * Frame::clear (front/src/lib.rs:103-120) into the call: the colour buffer is
 * first filled with color_clear, the depth buffer with 1 / depth_clear. */
typedef struct rf_ctx {
    int32_t depth_test;
    int32_t face_cull;
    int32_t color_write;
    int32_t depth_write;
    int32_t depth_sort;
    int32_t clear_color;
    uint8_t color_clear[4]; /* Color4 r, g, b, a */
    int32_t clear_depth;
    float depth_clear;      /* default +inf: the buffer is filled with 1/inf = 0 */
} rf_ctx;

/* Stats (stats.rs:16-40). `seconds` is the wall time of the call. */
typedef struct rf_stats {
    uint64_t calls;
    uint64_t prims_i, prims_o;
    uint64_t verts_i, verts_o;
    uint64_t frags_i, frags_o;
    double seconds;
    /* Stats::objs (stats.rs): draws that asked for the bounding-box test
     * (rf_mesh_draw.cull) and those of them that it kept. rf_render leaves
     * them as they are, as render() does. */
    uint64_t objs_i, objs_o;
} rf_stats;

/* A render target: a colour buffer and an optional f32 depth buffer holding
 * 1/w (target.rs:35-44). Pointers are host memory for rf_render and device
 * memory (from rf_framebuf_create) for the batched and serving paths. */
typedef struct rf_target {
    void* color;
    uint32_t width, height;
    uint32_t color_stride;   /* bytes per row */
    uint32_t fmt;            /* rf_pixfmt */
    float* depth;            /* NULL: colour-only target, no depth test */
    uint32_t depth_stride;   /* floats per row */
} rf_target;

/* One render() call with host-resident data: vertices are position (x,y,z)
 * plus an attribute of 0 (flat), 3 (Gouraud, Color3f) or 2 (textured,
 * TexCoord) floats. */
/* Primitive types (the Render impls for Tri<usize> and Edge<usize>,
 * prim.rs:11-60). */
typedef enum rf_prim {
    RF_PRIM_TRI = 0,    /* index triples; tri_fill (raster.rs:185-302)          */
    RF_PRIM_EDGE = 1    /* index pairs; edge clipping (clip.rs:311-348), no
                         * backfaces, raster::line (raster.rs:122-177)          */
} rf_prim;

typedef struct rf_draw {
    uint32_t kind;                  /* rf_shader_kind */
    const uint32_t* tris;           /* n_tris x 3 indices (x 2 for RF_PRIM_EDGE) */
    uint64_t n_tris;                /* primitives */
    const float* pos;               /* n_verts x 3 */
    const float* attr;              /* n_verts x attribute floats */
    uint64_t n_verts;
    float mvp[16];
    float viewport[16];
    uint8_t flat_color[4];          /* RF_SHADER_FLAT: Color4 r, g, b, a */
    const uint8_t* tex;             /* textured: Color4 texels, row-major */
    uint32_t tex_w, tex_h, tex_stride; /* stride in texels */
    uint32_t prim;                  /* rf_prim */
} rf_draw;

typedef struct rf_device rf_device;
typedef struct rf_mesh rf_mesh;
typedef struct rf_texture rf_texture;
typedef struct rf_framebuf rf_framebuf;
typedef struct rf_serve rf_serve;

/* ---- device ----------------------------------------------------------
 * No counterpart in the reference (its render() needs no device): one handle
 * per GPU, holding its stream and the scratch of rf_render. */
rf_status rf_device_open(int ordinal, rf_device** out);
rf_status rf_device_close(rf_device* dev);
const char* rf_status_str(rf_status s);
/* Bytes of device memory the library holds for scratch and residency. */
uint64_t rf_device_bytes(const rf_device* dev);
/* Rasterizer tile width (16, 32 or 64 pixel columns; 0 = the default, 32) used by
 * rf_render on this device. No counterpart in the reference (a tuning knob of
 * this implementation); results are identical for every width. */
rf_status rf_device_set_tile_width(rf_device* dev, uint32_t tile_w);
/* Rasterizer span chunk (6 or 8 pixels a lane applies per round; 0 = the
 * default, 8) used by rf_render and by batches created with a forced tile
 * width. No counterpart in the reference (a tuning knob of this
 * implementation); results are identical for both. */
rf_status rf_device_set_raster_chunk(rf_device* dev, uint32_t chunk);
/* Default rasterizer tile height (rows) of rf_render and of served launches.
 * No counterpart in the reference (an implementation parameter, reported by
 * benchmarks). */
uint32_t rf_tile_height(void);
/* Rasterizer tile height (16, 32 or 64 rows; 0 = the default, 32) used by
 * rf_render and by batches created with a forced tile width on this device.
 * No counterpart in the reference (a tuning knob of this implementation);
 * results are identical for every height. */
rf_status rf_device_set_tile_height(rf_device* dev, uint32_t tile_h);

/* ---- plain path: the drop-in for render() ------------------------------
 * Synchronous, like the reference call: uploads the draw's data and the
 * target, runs the pipeline, reads colour (and depth) back before returning.
 * Replaces render() (core/src/render.rs:134-207) with the built-in shader of
 * `draw->kind`, `uniform` = draw->mvp and `to_screen` = draw->viewport.
 * Stats are added to *stats (may be NULL). */
rf_status rf_render(rf_device* dev, const rf_draw* draw, rf_target* target, const rf_ctx* ctx,
                    rf_stats* stats);

/* rf_render without device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and errors as rf_render, but every device buffer (the
 * draw's vertices, indices and texels, the target, descriptors and pipeline
 * scratch) comes from one static device array of 64 MiB per GPU, held from
 * the first rf_device_open until the process ends; the call makes no
 * hipMalloc or hipFree. A draw whose buffers do not fit the array returns
 * RF_E_BUDGET before any device work (a 1920x1080 RGBA8 target with depth
 * takes 16.6 MB of it); ctx->depth_sort returns RF_E_UNSUPPORTED (its sort
 * needs temporary storage: use rf_render). Calls on one GPU share the
 * array: calls that overlap in time on one GPU (from several threads) race. */
rf_status rf_render_noalloc(rf_device* dev, const rf_draw* draw, rf_target* target, const rf_ctx* ctx,
                            rf_stats* stats);

/* ---- residency ---------------------------------------------------------
 * Device-resident meshes and textures (geom/mesh.rs:24-30, tex.rs:33-37).
 * n_attr is the attribute float count of the kinds the mesh is drawn with. */
rf_status rf_mesh_create(rf_device* dev, const float* pos, const float* attr, uint32_t n_attr,
                         uint64_t n_verts, const uint32_t* tris, uint64_t n_tris, rf_mesh** out);
/* A resident mesh of edges (Edge<usize> index pairs, as debug.rs batches
 * hold): draws over it render lines. */
rf_status rf_mesh_create_edges(rf_device* dev, const float* pos, const float* attr, uint32_t n_attr,
                               uint64_t n_verts, const uint32_t* edges, uint64_t n_edges, rf_mesh** out);
rf_status rf_mesh_free(rf_device* dev, rf_mesh* mesh);
rf_status rf_texture_create(rf_device* dev, const uint8_t* rgba, uint32_t w, uint32_t h,
                            uint32_t stride_texels, rf_texture** out);
/* A sub-texture of a resident texture, sharing its texels and row stride:
 * Atlas::get (tex.rs:157-172) for a Grid layout cell, or any rectangle. The
 * view must be freed before its texture. */
rf_status rf_texture_view(rf_device* dev, const rf_texture* tex, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                          rf_texture** out);
rf_status rf_texture_free(rf_device* dev, rf_texture* tex);
/* A device-resident target; rf_framebuf_target fills an rf_target with its
 * device pointers. */
rf_status rf_framebuf_create(rf_device* dev, uint32_t w, uint32_t h, uint32_t fmt, int with_depth,
                             rf_framebuf** out);
rf_status rf_framebuf_free(rf_device* dev, rf_framebuf* fb);
rf_status rf_framebuf_target(const rf_framebuf* fb, rf_target* out);
rf_status rf_framebuf_read(rf_device* dev, const rf_framebuf* fb, void* color, float* depth);

/* ---- batched path -------------------------------------------------------
 * Many frames of uneven resolution, each an ordered list of draws over
 * resident meshes, go through one pipeline launch. Each draw is one render()
 * call (render.rs:134-207) as the frame loops of Batch::render
 * (batch.rs:127-146) and Camera::render (cam.rs:229-254) make them; the
 * batch itself has no counterpart in the reference. */
typedef struct rf_mesh_draw {
    const rf_mesh* mesh;
    uint32_t kind;
    float mvp[16];
    float viewport[16];
    uint8_t flat_color[4];
    const rf_texture* tex;
    uint64_t first_tri;      /* draw primitives [first_tri, first_tri + n_tris) of the mesh */
    uint64_t n_tris;         /* 0: to the end of the mesh */
    /* 1: whole-object culling on the device, as callers do before render()
     * (crates.rs:118-122): the draw is skipped, and counts in no stats but
     * objs_i, when BBox::visibility(mvp) is Hidden for the mesh's bounding
     * box (scene.rs:37-39, 81-87). 0: always drawn. */
    uint32_t cull;
} rf_mesh_draw;

typedef struct rf_frame {
    rf_framebuf* target;
    rf_ctx ctx;
    const rf_mesh_draw* draws;
    uint32_t n_draws;
} rf_frame;

typedef struct rf_batch rf_batch;
/* Plans a batch: sizes scratch from one eager run, picks the rasterizer's
 * tile shape and span chunk (tile_w = 0: autotune over tile widths 16, 32,
 * 64, tile heights 16, 32, 64 and chunks 6, 8, each timed on the batch; 16,
 * 32 or 64 to force the width, with the device's tile height and chunk, see
 * rf_device_set_tile_height and rf_device_set_raster_chunk) and captures the
 * launch sequence
 * into a HIP graph. The frames' descriptors are copied; the meshes, textures
 * and framebuffers must stay alive while the batch exists. */
rf_status rf_batch_create(rf_device* dev, const rf_frame* frames, uint32_t n_frames,
                          uint32_t tile_w, rf_batch** out);
/* Replays the batch; stats (n_frames entries, may be NULL) are per frame.
 * `sync` = 0 returns without waiting (rf_batch_wait then waits). */
rf_status rf_batch_run(rf_batch* b, int sync);
rf_status rf_batch_wait(rf_batch* b, rf_stats* stats);
rf_status rf_batch_free(rf_batch* b);
uint32_t rf_batch_tile_width(const rf_batch* b);
uint32_t rf_batch_raster_chunk(const rf_batch* b);
uint32_t rf_batch_tile_height(const rf_batch* b);
/* The autotuner's measurements: one entry per candidate it timed (none when
 * the tile width was forced). Copies up to `max` entries and returns their
 * number. */
typedef struct rf_tune_entry {
    uint32_t tile_w, tile_h, chunk;
    float ms;   /* mean pipeline milliseconds of the batch (eager launches) */
} rf_tune_entry;
uint32_t rf_batch_tune_results(const rf_batch* b, rf_tune_entry* out, uint32_t max);
/* Kernel-level timing of the last replay (HIP events on the batch stream):
 * total milliseconds and the rasterizer's share. */
rf_status rf_batch_last_ms(const rf_batch* b, float* total_ms, float* raster_ms);
/* Runs the pipeline `iters` times outside the graph with an event after each
 * stage on the batch stream; ms[0..5] receive the mean milliseconds of:
 * vertex (with counter reset), setup, scan, bin, bin sort, raster. */
rf_status rf_batch_profile(rf_batch* b, int iters, float* ms);

/* ---- serving ------------------------------------------------------------
 * A request loop that keeps meshes and textures resident, admits frames by a
 * device-memory budget and batches what is admitted while earlier work is
 * still on the GPU (continuous batching over `n_streams` streams; a launch
 * groups requests of similar cost). No counterpart in the reference; every
 * request is a frame of render() calls cleared first, as the frontends' loops
 * do (front/src/lib.rs:103-120). rf_serve_open allocates all the loop's
 * memory: the device budget and a pinned read-back block of budget / 4 (64
 * MiB to 1 GiB). A result's pixels occupy that block until rf_serve_poll
 * returns them, so results left unpolled hold back new launches. */
typedef struct rf_request {
    uint64_t id;
    const rf_mesh_draw* draws;   /* copied at submit */
    uint32_t n_draws;
    uint32_t width, height, fmt;
    int32_t with_depth;
    rf_ctx ctx;
} rf_request;

typedef struct rf_result {
    uint64_t id;
    rf_status status;
    uint32_t width, height, fmt;
    double submit_s, start_s, done_s;   /* monotonic clock seconds */
    rf_stats stats;
} rf_result;

rf_status rf_serve_open(rf_device* dev, uint64_t budget_bytes, uint32_t max_batch,
                        uint32_t n_streams, rf_serve** out);
/* Non-blocking; the result is delivered through rf_serve_poll. */
rf_status rf_serve_submit(rf_serve* s, const rf_request* req);
/* Copies up to `max` finished results; with `color` non-NULL, result i's
 * pixels are copied to color[i] (width * height * bpp bytes). Blocks up to
 * timeout_ms for at least one result (0 = do not block). */
rf_status rf_serve_poll(rf_serve* s, rf_result* out, void** color, uint32_t max,
                        uint32_t* n_out, uint32_t timeout_ms);
rf_status rf_serve_close(rf_serve* s);
double rf_now_s(void);

#ifdef __cplusplus
}
#endif

#endif
