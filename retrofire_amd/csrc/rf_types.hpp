// Device-side descriptors shared by the host runtime and the kernels.
//
// A *batch* is one or more frames; a frame is one render target, one context
// and an ordered list of draws; a draw is one `render()` call of the
// reference (core/src/render.rs:134-207): one mesh, one built-in shader kind,
// one uniform (the mvp matrix) and one `to_screen` matrix. Primitive
// submission order -- which decides depth-test ties -- is draw order, then
// triangle order within the draw, then clip-fan order (clip.rs:360-398).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

// The rasterizer's tiles are split into this many contiguous partitions, one
// per XCD (accelerator complex die, each with its own L2) of the MI355X.
constexpr uint32_t TILE_PARTS = 8;


// Shader kinds (the built-in vertex/fragment shader pairs).
enum ShaderKind : uint32_t {
    KIND_FLAT = 0,            // constant Color4, varying ()
    KIND_GOURAUD = 1,         // Color3f varying -> Color3f::to_color4
    KIND_TEX_CLAMP = 2,       // TexCoord varying (z-divided) -> SamplerClamp
    KIND_TEX_REPEAT_POT = 3,  // TexCoord varying (z-divided) -> SamplerRepeatPot
    KIND_TEX_ONCE = 4,        // TexCoord varying (z-divided) -> SamplerOnce
};
constexpr uint32_t KIND_LAST = KIND_TEX_ONCE;

// Primitive types of a draw (Render impls, prim.rs:11-60).
enum PrimType : uint32_t {
    PRIM_TRI = 0,   // Tri<usize>: index triples, tri_fill
    PRIM_EDGE = 1,  // Edge<usize>: index pairs, raster::line
};
// ScrTri::kind bit of a line (Edge) primitive: then v[0] is the line's first
// end (its z and varyings hold for every pixel) and v[1] its pixel walk
// (see k_setup).
constexpr uint32_t KIND_LINE = 0x100;

__host__ __device__ inline int kind_attrs(uint32_t k) {
    return k == KIND_FLAT ? 0 : (k == KIND_GOURAUD ? 3 : 2);
}
__host__ __device__ inline bool kind_persp(uint32_t k) { return k >= KIND_TEX_CLAMP; }

// Colour buffer formats.
enum PixFmt : uint32_t {
    PIX_RGBA8 = 0,      // [u8; 4] r, g, b, a  (Buf2<Color4>, Rgba8888 bytes)
    PIX_RGB8 = 1,       // [u8; 3] r, g, b     (Buf2<Color3>)
    PIX_U32_RGBA = 2,   // u32 0xRRGGBBAA      (Rgba8888 into u32)
    PIX_U32_XRGB = 3,   // u32 0x00RRGGBB      (Xrgb8888)
    PIX_U32_ARGB = 4,   // u32 0xAARRGGBB      (Argb8888 into u32; Bgra8888 bytes)
    PIX_U32_BGRA = 5,   // u32 0xBBGGRRAA      (Bgra8888 into u32; Argb8888 bytes)
    PIX_RGBA4444 = 6,   // u16 rrrrggggbbbbaaaa (Rgba4444)
    PIX_RGB565 = 7,     // u16 rrrrrggggggbbbbb (Rgb565)
};
constexpr uint32_t PIX_LAST = PIX_RGB565;
__host__ __device__ inline uint32_t pixfmt_bytes(uint32_t fmt) {
    return fmt == PIX_RGB8 ? 3u : (fmt >= PIX_RGBA4444 ? 2u : 4u);
}

enum DepthTest : int32_t { DT_NONE = 0, DT_LESS = 1, DT_EQUAL = 2, DT_GREATER = 3 };
enum FaceCull : int32_t { FC_NONE = 0, FC_BACK = 1, FC_FRONT = 2 };

// Error bits raised by kernels (OR-ed into one device word per batch).
enum ErrBits : uint32_t {
    ERRB_INDEX = 1u << 0,      // a vertex index >= vertex count (prim.rs:17-19 panics)
    ERRB_BOUNDS = 1u << 1,     // a span outside the target (target.rs:173-174 panics)
    ERRB_MARGIN = 1u << 2,     // a span left the tiles its triangle was binned to
    ERRB_EXTRA = 1u << 3,      // clip-fan storage exhausted
    ERRB_RUNS = 1u << 4,       // bin-run storage exhausted
    ERRB_NONFINITE = 1u << 5,  // a screen-space vertex is NaN or infinite
    ERRB_NEGY = 1u << 6,       // a screen-space vertex lies above row -0.5
    ERRB_CLIP = 1u << 7,       // clipping produced more than 9 vertices
    ERRB_TEXEL = 1u << 8,      // SamplerOnce indexed outside the texture (buf.rs:731-733 panics)
};

struct DrawDesc {
    const float* pos;       // n_verts x (x, y, z)
    const float* attr;      // n_verts x kind_attrs(kind)
    const uint32_t* tris;   // n_tris x (i0, i1, i2), or x (i0, i1) for edges
    const uint8_t* tex;     // Color4 texels (textured kinds)
    uint32_t n_verts, n_tris;
    uint32_t kind, frame;
    uint32_t tex_w, tex_h, tex_stride;  // stride in texels
    uint32_t flat_rgba;                 // r | g << 8 | b << 16 | a << 24
    uint32_t attr_stride;               // floats per vertex in `attr`
    uint32_t cull;                      // 1: skipped when BBox::visibility(mvp) is Hidden (scene.rs:81-87)
    uint32_t prim;                      // PrimType; n_tris counts primitives of that type
    uint32_t pre;                       // 1: vertex slots already filled (depth-sorted clip fans)
    uint32_t sort;                      // depth sort of this render() call: 0 none, 1 front to back, 2 back to front
    uint32_t in_tris, in_verts;         // a pre draw's render() input counts (prims.i, verts.i), for culled draws
    uint64_t vert_base;                 // first slot in the batch vertex buffers
    uint64_t tri_base;                  // first global triangle index
    float mvp[16];                      // row-major ProjMat3 (the uniform)
    float vp[16];                       // row-major Mat4<Ndc, Screen>
    float bbox[6];                      // the mesh's BBox::of (scene.rs:37-39): low x, y, z, high x, y, z
};

struct FrameDesc {
    uint8_t* color;          // device colour buffer
    float* depth;            // device depth buffer or null (colour-only target)
    uint64_t color_stride;   // bytes per row
    uint64_t depth_stride;   // floats per row
    uint32_t width, height, fmt;
    int32_t depth_test, face_cull, color_write, depth_write;
    int32_t clear_color, clear_depth;  // fused Frame::clear (front/src/lib.rs:103-120)
    uint32_t clear_pixel;              // clear colour already in the target format
    float clear_depth_value;           // 1 / depth_clear
    uint32_t ntx, nty;                 // tile grid
    uint32_t tile_base;                // first global tile index of this frame
    uint32_t pad;
};

// Per-frame statistics (stats.rs:16-40), accumulated by the kernels.
// prims.o is counted in this many partial counters per frame, each on its
// own 128-byte line: atomics on one address serialise, and a frame of 3.2M
// small triangles has 50,000 setup waves adding to it.
constexpr uint32_t PRIM_PARTS = 16;
constexpr uint32_t PRIM_PART_STRIDE = 16;  // u64 words between partial counters

struct FrameStats {
    unsigned long long prims_o, verts_o, frags_i, frags_o;
    unsigned long long hidden_draws, hidden_prims, hidden_verts;  // draws skipped by the bounding-box test
};

// Screen-space triangle as stored by the setup kernel: three vertices of
// (x, y, z, a0, a1, a2) plus the draw index.
struct ScrTri {
    float v[3][6];
    uint32_t draw;
    uint32_t kind;   // the draw's shader kind
};

// A bin run: the triangles of group `group` (64 consecutive global triangle
// indices) that touch one tile, as a bit mask.
struct Run {
    uint32_t group;
    uint32_t multi;  // 1: some triangle of the group has several clip fans (read its header)
    unsigned long long mask;
};

struct BatchDesc {
    const DrawDesc* draws;
    const FrameDesc* frames;
    FrameStats* stats;
    uint32_t n_draws, n_frames;
    uint64_t n_verts;   // total vertex slots
    uint64_t n_tris;    // total input triangles
    uint32_t n_groups;  // ceil(n_tris / 64)
    uint32_t n_tiles;   // total tiles over all frames
    uint32_t tile_w, tile_h;  // rasterizer tile: 16, 32 or 64 columns x 16, 32 or 64 rows
    uint32_t chunk;     // rasterizer span chunk: 6 or 8 pixels (anything else: 8)
    uint32_t sort_tpw;  // tiles per k_binsort wave (1..64; 0: the default, 8)
    // scratch
    float4* cv;          // clip-space position per vertex slot
    float4* ca;          // clip attributes (a0, a1, a2, outcode bits)
    float* sv;           // screen vertex (x, y, z, a0, a1, a2) per slot
    uint32_t* hdr;       // per triangle: surviving fan count | extra base << 3
    ScrTri* scr;         // per triangle: first surviving fan
    ScrTri* extra;       // further fans of clipped triangles
    uint32_t* extra_count;
    uint32_t extra_cap;
    uint32_t* rect;      // per triangle: tile rect (tx0 | ty0 << 16, tx1 | ty1 << 16), 2 words
    uint32_t* run_count; // per tile
    uint32_t* run_off;   // per tile + 1 (exclusive scan)
    uint32_t* run_cursor;
    uint32_t* scan_blk;  // per 1024 tiles: run total, then its offset
    uint32_t* tile_frame;  // per tile: its frame
    uint32_t* tile_list;   // tiles with runs, in tile order
    uint32_t* busy_count;  // their number
    uint32_t clear_in_raster;  // 1: the rasterizer clears the tiles without runs itself
    uint32_t* group_draw;   // per 64-triangle group: the draw of its first triangle
    uint32_t* vblock_draw;  // per 256 vertex slots: the draw of the first
    Run* runs;
    uint32_t runs_cap;
    uint32_t* tile_counter;  // persistent raster work counters, one per tile partition (TILE_PARTS)
    uint32_t* err;
    unsigned long long* prims_part;  // per frame: PRIM_PARTS partial prims.o counters
    uint32_t* draw_vis;  // per draw: 0 if the bounding-box test hid it, else 1
    uint32_t any_cull;   // some draw asks for the bounding-box test
    // depth sort (render.rs:180-183, 209-219): clip fans in clip space, their keys
    uint32_t depth_sort;              // 0 none, 1 front to back, 2 back to front
    uint32_t fan_cap;
    float* fans;                      // per fan: 3 x (x, y, z, w, a0, a1, a2)
    unsigned long long* fan_keys;     // per fan: order key << 32 | triangle << 3 | fan
    uint32_t* fan_vals;               // per fan: its index in `fans`
    uint32_t* fan_count;
    uint32_t* fan_draw;               // per fan: its draw
    uint32_t* fan_cnt;                // per draw: its fans
};
constexpr int FAN_FLOATS = 21;

}  // namespace rf
