// Scalar numerics of the render path, written once for host and device.
//
// The translation unit must be compiled with -ffp-contract=off and without
// fast-math: every expression below is a sequence of individually rounded
// IEEE-754 binary32 operations, in the order the reference performs them,
// so the GPU produces the reference's bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

#define RF_HD __host__ __device__ __forceinline__

RF_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
RF_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// Rust `as` casts from f32 (truncate toward zero, saturate, NaN -> 0).
RF_HD uint32_t as_u32(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
RF_HD uint64_t as_usize(float x) {
    if (!(x > 0.0f)) return 0ull;
    if (x >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)x;
}
RF_HD int32_t as_i32(float x) {
    if (x != x) return 0;
    if (x <= -2147483648.0f) return INT32_MIN;
    if (x >= 2147483648.0f) return INT32_MAX;
    return (int32_t)x;
}
RF_HD uint32_t as_u8(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 255.0f) return 255u;
    return (uint32_t)x;
}

// f32::total_cmp key (signed integer order == total order).
RF_HD int32_t total_key(float x) {
    int32_t b = (int32_t)f2u(x);
    return b ^ (int32_t)(((uint32_t)(b >> 31)) >> 1);
}

// round_up_to_half (raster.rs:305-307).
RF_HD float round_up_to_half(float x) { return floorf(x + 0.5f) + 0.5f; }

// `dot` as a left fold from +0.0 (vec.rs:231-238), 4 terms.
RF_HD float dot4(const float* a, float b0, float b1, float b2, float b3) {
    float acc = 0.0f;
    acc = acc + a[0] * b0;
    acc = acc + a[1] * b1;
    acc = acc + a[2] * b2;
    acc = acc + a[3] * b3;
    return acc;
}

// Signed distance to frustum plane p (clip.rs:101-103 over PLANES, clip.rs:215-222),
// written with the plane constants so the operation sequence is the dot's.
RF_HD float plane_dist(int p, float x, float y, float z, float w) {
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    switch (p) {
        case 0: nz = -1.0f; break;  // near
        case 1: nz = 1.0f; break;   // far
        case 2: nx = -1.0f; break;  // left
        case 3: nx = 1.0f; break;   // right
        case 4: ny = -1.0f; break;  // bottom
        default: ny = 1.0f; break;  // top
    }
    float acc = 0.0f;
    acc = acc + nx * x;
    acc = acc + ny * y;
    acc = acc + nz * z;
    acc = acc + (-1.0f) * w;
    return acc;
}

// view_frustum::outcode (clip.rs:240-242).
RF_HD uint32_t outcode(float x, float y, float z, float w) {
    uint32_t c = 0;
    for (int p = 0; p < 6; ++p) c |= (plane_dist(p, x, y, z, w) > 0.0f) ? (1u << p) : 0u;
    return c;
}

// Colour of a fragment in the target pixel format from Color4 bytes.
RF_HD uint32_t pack_pixel(uint32_t fmt, uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    switch (fmt) {
        case 0: return r | (g << 8) | (b << 16) | (a << 24);  // [u8;4] rgba
        case 1: return r | (g << 8) | (b << 16);              // [u8;3] rgb
        case 2: return (r << 24) | (g << 16) | (b << 8) | a;  // u32 0xRRGGBBAA
        case 3: return (r << 16) | (g << 8) | b;              // u32 0x00RRGGBB
        case 4: return (a << 24) | (r << 16) | (g << 8) | b;  // u32 0xAARRGGBB
        case 5: return (b << 24) | (g << 16) | (r << 8) | a;  // u32 0xBBGGRRAA
        case 6: return ((r >> 4) << 12) | ((g >> 4) << 8) | ((b >> 4) << 4) | (a >> 4);  // u16 4444
        default: return (((r >> 3) & 0x1F) << 11) | (((g >> 2) & 0x3F) << 5) | ((b >> 3) & 0x1F);  // u16 565
    }
}

}  // namespace rf
