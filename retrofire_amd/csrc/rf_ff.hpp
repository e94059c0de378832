// Exact fast-forward of a serial float accumulation.
//
// The reference steps every varying by repeated addition (vary::Iter::next,
// core/src/math/vary.rs:146-154; left/right edges in ScanlineIter::next,
// core/src/render/raster.rs:84-113). Pixel j of a span is therefore j serial
// binary32 additions away from the span start, and a closed form v0 + j*dv
// gives different bits. `ff_add(s, d, n)` returns exactly the value of
//
//     for (i = 0; i < n; ++i) s = s + d;      // round-to-nearest-even
//
// in O(number of binades the sequence crosses) steps instead of O(n).
//
// Why it is exact. Take s > 0 in the binade B = [2^e, 2^(e+1)); s is a
// multiple of u = 2^(e-23), s = M u. If the exact sum s + d lies in B, then
// fl(s + d) is the multiple of u nearest to s + d: s + u * rne(d / u), one
// increment for every such s -- except when d / u = m + 1/2 (a tie), where
// the result is whichever of M + m, M + m + 1 is even. After one tie step
// that rounds on B's grid, M is even, and from then on M + m or M + m + 1
// (whichever is even) is the same choice every step. So: take one serial
// step s0 -> s1; if s0 and s1 are both in B, the rounding of that step was on
// B's grid (a sum below 2^e that rounds up to 2^e lands on an even multiple
// too), so from s1 on the increment inc = fl(s1 + d) - s1 repeats for as long
// as each exact sum s1 + j*inc + d stays in B, and s1 + J*inc is a float
// (see ff_run for how J is found without rounding). A step whose result
// equals its input is a fixed point.
// Negative values are the mirror image, fl(s + d) = -fl(-s + -d). Zero,
// subnormal, huge and non-finite values take single serial steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

// One binade run of ff_add, in float arithmetic only. From s0 take one
// serial step s0 -> s1; if s0, s1 and s2 = fl(s1 + d) share sign and binade
// B = [lo, hi), every further step whose exact sum s1 + j*inc + d lies in B
// adds the same increment inc = s2 - s1 (exact: same binade). The sums are
// linear in j; J, the number of such steps (at most n - 1), is estimated by
// one division and settled by exact tests: y = fmaf(j, inc, s1) is exact
// whenever its value lies in B (and lands outside B otherwise), and y - lo
// and hi - y are floats for y in B (Sterbenz), so "lo <= y + d < hi" is
// decided without rounding. Returns the number of steps taken (1 + J) and
// the value after them; 1 step when the conditions do not hold.
__host__ __device__ inline uint32_t ff_run(float s0, float d, uint32_t n, float& out) {
    const float s1 = s0 + d;
    out = s1;
    if (n < 3) return 1;
    const float s2 = s1 + d;
    uint32_t b0, b1, b2;
    __builtin_memcpy(&b0, &s0, 4);
    __builtin_memcpy(&b1, &s1, 4);
    __builtin_memcpy(&b2, &s2, 4);
    const uint32_t ex = (b1 >> 23) & 0xFFu;
    // same sign and exponent field; magnitudes in [2^-90, 2^100); d finite; a real step
    if (((b0 ^ b1) | (b1 ^ b2)) >= (1u << 23) || ex < 37u || ex > 226u || !(d - d == 0.0f) || s1 == s2) return 1;
    const bool neg = (b1 >> 31) != 0;
    const float a1 = neg ? -s1 : s1, da = neg ? -d : d;
    const float inc = neg ? s1 - s2 : s2 - s1;
    const uint32_t blo = ex << 23;
    float lo;
    __builtin_memcpy(&lo, &blo, 4);
    const float hi = 2.0f * lo;
    const uint32_t m = n - 1;  // steps left after s1
    auto ok = [&](uint32_t j) {
        const float y = fmaf((float)j, inc, a1);
        return y >= lo && y < hi && y - lo >= -da && hi - y > da;
    };
    uint32_t J;
    if (ok(m - 1)) {
        J = ok(0) ? m : 0;  // linear: both ends in B, or none
    } else {
        const float t = inc > 0.0f ? ((hi - a1) - da) / inc : ((a1 - lo) + da) / -inc;
        J = t >= 0.0f ? (t < (float)m ? (uint32_t)t : m - 1) : 0u;
        while (J > 0 && !ok(J - 1)) --J;
        while (J < m && ok(J)) ++J;
    }
    const float r = fmaf((float)J, inc, a1);
    out = neg ? -r : r;
    return 1 + J;
}

// Exact fast-forward: the value of n serial additions s = s + d.
__host__ __device__ inline float ff_add(float s, float d, uint32_t n) {
    while (n > 0) {
        float r;
        const uint32_t k = ff_run(s, d, n, r);
        if (r == s || r != r) return r;  // a fixed point or NaN repeats forever
        s = r;
        n -= k;
    }
    return s;
}

}  // namespace rf
