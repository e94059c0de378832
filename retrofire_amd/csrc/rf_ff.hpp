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
// as each exact sum s1 + j*inc + d stays in B. The largest such J is found in
// double precision, where every quantity involved is exact (s, d, inc are
// floats within a few binades of each other, J < 2^24), and s1 + J*inc is a
// float. A step whose result equals its input (inc = 0) is a fixed point.
// Negative values are the mirror image, fl(s + d) = -fl(-s + -d). Zero,
// subnormal, huge and non-finite values take single serial steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

__host__ __device__ inline int ff_binade(double a) {
    int e;
    (void)frexp(a, &e);  // a = m * 2^e, m in [0.5, 1)
    return e;
}

// Float-only fast path of ff_add: after one serial step s0 -> s1, if s0, s1
// and s2 = fl(s1 + d) share sign and binade B = [lo, hi) and every exact sum
// s1 + j*inc + d (j = 0 .. n-2) stays in B, the result is s1 + (n-1)*inc,
// which fmaf computes exactly (it is a float of B or hi). The two end checks
// are exact: y - lo and hi - y are floats for y in B (Sterbenz), and
// fmaf(j, inc, s1) is exact whenever its value lies in B. Returns false
// (leaving `out` unset) when the conditions do not hold.
__host__ __device__ inline bool ff_add_fast(float s0, float d, uint32_t n, float& out) {
    const float s1 = s0 + d, s2 = s1 + d;
    uint32_t b0, b1, b2;
    __builtin_memcpy(&b0, &s0, 4);
    __builtin_memcpy(&b1, &s1, 4);
    __builtin_memcpy(&b2, &s2, 4);
    const uint32_t ex = (b1 >> 23) & 0xFFu;
    // same sign and exponent field; magnitudes in [2^-90, 2^100); d finite; a real step
    if (((b0 ^ b1) | (b1 ^ b2)) >= (1u << 23) || ex < 37u || ex > 226u || !(d - d == 0.0f) || s1 == s2)
        return false;
    const bool neg = (b1 >> 31) != 0;
    const float a1 = neg ? -s1 : s1, da = neg ? -d : d;
    const float inc = neg ? s1 - s2 : s2 - s1;  // exact: same binade
    const uint32_t blo = ex << 23;
    float lo;
    __builtin_memcpy(&lo, &blo, 4);
    const float hi = 2.0f * lo;
    const float yl = fmaf((float)(n - 2), inc, a1);  // s_(n-1), exact when in B
    if (!(yl >= lo && yl < hi && yl - lo >= -da && hi - yl > da)) return false;
    const float r = fmaf((float)(n - 1), inc, a1);
    out = neg ? -r : r;
    return true;
}

__host__ __device__ inline float ff_add(float s, float d, uint32_t n) {
    if (n > 3) {
        float r;
        if (ff_add_fast(s, d, n, r)) return r;
    }
    while (n > 0) {
        if (n <= 3) {
            for (; n > 0; --n) s = s + d;
            return s;
        }
        const float s0 = s;
        s = s + d;
        --n;
        const float t = s + d;
        if (t == s) return s;  // fixed point: every further step repeats this one
        if (t != t) return t;  // NaN: every further sum is this NaN
        const float a0 = s0 < 0.0f ? -s0 : s0, a1 = s < 0.0f ? -s : s, at = t < 0.0f ? -t : t;
        // same sign, ordinary magnitudes, finite step
        if (!((s0 < 0.0f) == (s < 0.0f) && (s < 0.0f) == (t < 0.0f)) || !(a0 >= 1.0e-30f) ||
            !(a1 >= 1.0e-30f) || !(at >= 1.0e-30f) || !(a0 < 1.0e37f) || !(a1 < 1.0e37f) ||
            !(at < 1.0e37f) || !(d - d == 0.0f))
            continue;
        const int e = ff_binade((double)a1);
        if (ff_binade((double)a0) != e || ff_binade((double)at) != e) continue;
        const bool neg = s < 0.0f;
        const double ss = neg ? -(double)s : (double)s;
        const double dd = neg ? -(double)d : (double)d;
        const double tt = neg ? -(double)t : (double)t;
        const double lo = ldexp(1.0, e - 1), hi = ldexp(1.0, e);
        const double inc = tt - ss;  // exact, same binade
        // J = number of steps j = 0, 1, ... whose exact sum ss + j*inc + dd
        // lies in [lo, hi)
        auto ok = [&](uint32_t j) {
            const double x = ss + (double)j * inc + dd;
            return x >= lo && x < hi;
        };
        // common case: every remaining step stays in the binade (the exact
        // sums are linear in j, so the first and the last decide)
        if (ok(0) && ok(n - 1)) {
            const double r = ss + (double)n * inc;
            return neg ? -(float)r : (float)r;
        }
        double jm;
        if (inc > 0.0)
            jm = ceil((hi - ss - dd) / inc);
        else
            jm = floor((ss + dd - lo) / -inc) + 1.0;
        if (!(jm >= 0.0)) jm = 0.0;
        if (jm > (double)n) jm = (double)n;
        uint32_t J = (uint32_t)jm;
        while (J > 0 && !ok(J - 1)) --J;  // correct the rounded quotient
        while (J < n && ok(J)) ++J;
        if (J == 0) continue;
        const double r = ss + (double)J * inc;  // a float: a multiple of u in [lo, hi]
        s = neg ? -(float)r : (float)r;
        n -= J;
    }
    return s;
}

}  // namespace rf
