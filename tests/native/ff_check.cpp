// Checks rf::ff_add (retrofire_amd/csrc/rf_ff.hpp) against the serial loop it
// replaces, bit for bit, over random and adversarial cases. Built as host
// code by tests/test_ff.py; prints "cases mismatches" and exits 1 on any
// mismatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../retrofire_amd/csrc/rf_ff.hpp"

static float serial(float s, float d, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) s = s + d;
    return s;
}
static uint32_t bits(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return b;
}

int main(int argc, char** argv) {
    long cases = argc > 1 ? atol(argv[1]) : 200000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, total = 0;
    auto check = [&](float s, float d, uint32_t n) {
        const float a = rf::ff_add(s, d, n), b = serial(s, d, n);
        ++total;
        if (bits(a) != bits(b) && !(a != a && b != b)) {
            if (bad < 10) printf("MISMATCH s=%.9g d=%.9g n=%u ff=%.9g serial=%.9g\n", s, d, n, a, b);
            ++bad;
        }
    };
    // 1. random magnitudes and ratios
    for (long i = 0; i < cases; ++i) {
        const double mag = std::pow(10.0, -3.0 + 7.0 * U(rng));
        const double ratio = std::pow(10.0, -9.0 + 10.0 * U(rng));
        float s = (float)(mag * (U(rng) < 0.5 ? -1 : 1));
        float d = (float)(mag * ratio * (U(rng) < 0.5 ? -1 : 1));
        uint32_t n = (uint32_t)(U(rng) < 0.5 ? U(rng) * 64 : U(rng) * 4000);
        check(s, d, n);
    }
    // 2. ties: d = (m + 1/2) * ulp(s) and similar round values
    for (long i = 0; i < cases / 4; ++i) {
        const int e = (int)(U(rng) * 30) - 10;
        const float s = std::ldexp(1.0f + (float)U(rng), e) * (U(rng) < 0.5 ? -1.0f : 1.0f);
        const float u = std::ldexp(1.0f, e - 23);
        const int m = (int)(U(rng) * 40) - 20;
        const float d = ((float)m + 0.5f) * u * (float)(1 << (int)(U(rng) * 3));
        check(s, d, (uint32_t)(U(rng) * 3000));
        check(s, std::ldexp(1.0f, e - 1 - (int)(U(rng) * 6)), (uint32_t)(U(rng) * 3000));
        check(s, 0.5f, (uint32_t)(U(rng) * 3000));
        check(s, -0.25f, (uint32_t)(U(rng) * 3000));
    }
    // 3. starting at binade boundaries, crossing zero, values like pixel
    //    coordinates, colours and 1/w
    for (long i = 0; i < cases / 4; ++i) {
        const int e = (int)(U(rng) * 20) - 8;
        const float b = std::ldexp(1.0f, e);
        const float s = b + (float)(U(rng) - 0.5) * std::ldexp(1.0f, e - 20);
        check(s, -(float)(U(rng) * 0.01), (uint32_t)(U(rng) * 2000));
        check(-s, (float)(U(rng) * 0.01), (uint32_t)(U(rng) * 2000));
        check((float)(U(rng) * 2 - 1), (float)(U(rng) * 0.02 - 0.01), (uint32_t)(U(rng) * 2000));
        check((float)(U(rng) * 1920), (float)(U(rng) * 4 - 2), (uint32_t)(U(rng) * 1100));
        check((float)(U(rng) * 1e-3), (float)(U(rng) * 1e-6), (uint32_t)(U(rng) * 2000));
    }
    // 4. every prefix length of a few sequences
    for (int k = 0; k < 40; ++k) {
        const float s = (float)(U(rng) * 10 - 5), d = (float)((U(rng) - 0.5) * std::pow(10.0, -6 * U(rng)));
        for (uint32_t n = 0; n < 600; ++n) check(s, d, n);
    }
    // 5. sequences that arrive within a few steps of a binade boundary:
    //    s starts J*m + o ulps below 2^(e+1) (or above 2^e, stepping down), so
    //    the step that leaves the binade has an exact sum close to the edge
    for (long i = 0; i < cases / 2; ++i) {
        const int e = (int)(U(rng) * 30) - 10;
        const float lo = std::ldexp(1.0f, e), hi = 2.0f * lo, u = std::ldexp(1.0f, e - 23);
        const int m = 1 + (int)(U(rng) * 2000);
        static const float fr[] = {0.0f, 0.5f, -0.5f, 0.25f, -0.25f, 0.375f, -0.375f, 0.125f};
        const float d = ((float)m + fr[(int)(U(rng) * 8)]) * u;
        const uint32_t J = 1 + (uint32_t)(U(rng) * 3000);
        const float o = (float)(int)(U(rng) * 2 * m);
        const float up = hi - ((float)J * (float)m + o) * u, dn = lo + ((float)J * (float)m + o) * u;
        const uint32_t n = J + (uint32_t)(U(rng) * 12);
        if (up >= lo) {
            check(up, d, n);
            check(-up, -d, n);
        }
        if (dn < hi) {
            check(dn, -d, n);
            check(-dn, d, n);
        }
    }
    // 6. specials
    const float inf = INFINITY, nan = NAN;
    const float sp[] = {0.0f, -0.0f, 1e-40f, -1e-40f, 1e-38f, 3e38f, -3e38f, inf, -inf, nan, 1.0f, -1.0f};
    for (float s : sp)
        for (float d : sp)
            for (uint32_t n : {0u, 1u, 2u, 3u, 4u, 5u, 17u, 100u, 1000u}) check(s, d, n);
    printf("%ld %ld\n", total, bad);
    return bad ? 1 : 0;
}
