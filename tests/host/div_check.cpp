// The traversal's binary32 quotient from a binary64 reciprocal,
// RN32(RN64(n * r)) with r = RN64(1/d) or up to 3 binary64 ulps off it,
// against IEEE binary32 division RN32(n/d), on
// the host (SSE arithmetic is IEEE binary32/binary64, no contraction):
// random operands over the traversal's ranges, operands next to each other
// (quotients near 1), quotients next to binary32 midpoints, powers of two,
// signed zeros and the divisor 1 of a still axis. Prints "div ok" and the
// count, or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static float via_reciprocal(float n, float d, int ulps = 0) {
    volatile double r = 1.0 / (double)d;
    // The device's reciprocal is an estimate refined by Newton steps, within
    // an ulp or so of RN64(1/d): the quotient must not depend on which.
    for (int k = 0; k < (ulps < 0 ? -ulps : ulps); ++k) r = std::nextafter((double)r, ulps < 0 ? -1e300 : 1e300);
    volatile double p = (double)n * r;
    return (float)p;
}
static uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static float val(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

static long checked = 0;
static bool check(float n, float d) {
    volatile float q = n / d;
    ++checked;
    for (int ulps = -3; ulps <= 3; ++ulps) {
        float r = via_reciprocal(n, d, ulps);
        if (bits(q) != bits(r)) {
            std::printf("mismatch n=%.9g (%08x) d=%.9g (%08x): div %.9g (%08x) reciprocal %+d ulps %.9g (%08x)\n", n, bits(n), d,
                        bits(d), (float)q, bits(q), ulps, r, bits(r));
            return false;
        }
    }
    return true;
}

int main(int argc, char **argv) {
    long count = argc > 1 ? std::atol(argv[1]) : 20000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<float> num(-4096.0f, 4096.0f), unit(-1.0f, 1.0f);
    std::uniform_int_distribution<int> expo(-27, 0);
    for (long i = 0; i < count; ++i) {
        float d = unit(rng);
        if (std::fabs(d) <= 1e-8f) continue;
        // Numerators as the traversal forms them: plane - position, planes
        // on a lattice of 1/4/16 voxels, positions anywhere.
        float n = (i & 3) == 0 ? num(rng) : std::ldexp(unit(rng), expo(rng) + 12);
        if (!check(n, d)) return 1;
        // Tiny directions (|d| just above 1e-8).
        float t = std::ldexp(unit(rng), -26);
        if (std::fabs(t) > 1e-8f && !check(n, t)) return 1;
    }
    // The light tree's quotients: a node's offset and radius over the
    // distance to it (at least 1e-5, at most a few thousand voxels).
    std::uniform_real_distribution<float> logd(-16.7f, 11.5f);
    for (long i = 0; i < count / 4; ++i) {
        float d = std::exp2(logd(rng));
        float n = (i & 1) ? num(rng) : std::ldexp(unit(rng), expo(rng));
        if (!check(n, d)) return 1;
    }
    // Quotients next to binary32 midpoints: n = RN32(m * d) for a midpoint m
    // of two neighbouring floats, and its neighbours.
    std::uniform_int_distribution<uint32_t> mant(0, (1u << 23) - 1);
    for (long i = 0; i < count / 4; ++i) {
        float d = unit(rng);
        if (std::fabs(d) <= 1e-8f) continue;
        float lo = val((uint32_t)(127 + expo(rng) + 10) << 23 | mant(rng));
        double mid = ((double)lo + (double)std::nextafter(lo, 1e30f)) * 0.5;
        float n = (float)(mid * (double)d);
        for (int k = -2; k <= 2; ++k) {
            float nk = n;
            for (int s = 0; s < (k < 0 ? -k : k); ++s) nk = std::nextafter(nk, k < 0 ? -1e30f : 1e30f);
            if (!check(nk, d)) return 1;
        }
    }
    // Operands next to each other, powers of two, zeros, the still axis's 1.
    for (long i = 0; i < count / 4; ++i) {
        float n = num(rng);
        float d = std::nextafter(n, (i & 1) ? 1e30f : -1e30f);
        if (std::fabs(d) > 1e-8f && std::fabs(d) <= 4096.0f && !check(n, d)) return 1;
    }
    for (int e = -26; e <= 0; ++e)
        for (float n : {0.0f, -0.0f, 1.0f, -3.0f, 0.001f, 4095.999f, 1e-7f})
            for (float s : {1.0f, -1.0f}) {
                if (!check(n, s * std::ldexp(1.0f, e))) return 1;
                if (!check(n, s * std::ldexp(1.5f, e))) return 1;
            }
    for (float n : {0.0f, -0.0f, 123.25f, -7.5f, 1e-30f}) if (!check(n, 1.0f)) return 1;
    std::printf("div ok %ld\n", checked);
    return 0;
}
