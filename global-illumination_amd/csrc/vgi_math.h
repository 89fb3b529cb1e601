// Device math shared by the lighting kernels: integer hashes, the quasi-random
// sampler, this project's own sin/cos/exp, and the cosine hemisphere.
//
// Every float expression is written as single IEEE binary32 operations in the
// order the reference WGSL evaluates them; the library is compiled with
// -ffp-contract=off and correctly rounded division and square root, so no
// multiply-add is fused and results are reproducible bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgi {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t AIR = 256u;

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float get(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// WGSL min/max/clamp on non-NaN operands.
__device__ __forceinline__ float fmn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmx(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fclamp(float v, float lo, float hi) { return fmn(fmx(v, lo), hi); }
// Correctly rounded square root and division: the library is built with
// -fhip-fp32-correctly-rounded-divide-sqrt, so `/` and llvm.sqrt.f32 are IEEE.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float len3(f3 a) { return fsqrt(dot3(a, a)); }
__device__ __forceinline__ f3 normalize3(f3 a) {
    float inv = 1.0f / fsqrt(dot3(a, a));
    return a * inv;
}
__device__ __forceinline__ float max3(f3 v) { return fmx(v.x, fmx(v.y, v.z)); }

// sampling.rs:5-9 / voxel_traversal.wgsl:24-28 (PCG RXS-M-XS 32)
__device__ __forceinline__ uint32_t pcg(uint32_t value) {
    uint32_t state = value * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ uint32_t hash_voxel(int x, int y, int z) {
    return pcg(pcg(pcg((uint32_t)x) ^ (uint32_t)y) ^ (uint32_t)z);
}

// sin/cos: Cody-Waite reduction by pi/2 in three parts (PIO2_A has 8
// significant bits and PIO2_B 12, so their products with the quadrant are
// exact), then Taylor polynomials on [-pi/4, pi/4]: sin to r^9, cos to r^10.
__device__ __forceinline__ float reduce_pio2(float x, int &quadrant) {
    float q = floorf(x * 0.636619772367581343f + 0.5f);
    quadrant = (int)q;
    float r = x - q * 1.5703125f;
    r = r - q * 4.8387050628662109375e-4f;
    r = r - q * -4.37113882867379e-8f;
    return r;
}
__device__ __forceinline__ float sin_poly(float r) {
    float r2 = r * r;
    float p = 2.75573192239858906526e-6f;
    p = p * r2 + -1.98412698412698412698e-4f;
    p = p * r2 + 8.33333333333333333333e-3f;
    p = p * r2 + -1.66666666666666666667e-1f;
    return r + (r * r2) * p;
}
__device__ __forceinline__ float cos_poly(float r) {
    float r2 = r * r;
    float p = -2.75573192239858906526e-7f;
    p = p * r2 + 2.48015873015873015873e-5f;
    p = p * r2 + -1.38888888888888888889e-3f;
    p = p * r2 + 4.16666666666666666667e-2f;
    p = p * r2 + -0.5f;
    return 1.0f + r2 * p;
}
__device__ __forceinline__ void sincos_own(float x, float &s, float &c) {
    int k;
    float r = reduce_pio2(x, k);
    float sp = sin_poly(r), cp = cos_poly(r);
    switch (k & 3) {
    case 0: s = sp; c = cp; break;
    case 1: s = cp; c = -sp; break;
    case 2: s = -sp; c = -cp; break;
    default: s = -cp; c = sp; break;
    }
}
// exp: Cody & Waite (1980) split of ln2, Taylor to r^7, 2^k from exponent bits.
__device__ __forceinline__ float exp_own(float x) {
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return __int_as_float(0x7f800000);
    float k = floorf(x * 1.44269504088896341f + 0.5f);
    float r = x - k * 0.693359375f;
    r = r - k * -2.12194440054690583e-4f;
    float p = 1.98412698412698412698e-4f;
    p = p * r + 1.38888888888888888889e-3f;
    p = p * r + 8.33333333333333333333e-3f;
    p = p * r + 4.16666666666666666667e-2f;
    p = p * r + 1.66666666666666666667e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int e = (int)k;
    int e1 = e / 2, e2 = e - e1;
    return (p * __int_as_float((127 + e1) << 23)) * __int_as_float((127 + e2) << 23);
}

// voxel_path.wgsl:178-186 (hash-based Owen scramble, Vegdahl's sobol_burley,
// MIT: NOTICE-sobol_burley.txt)
__device__ __forceinline__ uint32_t scramble_reverse(uint32_t input, uint32_t seed) {
    uint32_t value = input;
    value ^= value * 0x3d20adeau;
    value += seed;
    value *= (seed >> 16u) | 1u;
    value ^= value * 0x05526c56u;
    value ^= value * 0x53a22864u;
    return value;
}
// The Halton loop's running scale is fl(...fl(1/b)/b...): the same values for
// every call, so they are folded at compile time (IEEE division in the
// compiler's constant evaluator, identical to the run-time divide), and the
// integer divide and modulo take compile-time divisors.
template <uint32_t B>
struct HaltonScales {
    float s[26];
    constexpr HaltonScales() : s() {
        float v = 1.0f;
        for (int k = 0; k < 26; ++k) {
            v = v / (float)B;
            s[k] = v;
        }
    }
};
template <uint32_t B>
__device__ __forceinline__ float halton(uint32_t sequence) {
    constexpr HaltonScales<B> scales;
    uint32_t index = (sequence & 0x00ffffffu) + 1u;
    float value = 0.0f;
    int k = 0;
    while (index != 0u) {
        value = value + (float)(index % B) * scales.s[k];
        index = index / B;
        ++k;
    }
    return value;
}
__device__ __forceinline__ float radical_inverse(uint32_t dimension, uint32_t sequence) {
    switch (dimension) {
    case 0: return halton<2>(sequence);
    case 1: return halton<3>(sequence);
    case 2: return halton<5>(sequence);
    case 3: return halton<7>(sequence);
    case 4: return halton<11>(sequence);
    case 5: return halton<13>(sequence);
    case 6: return halton<17>(sequence);
    case 7: return halton<19>(sequence);
    case 8: return halton<23>(sequence);
    case 9: return halton<29>(sequence);
    case 10: return halton<31>(sequence);
    default: return halton<37>(sequence);
    }
}
// voxel_path.wgsl:187-213: Halton with Cranley-Patterson rotation below
// dimension 12, hash-Owen-scrambled Sobol pairs from 12 on.
__device__ __forceinline__ float sample_1d(uint32_t stream, uint32_t sequence, uint32_t dimension) {
    if (dimension < 12u) {
        float value = radical_inverse(dimension, sequence);
        float rotation = (float)(pcg(stream ^ pcg(dimension)) >> 9u) / 8388608.0f;
        float x = value + rotation;
        return x - floorf(x);
    }
    uint32_t seed = pcg(stream ^ pcg(dimension / 2u));
    uint32_t index = __brev(scramble_reverse(__brev(sequence), pcg(seed ^ 0x79c68e4au)));
    if ((dimension & 1u) != 0u) {
        index ^= (index >> 1u) & 0x55555555u;
        index ^= (index >> 2u) & 0x33333333u;
        index ^= (index >> 4u) & 0x0f0f0f0fu;
        index ^= (index >> 8u) & 0x00ff00ffu;
        index ^= (index >> 16u) & 0x0000ffffu;
    }
    uint32_t value = __brev(scramble_reverse(index, pcg(seed ^ pcg(dimension))));
    return (float)(value >> 9u) / 8388608.0f;
}
// voxel_path.wgsl:218-224
__device__ __forceinline__ float continuation_sample(uint32_t stream, uint32_t sequence, uint32_t bounce,
                                                     uint32_t dimension) {
    if (bounce < 3u) {
        uint32_t legacy = dimension == 2u ? 8u + bounce : 2u + bounce * 2u + dimension;
        return sample_1d(stream, sequence, legacy);
    }
    return sample_1d(stream, sequence, 16u + (bounce - 3u) * 3u + dimension);
}

// voxel_path.wgsl:226-230
__device__ __forceinline__ f3 face_normal(uint32_t face) {
    float s = (face & 1u) != 0u ? -1.0f : 1.0f;
    uint32_t a = face / 2u;
    return mk(a == 0u ? s : 0.0f, a == 1u ? s : 0.0f, a == 2u ? s : 0.0f);
}
// voxel_traversal.wgsl:17-22
__device__ __forceinline__ uint32_t face_index_of(f3 n) {
    if (n.x != 0.0f) return n.x < 0.0f ? 1u : 0u;
    if (n.y != 0.0f) return n.y < 0.0f ? 3u : 2u;
    if (n.z != 0.0f) return n.z < 0.0f ? 5u : 4u;
    return 6u;
}
// voxel_path.wgsl:231-238
__device__ __forceinline__ f3 hemisphere(f3 n, float u, float v) {
    f3 tangent;
    if (fabsf(n.z) < 0.999f) tangent = normalize3(cross3(n, mk(0.0f, 0.0f, 1.0f)));
    else tangent = normalize3(cross3(n, mk(0.0f, 1.0f, 0.0f)));
    f3 bitangent = cross3(n, tangent);
    float phi = 6.28318530718f * u;
    float sv = fsqrt(v);
    float s1 = fsqrt(1.0f - v);
    float sp, cp;
    sincos_own(phi, sp, cp);
    return tangent * (cp * sv) + bitangent * (sp * sv) + n * s1;
}
// voxel_path.wgsl:326-328
__device__ __forceinline__ float cosine_pdf(f3 normal, f3 direction) {
    return fmx(dot3(normal, direction), 0.0f) / 3.14159265359f;
}

}  // namespace vgi
