// Learned radiance cache kernels for gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16,
// FP32 accumulate), wave64.
//
// Formulation: each wavefront carries 16 samples and computes Y^T = W X^T, so
// an MFMA's A operand is a weight tile and B the activations. The 16x16
// accumulator layout (lane l: column n = l % 16, rows 4 (l / 16) + j, j =
// 0..3) of two consecutive output tiles 2s and 2s + 1 is used, unchanged, as
// the 8-value B operand of the next layer's 32-wide k-step s: lane group g =
// l / 16 holds logical k = 32 s + 4 g + j (j < 4, tile 2s) and 32 s + 16 + 4 g
// + j - 4 (j >= 4, tile 2s + 1). An MFMA multiplies the A and B values at
// the same (lane group, j) position, so the weight fragments are packed with
// the same k order and activations chain through the five layers in
// registers. Weights sit in LDS as pre-packed fragments.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "vgi_math.h"
#include "vgi_mlp.h"

namespace vgi {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even float -> bf16 bits (no NaNs arise here).
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// 16x16x16 (k = 4 (l / 16) + j): only the 3-output W5^T step of the backward pass.
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// 16x16x32: the CDNA4 double-k BF16 MFMA.
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
// Two 4-value fragments as the 8-value B operand of one 32-wide k-step.
__device__ __forceinline__ s16x8 join(s16x4 lo, s16x4 hi) { return s16x8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w}; }
__device__ __forceinline__ s16x4 low4(s16x8 v) { return s16x4{v.s0, v.s1, v.s2, v.s3}; }
__device__ __forceinline__ s16x4 pack4(float a, float b, float c, float d) {
    s16x4 r;
    r.x = (short)to_bf16(a);
    r.y = (short)to_bf16(b);
    r.z = (short)to_bf16(c);
    r.w = (short)to_bf16(d);
    return r;
}

// Fragment slots (see vgi_mlp.h).
__device__ __host__ constexpr int fwd_slot(int layer, int t, int s) {
    return layer == 1 ? t : (layer <= 4 ? 4 + 8 * (layer - 2) + 2 * t + s : 28 + s);
}
__device__ __host__ constexpr int bwd_slot(int layer, int t, int s) {
    return MLP_FWD_SLOTS + (layer == 1 ? 24 + s : layer <= 4 ? 8 * (layer - 2) + 2 * t + s : 26 + t);
}
// The logical k (input index) of lane group g, value j of 32-wide k-step s.
__device__ __host__ constexpr int chained_k(int g, int j, int s) { return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }

// Packs FP32 master weights into BF16 A-operand slots: forward slot (layer,
// t, s) lane l = m + 16 g value j = W[16 t + m][chained_k(g, j, s)];
// transposed slot (layer, t, s) = W[chained_k(g, j, s)][16 t + m]; the W5^T
// slots (16x16x16) value j < 4 = W5[4 g + j][16 t + m]. Entries beyond a
// matrix are zero.
__global__ void mlp_pack_kernel(const float *__restrict__ w, uint16_t *__restrict__ frags) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MLP_FRAG_VALUES) return;
    int slot = e >> 9, lane = (e >> 3) & 63, j = e & 7;
    int m = lane & 15, g = lane >> 4;
    int layer, t, s;
    bool transposed = slot >= MLP_FWD_SLOTS;
    int local = transposed ? slot - MLP_FWD_SLOTS : slot;
    int row, col;
    if (!transposed) {
        if (local < 4) { layer = 1; t = local; s = 0; }
        else if (local < 28) { layer = 2 + (local - 4) / 8; t = ((local - 4) % 8) / 2; s = (local - 4) % 2; }
        else { layer = 5; t = 0; s = local - 28; }
        row = 16 * t + m;
        col = chained_k(g, j, s);
    } else if (local < 26) {
        if (local < 24) { layer = 2 + local / 8; t = (local % 8) / 2; s = local % 2; }
        else { layer = 1; t = 0; s = local - 24; }
        row = chained_k(g, j, s);
        col = 16 * t + m;
    } else {
        layer = 5; t = local - 26;
        row = j < 4 ? 4 * g + j : 1 << 20;
        col = 16 * t + m;
    }
    int in = layer == 1 ? MLP_IN : MLP_HIDDEN;
    int out = layer == 5 ? MLP_OUT : MLP_HIDDEN;
    int base = layer == 1 ? MLP_W1 : layer == 2 ? MLP_W2 : layer == 3 ? MLP_W3 : layer == 4 ? MLP_W4 : MLP_W5;
    float v = (row < out && col < in) ? w[base + row * in + col] : 0.0f;
    frags[e] = to_bf16(v);
}

// The face centre in scene units (0..1 across the largest extent).
__device__ __forceinline__ f3 query_position(const MlpScene &sc, const MlpQuery &q, f3 n) {
    return mk(((float)q.vx + 0.5f + n.x * 0.5f - sc.lo_x) * sc.inv_extent, ((float)q.vy + 0.5f + n.y * 0.5f - sc.lo_y) * sc.inv_extent,
              ((float)q.vz + 0.5f + n.z * 0.5f - sc.lo_z) * sc.inv_extent);
}

// One level of the hash grid at position u: the 8 corner entries and
// trilinear weights (corner c = x + 2y + 4z bits), in the order the features
// sum them.
struct HashCorners {
    uint32_t index[8];
    float weight[8];
};
__device__ __forceinline__ HashCorners hash_corners(uint32_t res, f3 u) {
    HashCorners h;
    float px = u.x * (float)res, py = u.y * (float)res, pz = u.z * (float)res;
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    float wx = px - fx, wy = py - fy, wz = pz - fz;
    uint32_t x0 = (uint32_t)(int)fx, y0 = (uint32_t)(int)fy, z0 = (uint32_t)(int)fz;
    const uint32_t side = res + 2u;
    const bool dense = side * side * side <= MLP_HASH_SIZE;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        uint32_t x = x0 + (uint32_t)(c & 1), y = y0 + (uint32_t)((c >> 1) & 1), z = z0 + (uint32_t)(c >> 2);
        uint32_t i = dense ? x + side * (y + side * z) : x ^ (y * 2654435761u) ^ (z * 805459861u);
        h.index[c] = i & (MLP_HASH_SIZE - 1u);
        h.weight[c] = ((c & 1) ? wx : 1.0f - wx) * (((c >> 1) & 1) ? wy : 1.0f - wy) * ((c >> 2) ? wz : 1.0f - wz);
    }
    return h;
}

// A level's two features: the trilinear interpolation of its corners.
__device__ __forceinline__ float2 hash_level(const MlpScene &sc, int level, f3 u) {
    HashCorners h = hash_corners(sc.res[level], u);
    const float2 *t = sc.table + (size_t)level * MLP_HASH_SIZE;
    float a = 0.0f, b = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float2 e = t[h.index[c]];
        a = a + h.weight[c] * e.x;
        b = b + h.weight[c] * e.y;
    }
    return make_float2(a, b);
}

// sin (cosine = false) or cos of 2^band pi u along `axis`.
__device__ __forceinline__ float frequency_feature(f3 u, int axis, int band, bool cosine) {
    float x = get(u, axis) * (float)(1 << band) * 3.14159265358979f;
    float sv, cv;
    sincos_own(x, sv, cv);
    return cosine ? cv : sv;
}

// Feature f < 16 of the frequency encoding: bands 1 and 2 for every axis,
// then band 3 for x and y.
__device__ __forceinline__ float frequency_low(f3 u, int f) {
    if (f < 12) return frequency_feature(u, (f % 6) >> 1, 1 + f / 6, (f & 1) != 0);
    return frequency_feature(u, (f - 12) >> 1, 3, (f & 1) != 0);
}

// Features 16..31 of a query: sin and cos of pi u per axis, the normal, the
// albedo, the constant (and band 3 of z in 29..30 with the frequency encoding).
__device__ __forceinline__ float encode_tail(const MlpScene &sc, const MlpQuery &q, int f) {
    uint32_t face = q.face_material & 0xffu, material = (q.face_material >> 8) & 0xffu;
    f3 n = face_normal(face);
    if (f < 22) return frequency_feature(query_position(sc, q, n), (f - 16) >> 1, 0, (f & 1) != 0);
    if ((f == 29 || f == 30) && sc.encoding == MLP_ENCODING_FREQUENCY)
        return frequency_feature(query_position(sc, q, n), 2, 3, f == 30);
    if (f < 25) return get(n, f - 22);
    if (f < 28) {
        float4 d = sc.materials[material * 2u];
        return f == 25 ? d.x : f == 26 ? d.y : d.z;
    }
    return f == 28 ? 1.0f : 0.0f;
}

// Forward of one 16-sample tile held by a wavefront. x0: input fragments;
// h: per-layer ReLU masks (bit (t*4+j) of masks[layer-1]); returns the output
// accumulator (lanes 0..15 hold outputs 0..2 of their sample in j = 0..2).
__device__ __forceinline__ f32x4 mlp_forward_tile(const s16x8 *__restrict__ lds, s16x4 x0[2], uint32_t masks[4], s16x4 acts[4][4]) {
    const int lane = threadIdx.x & 63;
    f32x4 acc[4];
    // layer 1 (K = 32: one k-step, features 4g..4g+3 and 16+4g..16+4g+3)
    const s16x8 b0 = join(x0[0], x0[1]);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma32(lds[fwd_slot(1, t, 0) * 64 + lane], b0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    s16x4 h[4];
#pragma unroll
    for (int layer = 2; layer <= 5; ++layer) {
        uint32_t mask = 0u;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float a = fmaxf(acc[t].x, 0.0f), b = fmaxf(acc[t].y, 0.0f), c = fmaxf(acc[t].z, 0.0f), d = fmaxf(acc[t].w, 0.0f);
            mask |= (acc[t].x > 0.0f ? 1u : 0u) << (t * 4) | (acc[t].y > 0.0f ? 1u : 0u) << (t * 4 + 1) |
                    (acc[t].z > 0.0f ? 1u : 0u) << (t * 4 + 2) | (acc[t].w > 0.0f ? 1u : 0u) << (t * 4 + 3);
            h[t] = pack4(a, b, c, d);
            acts[layer - 2][t] = h[t];
        }
        masks[layer - 2] = mask;
        const s16x8 k0 = join(h[0], h[1]), k1 = join(h[2], h[3]);
        if (layer < 5) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = mfma32(lds[fwd_slot(layer, t, 0) * 64 + lane], k0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                acc[t] = mfma32(lds[fwd_slot(layer, t, 1) * 64 + lane], k1, acc[t]);
            }
        } else {
            acc[0] = mfma32(lds[fwd_slot(5, 0, 0) * 64 + lane], k0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            acc[0] = mfma32(lds[fwd_slot(5, 0, 1) * 64 + lane], k1, acc[0]);
        }
    }
    return acc[0];
}

// The lane group g of a wavefront's 16-sample tile holds input rows 4g..4g+3
// (hash levels 2g and 2g + 1) and 16 + 4g..16 + 4g + 3.
__device__ __forceinline__ void load_input(const MlpScene &sc, const MlpQuery &q, bool valid, s16x4 x0[2]) {
    const int g = (threadIdx.x & 63) >> 4;
    float2 a = make_float2(0.0f, 0.0f), b = make_float2(0.0f, 0.0f);
    if (valid) {
        f3 u = query_position(sc, q, face_normal(q.face_material & 0xffu));
        if (sc.encoding == MLP_ENCODING_FREQUENCY) {
            a = make_float2(frequency_low(u, 4 * g), frequency_low(u, 4 * g + 1));
            b = make_float2(frequency_low(u, 4 * g + 2), frequency_low(u, 4 * g + 3));
        } else {
            a = hash_level(sc, 2 * g, u);
            b = hash_level(sc, 2 * g + 1, u);
        }
    }
    x0[0] = pack4(a.x, a.y, b.x, b.y);
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = valid ? encode_tail(sc, q, 16 + 4 * g + j) : 0.0f;
    x0[1] = pack4(v[0], v[1], v[2], v[3]);
}

// Adds dL/d(feature) of a lane's two hash levels (rows 4g..4g+3 of dX0) into
// the table gradient, spread over each level's corners by their weights, and
// records the entries it touched (level * MLP_HASH_SIZE + index, 8 per level)
// in the sample's row of MLP_TOUCHED words for the table's Adam step.
__device__ __forceinline__ void scatter_hash_grad(const MlpScene &sc, const MlpQuery &q, f32x4 dx, float *__restrict__ table_grad,
                                                  uint32_t *__restrict__ touched) {
    const int g = (threadIdx.x & 63) >> 4;
    f3 u = query_position(sc, q, face_normal(q.face_material & 0xffu));
    float d[4] = {dx.x, dx.y, dx.z, dx.w};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int level = 2 * g + k;
        HashCorners h = hash_corners(sc.res[level], u);
        float *t = table_grad + (size_t)level * MLP_HASH_SIZE * 2u;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            atomicAdd(&t[h.index[c] * 2u], h.weight[c] * d[2 * k]);
            atomicAdd(&t[h.index[c] * 2u + 1u], h.weight[c] * d[2 * k + 1]);
            touched[level * 8 + c] = (uint32_t)level * MLP_HASH_SIZE + h.index[c];
        }
    }
}

__device__ __forceinline__ void stage_weights(const uint16_t *__restrict__ frags, s16x8 *lds, int slots) {
    const s16x8 *src = reinterpret_cast<const s16x8 *>(frags);
    for (int i = threadIdx.x; i < slots * 64; i += blockDim.x) lds[i] = src[i];
    __syncthreads();
}

// Inference over `count` queries: persistent blocks loop over 16-query tiles
// (one per wavefront); invalid queries give 0.
__global__ __launch_bounds__(256) void mlp_infer_kernel(const uint16_t *__restrict__ frags, MlpScene sc,
                                                        const MlpQuery *__restrict__ queries, uint32_t count,
                                                        float *__restrict__ pred) {
    __shared__ s16x8 lds[MLP_FWD_SLOTS * 64];
    stage_weights(frags, lds, MLP_FWD_SLOTS);
    const int lane = threadIdx.x & 63;
    const uint32_t tiles = (count + 15u) / 16u;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    for (uint32_t tile = wave; tile < tiles; tile += waves) {
        uint32_t n = tile * 16u + (uint32_t)(lane & 15);
        MlpQuery q = {0, 0, 0, NONE};
        if (n < count) q = queries[n];
        bool valid = n < count && q.face_material != NONE;
        s16x4 x0[2];
        load_input(sc, q, valid, x0);
        uint32_t masks[4];
        s16x4 acts[4][4];
        f32x4 out = mlp_forward_tile(lds, x0, masks, acts);
        if (lane < 16 && n < count) {
            float *p = pred + (size_t)n * 3u;
            p[0] = valid ? out.x : 0.0f;
            p[1] = valid ? out.y : 0.0f;
            p[2] = valid ? out.z : 0.0f;
        }
    }
}

// Training: forward, relative L2 loss (pred - y)^2 / (pred^2 + 0.01) with the
// denominator held constant, and the backward data pass, one tile per wave.
// Saves X_i (layer inputs, bf16, sample-major rows) and dY_i (layer output
// gradients, bf16) for the weight-gradient GEMM.
// acts layout: layer 1 [n][32] at 0, layers 2..5 [n][64] at 32N + 64N (i-2);
// grads layout: layers 1..4 [n][64] at 64N (i-1), layer 5 [n][16] at 256N.
__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(const uint16_t *__restrict__ frags, MlpScene sc,
                                                          const MlpQuery *__restrict__ queries, const float *__restrict__ targets,
                                                          uint32_t count, uint16_t *__restrict__ acts_out,
                                                          uint16_t *__restrict__ grads_out, float *__restrict__ loss,
                                                          float *__restrict__ pred, float *__restrict__ table_grad,
                                                          uint32_t *__restrict__ touched) {
    __shared__ s16x8 lds[MLP_SLOTS * 64];
    stage_weights(frags, lds, MLP_SLOTS);
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const uint32_t tile = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    if (tile * 16u >= count) return;
    const uint32_t n = tile * 16u + (uint32_t)col;
    const size_t N = count;
    MlpQuery q = queries[n];
    bool valid = q.face_material != NONE;
    s16x4 x0[2];
    load_input(sc, q, valid, x0);
    uint32_t masks[4];
    s16x4 acts[4][4];
    f32x4 out = mlp_forward_tile(lds, x0, masks, acts);
    // Save layer inputs (8 bytes per lane per 16x16 fragment row group).
    for (int s = 0; s < 2; ++s) *reinterpret_cast<s16x4 *>(acts_out + (size_t)n * 32 + 16 * s + 4 * g) = x0[s];
    for (int layer = 2; layer <= 5; ++layer)
        for (int t = 0; t < 4; ++t)
            *reinterpret_cast<s16x4 *>(acts_out + 32 * N + 64 * N * (layer - 2) + (size_t)n * 64 + 16 * t + 4 * g) = acts[layer - 2][t];
    // Loss gradient on lanes 0..15 (rows 0..2 of the output tile).
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, l = 0.0f;
    if (g == 0) {
        const float *y = targets + (size_t)n * 4u;
        float w = y[3] / (float)count;
        float p[3] = {out.x, out.y, out.z};
        float dd[3];
        for (int c = 0; c < 3; ++c) {
            float den = p[c] * p[c] + 0.01f;
            float r = p[c] - y[c];
            l += w * r * r / den;
            dd[c] = w * 2.0f * r / den;
        }
        d0 = dd[0]; d1 = dd[1]; d2 = dd[2];
        float *pp = pred + (size_t)n * 3u;
        pp[0] = p[0]; pp[1] = p[1]; pp[2] = p[2];
    }
    // Wave sum of the loss, one atomic per wave.
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (lane == 0) atomicAdd(loss, l);
    s16x4 dy5 = pack4(d0, d1, d2, 0.0f);  // rows 4g+j: nonzero only for g = 0
    *reinterpret_cast<s16x4 *>(grads_out + 256 * N + (size_t)n * 16 + 4 * g) = dy5;
    // dH4^T = W5^T dY5^T (k = the 3 outputs: one 16x16x16 step), masked by layer 4's ReLU.
    s16x4 dy[4];
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = mfma16(low4(lds[bwd_slot(5, t, 0) * 64 + lane]), dy5, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    for (int layer = 4; layer >= 1; --layer) {
        uint32_t mask = masks[layer - 1];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float a = (mask >> (t * 4)) & 1u ? acc[t].x : 0.0f;
            float b = (mask >> (t * 4 + 1)) & 1u ? acc[t].y : 0.0f;
            float c = (mask >> (t * 4 + 2)) & 1u ? acc[t].z : 0.0f;
            float d = (mask >> (t * 4 + 3)) & 1u ? acc[t].w : 0.0f;
            dy[t] = pack4(a, b, c, d);
            *reinterpret_cast<s16x4 *>(grads_out + 64 * N * (layer - 1) + (size_t)n * 64 + 16 * t + 4 * g) = dy[t];
        }
        const s16x8 k0 = join(dy[0], dy[1]), k1 = join(dy[2], dy[3]);
        if (layer > 1) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                acc[t] = mfma32(lds[bwd_slot(layer, t, 0) * 64 + lane], k0, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                acc[t] = mfma32(lds[bwd_slot(layer, t, 1) * 64 + lane], k1, acc[t]);
            }
        }
    }
    // dX0 rows 0..15 = (W1^T)[0..15] dY1: the hash features' gradient, lane
    // group g holding the rows of the levels it encoded.
    f32x4 dx = mfma32(lds[bwd_slot(1, 0, 0) * 64 + lane], join(dy[0], dy[1]), f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    dx = mfma32(lds[bwd_slot(1, 0, 1) * 64 + lane], join(dy[2], dy[3]), dx);
    if (sc.encoding == MLP_ENCODING_HASH) {
        uint32_t *row = touched + (size_t)n * MLP_TOUCHED;
        if (valid) {
            scatter_hash_grad(sc, q, dx, table_grad, row);
        } else {
            for (int k = 0; k < 16; ++k) row[16 * g + k] = NONE;
        }
    }
}

// dW_i[o][in] = sum_n dY_i[n][o] X_i[n][in]: one wavefront per (16x16 weight
// tile, chunk of 1024 samples); A = dY^T (o x samples), B = X (samples x
// in), 32 samples per 16x16x32 MFMA (lane group g, value j: sample 8 g + j).
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const uint16_t *__restrict__ acts, const uint16_t *__restrict__ grads,
                                                        uint32_t count, float *__restrict__ grad) {
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const uint32_t wave = blockIdx.x * (blockDim.x / 64u) + (threadIdx.x >> 6);
    const uint32_t chunks = (count + 1023u) / 1024u;
    const uint32_t tile = wave / chunks, chunk = wave % chunks;
    if (tile >= (uint32_t)MLP_GRAD_TILES) return;
    int layer, t, s;
    if (tile < 8) { layer = 1; t = tile / 2; s = tile % 2; }
    else if (tile < 56) { layer = 2 + (tile - 8) / 16; t = ((tile - 8) % 16) / 4; s = (tile - 8) % 4; }
    else { layer = 5; t = 0; s = tile - 56; }
    const size_t N = count;
    const int in = layer == 1 ? MLP_IN : MLP_HIDDEN;
    const int gw = layer == 5 ? 16 : 64;  // row length of dY_i
    const uint16_t *X = acts + (layer == 1 ? 0 : 32 * N + 64 * N * (layer - 2));
    const uint16_t *DY = grads + (layer == 5 ? 256 * N : 64 * N * (layer - 1));
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t begin = chunk * 1024u, end = min(count, begin + 1024u);
    for (uint32_t n0 = begin; n0 < end; n0 += 32u) {
        s16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            size_t n = n0 + 8 * g + j;
            bool in_range = n < end;
            a[j] = in_range ? (short)DY[n * gw + 16 * t + col] : (short)0;
            b[j] = in_range ? (short)X[n * in + 16 * s + col] : (short)0;
        }
        acc = mfma32(a, b, acc);
    }
    // C layout: lane holds rows o = 16t + 4g + j, column in = 16s + col.
    const int out = layer == 5 ? MLP_OUT : MLP_HIDDEN;
    const int base = layer == 1 ? MLP_W1 : layer == 2 ? MLP_W2 : layer == 3 ? MLP_W3 : layer == 4 ? MLP_W4 : MLP_W5;
    float v[4] = {acc.x, acc.y, acc.z, acc.w};
    for (int j = 0; j < 4; ++j) {
        int o = 16 * t + 4 * g + j;
        if (o < out) atomicAdd(&grad[base + o * in + 16 * s + col], v[j]);
    }
}

// Adam (Kingma & Ba 2015), beta1 0.9, beta2 0.99, eps 1e-10; clears the
// gradient.
__global__ void mlp_adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, float *__restrict__ g,
                                float lr, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_PARAMS) return;
    float gi = g[i];
    float mi = 0.9f * m[i] + 0.1f * gi;
    float vi = 0.99f * v[i] + 0.01f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-10f);
    g[i] = 0.0f;
}

// Adam over the hash table (eps 1e-15, as the hash-grid encoding is trained
// in its paper), lazily: only entries the step's samples touched, each once
// (the first of its duplicates in the touched list takes the gradient and
// clears it; the rest read 0 and skip); untouched entries keep their moments.
// Its cost follows the samples, not the table's size.
__global__ void mlp_table_adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, float *__restrict__ g,
                                      const uint32_t *__restrict__ touched, uint32_t n, float lr, float c1, float c2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t entry = touched[i];
    if (entry == NONE) return;
#pragma unroll
    for (uint32_t f = 0; f < 2u; ++f) {
        size_t k = (size_t)entry * 2u + f;
        float gi = atomicExch(&g[k], 0.0f);
        if (gi == 0.0f) continue;
        float mi = 0.9f * m[k] + 0.1f * gi;
        float vi = 0.99f * v[k] + 0.01f * gi * gi;
        m[k] = mi;
        v[k] = vi;
        w[k] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-15f);
    }
}

void mlp_table_adam(float *table, float *m, float *v, float *grad, const uint32_t *touched, uint32_t samples, uint32_t step, float lr,
                    hipStream_t stream) {
    if (samples == 0) return;
    float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.99f, (float)step);
    uint32_t n = samples * (uint32_t)MLP_TOUCHED;
    hipLaunchKernelGGL(mlp_table_adam_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, table, m, v, grad, touched, n, lr, c1, c2);
}

void mlp_pack(const float *master, uint16_t *frags, hipStream_t stream) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((MLP_FRAG_VALUES + 255) / 256), dim3(256), 0, stream, master, frags);
}
void mlp_infer(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, uint32_t count, float *pred,
               hipStream_t stream) {
    if (count == 0) return;
    uint32_t tiles = (count + 15u) / 16u;
    uint32_t blocks = std::min<uint32_t>((tiles + 3u) / 4u, 4096u);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(256), 0, stream, frags, scene, queries, count, pred);
}
void mlp_forward_backward(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, const float *targets,
                          uint32_t count, uint16_t *acts, uint16_t *grads, float *loss, float *pred, float *table_grad,
                          uint32_t *touched, hipStream_t stream) {
    if (count == 0) return;
    uint32_t tiles = count / 16u;
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3((tiles + 3u) / 4u), dim3(256), 0, stream, frags, scene, queries, targets, count,
                       acts, grads, loss, pred, table_grad, touched);
}
void mlp_weight_grad(const uint16_t *acts, const uint16_t *grads, uint32_t count, float *grad, hipStream_t stream) {
    if (count == 0) return;
    uint32_t chunks = (count + 1023u) / 1024u;
    uint32_t waves = (uint32_t)MLP_GRAD_TILES * chunks;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((waves + 3u) / 4u), dim3(256), 0, stream, acts, grads, count, grad);
}
void mlp_adam(float *master, float *m, float *v, float *grad, uint32_t step, float lr, hipStream_t stream) {
    float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.99f, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, stream, master, m, v, grad, lr, c1, c2);
}

}  // namespace vgi
