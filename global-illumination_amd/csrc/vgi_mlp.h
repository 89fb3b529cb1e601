// The learned radiance cache: a 32 -> 64 -> 64 -> 64 -> 64 -> 3 ReLU MLP that
// predicts a face's incident radiance (the quantity the world cache stores,
// voxel_path.wgsl:63-87) from an encoding of its position, normal and albedo.
// BF16 weights and activations on MFMA with FP32 accumulation; FP32 master
// weights with Adam.
//
// Position encoding: a multiresolution hash grid (Mueller et al. 2022,
// "Instant neural graphics primitives"): 8 levels of 2 trainable features,
// level l a grid of res_l cells across the scene's largest extent (res_l =
// floor(16 b^l), b = (extent / 16)^(1/7)), the features trilinearly
// interpolated from the cell's 8 corners; a level whose (res_l + 2)^3 corners
// fit the 2^22-entry table indexes them densely, finer levels by the spatial
// hash x ^ 2654435761 y ^ 805459861 z (mod 2^22). Features 0..15: the grid;
// 16..21: sin and cos of pi u per axis; 22..24: the normal; 25..27: the
// albedo; 28: 1 (the bias); 29..31: 0.
//
// Frequency encoding (MLP_ENCODING_FREQUENCY, the network the 512^3 cave
// configuration states: a 32-dim frequency encoding of position, normal and
// albedo): sin and cos of 2^k pi u per axis for k = 0..3 in place of the
// grid -- features 0..11: k = 1, 2 (axis-major pairs per band), 12..15: k = 3
// for x and y, 29..30: k = 3 for z -- and no hash table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vgi {

constexpr int MLP_IN = 32;
constexpr int MLP_HIDDEN = 64;
constexpr int MLP_OUT = 3;
// Offsets of the five weight matrices (row-major out x in) in the flat
// parameter vector; 14,528 parameters in all, no biases (feature 30 is 1).
constexpr int MLP_W1 = 0;
constexpr int MLP_W2 = MLP_W1 + MLP_HIDDEN * MLP_IN;
constexpr int MLP_W3 = MLP_W2 + MLP_HIDDEN * MLP_HIDDEN;
constexpr int MLP_W4 = MLP_W3 + MLP_HIDDEN * MLP_HIDDEN;
constexpr int MLP_W5 = MLP_W4 + MLP_HIDDEN * MLP_HIDDEN;
constexpr int MLP_PARAMS = MLP_W5 + MLP_OUT * MLP_HIDDEN;
// BF16 MFMA A-operand fragments, one slot per 16-row x 32-k tile (64 lanes x
// 8 values, the v_mfma_f32_16x16x32_bf16 operand): forward W tiles (W1 4x1,
// W2..W4 4x2, W5 1x2 with rows padded to 16), then transposed tiles for the
// backward data pass (W2..W4 4x2, W1^T's first row 1x2: the gradient of the
// hash features; W5^T 4 tiles of 16x16, values 0..3 of a slot, for the
// 16x16x16 step over the 3 outputs).
constexpr int MLP_FWD_SLOTS = 4 + 8 * 3 + 2;
constexpr int MLP_BWD_SLOTS = 8 * 3 + 2 + 4;
constexpr int MLP_SLOTS = MLP_FWD_SLOTS + MLP_BWD_SLOTS;
constexpr int MLP_FRAG_VALUES = MLP_SLOTS * 64 * 8;  // bf16 values in the fragment buffer
// The 16x16 tiles of the weight matrices (W1 4x2, W2..W4 4x4, W5 1x4): the
// weight-gradient pass computes one per wavefront.
constexpr int MLP_GRAD_TILES = 8 + 16 * 3 + 4;

// One query of the cache: the face whose incident radiance is wanted.
struct MlpQuery {
    int32_t vx, vy, vz;
    uint32_t face_material;  // face | material << 8; 0xffffffff: no query
};

constexpr int MLP_HASH_LEVELS = 8;
constexpr int MLP_HASH_FEATURES = 2;
#ifndef VGI_MLP_HASH_LOG2
#define VGI_MLP_HASH_LOG2 22
#endif
constexpr uint32_t MLP_HASH_SIZE = 1u << VGI_MLP_HASH_LOG2;  // entries per level
constexpr uint32_t MLP_TABLE = (uint32_t)MLP_HASH_LEVELS * MLP_HASH_SIZE * MLP_HASH_FEATURES;
// Table entries a training sample touches (8 corners per level).
constexpr uint32_t MLP_TOUCHED = (uint32_t)MLP_HASH_LEVELS * 8u;

constexpr uint32_t MLP_ENCODING_HASH = 0u;
constexpr uint32_t MLP_ENCODING_FREQUENCY = 1u;

struct MlpScene {
    float lo_x, lo_y, lo_z;  // scene minimum corner (voxels)
    float inv_extent;        // 1 / largest scene extent
    uint32_t encoding;       // MLP_ENCODING_HASH or MLP_ENCODING_FREQUENCY
    const float4 *materials; // (diffuse rgb, emission power), (emission rgb, 0)
    const float2 *table;     // level-major, MLP_HASH_SIZE entries of 2 features per level
    uint32_t res[MLP_HASH_LEVELS];
};

// Cells across the largest extent at each level (host and oracle alike).
inline uint32_t mlp_hash_res(int level, int extent) {
    double top = extent > 16 ? (double)extent : 16.0;
    double r = 16.0 * __builtin_exp((double)level * __builtin_log(top / 16.0) / (double)(MLP_HASH_LEVELS - 1));
    return (uint32_t)__builtin_floor(r + 1e-3);
}

// Launchers (vgi_mlp.hip).
void mlp_pack(const float *master, uint16_t *frags, hipStream_t stream);
void mlp_infer(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, uint32_t count, float *pred,
               hipStream_t stream);
// Fused forward + loss + backward-data over `count` samples (count a multiple
// of 16): writes per-layer inputs X_i and output gradients dY_i (bf16) for the
// weight-gradient pass, and the summed loss.
// `table_grad` gets the hash table's gradient added (FP32 atomics) and
// `touched` (count x MLP_TOUCHED words) the entries each sample touched
// (0xffffffff for a sample without a query); hash encoding only.
void mlp_forward_backward(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, const float *targets,
                          uint32_t count, uint16_t *acts, uint16_t *grads, float *loss, float *pred, float *table_grad,
                          uint32_t *touched, hipStream_t stream);
// dW_i = sum over samples of dY_i^T X_i on MFMA (split over sample chunks,
// FP32 atomics into `grad`).
void mlp_weight_grad(const uint16_t *acts, const uint16_t *grads, uint32_t count, float *grad, hipStream_t stream);
void mlp_adam(float *master, float *m, float *v, float *grad, uint32_t step, float lr, hipStream_t stream);
// Lazy Adam over the hash-table entries listed in `touched` (samples x
// MLP_TOUCHED words; eps 1e-15); clears their gradient.
void mlp_table_adam(float *table, float *m, float *v, float *grad, const uint32_t *touched, uint32_t samples, uint32_t step, float lr,
                    hipStream_t stream);

}  // namespace vgi
