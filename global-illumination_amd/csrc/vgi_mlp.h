// The learned radiance cache: a 32 -> 64 -> 64 -> 64 -> 64 -> 3 ReLU MLP that
// predicts a face's incident radiance (the quantity the world cache stores,
// voxel_path.wgsl:63-87) from an encoding of its position, normal and albedo.
// BF16 weights and activations on MFMA with FP32 accumulation; FP32 master
// weights with Adam.
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
// BF16 MFMA fragments (64 lanes x 4 values per 16x16 tile): forward W tiles
// (W1 4x2, W2..W4 4x4, W5 1x4 with rows padded to 16) then transposed tiles
// for the backward data pass (W2..W4 4x4, W5^T 4x1).
constexpr int MLP_FWD_TILES = 8 + 16 * 3 + 4;
constexpr int MLP_BWD_TILES = 16 * 3 + 4;
constexpr int MLP_FRAG = MLP_FWD_TILES + MLP_BWD_TILES;

// One query of the cache: the face whose incident radiance is wanted.
struct MlpQuery {
    int32_t vx, vy, vz;
    uint32_t face_material;  // face | material << 8; 0xffffffff: no query
};

struct MlpScene {
    float lo_x, lo_y, lo_z;  // scene minimum corner (voxels)
    float inv_extent;        // 1 / largest scene extent
    const float4 *materials; // (diffuse rgb, emission power), (emission rgb, 0)
};

// Launchers (vgi_mlp.hip).
void mlp_pack(const float *master, uint16_t *frags, hipStream_t stream);
void mlp_infer(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, uint32_t count, float *pred,
               hipStream_t stream);
// Fused forward + loss + backward-data over `count` samples (count a multiple
// of 16): writes per-layer inputs X_i and output gradients dY_i (bf16) for the
// weight-gradient pass, and the summed loss.
void mlp_forward_backward(const uint16_t *frags, const MlpScene &scene, const MlpQuery *queries, const float *targets,
                          uint32_t count, uint16_t *acts, uint16_t *grads, float *loss, float *pred, hipStream_t stream);
// dW_i = sum over samples of dY_i^T X_i on MFMA (split over sample chunks,
// FP32 atomics into `grad`).
void mlp_weight_grad(const uint16_t *acts, const uint16_t *grads, uint32_t count, float *grad, hipStream_t stream);
void mlp_adam(float *master, float *m, float *v, float *grad, uint32_t step, float lr, hipStream_t stream);

}  // namespace vgi
