// Warm-start MLP: parameter packing and device layout (pbf_learned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

constexpr int kIn = 16, kHid = 64;
// FP32 master parameters, packed: W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kHid;
constexpr int kMlpParams = kOffB4 + 1;  // 9473

// What the kernels read: bf16 weights (and the transposes the backward pass needs), fp32
// biases, the output weights rounded to bf16, and the feature / output normalisation.
struct MlpModel {
    __bf16 W1[kHid * kIn];
    __bf16 W2[kHid * kHid];
    __bf16 W3[kHid * kHid];
    __bf16 W2T[kHid * kHid];
    __bf16 W3T[kHid * kHid];
    float b1[kHid], b2[kHid], b3[kHid], w4[kHid];
    float b4;
    float fmean[kIn], finv[kIn];
    float yscale, yshift;  // multiplier / h^2 = y * yscale + yshift
};

// Workgroups of a gradient launch: one per 64 samples, at most kGradMaxBlocks (each loops over
// further chunks); each writes one partial row of kMlpParams + 1 floats.
constexpr int kGradMaxBlocks = 256;
int mlp_grad_blocks(int nb);
void launch_mlp_infer(hipStream_t st, const MlpModel* M, const float* feat, int n, float4* pos, float* y_out,
                      float lam_scale);
// idx may be NULL (samples 0..nb-1 in order); partial holds mlp_grad_blocks(nb) rows of
// kMlpParams + 1 floats; grad receives kMlpParams + 1 floats (the last: sum of squared errors).
void launch_mlp_grad(hipStream_t st, const MlpModel* M, const float* X, const float* T, const uint32_t* idx, int nb,
                     float* partial, float* grad);
void launch_adam(hipStream_t st, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps,
                 float c1, float c2);
void launch_mlp_pack(hipStream_t st, const float* p, MlpModel* M);
void launch_record(hipStream_t st, int n, const uint32_t* ids, const float* feat, const float* lam_sum, float inv_h2,
                   int stride, int phase, size_t base, size_t cap, float* X, float* T);
void launch_batch_indices(hipStream_t st, uint64_t seed, uint64_t step, int nb, uint64_t nsamples, uint32_t* idx);

}  // namespace pbf
