// Learned warm start: a per-particle MLP 16 -> 64 -> 64 -> 64 -> 1 (hard-swish) that predicts the
// constraint multiplier a plain solve accumulates over its Jacobi iterations, so the first
// delta-p pass can start from it.  No counterpart in the reference.
//
// Numerics: BF16 weights and activations, FP32 accumulation on gfx950 MFMA
// (v_mfma_f32_32x32x16_bf16), FP32 biases, FP32 master weights and Adam state.
//
// Tiling: the GEMMs are computed transposed, Z^T[feature][sample] = W[feature][k] * X^T[k][sample],
// one 32 x 32 tile per wavefront; both MFMA operands are read "K-contiguous" (row-major W and
// sample-major activations), so every operand fetch is one 16-byte read per lane.  A tile's
// accumulator holds, in lane l, element 4i+j at row 4*(l/32) + 8i + j, column l % 32.
//
// A workgroup (4 wavefronts) handles 64 samples per chunk: wave w owns feature tile w & 1 of
// sample half w >> 1 in the forward and input-gradient GEMMs, and one 32 x 32 tile of each
// weight gradient.  Weight gradients are kept in registers across chunks, written once per
// workgroup, and summed over workgroups in a fixed order (deterministic).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "pbf_kernels.h"
#include "pbf_learned.h"

namespace pbf {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kB = 256;
constexpr int LD64 = 72;  // LDS row pitch (bf16) of 64-wide rows: 144 B spreads the banks
constexpr int LD16 = 24;  // ... of 16-wide rows

// Hard-swish, h = z * clamp(z + 3, 0, 6) / 6 (A. Howard et al., "Searching for MobileNetV3",
// 2019): SiLU's shape from full-rate operations (an add, a clamp, two multiplies).  SiLU's exp and
// reciprocal issue at a quarter of the rate and took about half of the inference kernel's issue
// cycles.  The derivative is torch's: 0 up to and including -3, z / 3 + 1/2 between, 1 from 3 on.
__device__ __forceinline__ float actf(float z) {
    return (z * __builtin_amdgcn_fmed3f(z + 3.0f, 0.0f, 6.0f)) * (1.0f / 6.0f);
}
__device__ __forceinline__ float dactf(float z) {
    return z <= -3.0f ? 0.0f : (z < 3.0f ? z * (1.0f / 3.0f) + 0.5f : 1.0f);
}

// acc += A[32 x K] * B[32 x K]^T, rows of A at a + r*lda, rows of B at b + r*ldb (bf16,
// K contiguous, 16-byte aligned), K a multiple of 16.
__device__ __forceinline__ f32x16 mma_nt(const bf16* a, int lda, const bf16* b, int ldb, int K, f32x16 acc) {
    const int l = threadIdx.x & 63;
    const int r = l & 31, kb = (l >> 5) * 8;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 av = *reinterpret_cast<const bf16x8*>(a + r * lda + k0 + kb);
        bf16x8 bv = *reinterpret_cast<const bf16x8*>(b + r * ldb + k0 + kb);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ int acc_row(int e) { return 4 * ((threadIdx.x & 63) >> 5) + 8 * (e >> 2) + (e & 3); }
__device__ __forceinline__ int acc_col() { return threadIdx.x & 31; }

__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
    for (int e = 0; e < 16; ++e) z[e] = 0.0f;
    return z;
}

// The accumulator of feature tile ft preset to the layer's biases, so that the MFMAs produce
// W x + b (no add per element afterwards).
__device__ __forceinline__ f32x16 bias16(const float* __restrict__ bias, int ft) {
    f32x16 z;
    for (int e = 0; e < 16; ++e) z[e] = bias[ft * 32 + acc_row(e)];
    return z;
}

// Standardised, clipped, bf16 features of 64 samples into Xs[64][LD16] and XsT[32][LD64]
// (rows 16..31 of XsT are zero).  src(n) gives the raw feature row of sample n or NULL.
template <class Src>
__device__ void load_features(const MlpModel* __restrict__ M, Src src, bf16* Xs, bf16* XsT) {
    const int t = threadIdx.x;
    const int n = t >> 2, f0 = (t & 3) * 4;
    const float* row = src(n);
    for (int q = 0; q < 4; ++q) {
        int f = f0 + q;
        float v = row ? (row[f] - M->fmean[f]) * M->finv[f] : 0.0f;
        v = fminf(fmaxf(v, -8.0f), 8.0f);
        bf16 b = (bf16)v;
        Xs[n * LD16 + f] = b;
        XsT[f * LD64 + n] = b;
        XsT[(f + 16) * LD64 + n] = (bf16)0.0f;
    }
}

// Stores one thread's 4 raw features of sample n (valid) standardised, clipped and in BF16, as
// load_features does; inference fetches the raw values a chunk ahead.
__device__ __forceinline__ void store_features(float4 raw, bool valid, const float* mean, const float* inv, bf16* Xs,
                                               bf16* XsT) {
    const int t = threadIdx.x;
    const int n = t >> 2, f0 = (t & 3) * 4;
    const float r[4] = {raw.x, raw.y, raw.z, raw.w};
    for (int q = 0; q < 4; ++q) {
        int f = f0 + q;
        float v = valid ? (r[q] - mean[q]) * inv[q] : 0.0f;
        v = fminf(fmaxf(v, -8.0f), 8.0f);
        bf16 b = (bf16)v;
        Xs[n * LD16 + f] = b;
        XsT[f * LD64 + n] = b;
        XsT[(f + 16) * LD64 + n] = (bf16)0.0f;
    }
}

// One hidden layer for this wave's tile: z = W[ft rows] * in^T[sh cols] + b.  Writes
// bf16(act(z)) sample-major into out[64][LD64] and, if outT, feature-major into outT[64][LD64].
__device__ __forceinline__ f32x16 hidden_layer(const bf16* W, int ldw, int K, const float* bias, const bf16* in, int ldin,
                                               bf16* out, bf16* outT, int ft, int sh) {
    f32x16 z = mma_nt(W + ft * 32 * ldw, ldw, in + sh * 32 * ldin, ldin, K, bias16(bias, ft));
    const int n = sh * 32 + acc_col();
    for (int e = 0; e < 16; ++e) {
        int o = ft * 32 + acc_row(e);
        bf16 h = (bf16)actf(z[e]);
        out[n * LD64 + o] = h;
        if (outT) outT[o * LD64 + n] = h;
    }
    return z;
}

// ---- inference: Lambda-hat into pos.w ------------------------------------------------------
__global__ __launch_bounds__(kB) void k_mlp_infer(const MlpModel* __restrict__ M, const float* __restrict__ feat, int n,
                                                  float4* __restrict__ pos, float* __restrict__ y_out, float lam_scale) {
    __shared__ __attribute__((aligned(16))) bf16 Xs[64 * LD16];
    __shared__ __attribute__((aligned(16))) bf16 XsT[32 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H1[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H2[64 * LD64];
    __shared__ float ypart[2][64];
    const int w = threadIdx.x >> 6, ft = w & 1, sh = w >> 1;
    // this thread's 4 features of sample threadIdx.x / 4: the standardisation in registers, and
    // the raw values fetched one chunk ahead so that the load overlaps the previous chunk's layers
    const int fs = threadIdx.x >> 2, f0 = (threadIdx.x & 3) * 4;
    float mean[4], inv[4];
    for (int q = 0; q < 4; ++q) {
        mean[q] = M->fmean[f0 + q];
        inv[q] = M->finv[f0 + q];
    }
    auto fetch = [&](int base) -> float4 {
        const int i = base + fs;
        return i < n ? *reinterpret_cast<const float4*>(feat + (size_t)i * kNumFeatures + f0) : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    float4 next = fetch(blockIdx.x * 64);
    for (int base = blockIdx.x * 64; base < n; base += gridDim.x * 64) {
        const float4 cur = next;
        next = fetch(base + gridDim.x * 64);
        store_features(cur, base + fs < n, mean, inv, Xs, XsT);
        __syncthreads();
        hidden_layer(M->W1, 16, 16, M->b1, Xs, LD16, H1, nullptr, ft, sh);
        __syncthreads();
        hidden_layer(M->W2, 64, 64, M->b2, H1, LD64, H2, nullptr, ft, sh);
        __syncthreads();
        f32x16 z3 = mma_nt(M->W3 + ft * 32 * 64, 64, H2 + sh * 32 * LD64, LD64, 64, bias16(M->b3, ft));
        float part = 0.0f;
        for (int e = 0; e < 16; ++e) {
            int o = ft * 32 + acc_row(e);
            float h = (float)(bf16)actf(z3[e]);
            part = part + M->w4[o] * h;
        }
        part = part + __shfl_xor(part, 32);
        if ((threadIdx.x & 63) < 32) ypart[ft][sh * 32 + acc_col()] = part;
        __syncthreads();
        if (threadIdx.x < 64) {
            int i = base + threadIdx.x;
            if (i < n) {
                float y = ypart[0][threadIdx.x] + ypart[1][threadIdx.x] + M->b4;
                if (y_out) y_out[i] = y;
                if (pos) reinterpret_cast<float*>(pos + i)[3] = (y * M->yscale + M->yshift) * lam_scale;
            }
        }
        __syncthreads();
    }
}

// ---- training: loss and gradients of a batch -----------------------------------------------
// partial[blockIdx.x][kMlpParams + 1]: this workgroup's gradient sums (packed as the parameter
// vector) and its sum of squared errors; every entry is written.  Sample i of the batch is row
// idx[i] of (X, T), or row i when idx is NULL.  inv_b = 1 / batch size.
__global__ __launch_bounds__(kB) void k_mlp_grad(const MlpModel* __restrict__ M, const float* __restrict__ X,
                                                 const float* __restrict__ T, const uint32_t* __restrict__ idx, int nb,
                                                 float inv_b, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) bf16 Xs[64 * LD16];
    __shared__ __attribute__((aligned(16))) bf16 XsT[32 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H1[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H1T[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H2[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 H2T[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 D[64 * LD64];
    __shared__ __attribute__((aligned(16))) bf16 DT[64 * LD64];
    __shared__ float ypart[2][64];
    __shared__ float dy[64];
    const int tid = threadIdx.x, w = tid >> 6, ft = w & 1, sh = w >> 1;
    const int gm = w >> 1, gn = w & 1;  // weight-gradient tile owned by this wave
    f32x16 gW1 = zero16(), gW2 = zero16(), gW3 = zero16();
    float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, gw4 = 0.0f, gb4 = 0.0f, sse = 0.0f;

    for (int base = blockIdx.x * 64; base < nb; base += gridDim.x * 64) {
        load_features(M, [&](int i) -> const float* {
                          return base + i < nb ? X + (size_t)(idx ? idx[base + i] : (uint32_t)(base + i)) * kNumFeatures
                                               : nullptr;
                      },
                      Xs, XsT);
        __syncthreads();
        f32x16 z1 = hidden_layer(M->W1, 16, 16, M->b1, Xs, LD16, H1, H1T, ft, sh);
        __syncthreads();
        f32x16 z2 = hidden_layer(M->W2, 64, 64, M->b2, H1, LD64, H2, H2T, ft, sh);
        __syncthreads();
        f32x16 z3 = mma_nt(M->W3 + ft * 32 * 64, 64, H2 + sh * 32 * LD64, LD64, 64, bias16(M->b3, ft));
        float part = 0.0f;
        const int ncol = sh * 32 + acc_col();
        for (int e = 0; e < 16; ++e) {
            int o = ft * 32 + acc_row(e);
            bf16 hb = (bf16)actf(z3[e]);
            D[ncol * LD64 + o] = hb;  // h3, sample-major, for the W4 gradient
            part = part + M->w4[o] * (float)hb;
        }
        part = part + __shfl_xor(part, 32);
        if ((tid & 63) < 32) ypart[ft][ncol] = part;
        __syncthreads();
        if (tid < 64) {
            float d = 0.0f;
            if (base + tid < nb) {
                float y = ypart[0][tid] + ypart[1][tid] + M->b4;
                float r = y - T[idx ? idx[base + tid] : (uint32_t)(base + tid)];
                sse = sse + r * r;
                d = 2.0f * r * inv_b;
            }
            dy[tid] = d;
            gb4 = gb4 + d;
        }
        __syncthreads();
        if (tid < 64) {  // W4 gradient: sum_n dy[n] * h3[n][tid]
            float s = 0.0f;
            for (int q = 0; q < 64; ++q) s = s + dy[q] * (float)D[q * LD64 + tid];
            gw4 = gw4 + s;
        }
        __syncthreads();
        // dz3 = dy * w4 * act'(z3)
        for (int e = 0; e < 16; ++e) {
            int o = ft * 32 + acc_row(e);
            bf16 g = (bf16)(dy[ncol] * M->w4[o] * dactf(z3[e]));
            D[ncol * LD64 + o] = g;
            DT[o * LD64 + ncol] = g;
        }
        __syncthreads();
        if (tid < 64) {
            float s = 0.0f;
            for (int q = 0; q < 64; ++q) s = s + (float)DT[tid * LD64 + q];
            gb3 = gb3 + s;
        }
        gW3 = mma_nt(DT + gm * 32 * LD64, LD64, H2T + gn * 32 * LD64, LD64, 64, gW3);
        f32x16 d2 = mma_nt(M->W3T + ft * 32 * 64, 64, D + sh * 32 * LD64, LD64, 64, zero16());
        __syncthreads();
        for (int e = 0; e < 16; ++e) {
            int o = ft * 32 + acc_row(e);
            bf16 g = (bf16)(d2[e] * dactf(z2[e]));
            D[ncol * LD64 + o] = g;
            DT[o * LD64 + ncol] = g;
        }
        __syncthreads();
        if (tid < 64) {
            float s = 0.0f;
            for (int q = 0; q < 64; ++q) s = s + (float)DT[tid * LD64 + q];
            gb2 = gb2 + s;
        }
        gW2 = mma_nt(DT + gm * 32 * LD64, LD64, H1T + gn * 32 * LD64, LD64, 64, gW2);
        f32x16 d1 = mma_nt(M->W2T + ft * 32 * 64, 64, D + sh * 32 * LD64, LD64, 64, zero16());
        __syncthreads();
        for (int e = 0; e < 16; ++e) {
            int o = ft * 32 + acc_row(e);
            DT[o * LD64 + ncol] = (bf16)(d1[e] * dactf(z1[e]));
        }
        __syncthreads();
        if (tid < 64) {
            float s = 0.0f;
            for (int q = 0; q < 64; ++q) s = s + (float)DT[tid * LD64 + q];
            gb1 = gb1 + s;
        }
        if (w < 2) gW1 = mma_nt(DT + w * 32 * LD64, LD64, XsT, LD64, 64, gW1);
        __syncthreads();
    }
    // write this workgroup's sums, packed like the parameters
    float* P = partial + (size_t)blockIdx.x * (kMlpParams + 1);
    const int col = acc_col();
    for (int e = 0; e < 16; ++e) {
        int r = acc_row(e);
        P[kOffW3 + (gm * 32 + r) * 64 + gn * 32 + col] = gW3[e];
        P[kOffW2 + (gm * 32 + r) * 64 + gn * 32 + col] = gW2[e];
        if (w < 2 && col < 16) P[kOffW1 + (w * 32 + r) * 16 + col] = gW1[e];
    }
    if (tid < 64) {
        P[kOffB1 + tid] = gb1;
        P[kOffB2 + tid] = gb2;
        P[kOffB3 + tid] = gb3;
        P[kOffW4 + tid] = gw4;
    }
    __shared__ float red[64];
    if (tid < 64) red[tid] = gb4;
    __syncthreads();
    if (tid == 0) {
        float s = 0.0f;
        for (int q = 0; q < 64; ++q) s = s + red[q];
        P[kOffB4] = s;
    }
    __syncthreads();
    if (tid < 64) red[tid] = sse;
    __syncthreads();
    if (tid == 0) {
        float s = 0.0f;
        for (int q = 0; q < 64; ++q) s = s + red[q];
        P[kMlpParams] = s;
    }
}

// Sum of the per-workgroup partials in workgroup order.
__global__ __launch_bounds__(kB) void k_grad_reduce(const float* __restrict__ partial, int nparts,
                                                    float* __restrict__ grad) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i > kMlpParams) return;
    float s = 0.0f;
    for (int p = 0; p < nparts; ++p) s = s + partial[(size_t)p * (kMlpParams + 1) + i];
    grad[i] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ __launch_bounds__(kB) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                             const float* __restrict__ g, float lr, float b1, float b2, float eps,
                                             float c1, float c2) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i >= kMlpParams) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Packs the FP32 master parameters into the kernels' layout (bf16 weights, transposes).
__global__ __launch_bounds__(kB) void k_mlp_pack(const float* __restrict__ p, MlpModel* __restrict__ M) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i < kHid * kIn) M->W1[i] = (bf16)p[kOffW1 + i];
    if (i < kHid * kHid) {
        int o = i / kHid, c = i % kHid;
        bf16 a = (bf16)p[kOffW2 + i], b = (bf16)p[kOffW3 + i];
        M->W2[i] = a;
        M->W3[i] = b;
        M->W2T[c * kHid + o] = a;
        M->W3T[c * kHid + o] = b;
    }
    if (i < kHid) {
        M->b1[i] = p[kOffB1 + i];
        M->b2[i] = p[kOffB2 + i];
        M->b3[i] = p[kOffB3 + i];
        M->w4[i] = (float)(bf16)p[kOffW4 + i];
    }
    if (i == 0) M->b4 = p[kOffB4];
}

// Appends (features, accumulated multiplier / h^2) of every particle whose id satisfies
// id % stride == phase, at row base + id / stride (deterministic placement).
__global__ __launch_bounds__(kB) void k_record(int n, const uint32_t* __restrict__ ids, const float* __restrict__ feat,
                                               const float* __restrict__ lam_sum, float inv_h2, int stride, int phase,
                                               size_t base, size_t cap, float* __restrict__ X, float* __restrict__ T) {
    int s = blockIdx.x * kB + threadIdx.x;
    if (s >= n) return;
    uint32_t id = ids[s];
    if ((int)(id % (uint32_t)stride) != phase) return;
    size_t row = base + id / (uint32_t)stride;
    if (row >= cap) return;
    const float4* f = reinterpret_cast<const float4*>(feat + (size_t)s * kNumFeatures);
    float4* o = reinterpret_cast<float4*>(X + row * kNumFeatures);
    o[0] = f[0];
    o[1] = f[1];
    o[2] = f[2];
    o[3] = f[3];
    T[row] = lam_sum[s] * inv_h2;
}

// Batch sample indices: SplitMix64 of (seed, step, i) modulo the sample count.
__global__ __launch_bounds__(kB) void k_batch_indices(uint64_t seed, uint64_t step, int nb, uint64_t nsamples,
                                                      uint32_t* __restrict__ idx) {
    int i = blockIdx.x * kB + threadIdx.x;
    if (i >= nb) return;
    uint64_t z = seed + (step * (uint64_t)nb + (uint64_t)i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    idx[i] = (uint32_t)(z % nsamples);
}

// ---- launch wrappers ----------------------------------------------------------------------
int mlp_grad_blocks(int nb) {
    int b = (nb + 63) / 64;
    return b < 1 ? 1 : (b > kGradMaxBlocks ? kGradMaxBlocks : b);
}

void launch_mlp_infer(hipStream_t st, const MlpModel* M, const float* feat, int n, float4* pos, float* y_out,
                      float lam_scale) {
    if (n <= 0) return;
    // The kernel hoists the weight fragments, biases and output weights of its wavefront into
    // registers once (about 100 VGPRs, 3 waves per SIMD); a grid of about 6 workgroups per CU,
    // each looping over several 64-sample chunks, reuses them instead of reloading them per
    // chunk (measured: 5 us less per learned step at 262k than one workgroup per chunk).
    int blocks = std::min((n + 63) / 64, 1536);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(kB), 0, st, M, feat, n, pos, y_out, lam_scale);
}

void launch_mlp_grad(hipStream_t st, const MlpModel* M, const float* X, const float* T, const uint32_t* idx, int nb,
                     float* partial, float* grad) {
    int blocks = mlp_grad_blocks(nb);
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kB), 0, st, M, X, T, idx, nb, 1.0f / (float)(nb > 0 ? nb : 1),
                       partial);
    hipLaunchKernelGGL(k_grad_reduce, dim3((kMlpParams + 1 + kB - 1) / kB), dim3(kB), 0, st, partial, blocks, grad);
}

void launch_adam(hipStream_t st, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps,
                 float c1, float c2) {
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + kB - 1) / kB), dim3(kB), 0, st, p, m, v, g, lr, b1, b2, eps, c1, c2);
}

void launch_mlp_pack(hipStream_t st, const float* p, MlpModel* M) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((kHid * kHid + kB - 1) / kB), dim3(kB), 0, st, p, M);
}

void launch_batch_indices(hipStream_t st, uint64_t seed, uint64_t step, int nb, uint64_t nsamples, uint32_t* idx) {
    if (nb <= 0 || nsamples == 0) return;
    hipLaunchKernelGGL(k_batch_indices, dim3((nb + kB - 1) / kB), dim3(kB), 0, st, seed, step, nb, nsamples, idx);
}

void launch_record(hipStream_t st, int n, const uint32_t* ids, const float* feat, const float* lam_sum, float inv_h2,
                   int stride, int phase, size_t base, size_t cap, float* X, float* T) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_record, dim3((n + kB - 1) / kB), dim3(kB), 0, st, n, ids, feat, lam_sum, inv_h2, stride, phase,
                       base, cap, X, T);
}

}  // namespace pbf
