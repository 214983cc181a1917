// Device-side math of one PBF step: smoothing kernels, boundary clamp with jitter and grid
// hashing.  The formulas are those of the reference's kernels/cudaPBFSolver.cuh:99-181; the
// arithmetic is fixed so that a CPU build of the same operations gives the same bits:
//   * fp32 +,-,*,/ and sqrt (correctly rounded; the library is built with -ffp-contract=off and
//     -fhip-fp32-correctly-rounded-divide-sqrt), left to right, plus fused multiply-adds
//     (correctly rounded, fmaf on the CPU) written out at fixed places: r.r and the sums over
//     neighbours in the constraint, curl and vorticity passes, and the Newton steps of pbf_rsqrt;
//   * 1/|r| of the spiky gradient from pbf_rsqrt (bit-pattern seed + Newton steps in those
//     operations), not from sqrt and a division;
//   * integer powers by repeated multiplication;
//   * the jitter's sin is pbf_sin below, not the device libm's.
// The fast mode (StepConsts::fast, PBF_FLAG_FAST_MATH) leaves that contract in one place: the
// spiky gradient's 1/|r| comes from the hardware reciprocal square root (spiky_scale_t<true>);
// its results are held to tolerances instead of bits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

// Grid and clamp box of one scene.  A batch concatenates scenes: scene s owns cells
// [coff, coff + ncells] (the last one collects its out-of-grid particles) and a fixed slot range.
struct SceneGeom {
    float gmin[3];
    float bmin[3], bmax[3];
    int dims[3];
    int coff;
    int ncells;
    int rowoff;  // first entry of the scene's row order in StepConsts::rowmap
};

// Everything a step kernel needs, passed by value as one kernel argument.
struct StepConsts {
    float h, h2, c6, cs, wdq, inv_wdq, k, scorr_n, restRho, eps, dt, visc_c, vort_eps, jitter, rmin;
    float rmin2;  // rmin * rmin (the fast mode's cut-off test)
    int fast;     // PBF_FLAG_FAST_MATH: the constraint kernels' FAST instances
    int lpf;      // the neighbour passes' PF instances (list entries loaded a group ahead)
    int n_int;  // exponent of s_corr when it is a small non-negative integer, else -1
    float g[3];
    float gmin[3];
    float invH;
    int dims[3];
    int ncells;  // dims[0]*dims[1]*dims[2]; bucket ncells holds out-of-grid particles
    float bmin[3], bmax[3];
    int cap;     // neighbour cap (the reference's MAX_NEIGHBOR = 63)
    int n;       // particles
    int nstride; // row stride of the neighbour table (>= n, multiple of 64)
    int nbuckets;  // cell counters of the whole step (ncells + 1, or the sum over a batch)
    // Order of the cell rows (the dims[0] cells of one (y, z)) in the cell table: NULL = z, then
    // y (x-fastest throughout); else rowmap[rowoff + z * dims[1] + y] is the row's place.  Cells
    // stay x-fastest inside a row either way, and the neighbour lists and sums do not depend on
    // the order of the rows.
    const int* rowmap;
};

// The geometry of a single-scene solver, from its constants.
__device__ __forceinline__ SceneGeom geom_single(const StepConsts& K) {
    SceneGeom G;
    for (int a = 0; a < 3; ++a) {
        G.gmin[a] = K.gmin[a];
        G.bmin[a] = K.bmin[a];
        G.bmax[a] = K.bmax[a];
        G.dims[a] = K.dims[a];
    }
    G.coff = 0;
    G.ncells = K.ncells;
    G.rowoff = 0;
    return G;
}

// sin for the boundary jitter: Cody & Waite (1980) reduction by pi/2 in three parts
// (P1: 8 significant bits, P2: 12) and the Taylor series of sin / cos on [-pi/4, pi/4],
// truncated after x^9 / x^10.
__device__ __forceinline__ float pbf_sin(float x) {
    if (!(fabsf(x) < 1.0e6f)) return x - x;
    const float two_over_pi = 0x1.45f306p-1f;
    const float P1 = 0x1.92p+0f;
    const float P2 = 0x1.fb6p-12f;
    const float P3 = -0x1.777a5cp-25f;
    float t = x * two_over_pi;
    int k = (int)(t + (t >= 0.0f ? 0.5f : -0.5f));
    float kf = (float)k;
    float r = ((x - kf * P1) - kf * P2) - kf * P3;
    float r2 = r * r;
    float s = r + r * r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f)));
    float c = 1.0f + r2 * (-0.5f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f))));
    float v = (k & 1) ? c : s;
    return (k & 2) ? -v : v;
}

// frac / randomOffset, cuh:112-119.
__device__ __forceinline__ float random_offset(float u, float v) {
    float a = pbf_sin(2.0f * (u * 12.9898f + v * 78.233f)) * 43758.5453f;
    return fabsf(a - (float)(int)a);
}

// BoundParticle, cuh:121-140, including the y-max case taking its jitter from (y, z).
__device__ __forceinline__ void bound_particle(float& x, float& y, float& z, const StepConsts& K, const SceneGeom& G) {
    if (x < G.bmin[0]) x = G.bmin[0] + K.jitter * random_offset(y, z);
    if (x > G.bmax[0]) x = G.bmax[0] - K.jitter * random_offset(y, z);
    if (y < G.bmin[1]) y = G.bmin[1] + K.jitter * random_offset(x, z);
    if (y > G.bmax[1]) y = G.bmax[1] - K.jitter * random_offset(y, z);
    if (z < G.bmin[2]) z = G.bmin[2] + K.jitter * random_offset(x, y);
    if (z > G.bmax[2]) z = G.bmax[2] - K.jitter * random_offset(x, y);
}

// Wpoly6, cuh:99-103, on r.r.  Outside the support h^2 - r.r < 0 is clamped to +0, which gives
// W = +0 exactly as the reference's early return does (no compare and select).
__device__ __forceinline__ float poly6(float r2, const StepConsts& K) {
    const float d = fmaxf(K.h2 - r2, 0.0f);
    return K.c6 * (d * d * d);
}

// 1/sqrt(x): Lomont's refined bit-pattern seed 0x5f375a86 (C. Lomont, "Fast Inverse Square
// Root", 2003) and three Newton steps y <- y (1.5 - ((x/2) y) y), each 1.5 - ((x/2) y) y one
// fused multiply-add, in IEEE fp32, so the oracle computes the same bits.  Relative error below
// 2e-7 for normal x > 0.
__device__ __forceinline__ float pbf_rsqrt(float x) {
    float y = __uint_as_float(0x5f375a86u - (__float_as_uint(x) >> 1));
    const float hx = 0.5f * x;
    y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
    y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
    y = y * __builtin_fmaf(-(hx * y), y, 1.5f);
    return y;
}

// WspikyGrad, cuh:105-110, returned as the scalar t with grad = r * t (no branch):
// |r| = r.r * rsqrt(r.r) and t = cs (h - |r|)^2 * rsqrt(r.r).  For |r| > h, h - |r| is clamped to
// +0, so t is +-0 instead of the reference's 0: every use multiplies t into a term of a sum that
// starts at +0 and so can never be -0, where adding +-0 changes nothing.  |r| < rmin selects 0.
__device__ __forceinline__ float spiky_scale(float r2, const StepConsts& K) {
    const float inv = pbf_rsqrt(r2);
    const float len = r2 * inv;
    const float hl = fmaxf(K.h - len, 0.0f);
    const float t = (K.cs * (hl * hl)) * inv;
    return len < K.rmin ? 0.0f : t;
}

// a * b + c as one fused multiply-add: a neighbour term into its sum.
__device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// r.r of a neighbour pair as x x, then two fused multiply-adds.  (The neighbour search tests its
// candidates with r.r left to right, as the reference's buildNeighborhood.)
__device__ __forceinline__ float dot3f(float x, float y, float z) {
    return __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
}

// spiky_scale with the hardware reciprocal square root (v_rsq_f32, about 1 ulp) in FAST mode.
// The cut-off |r| < rmin is tested as r.r < rmin^2, which also keeps r.r = 0 (1/0 = inf) out of
// the result.
template <bool FAST>
__device__ __forceinline__ float spiky_scale_t(float r2, const StepConsts& K) {
    if constexpr (FAST) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float len = r2 * inv;
        const float hl = fmaxf(K.h - len, 0.0f);
        const float t = (K.cs * (hl * hl)) * inv;
        return r2 < K.rmin2 ? 0.0f : t;
    }
    return spiky_scale(r2, K);
}

// b^e for 0 <= e < 128: the multiplications of square-and-multiply from the low bit up (the
// exponent is uniform, so the branches are scalar).
__device__ __forceinline__ float powi(float b, int e) {
    float r = 1.0f;
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) {
        if (e & (1 << bit)) r = r * b;
        b = b * b;
    }
    return r;
}

// Scorr, cuh:142-144, given W(r): -k (W / W(dq))^n with the ratio as W * (1 / W(dq)).
__device__ __forceinline__ float scorr(float w, const StepConsts& K) {
    float ratio = w * K.inv_wdq;
    float p = K.n_int >= 0 ? powi(ratio, K.n_int) : powf(ratio, K.scorr_n);
    return -K.k * p;
}

// scorr for n = 4 (the configured exponent): powi's multiplications for e = 4 are
// 1 * ((b * b) * (b * b)), and 1 * x == x, so this is the same value with the exponent fixed.
__device__ __forceinline__ float scorr4(float w, const StepConsts& K) {
    float ratio = w * K.inv_wdq;
    float b2 = ratio * ratio;
    return -K.k * (b2 * b2);
}

// computeDimID, cuh:156-160: truncation toward zero of (p - pMin) * invH, clamped to
// [-3, dim+2] first (no in-range cell or 27-cell search changes).  NaN gives 0, what the CUDA
// device's float-to-int conversion returns for NaN: the particle is counted and sorted into the
// first cell of that axis and, since every distance to it compares false, it neither has nor is
// a neighbour.
__device__ __forceinline__ int dim_of(float p, float pmin, float invH, int dim) {
    float f = (p - pmin) * invH;
    if (!(f == f)) return 0;
    f = f < -3.0f ? -3.0f : f;
    f = f > (float)(dim + 2) ? (float)(dim + 2) : f;
    return (int)f;
}

// computeGid, cuh:172-181, offset by the scene's first cell; out-of-grid particles go to the
// scene's bucket coff + ncells.
// Place of row (y, z), 0 <= y < dims[1], 0 <= z < dims[2], in the scene's cell table.
__device__ __forceinline__ int row_of(int y, int z, const StepConsts& K, const SceneGeom& G) {
    const int r = z * G.dims[1] + y;
    return K.rowmap ? K.rowmap[G.rowoff + r] : r;
}

// The cell counter (bucket) of a position: its cell's place in the table, or the scene's
// out-of-grid bucket.
__device__ __forceinline__ int cell_of(float x, float y, float z, const StepConsts& K, const SceneGeom& G) {
    int d0 = dim_of(x, G.gmin[0], K.invH, G.dims[0]);
    int d1 = dim_of(y, G.gmin[1], K.invH, G.dims[1]);
    int d2 = dim_of(z, G.gmin[2], K.invH, G.dims[2]);
    bool out = d0 < 0 || d0 >= G.dims[0] || d1 < 0 || d1 >= G.dims[1] || d2 < 0 || d2 >= G.dims[2];
    return G.coff + (out ? G.ncells : d0 + row_of(out ? 0 : d1, out ? 0 : d2, K, G) * G.dims[0]);
}

// computeGid (cuh:156-160): the reference's cell id d0 + dims0 * (d1 + dims1 * d2), -1 outside.
__device__ __forceinline__ int cell_id_linear(float x, float y, float z, const StepConsts& K, const SceneGeom& G) {
    int d0 = dim_of(x, G.gmin[0], K.invH, G.dims[0]);
    int d1 = dim_of(y, G.gmin[1], K.invH, G.dims[1]);
    int d2 = dim_of(z, G.gmin[2], K.invH, G.dims[2]);
    bool out = d0 < 0 || d0 >= G.dims[0] || d1 < 0 || d1 >= G.dims[1] || d2 < 0 || d2 >= G.dims[2];
    return out ? -1 : d0 + (d2 * G.dims[1] + d1) * G.dims[0];
}

}  // namespace pbf
