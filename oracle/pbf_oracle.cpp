// PBF oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++17/OpenMP restatement of the reference's GPU solver step
// (kernels/cudaPBFSolver.cu:4-239, kernels/cudaPBFSolver.cuh:99-181, stage order of
// kernels/cudaPBF.cu:50-114).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker / CPU baseline.  The
// product (pbf_amd/) never links, loads or calls it.
//
// Arithmetic contract (shared with the HIP kernels, which must reproduce it bit for bit):
//   * IEEE fp32 +,-,*,/ and sqrt, evaluated left to right (built with -ffp-contract=off, so the
//     compiler fuses nothing), plus fused multiply-adds (std::fma, one rounding) written out at
//     fixed places: a pair's r.r in the constraint, curl and vorticity passes (dot3f), the sums
//     over neighbours there, and the Newton steps of pbf_rsqrt.  The neighbour search tests r.r
//     left to right;
//   * pow() of the reference is replaced by repeated multiplication for integer exponents;
//   * the boundary jitter's sin() is the self-written pbf_sin below (Cody-Waite reduction +
//     truncated Taylor series), not libm's, so it is identical on both sides;
//   * the spiky gradient is evaluated as r * (cs*(h-|r|)^2 * rsqrt(r.r)) with |r| = r.r *
//     rsqrt(r.r) and rsqrt from a bit-pattern seed and three Newton steps (pbf_rsqrt) -- the
//     reference (cuh:105-110) evaluates (cs*(h-|r|)^2) * (r/|r|) with sqrt; they agree to ~2e-7;
//   * s_corr's ratio W(r)/W(dq) (cuh:143) is evaluated as W(r) * (1/W(dq));
//   * particles inside one cell are ordered by ascending particle index (the reference's
//     atomicSub scatter, cu:112-122, leaves that order undefined).
//
// Pinning: tests/test_oracle.py checks every function here against closed forms, brute force,
// the reference's own scan test (test/testPrefixSum.cpp) and the reference CPU solver
// (src/solver.cpp) built from its sources into oracle/_ref.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <climits>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Same field order and meaning as the reference SolverParameter (include/solverBase.h:16-40),
// followed by the fields the reference hard-codes (gravity cu:9, jitter coefficient
// solverBase.h:5, the spiky-gradient lower cut-off cuh:107, grid dims solverBase.h:11-13,
// neighbour cap solverBase.h:6).
struct orc_params {
    float h;
    int32_t nParticles;
    float restRho;
    float epsilon;
    float k;
    float n;
    float deltaq;
    float c;
    float dt;
    float invDt;
    int32_t iterations;
    float vorticityConfinementEpilon;
    float radius;
    float gravity[3];
    float jitter_coeff;
    float grad_rmin;
    int32_t grid_dims[3];
    int32_t max_neighbors;
    int32_t flags;
};

}  // extern "C"

namespace {

const double kInvPi = 0.31830988618;  // INVPI, include/solverBase.h:4 (a double literal)

struct V3 {
    float x, y, z;
};

inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// r.r of a neighbour pair in the passes: x x, then two fused multiply-adds.
inline float dot3f(const V3& r) { return std::fma(r.z, r.z, std::fma(r.y, r.y, r.x * r.x)); }
inline V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// sin(x) for the boundary jitter.  Cody & Waite (1980) argument reduction by pi/2 split into
// three parts (P1 has 8 significant bits, P2 12, so k*P1 and k*P2 are exact for |k| < 2^12),
// then the Taylor series of sin/cos on [-pi/4, pi/4] (degree 9 / 10, truncation error below
// 2e-9).  Large or non-finite arguments take a fixed path so both builds agree.
float pbf_sin(float x) {
    if (!(std::fabs(x) < 1.0e6f)) return x - x;  // NaN for NaN/inf, 0 for huge finite
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
    switch (k & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}

// frac / randomOffset, kernels/cudaPBFSolver.cuh:112-119.
inline float frac(float v) { return v - (float)(int)v; }
float random_offset(float u, float v) {
    float noise = frac(pbf_sin(2.0f * (u * 12.9898f + v * 78.233f)) * 43758.5453f);
    return std::fabs(noise);
}

// BoundParticle, kernels/cudaPBFSolver.cuh:121-140 (including the y-max case drawing its
// jitter from (y, z), cuh:132).
void bound_particle(V3& p, const float* bmin, const float* bmax, float coeff) {
    if (p.x < bmin[0]) p.x = bmin[0] + coeff * random_offset(p.y, p.z);
    if (p.x > bmax[0]) p.x = bmax[0] - coeff * random_offset(p.y, p.z);
    if (p.y < bmin[1]) p.y = bmin[1] + coeff * random_offset(p.x, p.z);
    if (p.y > bmax[1]) p.y = bmax[1] - coeff * random_offset(p.y, p.z);
    if (p.z < bmin[2]) p.z = bmin[2] + coeff * random_offset(p.x, p.y);
    if (p.z > bmax[2]) p.z = bmax[2] - coeff * random_offset(p.x, p.y);
}

// Kernel constants, kernels/cudaPBFSolver.cuh:99-110: the reference evaluates
// 315.0f*INVPI/(64.0f*powf(h,9)) and -45.0f*INVPI/powf(h,6) in double and rounds to float.
struct Consts {
    float h, h2, c6, cs, wdq, inv_wdq, rmin, k, n, restRho, eps;
    int n_int;  // -1 if n is not a small non-negative integer
};

inline float powi(float b, int e) {
    float r = 1.0f;
    while (e) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

inline float poly6(float r2, const Consts& K) {
    if (r2 > K.h2) return 0.0f;
    float d = K.h2 - r2;
    return K.c6 * (d * d * d);
}

// 1/sqrt(x): Lomont's refined bit-pattern seed 0x5f375a86 (C. Lomont, "Fast Inverse Square
// Root", 2003) and three Newton steps y <- y (1.5 - ((x/2) y) y), each 1.5 - ((x/2) y) y one fused
// multiply-add (the kernels use the same steps).
inline float pbf_rsqrt(float x) {
    uint32_t i;
    std::memcpy(&i, &x, sizeof(i));
    i = 0x5f375a86u - (i >> 1);
    float y;
    std::memcpy(&y, &i, sizeof(y));
    const float hx = 0.5f * x;
    y = y * std::fma(-(hx * y), y, 1.5f);
    y = y * std::fma(-(hx * y), y, 1.5f);
    y = y * std::fma(-(hx * y), y, 1.5f);
    return y;
}

// WspikyGrad, cuh:105-110: |r| and 1/|r| from pbf_rsqrt(r.r).
inline V3 spiky_grad(const V3& r, float r2, const Consts& K) {
    const float inv = pbf_rsqrt(r2);
    const float len = r2 * inv;
    if (len > K.h || len < K.rmin) return {0.0f, 0.0f, 0.0f};
    float hl = K.h - len;
    float t = (K.cs * (hl * hl)) * inv;
    return {r.x * t, r.y * t, r.z * t};
}

inline float scorr(float w, const Consts& K) {
    float ratio = w * K.inv_wdq;
    float p = K.n_int >= 0 ? powi(ratio, K.n_int) : std::pow(ratio, K.n);
    return -K.k * p;
}

Consts make_consts(const orc_params& P) {
    Consts K;
    K.h = P.h;
    K.h2 = P.h * P.h;
    float h9 = (float)std::pow((double)P.h, 9.0);
    float h6 = (float)std::pow((double)P.h, 6.0);
    K.c6 = (float)((315.0 * kInvPi) / (double)(64.0f * h9));
    K.cs = (float)((-45.0 * kInvPi) / (double)h6);
    float dq2 = P.deltaq * P.deltaq + 0.0f * 0.0f + 0.0f * 0.0f;
    K.wdq = poly6(dq2, K);
    K.inv_wdq = 1.0f / K.wdq;
    K.rmin = P.grad_rmin;
    K.k = P.k;
    K.n = P.n;
    K.restRho = P.restRho;
    K.eps = P.epsilon;
    K.n_int = (P.n >= 0.0f && P.n <= 64.0f && P.n == (float)(int)P.n) ? (int)P.n : -1;
    return K;
}

// computeDimID, kernels/cudaPBFSolver.cuh:156-160: int((p - pMin) * invH), truncation toward
// zero.  Values are clamped to [-3, dim+2] before the conversion (which changes no in-range
// cell and no 27-cell search result).  NaN converts to 0, as the CUDA device's float-to-int
// conversion (cvt.rzi.s32.f32) returns for NaN, so a NaN coordinate lands in the first cell of
// its axis and the particle is counted and sorted like any other.
inline int dim_of(float p, float pmin, float invH, int dim) {
    float f = (p - pmin) * invH;
    if (!(f == f)) return 0;
    if (f < -3.0f) f = -3.0f;
    if (f > (float)(dim + 2)) f = (float)(dim + 2);
    return (int)f;
}

}  // namespace

extern "C" {

float orc_sin(float x) { return pbf_sin(x); }

void orc_rsqrt(const float* x, float* y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) y[i] = pbf_rsqrt(x[i]);
}
float orc_random_offset(float u, float v) { return random_offset(u, v); }

void orc_bound(float* p3, const float boundary[6], float coeff) {
    V3 p{p3[0], p3[1], p3[2]};
    bound_particle(p, boundary, boundary + 3, coeff);
    p3[0] = p.x;
    p3[1] = p.y;
    p3[2] = p.z;
}

// out[0..3] = c6, cs, W(deltaq), h*h
void orc_consts(const orc_params* P, float* out) {
    Consts K = make_consts(*P);
    out[0] = K.c6;
    out[1] = K.cs;
    out[2] = K.wdq;
    out[3] = K.h2;
}

float orc_poly6(const orc_params* P, const float* r) {
    Consts K = make_consts(*P);
    V3 v{r[0], r[1], r[2]};
    return poly6(dot3f(v), K);
}

void orc_spiky_grad(const orc_params* P, const float* r, float* g) {
    Consts K = make_consts(*P);
    V3 v{r[0], r[1], r[2]};
    V3 o = spiky_grad(v, dot3f(v), K);
    g[0] = o.x;
    g[1] = o.y;
    g[2] = o.z;
}

// Exclusive prefix sum, the contract of kernels/cudaPBFSolver.cu:42-96 (checked by the
// reference's own test/testPrefixSum.cpp:3-12 against a serial scan).
void orc_exclusive_scan(const int32_t* in, int64_t n, int32_t* out) {
    int32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        out[i] = acc;
        acc += in[i];
    }
}

// Cell id per particle (computeGid, cuh:172-181): -1 when any axis is outside the grid.
void orc_cell_ids(const float* pos, int32_t n, const float grid_min[3], float h, const int32_t dims[3],
                  int32_t* gid) {
    float invH = 1.0f / h;
    for (int32_t i = 0; i < n; ++i) {
        int d0 = dim_of(pos[3 * i + 0], grid_min[0], invH, dims[0]);
        int d1 = dim_of(pos[3 * i + 1], grid_min[1], invH, dims[1]);
        int d2 = dim_of(pos[3 * i + 2], grid_min[2], invH, dims[2]);
        bool out = d0 < 0 || d0 >= dims[0] || d1 < 0 || d1 >= dims[1] || d2 < 0 || d2 >= dims[2];
        gid[i] = out ? -1 : d0 + d1 * dims[0] + d2 * dims[0] * dims[1];
    }
}

struct orc_state {
    orc_params P;
    float grid_min[3];
    float bmin[3], bmax[3];
    int n;
    int ncells;
    std::vector<V3> x, v, xo, dp, curl, vort, visc;
    std::vector<float> lam, lam_first, err;
    std::vector<int32_t> gid, count, cell_start, sorted;
    std::vector<int32_t> nbr;  // n * cap, id space
    std::vector<int32_t> ncnt;
    int n_in_grid;
};

orc_state* orc_create(const orc_params* P, const float grid_space[6], const float boundary[6], const float* pos,
                      const float* vel, int32_t n) {
    orc_state* s = new orc_state();
    s->P = *P;
    s->P.nParticles = n;
    for (int a = 0; a < 3; ++a) {
        s->grid_min[a] = grid_space[a];
        s->bmin[a] = boundary[a];
        s->bmax[a] = boundary[3 + a];
    }
    s->n = n;
    s->ncells = P->grid_dims[0] * P->grid_dims[1] * P->grid_dims[2];
    s->x.resize(n);
    s->v.resize(n);
    for (int i = 0; i < n; ++i) {
        s->x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        s->v[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    }
    s->xo.resize(n);
    s->dp.resize(n);
    s->curl.resize(n);
    s->vort.resize(n);
    s->visc.resize(n);
    s->lam.assign(n, 0.0f);
    s->lam_first.assign(n, 0.0f);
    s->err.assign(n, 0.0f);
    s->gid.resize(n);
    s->count.assign(s->ncells, 0);
    s->cell_start.assign(s->ncells + 1, 0);
    s->sorted.assign(n, -1);
    s->nbr.assign((size_t)n * (size_t)std::max(1, P->max_neighbors), -1);
    s->ncnt.assign(n, 0);
    s->n_in_grid = 0;
    return s;
}

void orc_destroy(orc_state* s) { delete s; }

void orc_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// One solver step, the stage order of kernels/cudaPBF.cu:50-114.
void orc_step(orc_state* s, int32_t nsteps) {
    const orc_params& P = s->P;
    const Consts K = make_consts(P);
    const int n = s->n;
    const int cap = P.max_neighbors;
    const int* dims = P.grid_dims;
    const float dt = P.dt;
    const float invH = 1.0f / P.h;
    for (int step = 0; step < nsteps; ++step) {
        // updateParticles, cu:4-13
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            s->xo[i] = s->x[i];
            V3& v = s->v[i];
            v.x = v.x + dt * P.gravity[0];
            v.y = v.y + dt * P.gravity[1];
            v.z = v.z + dt * P.gravity[2];
            V3& x = s->x[i];
            x.x = x.x + dt * v.x;
            x.y = x.y + dt * v.y;
            x.z = x.z + dt * v.z;
            bound_particle(x, s->bmin, s->bmax, P.jitter_coeff);
        }
        // calculateParticlesGid + prefix sum + setGrid + countingSort, cu:31-122.  The sort is a
        // stable counting sort: ascending particle index inside a cell, out-of-grid (gid -1)
        // particles omitted (cu:118).
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const V3& x = s->x[i];
            int d0 = dim_of(x.x, s->grid_min[0], invH, dims[0]);
            int d1 = dim_of(x.y, s->grid_min[1], invH, dims[1]);
            int d2 = dim_of(x.z, s->grid_min[2], invH, dims[2]);
            bool out = d0 < 0 || d0 >= dims[0] || d1 < 0 || d1 >= dims[1] || d2 < 0 || d2 >= dims[2];
            s->gid[i] = out ? -1 : d0 + d1 * dims[0] + d2 * dims[0] * dims[1];
        }
        std::fill(s->count.begin(), s->count.end(), 0);
        for (int i = 0; i < n; ++i)
            if (s->gid[i] >= 0) s->count[s->gid[i]]++;
        orc_exclusive_scan(s->count.data(), s->ncells, s->cell_start.data());
        s->n_in_grid = s->ncells ? s->cell_start[s->ncells - 1] + s->count[s->ncells - 1] : 0;
        s->cell_start[s->ncells] = s->n_in_grid;
        {
            std::vector<int32_t> fill(s->cell_start.begin(), s->cell_start.end() - 1);
            for (int i = 0; i < n; ++i) {
                int g = s->gid[i];
                if (g >= 0) s->sorted[fill[g]++] = i;
            }
            for (int i = s->n_in_grid; i < n; ++i) s->sorted[i] = -1;
        }
        // buildNeighborhood, cu:124-154: z-outer / y / x-inner cell order, cell members in sorted
        // order, self excluded, r.r <= h*h, at most cap hits.  Evaluated at predicted positions
        // and frozen for the rest of the step.
        const float h2 = K.h2;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const V3 xi = s->x[i];
            int d[3] = {dim_of(xi.x, s->grid_min[0], invH, dims[0]), dim_of(xi.y, s->grid_min[1], invH, dims[1]),
                        dim_of(xi.z, s->grid_min[2], invH, dims[2])};
            int cnt = 0;
            int32_t* out = &s->nbr[(size_t)i * cap];
            for (int j = 0; j < 27 && cnt < cap; ++j) {
                int c0 = d[0] + j % 3 - 1, c1 = d[1] + (j / 3) % 3 - 1, c2 = d[2] + j / 9 - 1;
                if (c0 < 0 || c0 >= dims[0] || c1 < 0 || c1 >= dims[1] || c2 < 0 || c2 >= dims[2]) continue;
                int g = c0 + c1 * dims[0] + c2 * dims[0] * dims[1];
                for (int k = s->cell_start[g]; k < s->cell_start[g + 1] && cnt < cap; ++k) {
                    int nid = s->sorted[k];
                    if (nid == i) continue;
                    V3 r = sub(xi, s->x[nid]);
                    if (dot3(r, r) <= h2) out[cnt++] = nid;
                }
            }
            s->ncnt[i] = cnt;
        }
        for (int it = 0; it < P.iterations; ++it) {
            // calculateLambda, cu:156-180
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                float rho = 0.0f, gradMagSum = 0.0f;
                V3 gradSum{0.0f, 0.0f, 0.0f};
                const int32_t* L = &s->nbr[(size_t)i * cap];
                for (int j = 0; j < s->ncnt[i]; ++j) {
                    V3 r = sub(s->x[i], s->x[L[j]]);
                    float r2 = dot3f(r);
                    rho = rho + poly6(r2, K);
                    V3 g = spiky_grad(r, r2, K);
                    gradSum.x = gradSum.x + g.x;
                    gradSum.y = gradSum.y + g.y;
                    gradSum.z = gradSum.z + g.z;
                    gradMagSum = std::fma(g.z, g.z, std::fma(g.y, g.y, std::fma(g.x, g.x, gradMagSum)));
                }
                float Ci = rho / K.restRho - 1.0f;
                gradMagSum = gradMagSum + dot3(gradSum, gradSum);
                gradMagSum = gradMagSum / (K.restRho * K.restRho);
                s->lam[i] = -Ci / (gradMagSum + K.eps);
            }
            if (it == 0) s->lam_first = s->lam;
            // calculateOffset, cu:182-199
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                V3 acc{0.0f, 0.0f, 0.0f};
                const int32_t* L = &s->nbr[(size_t)i * cap];
                const float li = s->lam[i];
                for (int j = 0; j < s->ncnt[i]; ++j) {
                    int nid = L[j];
                    V3 r = sub(s->x[i], s->x[nid]);
                    float r2 = dot3f(r);
                    float w = poly6(r2, K);
                    V3 g = spiky_grad(r, r2, K);
                    float a = li + s->lam[nid] + scorr(w, K);
                    acc.x = std::fma(a, g.x, acc.x);
                    acc.y = std::fma(a, g.y, acc.y);
                    acc.z = std::fma(a, g.z, acc.z);
                }
                s->dp[i] = {acc.x / K.restRho, acc.y / K.restRho, acc.z / K.restRho};
            }
            // updatePosition, cu:15-22
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                V3& x = s->x[i];
                x.x = x.x + s->dp[i].x;
                x.y = x.y + s->dp[i].y;
                x.z = x.z + s->dp[i].z;
                bound_particle(x, s->bmin, s->bmax, P.jitter_coeff);
                s->v[i] = {(x.x - s->xo[i].x) / dt, (x.y - s->xo[i].y) / dt, (x.z - s->xo[i].z) / dt};
            }
        }
        // with no iteration (the loop of cudaPBF.cu:89-100 does not run) v keeps its predicted value
        // calculateCurl, cu:201-215 (also the density error at the final positions)
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3 w{0.0f, 0.0f, 0.0f};
            float rho = 0.0f;
            const int32_t* L = &s->nbr[(size_t)i * cap];
            for (int j = 0; j < s->ncnt[i]; ++j) {
                int nid = L[j];
                V3 vij = sub(s->v[nid], s->v[i]);
                V3 r = sub(s->x[i], s->x[nid]);
                float r2 = dot3f(r);
                V3 g = spiky_grad(r, r2, K);
                // w += vij x g, each component's two products fused into the sum
                w.x = std::fma(vij.y, g.z, std::fma(-vij.z, g.y, w.x));
                w.y = std::fma(vij.z, g.x, std::fma(-vij.x, g.z, w.y));
                w.z = std::fma(vij.x, g.y, std::fma(-vij.y, g.x, w.z));
                rho = rho + poly6(r2, K);
            }
            s->curl[i] = w;
            s->err[i] = rho / K.restRho - 1.0f;
        }
        // calculateVorticityAndViscosity, cu:217-239 + fVorticity cuh:146-148
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3 pv{0.0f, 0.0f, 0.0f}, eta{0.0f, 0.0f, 0.0f};
            const int32_t* L = &s->nbr[(size_t)i * cap];
            for (int j = 0; j < s->ncnt[i]; ++j) {
                int nid = L[j];
                V3 vij = sub(s->v[i], s->v[nid]);
                V3 r = sub(s->x[i], s->x[nid]);
                float r2 = dot3f(r);
                float w = poly6(r2, K);
                pv.x = std::fma(-vij.x, w, pv.x);
                pv.y = std::fma(-vij.y, w, pv.y);
                pv.z = std::fma(-vij.z, w, pv.z);
                float wl = std::sqrt(dot3(s->curl[nid], s->curl[nid]));
                V3 g = spiky_grad(r, r2, K);
                eta.x = std::fma(wl, g.x, eta.x);
                eta.y = std::fma(wl, g.y, eta.y);
                eta.z = std::fma(wl, g.z, eta.z);
            }
            float el = std::sqrt(dot3(eta, eta));
            if ((double)std::fabs(el) > 1e-3) eta = {eta.x / el, eta.y / el, eta.z / el};
            V3 cr = cross3(eta, s->curl[i]);
            const float ev = P.vorticityConfinementEpilon;
            s->vort[i] = {ev * cr.x, ev * cr.y, ev * cr.z};
            s->visc[i] = {P.c * pv.x, P.c * pv.y, P.c * pv.z};
        }
        // applyVorticityAndViscosity, cu:24-29
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            V3& v = s->v[i];
            v.x = v.x + (dt * s->vort[i].x + s->visc[i].x);
            v.y = v.y + (dt * s->vort[i].y + s->visc[i].y);
            v.z = v.z + (dt * s->vort[i].z + s->visc[i].z);
        }
    }
}

void orc_get_state(const orc_state* s, float* pos, float* vel) {
    for (int i = 0; i < s->n; ++i) {
        if (pos) {
            pos[3 * i] = s->x[i].x;
            pos[3 * i + 1] = s->x[i].y;
            pos[3 * i + 2] = s->x[i].z;
        }
        if (vel) {
            vel[3 * i] = s->v[i].x;
            vel[3 * i + 1] = s->v[i].y;
            vel[3 * i + 2] = s->v[i].z;
        }
    }
}

void orc_set_state(orc_state* s, const float* pos, const float* vel) {
    for (int i = 0; i < s->n; ++i) {
        if (pos) s->x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        if (vel) s->v[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    }
}

// Grid of the last step: gid[n], count[ncells], cell_start[ncells+1], sorted[n] (ids in cell
// order; -1 past n_in_grid).  Returns n_in_grid.
int32_t orc_get_grid(const orc_state* s, int32_t* gid, int32_t* count, int32_t* cell_start, int32_t* sorted) {
    if (gid) std::memcpy(gid, s->gid.data(), sizeof(int32_t) * s->n);
    if (count) std::memcpy(count, s->count.data(), sizeof(int32_t) * s->ncells);
    if (cell_start) std::memcpy(cell_start, s->cell_start.data(), sizeof(int32_t) * (s->ncells + 1));
    if (sorted) std::memcpy(sorted, s->sorted.data(), sizeof(int32_t) * s->n);
    return s->n_in_grid;
}

// Neighbour lists of the last step, id space: nbr[n*cap], cnt[n].
void orc_get_neighbors(const orc_state* s, int32_t* nbr, int32_t* cnt) {
    if (nbr) std::memcpy(nbr, s->nbr.data(), sizeof(int32_t) * s->nbr.size());
    if (cnt) std::memcpy(cnt, s->ncnt.data(), sizeof(int32_t) * s->n);
}

// lambda of the first and of the last Jacobi iteration, density error rho/rho0-1 at the end of
// the constraint loop, and the curl, all of the last step.
void orc_get_fields(const orc_state* s, float* lam_first, float* lam_last, float* err, float* curl) {
    for (int i = 0; i < s->n; ++i) {
        if (lam_first) lam_first[i] = s->lam_first[i];
        if (lam_last) lam_last[i] = s->lam[i];
        if (err) err[i] = s->err[i];
        if (curl) {
            curl[3 * i] = s->curl[i].x;
            curl[3 * i + 1] = s->curl[i].y;
            curl[3 * i + 2] = s->curl[i].z;
        }
    }
}

// Vorticity-confinement force eps (N x omega) (cu:237, cuh:146-148) and XSPH term
// c sum_j -(v_i - v_j) W (cu:229-238) of the last step, before applyVorticityAndViscosity added
// dt * vort + visc to the velocities (cu:24-29), by particle id.
void orc_get_vort_visc(const orc_state* s, float* vort, float* visc) {
    for (int i = 0; i < s->n; ++i) {
        if (vort) {
            vort[3 * i] = s->vort[i].x;
            vort[3 * i + 1] = s->vort[i].y;
            vort[3 * i + 2] = s->vort[i].z;
        }
        if (visc) {
            visc[3 * i] = s->visc[i].x;
            visc[3 * i + 1] = s->visc[i].y;
            visc[3 * i + 2] = s->visc[i].z;
        }
    }
}

// delta-p of the last Jacobi iteration (before it was added to the positions), by particle id.
void orc_get_dp(const orc_state* s, float* dp) {
    for (int i = 0; i < s->n; ++i) {
        dp[3 * i] = s->dp[i].x;
        dp[3 * i + 1] = s->dp[i].y;
        dp[3 * i + 2] = s->dp[i].z;
    }
}

}  // extern "C"
