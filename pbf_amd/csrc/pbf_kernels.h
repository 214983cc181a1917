// Launch interface of the step kernels (pbf_kernels.hip) for the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"

namespace pbf {

// Device buffers of one solver; index [2] arrays alternate between steps / iterations.
// pos[i] and vel[i] hold n + 1 entries: slot n is a padding particle at (kFarPos, kFarPos,
// kFarPos) with zero velocity and multiplier, which no kernel writes.  The neighbour list pads
// each particle's last group of 8 entries with n.  pos[i] holds kSearchTail such padding slots
// (n .. n + kSearchTail - 1): the search's candidate loads run up to 31 slots past a run's end.
struct StepBuffers {
    float4* pos[2];      // .w: lambda during the constraint passes, |omega| after the curl pass
    float4* vel[2];
    uint32_t* ids[2];
    float4* xold;        // step-start positions; after the velocity update, the XSPH sums
    float4* xp;          // xp / vp: staging of host transfers (free during a step)
    float4* vp;
    float4* omega;
    uint32_t* cell;
    uint32_t* local;
    uint32_t* count;     // ncells + 1 (+ padding to 4)
    uint32_t* cstart;    // ncells + 2
    uint32_t* tile_sum;  // scan tiles
    uint32_t* bucket;
    uint32_t* bucket_id; // particle id of each bucket entry (ranking inside a cell)
    uint32_t* nbr;       // ceil(cap / 8) * 8 * nstride, blocks of 8 entries per slot, each the
                         // neighbour's byte offset 16 q in the float4 slot arrays; entries
                         // cnt .. ceil8(cnt) - 1 of a slot hold the padding slot n (16 n)
    uint8_t* ncnt;
    float* lam1;         // iteration-1 lambda per slot (a warm-start feature)
};

// Warm-start features per particle (raw fp32).  The learned step runs one plain pass first; the
// features describe the particle after it and are written by the pass-2 lambda (slot 10 by the
// first lambda):
//  0 C_i = rho/rho0 - 1 after pass 1              1 neighbour count
//  2 sum_j |grad_j C_i|^2 after pass 1            3 |grad_i C_i| after pass 1
//  4-6 v * dt / h (predicted velocity)            7 distance to the nearest wall / h (<= 4)
//  8 lambda of pass 2 / h^2                       9 lambda of pass 1 / h^2
//  10 C_i at the predicted positions              11-15 distance to the walls x-, x+, y-, y+, z- / h (<= 4)
constexpr int kNumFeatures = 16;

// Coordinate of the padding slot: its distance to any particle of a scene is ~1.7e18, so r.r
// (~3e36) stays finite, poly6 and the spiky gradient are 0, and every term it adds to a
// neighbour sum is +0 or -0.  Sums start at +0 and never become -0, so adding such a term leaves
// them unchanged bit for bit: a padded group is evaluated as if its tail were absent.
constexpr float kFarPos = 1e18f;
constexpr int kSearchTail = 32;

// Particles per solver (all scenes of a batch) stay below 2^28 - kSearchTail, so the byte offset
// of any float4 slot, padding slots included, fits in 32 bits (slot_at, neighbour-list entries).
constexpr int64_t kMaxSlots = ((int64_t)1 << 28) - kSearchTail;

// Optional outputs of the constraint kernels (all NULL on the plain path).
struct LearnArgs {
    float* feat;         // [n][kNumFeatures] raw features (neighbour pass)
    const float4* vel;   // predicted velocity (features)
    const float* lam1;   // iteration-1 lambda (features)
    float* lam_keep;     // pass-2 lambda, kept for the fallback of a rejected guess
    float* lam_sum;      // accumulated multiplier of this step (delta-p passes)
    float* resid;        // per-workgroup (sum max(C, 0), sum |C|) pairs (lambda passes)
    const int* skip;     // delta-p / lambda passes return at once when *skip != 0
    const float* lam_src;  // delta-p: multipliers from here (n + 1 entries) instead of pos.w
};

// Per-slot scene geometry of a batch (both NULL for a single scene).
struct Geometry {
    const SceneGeom* geoms;
    const uint8_t* sid;
};

// Launch shape of the neighbour kernels (results do not depend on it).
struct Tuning {
    int block;    // workgroup size: 64, 128, 256 or 512
    int grid;     // > 0: persistent grid of this many workgroups, else one per block
    int xcd_map;  // map workgroups to contiguous particle chunks per XCD
};

int scan_tiles(int m);
// Grids of at most kScanFusedMax tiles (1024 counters each) scan in two launches (tile totals,
// then a down-sweep that adds up the totals before its tile); larger grids in three.
constexpr int kScanFusedMax = 2048;
// Workgroups (= reduction partials) of a neighbour-kernel launch over n particles.
int reduce_blocks(const Tuning& T, int n);

// predict + clamp + hash + count (reads the state, writes cell / local / count)
void launch_predict_hash(hipStream_t st, const StepConsts& K, const Geometry& GE, const float4* pos, const float4* vel,
                         const StepBuffers& B);
// an empty kernel of the particle grid's shape (launch-chain calibration)
void launch_noop(hipStream_t st, int n);
// VALU issue-rate calibration: kValuProbeBlocks workgroups of 256 (8 waves per SIMD on 1024
// SIMDs), 16 * kValuProbeLinks wave64 VALU instructions per wavefront; sink is never written
constexpr int kValuProbeBlocks = 2048;
constexpr int kValuProbeLinks = 4096;
void launch_valu_probe(hipStream_t st, float* sink);
// exclusive scan of count into cstart; clears count for the next step
void launch_scan(hipStream_t st, const StepConsts& K, const StepBuffers& B);
// scatter into buckets, rank by particle id; the predicted state (recomputed from the old one),
// the start-of-step position and the id move to slot order
void launch_sort_reorder(hipStream_t st, const StepConsts& K, const Geometry& GE, const StepBuffers& B,
                         const float4* pos_in, const float4* vel_in, const uint32_t* ids_in, float4* pos_out,
                         float4* vel_out, uint32_t* ids_out, int32_t* dbg_gid);
void launch_neighbors_lambda(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, float4* pos,
                             const StepBuffers& B, float* dbg_lam, const LearnArgs& L);
// lambda over the frozen list; with L.feat also the pass-2 warm-start features and L.lam_keep
void launch_lambda(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, float4* pos,
                   const StepBuffers& B, const LearnArgs& L);
void launch_delta_apply(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, const float4* pos,
                        float4* pos_next, float4* vel, const StepBuffers& B, int last, int first, const LearnArgs& L);
void launch_curl(hipStream_t st, const StepConsts& K, const Tuning& T, float4* pos, const float4* vel,
                 const StepBuffers& B, float* err, float* partial);
// hash_next: also the next step's predict + hash + count on the final state (the next step then
// starts at the scan)
void launch_vort_apply(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, const float4* pos,
                       const float4* vel, float4* vel_out, const StepBuffers& B, bool hash_next);
// learned pass: accept / reject on the device
void launch_guess_gate(hipStream_t st, const float* r0, const float* rg, int np, int n, int* gate, float* stats);
void launch_velocity(hipStream_t st, const StepConsts& K, const float4* pos, float4* vel, const StepBuffers& B);
void launch_export(hipStream_t st, int n, const float4* pos, const float4* vel, const uint32_t* ids, float* pos_xyz,
                   float* vel_xyz);
void launch_import(hipStream_t st, int n, const float* pos_xyz, const float* vel_xyz, float4* pos, float4* vel,
                   uint32_t* ids);
void launch_fill_u32(hipStream_t st, uint32_t* p, size_t m, uint32_t v);
void launch_copy_u128(hipStream_t st, const void* src, void* dst, size_t bytes);

}  // namespace pbf
