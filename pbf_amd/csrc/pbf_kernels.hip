// gfx950 kernels of one PBF solver step.
//
// Data layout in HBM (all per-particle arrays in "slot" order, the cell-sorted order of the
// current step; ids[slot] is the caller's particle index):
//   pos4   float4 (x, y, z, lambda)   double-buffered across Jacobi iterations
//   vel4   float4 (vx, vy, vz, M)     double-buffered across the vorticity pass; M = the
//                                     multiplier accumulated last step (warm start only)
//   xold4  float4 (x_old)             position at the start of the step
//   omega4 float4 (wx, wy, wz, |w|)   curl
//   nbr    uint32 blocks of 8          neighbour slots as byte offsets into the float4 slot
//                                     arrays (slot q stored as 16 q): entry j of slot s at
//                                     ((j / 8) * nstride + s) * 8 + j % 8, so a particle's group
//                                     of 8 is 32 contiguous bytes and a wavefront's group is one
//                                     contiguous 2 KB range
//   ncnt   uint8  [n]                  neighbour count
//   count / cstart  uint32 [nbuckets] / [nbuckets+1]  cell histogram and its exclusive scan;
//                                     each scene's last bucket collects its out-of-grid particles
//
// A batch of independent scenes is one concatenated particle array: scene s owns cells
// [coff, coff + ncells] and, because the sort is by global cell, always the same slot range,
// so one counting sort + scan over all cells is the segmented sort.  geoms[scene_of_slot[s]]
// gives a slot's grid and clamp box; single-scene solvers pass geoms = NULL and use K.
//
// Neighbour kernels are templated on the workgroup size and can run persistent: a grid
// smaller than the particle blocks makes each workgroup walk a contiguous chunk of blocks.
//
// Reference kernels each stage replaces are cited at the kernel.

#include <algorithm>

#include "pbf_device.h"
#include "pbf_kernels.h"

namespace pbf {

constexpr int kBlock = 256;
constexpr int kXcds = 8;  // MI355X: workgroups are dispatched round-robin over 8 XCDs

// Workgroup -> chunk index: the workgroups one XCD receives (blockIdx.x = x, x+8, x+16, ...)
// take one contiguous range of chunks, so that range's neighbour gathers stay in that XCD's
// L2, and the sort kernels' cell counters, buckets and gathers of nearby slots are touched
// from one XCD rather than eight.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Particle blocks of BS slots this workgroup handles: one (full grid) or a contiguous chunk
// (persistent grid).
struct BlockRange {
    int first, last;
    __device__ __forceinline__ BlockRange(int nblocks, bool xcd_map) {
        const int g = xcd_map ? xcd_block() : (int)blockIdx.x;
        if ((int)gridDim.x >= nblocks) {
            first = g;
            last = g + 1;
        } else {
            const int per = (nblocks + gridDim.x - 1) / gridDim.x;
            first = g * per;
            last = min(nblocks, first + per);
        }
    }
};

__device__ __forceinline__ SceneGeom geom_of(const StepConsts& K, const SceneGeom* __restrict__ geoms,
                                             const uint8_t* __restrict__ sid, int s) {
    return geoms ? geoms[sid[s]] : geom_single(K);
}

// ---------------------------------------------------------------------------------------
// wavefront (64-lane) inclusive scan with DPP: Hillis-Steele inside each row of 16 lanes
// (row_shr 1, 2, 4, 8 with zero fill), then row_bcast:15 and row_bcast:31 carry the row
// totals across rows.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    return x;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    x = wave_inclusive_scan(x);
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// updateParticles (cu:4-13) + BoundParticle: v += dt g, x += dt v, clamp.  Evaluated by the hash
// pass for the cell and again by the reorder for the state it writes (same operations, same bits),
// so the predicted state is never stored in between.
__device__ __forceinline__ void predict(float4& x, float4& v, const StepConsts& K, const SceneGeom& G) {
    v.x = v.x + K.dt * K.g[0];
    v.y = v.y + K.dt * K.g[1];
    v.z = v.z + K.dt * K.g[2];
    x.x = x.x + K.dt * v.x;
    x.y = x.y + K.dt * v.y;
    x.z = x.z + K.dt * v.z;
    bound_particle(x.x, x.y, x.z, K, G);
}

// ---------------------------------------------------------------------------------------
// updateParticles (cu:4-13) + calculateParticlesGid (cu:31-40): predict, clamp, hash, count.
// The returned atomic slot gives an (unordered) place inside the cell bucket.
constexpr int kScanTile = 1024;  // counters per scan tile (256 threads x 4)

// Counts this lane's particle into counter c and returns its (unordered) place in the cell's
// bucket, with one atomic per run of equal cells in the wavefront (slots are in the previous
// step's cell order, so runs are long): the run head adds the run length and every lane takes
// base + its rank in the run.  Any order stays correct; runs only save atomics.  The active
// lanes must be a prefix of the wavefront (slot order, tail lanes inactive).
__device__ __forceinline__ uint32_t count_in_cell(uint32_t* __restrict__ count, int c) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(c, 1, 64);
    const bool head = lane == 0 || prev != c;
    const uint64_t heads = __ballot(head);
    const uint64_t active = __ballot(1);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int hpos = 63 - __clzll(heads & upto);                 // head of this lane's run
    const uint64_t later = heads & ~upto;                          // heads after this lane
    const int next = later ? (__ffsll((unsigned long long)later) - 1) : 64;
    const int run_end = min(next, 64 - __clzll(active));           // one past the run's last lane
    uint32_t base = 0;
    if (head) base = atomicAdd(&count[c], (uint32_t)(run_end - lane));
    base = __shfl(base, hpos, 64);
    return base + (uint32_t)(lane - hpos);
}

__global__ __launch_bounds__(kBlock) void k_predict_hash(StepConsts K, const SceneGeom* __restrict__ geoms,
                                                         const uint8_t* __restrict__ sid, const float4* __restrict__ pos,
                                                         const float4* __restrict__ vel, uint32_t* __restrict__ cell,
                                                         uint32_t* __restrict__ local, uint32_t* __restrict__ count) {
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= K.n) return;  // whole tail lanes leave; ballots below count the active lanes only
    const SceneGeom G = geom_of(K, geoms, sid, s);
    float4 x = pos[s];
    float4 v = vel[s];
    predict(x, v, K, G);
    const int c = cell_of(x.x, x.y, x.z, K, G);
    const uint32_t l = count_in_cell(count, c);
    cell[s] = (uint32_t)c;
    local[s] = l;
}

// ---------------------------------------------------------------------------------------
// Exclusive scan of count[0..m) into cstart[0..m], cstart[m] = total (calculatePrefixSum +
// addBlockSum + setGrid, cu:42-111).  Reduce-then-scan over tiles of 1024 counters: tile sums,
// one workgroup scanning the tile sums, then every tile rescanned with its offset.  The last
// pass also clears the counters for the next step.

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ count, int m,
                                                        uint32_t* __restrict__ tile_sum) {
    __shared__ uint32_t wsum[kBlock / 64];
    int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t t = 0;
    if (base + 3 < m) {
        uint4 c = *reinterpret_cast<const uint4*>(count + base);
        t = c.x + c.y + c.z + c.w;
    } else {
        for (int j = 0; j < 4; ++j)
            if (base + j < m) t += count[base + j];
    }
    t = wave_sum(t);
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[wave] = t;
    __syncthreads();
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* __restrict__ tile_sum, int ntiles,
                                                     uint32_t* __restrict__ cstart, int m) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < ntiles; base += 1024) {
        int i = base + threadIdx.x;
        uint32_t v = i < ntiles ? tile_sum[i] : 0u;
        uint32_t inc = wave_inclusive_scan(v);
        int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t off = carry_s;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (i < ntiles) tile_sum[i] = off + inc - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = off + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) cstart[m] = carry_s;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ count, int m,
                                                      const uint32_t* __restrict__ tile_off,
                                                      uint32_t* __restrict__ cstart) {
    __shared__ uint32_t wsum[kBlock / 64];
    int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t c[4];
    bool full = base + 3 < m;
    if (full) {
        uint4 q = *reinterpret_cast<const uint4*>(count + base);
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        c[3] = q.w;
    } else {
        for (int j = 0; j < 4; ++j) c[j] = base + j < m ? count[base + j] : 0u;
    }
    uint32_t t = c[0] + c[1] + c[2] + c[3];
    uint32_t inc = wave_inclusive_scan(t);
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = tile_off[blockIdx.x];
    for (int w = 0; w < wave; ++w) off += wsum[w];
    off += inc - t;
    uint32_t o[4];
    o[0] = off;
    o[1] = o[0] + c[0];
    o[2] = o[1] + c[1];
    o[3] = o[2] + c[2];
    if (full) {
        *reinterpret_cast<uint4*>(cstart + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(count + base) = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int j = 0; j < 4; ++j)
            if (base + j < m) {
                cstart[base + j] = o[j];
                count[base + j] = 0u;
            }
    }
}

// The down-sweep without a separate scan of the tile totals (grids of up to kScanFusedMax
// tiles): each workgroup adds up the totals k_scan_reduce wrote for the tiles before its own,
// scans its tile as k_scan_down does and clears its counters; the last tile writes the grand
// total.
__global__ __launch_bounds__(kBlock) void k_scan_fused(uint32_t* __restrict__ count, int m,
                                                       const uint32_t* __restrict__ tile_sum,
                                                       uint32_t* __restrict__ cstart) {
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t wpre[kBlock / 64];
    const int wave = threadIdx.x >> 6;
    uint32_t pre = 0;
    for (int j = threadIdx.x; j < (int)blockIdx.x; j += kBlock) pre += tile_sum[j];
    pre = wave_sum(pre);
    if ((threadIdx.x & 63) == 0) wpre[wave] = pre;
    int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    uint32_t c[4];
    bool full = base + 3 < m;
    if (full) {
        uint4 q = *reinterpret_cast<const uint4*>(count + base);
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        c[3] = q.w;
    } else {
        for (int j = 0; j < 4; ++j) c[j] = base + j < m ? count[base + j] : 0u;
    }
    uint32_t t = c[0] + c[1] + c[2] + c[3];
    uint32_t inc = wave_inclusive_scan(t);
    if ((threadIdx.x & 63) == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = wpre[0] + wpre[1] + wpre[2] + wpre[3];
    const uint32_t tile_start = off;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    off += inc - t;
    uint32_t o[4];
    o[0] = off;
    o[1] = o[0] + c[0];
    o[2] = o[1] + c[1];
    o[3] = o[2] + c[2];
    if (full) {
        *reinterpret_cast<uint4*>(cstart + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(count + base) = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int j = 0; j < 4; ++j)
            if (base + j < m) {
                cstart[base + j] = o[j];
                count[base + j] = 0u;
            }
    }
    if (threadIdx.x == 0 && (int)blockIdx.x == (int)gridDim.x - 1)
        cstart[m] = tile_start + wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// ---------------------------------------------------------------------------------------
// countingSort (cu:112-122) scatter into cell buckets, unordered inside a bucket.
__global__ __launch_bounds__(kBlock) void k_scatter(int n, const uint32_t* __restrict__ cell,
                                                    const uint32_t* __restrict__ local,
                                                    const uint32_t* __restrict__ cstart,
                                                    const uint32_t* __restrict__ ids_in, uint32_t* __restrict__ bucket,
                                                    uint32_t* __restrict__ bucket_id) {
    int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t d = cstart[cell[s]] + local[s];
    bucket[d] = (uint32_t)s;
    bucket_id[d] = ids_in[s];  // the cell's member ids side by side for the ranking
}

// Orders each bucket by ascending particle id (rank = members with a smaller id) and moves the
// particle's predicted position, velocity, start-of-step position and id to its final slot.
// Out-of-grid particles keep their bucket order: nothing reads their relative order.
__global__ __launch_bounds__(kBlock) void k_rank_reorder(StepConsts K, const SceneGeom* __restrict__ geoms,
                                                         const uint8_t* __restrict__ sid,
                                                         const uint32_t* __restrict__ bucket,
                                                         const uint32_t* __restrict__ bucket_id,
                                                         const uint32_t* __restrict__ cell,
                                                         const uint32_t* __restrict__ cstart,
                                                         const float4* __restrict__ vel_in,
                                                         const float4* __restrict__ pos_in, float4* __restrict__ pos_out,
                                                         float4* __restrict__ vel_out, float4* __restrict__ xold_out,
                                                         uint32_t* __restrict__ ids_out, int32_t* __restrict__ dbg_gid) {
    int t = xcd_block() * kBlock + threadIdx.x;
    if (t >= K.n) return;
    const uint32_t e = bucket[t];
    const uint32_t id = bucket_id[t];
    // the state gathers depend on e only: issued before the ranking's chain of loads
    const float4 xo = pos_in[e];
    float4 v = vel_in[e];
    const uint32_t c = cell[e];
    const SceneGeom G = geom_of(K, geoms, sid, (int)e);
    const bool inside = (int)c != G.coff + G.ncells;
    int d = t;
    if (inside) {
        const uint32_t st = cstart[c], en = cstart[c + 1];
        // four members per round trip: the loads are clamped into [st, en) (always in the
        // bucket) and the members past en are masked out of the count
        uint32_t rank = 0;
        for (uint32_t q = st; q < en; q += 4) {
            const uint32_t last = en - 1;
            const uint32_t a0 = bucket_id[q];
            const uint32_t a1 = bucket_id[min(q + 1, last)];
            const uint32_t a2 = bucket_id[min(q + 2, last)];
            const uint32_t a3 = bucket_id[min(q + 3, last)];
            rank += (a0 < id ? 1u : 0u) + (q + 1 < en && a1 < id ? 1u : 0u) + (q + 2 < en && a2 < id ? 1u : 0u) +
                    (q + 3 < en && a3 < id ? 1u : 0u);
        }
        d = (int)(st + rank);
    }
    float4 x = xo;
    predict(x, v, K, G);
    pos_out[d] = make_float4(x.x, x.y, x.z, 0.0f);
    vel_out[d] = v;
    xold_out[d] = xo;
    ids_out[d] = id;
    if (dbg_gid) dbg_gid[id] = cell_id_linear(x.x, x.y, x.z, K, G);  // the reference's id, whatever the row order
}

// ---------------------------------------------------------------------------------------
// Per-pair lambda terms (calculateLambda, cu:156-180).
struct LambdaAcc {
    float rho, gms, gx, gy, gz;
    __device__ __forceinline__ void init() { rho = gms = gx = gy = gz = 0.0f; }
    template <bool FAST>
    __device__ __forceinline__ void add(float rx, float ry, float rz, float r2, const StepConsts& K) {
        rho = rho + poly6(r2, K);
        float t = spiky_scale_t<FAST>(r2, K);
        float ax = rx * t, ay = ry * t, az = rz * t;
        gx = gx + ax;
        gy = gy + ay;
        gz = gz + az;
        gms = __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, __builtin_fmaf(ax, ax, gms)));
    }
    __device__ __forceinline__ float lambda(const StepConsts& K) const {
        float Ci = rho / K.restRho - 1.0f;
        float g = gms + (gx * gx + gy * gy + gz * gz);
        g = g / (K.restRho * K.restRho);
        return -Ci / (g + K.eps);
    }
};

__device__ __forceinline__ size_t nbr_index(int s, int j, int nstride) {
    return ((size_t)(j >> 3) * (size_t)nstride + (size_t)s) * 8 + (size_t)(j & 7);
}


// Neighbour loops fetch in groups of kGroup: all list entries of the group, then all gathered
// operands, then the group is processed in list order (the summation order of the reference's
// loop).  kGroup independent loads in flight per lane instead of one.  The search pads the last
// group with the padding slot n (kFarPos), whose terms are +-0 and leave every sum unchanged,
// so a group is straight-line code with no per-entry test.
constexpr int kGroup = 8;

// Entry q of a float4 slot array by its 32-bit byte offset qb = 16 q (q <= n + 31 < 2^28), the value
// a neighbour-list entry holds: the load takes the array's scalar base and one vector offset.
__device__ __forceinline__ const float4& slot_at(const float4* __restrict__ base, uint32_t qb) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + qb);
}

// The 8 list entries of group j0 (a multiple of 8, below cnt) of slot s.
__device__ __forceinline__ void load_group(const uint32_t* __restrict__ nbr, int s, int j0, int nstride,
                                           uint32_t q[8]) {
    const uint4* p = reinterpret_cast<const uint4*>(nbr + nbr_index(s, j0, nstride));
    const uint4 a = p[0], b = p[1];
    q[0] = a.x;
    q[1] = a.y;
    q[2] = a.z;
    q[3] = a.w;
    q[4] = b.x;
    q[5] = b.y;
    q[6] = b.z;
    q[7] = b.w;
}

// The largest list length among the wavefront's active lanes, as a wave-uniform value (counts are
// below 256; one ballot per bit).  Past it every active lane's entries are the padding slot, so
// the neighbour loops end there instead of at the end of the group: VALU instructions issue for
// the whole wavefront, and the padding terms (+-0 into sums that start at +0) change nothing.
__device__ __forceinline__ int wave_max_count(int cnt) {
    int m = 0;
#pragma unroll
    for (int b = 7; b >= 0; --b) {
        const int t = m | (1 << b);
        if (__ballot(cnt >= t)) m = t;
    }
    return m;
}

// g(0) .. g(k - 1) for one group, k = rem (the entries left before the wavefront's largest count)
// rounded up to even and at most 8.  The tests before entries 2, 4 and 6 are wave-uniform
// branches; a branch per entry, or a separate branch-free copy for whole groups, measured slower.
template <class G>
__device__ __forceinline__ void run_group(int rem, G&& g) {
    static_assert(kGroup == 8, "the tail below is written out for groups of 8");
    g(0);
    g(1);
    if (rem > 2) {
        g(2);
        g(3);
        if (rem > 4) {
            g(4);
            g(5);
            if (rem > 6) {
                g(6);
                g(7);
            }
        }
    }
}

// Calls f(qb, x_q) for the list entries j = 0 .. min(ceil8(cnt), ceil2(cmax)) - 1 of slot s in
// order, qb = 16 q the entry's byte offset and x_q = pos[q] (entries past cnt are the padding
// slot), cmax the wavefront's largest count.
// PF (grids of at most one wavefront per SIMD, where every memory round trip shows): the next
// group's entries load while the current group is summed, issued after the gathers and without a
// branch (the last group reloads itself), so that the wait for the gathers leaves them in flight.
// With more wavefronts the others hide the round trip and the prefetch costs time.
template <bool PF, class F>
__device__ __forceinline__ void for_neighbors(const uint32_t* __restrict__ nbr, int s, int nstride, int cnt,
                                              const float4* __restrict__ pos, F&& f) {
    const int cmax = wave_max_count(cnt);
    uint32_t qn[kGroup];
    if (PF && cnt > 0) load_group(nbr, s, 0, nstride, qn);
    for (int j0 = 0; j0 < cnt; j0 += kGroup) {
        uint32_t q[kGroup];
        if constexpr (PF) {
#pragma unroll
            for (int u = 0; u < kGroup; ++u) q[u] = qn[u];
        } else {
            load_group(nbr, s, j0, nstride, q);
        }
        float4 x[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) x[u] = slot_at(pos, q[u]);
        if constexpr (PF) load_group(nbr, s, min(j0 + kGroup, (cnt - 1) & ~(kGroup - 1)), nstride, qn);
        // The gathers are used here so that the compiler issues all of them before the tail
        // branches below instead of sinking each into its branch behind a wait of its own.
        asm volatile("" ::"v"(x[0].x), "v"(x[1].x), "v"(x[2].x), "v"(x[3].x), "v"(x[4].x), "v"(x[5].x),
                     "v"(x[6].x), "v"(x[7].x));
        run_group(cmax - j0, [&](int u) { f(q[u], x[u]); });
    }
}

// Block sums of two per-thread values in a fixed tree order into partial[2 * blockIdx.x + {0,1}].
template <int BS>
__device__ void block_sum2(float c, float d, float* __restrict__ partial) {
    __shared__ float sr[BS], sd[BS];
    int t = threadIdx.x;
    sr[t] = c;
    sd[t] = d;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (t < w) {
            sr[t] = sr[t] + sr[t + w];
            sd[t] = sd[t] + sd[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[2 * blockIdx.x] = sr[0];
        partial[2 * blockIdx.x + 1] = sd[0];
    }
}

// buildNeighborhood (cu:124-154): the 27 cells in z, y, x order (the three x-neighbours of a row
// are one contiguous slot run), self excluded, r.r <= h*h, at most cap hits, in that order.
// Eight lanes work on one particle: they test 8 consecutive candidates of a run at a time (8
// consecutive positions per load) and the hit masks give each hit its place in the list, so the
// list keeps the candidates' order.  Cell culling: a neighbour cell whose box lies farther than h
// from the particle cannot hold a neighbour and is left out of the run (the x-cells that remain
// are contiguous, so the order is unchanged).  Distances are taken to the particle's own cell
// faces less a margin of 1e-3 h, which covers the rounding of the cell index.
// The list is assembled in LDS (one row per particle) and then stored one group of 8 at a time,
// 32 contiguous bytes per subgroup, the last group padded with the padding slot n.
constexpr int kSub = 8;  // lanes per particle in the search

// b[R] = the value of lane R of this lane's group of 8 (R = 0..RMAX), for two values: a bit-mask
// swizzle inside 32 lanes, source lane = (lane & 0x18) | R.
template <int RMAX>
__device__ __forceinline__ void subgroup_bcast(uint32_t x, uint32_t y, uint32_t* bx, uint32_t* by) {
    bx[RMAX] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (RMAX << 5));
    by[RMAX] = (uint32_t)__builtin_amdgcn_ds_swizzle((int)y, 0x18 | (RMAX << 5));
    if constexpr (RMAX > 0) subgroup_bcast<RMAX - 1>(x, y, bx, by);
}

// Words per particle row of the search's LDS list.  A pass of 4 chunks starts with fewer than cap
// hits and adds at most 32, so a row of cap + 32 words or more (odd, so that the rows of a
// wavefront's eight particles start in different banks) takes every hit without a test against
// the cap; the hits past the cap are never read.  When such rows do not fit in 64 KiB, rows of
// cap words and a test per hit.
__host__ __device__ constexpr int search_lds_row_slack(int cap) { return ((cap + 7) & ~7) + 33; }
__host__ __device__ constexpr bool search_lds_slack(int cap, int ppb) {
    return ppb * search_lds_row_slack(cap) * 4 <= 65536;
}
__host__ __device__ constexpr int search_lds_row(int cap, int ppb) {
    return search_lds_slack(cap, ppb) ? search_lds_row_slack(cap) : ((cap + 7) & ~7);
}

// WIDE: the table's group stride nstride * 8 needs more than 24 bits (n > 2^21), so list
// offsets take a full 32-bit multiply instead of the 24-bit one.  SLACK: LDS rows with room for
// the hits past the cap (search_lds_row).
template <int BS, bool WIDE, bool SLACK>
__global__ __launch_bounds__(BS) void k_search(StepConsts K, const SceneGeom* __restrict__ geoms,
                                               const uint8_t* __restrict__ sid, const float4* __restrict__ pos,
                                               const uint32_t* __restrict__ cstart, uint32_t* __restrict__ nbr,
                                               uint8_t* __restrict__ ncnt, int xcd_map) {
    extern __shared__ uint32_t lds_list[];
    constexpr int PPB = BS / kSub;  // particles per workgroup pass
    const int row_words = search_lds_row(K.cap, PPB);
    const BlockRange R((K.n + PPB - 1) / PPB, xcd_map);
    const int lane8 = threadIdx.x & (kSub - 1);
    const int wsub = (threadIdx.x & 63) / kSub;  // subgroup inside the wavefront
    const uint32_t below = (1u << lane8) - 1u;
    uint32_t* const lrow = lds_list + (threadIdx.x / kSub) * row_words;
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * PPB + (int)threadIdx.x / kSub;
        if (s >= K.n) continue;  // whole subgroups leave together
        const SceneGeom G = geom_of(K, geoms, sid, s);
        const float4 xi = pos[s];
        const int d0 = dim_of(xi.x, G.gmin[0], K.invH, G.dims[0]);
        const int d1 = dim_of(xi.y, G.gmin[1], K.invH, G.dims[1]);
        const int d2 = dim_of(xi.z, G.gmin[2], K.invH, G.dims[2]);
        const int xlo = max(d0 - 1, 0), xhi = min(d0 + 1, G.dims[0] - 1);
        const bool inside = d0 >= 0 && d0 < G.dims[0] && d1 >= 0 && d1 < G.dims[1] && d2 >= 0 && d2 < G.dims[2];
        float fl[3] = {0.0f, 0.0f, 0.0f}, fh[3] = {0.0f, 0.0f, 0.0f};
        if (inside) {
            const float p3[3] = {xi.x, xi.y, xi.z};
            const int d3[3] = {d0, d1, d2};
            const float margin = 1e-3f * K.h;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float lo = G.gmin[a] + (float)d3[a] * K.h;
                fl[a] = fmaxf(p3[a] - lo - margin, 0.0f);
                fh[a] = fmaxf(lo + K.h - p3[a] - margin, 0.0f);
            }
        }
        // The 9 row runs (dz, dy) = (r / 3 - 1, r % 3 - 1), empty when out of the grid or culled.
        // Lane r of the subgroup finds the bounds of run r < 8 and every lane those of run 8; a
        // swizzle then hands run r's bounds from lane r to the whole subgroup, so the eight lanes
        // do not each repeat all nine.
        auto run_bounds = [&](int r, uint32_t& b, uint32_t& e) {
            const int dz = r / 3 - 1, dy = r - 3 * (r / 3) - 1;
            const int cz = d2 + dz, cy = d1 + dy;
            const float ez = dz < 0 ? fl[2] : (dz > 0 ? fh[2] : 0.0f);
            const float ey = dy < 0 ? fl[1] : (dy > 0 ? fh[1] : 0.0f);
            const float eyz = ey * ey + ez * ez;
            int rlo = xlo, rhi = xhi;
            bool use = cz >= 0 && cz < G.dims[2] && cy >= 0 && cy < G.dims[1] && xlo <= xhi;
            if (inside) {
                use = use && eyz <= K.h2;
                if (eyz + fl[0] * fl[0] > K.h2) rlo = d0;
                if (eyz + fh[0] * fh[0] > K.h2) rhi = d0;
            }
            // both indices stay inside the scene's cell table even for an unused run; cell
            // counts stay below 2^28, so their byte offsets fit in 32 bits
            const int row = use ? G.coff + row_of(use ? cy : 0, use ? cz : 0, K, G) * G.dims[0] : G.coff;
            const char* cb = reinterpret_cast<const char*>(cstart);
            const uint32_t lo = *reinterpret_cast<const uint32_t*>(cb + ((uint32_t)(use ? row + rlo : row) << 2));
            const uint32_t hi = *reinterpret_cast<const uint32_t*>(cb + ((uint32_t)(use ? row + rhi + 1 : row) << 2));
            b = use ? lo : 0u;
            e = use ? hi : 0u;
        };
        uint32_t mb, me, rb[9], re[9];
        run_bounds(lane8, mb, me);
        run_bounds(8, rb[8], re[8]);
        subgroup_bcast<7>(mb, me, rb, re);
        // Four chunks of 8 candidates per pass: their loads are issued together, at immediate
        // offsets from one address.  Loads past the run's end (at most 31 slots, since the pass
        // starts inside the run) read particles of later runs or the padding slots after slot n;
        // their lanes are masked out.  A chunk that holds no candidate for any particle of the
        // wavefront is skipped (a wave-uniform branch), so a run costs ceil(len / 8) chunks.
        constexpr int kChunks = 4;
        const char* const pb = reinterpret_cast<const char*>(pos);
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int e = (int)re[r];
            for (int q0 = (int)rb[r]; q0 < e && cnt < K.cap; q0 += kChunks * kSub) {
                const int ql = q0 + lane8;  // this lane's candidate in chunk 0
                const int left = e - ql;    // chunk u holds a candidate for this lane iff left > u * kSub
                const char* const pl = pb + (size_t)((uint32_t)ql << 4);
                float4 xs[kChunks];
#pragma unroll
                for (int u = 0; u < kChunks; ++u) xs[u] = *reinterpret_cast<const float4*>(pl + u * kSub * 16);
#pragma unroll
                for (int u = 0; u < kChunks; ++u) {
                    // wave masks of the compares (predicates: 5 ordered <=), combined in scalar
                    // registers; the run mask also decides the skip
                    const bool in_run = left > u * kSub;
                    const uint64_t m_run = __builtin_amdgcn_ballot_w64(in_run);
                    if (u > 0 && m_run == 0) break;
                    const float rx = xi.x - xs[u].x, ry = xi.y - xs[u].y, rz = xi.z - xs[u].z;
                    const float d2 = rx * rx + ry * ry + rz * rz;
                    // the particle itself is a candidate of its own row's run only (r = 4: dz = dy = 0)
                    const bool other = r != 4 || ql + u * kSub != s;
                    uint64_t m = m_run & __builtin_amdgcn_fcmpf(d2, K.h2, 5);
                    if (r == 4) m &= __builtin_amdgcn_ballot_w64(other);
                    const bool hit = in_run && other && d2 <= K.h2;
                    const uint32_t bits = (uint32_t)(m >> (kSub * wsub)) & 0xffu;
                    const int at = cnt + __popc(bits & below);
                    if (hit && (SLACK || at < K.cap)) lrow[at] = ((uint32_t)ql << 4) + (uint32_t)(u * kSub * 16);
                    cnt += __popc(bits);
                }
            }
        }
        // the list to its groups of 8; the lanes of a subgroup read their own particle's row
        const int c = min(cnt, K.cap);
        uint32_t* const lbase = nbr + (size_t)s * 8;  // entry j at lbase[(j / 8) * gstride + j % 8]
        const uint32_t gstride = (uint32_t)K.nstride * 8u;
        for (int g = 0; g * 8 < c; ++g) {
            const int j = g * 8 + lane8;
            const uint32_t v = j < c ? lrow[j] : (uint32_t)K.n << 4;
            lbase[(WIDE ? (uint32_t)g * gstride : __umul24((uint32_t)g, gstride)) + (uint32_t)lane8] = v;
        }
        if (lane8 == 0) ncnt[s] = (uint8_t)c;
    }
}

// The first calculateLambda (cu:156-180), over the list just built.  Lambda of iteration 1 goes
// to pos.w and to lam1 (a warm-start feature of pass 2).
// With L.feat, C at the predicted positions goes to feature slot 10; with L.resid the workgroup
// sums of max(C, 0) and |C| at the predicted positions.
template <int BS, bool PF, bool FAST>
__global__ __launch_bounds__(BS) void k_lambda_first(StepConsts K, float4* __restrict__ pos,
                                                     const uint32_t* __restrict__ nbr, const uint8_t* __restrict__ ncnt,
                                                     float* __restrict__ lam1, float* __restrict__ dbg_lam,
                                                     LearnArgs L, int xcd_map) {
    const BlockRange R((K.n + BS - 1) / BS, xcd_map);
    float comp = 0.0f, dev = 0.0f;
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * BS + threadIdx.x;
        if (s >= K.n) continue;
        const float4 xi = pos[s];
        const int cnt = ncnt[s];
        LambdaAcc A;
        A.init();
        for_neighbors<PF>(nbr, s, K.nstride, cnt, pos, [&](uint32_t, const float4& xj) {
            float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            A.template add<FAST>(rx, ry, rz, dot3f(rx, ry, rz), K);
        });
        const float lam = A.lambda(K);
        reinterpret_cast<float*>(pos + s)[3] = lam;
        lam1[s] = lam;
        if (dbg_lam) dbg_lam[s] = lam;
        const float C = A.rho / K.restRho - 1.0f;
        if (L.feat) L.feat[(size_t)s * kNumFeatures + 10] = C;  // the rest comes from the pass-2 lambda
        comp = comp + fmaxf(C, 0.0f);
        dev = dev + fabsf(C);
    }
    if (L.resid) block_sum2<BS>(comp, dev, L.resid);
}

// calculateLambda (cu:156-180) over the frozen list, iterations 2..I.  With L.resid, the
// workgroup sums of max(C_i, 0) and |C_i| at the current positions are written as well.
template <int BS, bool FEAT, bool PF, bool FAST>
__global__ __launch_bounds__(BS) void k_lambda(StepConsts K, const SceneGeom* __restrict__ geoms,
                                               const uint8_t* __restrict__ sid, float4* __restrict__ pos,
                                               const uint32_t* __restrict__ nbr, const uint8_t* __restrict__ ncnt,
                                               LearnArgs L, int xcd_map) {
    if (L.skip && *L.skip) return;
    const BlockRange R((K.n + BS - 1) / BS, xcd_map);
    float comp = 0.0f, dev = 0.0f;
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * BS + threadIdx.x;
        if (s >= K.n) continue;
        const float4 xi = pos[s];
        int cnt = ncnt[s];
        LambdaAcc A;
        A.init();
        for_neighbors<PF>(nbr, s, K.nstride, cnt, pos, [&](uint32_t, const float4& xj) {
            float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            A.template add<FAST>(rx, ry, rz, dot3f(rx, ry, rz), K);
        });
        const float lam = A.lambda(K);
        reinterpret_cast<float*>(pos + s)[3] = lam;
        const float C = A.rho / K.restRho - 1.0f;
        if constexpr (FEAT) {
            // the pass-2 lambda: warm-start features (layout in pbf_kernels.h; slot 10 was written
            // by the first lambda) and a copy of lambda for the fallback of a rejected guess
            const SceneGeom G = geom_of(K, geoms, sid, s);
            const float4 v = L.vel[s];
            const float ih = 1.0f / K.h, ih2 = ih * ih;
            float* f = L.feat + (size_t)s * kNumFeatures;
            float w[6];
            const float p3[3] = {xi.x, xi.y, xi.z};
            float dmin = 4.0f;
            for (int a = 0; a < 3; ++a) {
                w[2 * a] = fminf(fmaxf((p3[a] - G.bmin[a]) * ih, 0.0f), 4.0f);
                w[2 * a + 1] = fminf(fmaxf((G.bmax[a] - p3[a]) * ih, 0.0f), 4.0f);
                dmin = fminf(dmin, fminf(w[2 * a], w[2 * a + 1]));
            }
            f[0] = C;
            f[1] = (float)cnt;
            f[2] = A.gms / (K.restRho * K.restRho);
            f[3] = sqrtf(A.gx * A.gx + A.gy * A.gy + A.gz * A.gz) / K.restRho;
            f[4] = v.x * K.dt * ih;
            f[5] = v.y * K.dt * ih;
            f[6] = v.z * K.dt * ih;
            f[7] = dmin;
            f[8] = lam * ih2;
            f[9] = L.lam1[s] * ih2;
            f[11] = w[0];
            f[12] = w[1];
            f[13] = w[2];
            f[14] = w[3];
            f[15] = w[4];
            L.lam_keep[s] = lam;
        }
        comp = comp + fmaxf(C, 0.0f);
        dev = dev + fabsf(C);
    }
    if (L.resid) block_sum2<BS>(comp, dev, L.resid);
}

// calculateOffset (cu:182-199) + updatePosition (cu:15-22): delta-p with s_corr, then
// x + delta-p clamped into the next position buffer; the last iteration also writes
// v = (x - x_old) / dt.  With LA.lam_sum the multiplier of this iteration is accumulated.
// LS: the multipliers come from LA.lam_src (n + 1 entries, the padding slot's 0 included)
// instead of pos.w: the fallback pass after a rejected learned guess reads the kept lambda 2
// there, so pos.w (the guess) needs no restore.
template <int BS, int SN, bool LS, bool PF, bool FAST>
__global__ __launch_bounds__(BS) void k_delta_apply(StepConsts K, const SceneGeom* __restrict__ geoms,
                                                    const uint8_t* __restrict__ sid, const float4* __restrict__ pos,
                                                    float4* __restrict__ pos_next, const uint32_t* __restrict__ nbr,
                                                    const uint8_t* __restrict__ ncnt, const float4* __restrict__ xold,
                                                    float4* __restrict__ vel, int last, int first, LearnArgs LA,
                                                    int xcd_map) {
    if (LA.skip && *LA.skip) return;
    const BlockRange R((K.n + BS - 1) / BS, xcd_map);
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * BS + threadIdx.x;
        if (s >= K.n) continue;
        const float4 xi = pos[s];
        const float lam_i = LS ? LA.lam_src[s] : xi.w;
        int cnt = ncnt[s];
        if (LA.lam_sum) LA.lam_sum[s] = (first ? 0.0f : LA.lam_sum[s]) + lam_i;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for_neighbors<PF>(nbr, s, K.nstride, cnt, pos, [&](uint32_t q, const float4& xj) {
            float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            float r2 = dot3f(rx, ry, rz);
            float w = poly6(r2, K);
            float t = spiky_scale_t<FAST>(r2, K);
            const float lam_j = LS ? LA.lam_src[q >> 4] : xj.w;
            float a = (lam_i + lam_j) + (SN == 4 ? scorr4(w, K) : scorr(w, K));
            ax = madd(a, rx * t, ax);
            ay = madd(a, ry * t, ay);
            az = madd(a, rz * t, az);
        });
        float x = xi.x + ax / K.restRho;
        float y = xi.y + ay / K.restRho;
        float z = xi.z + az / K.restRho;
        const SceneGeom G = geom_of(K, geoms, sid, s);
        bound_particle(x, y, z, K, G);
        pos_next[s] = make_float4(x, y, z, 0.0f);
        if (last) {
            float4 o = xold[s];
            vel[s] = make_float4((x - o.x) / K.dt, (y - o.y) / K.dt, (z - o.z) / K.dt, 0.0f);
        }
    }
}

// calculateCurl (cu:201-215) and the XSPH sum of calculateVorticityAndViscosity (cu:217-239),
// which need the same neighbour velocities; with STATS also the density error rho/rho0 - 1 at the
// final positions into err and / or partial (per workgroup sum |e|, sum max(e,0), max e).
// Outputs for the vorticity pass: omega = (omega, |omega|), |omega| also in pos.w (pos.w is free
// after the last position update, and the vorticity pass then gathers one float4 per neighbour),
// and the XSPH sum sum_j (v_j - v_i) W(r) in xsph (the step-start positions, dead once the
// velocity is updated).  The XSPH term (v_j - v_i) W is the negated (v_i - v_j) W of the
// reference, and x - (-a) = x + a exactly, so the sum is the reference's bit for bit.  This pass
// reads only .xyz of its neighbours' positions.
template <int BS, bool STATS, bool FAST>
__global__ __launch_bounds__(BS) void k_curl(StepConsts K, float4* __restrict__ pos, const float4* __restrict__ vel,
                                             const uint32_t* __restrict__ nbr, const uint8_t* __restrict__ ncnt,
                                             float4* __restrict__ omega, float4* __restrict__ xsph,
                                             float* __restrict__ err, float* __restrict__ partial, int xcd_map) {
    const BlockRange R((K.n + BS - 1) / BS, xcd_map);
    float sa = 0.0f, sp = 0.0f, sm = -INFINITY;
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * BS + threadIdx.x;
        if (s >= K.n) continue;
        const float4 xi = pos[s];
        const float4 vi = vel[s];
        int cnt = ncnt[s];
        float wx = 0.0f, wy = 0.0f, wz = 0.0f, rho = 0.0f, px = 0.0f, py = 0.0f, pz = 0.0f;
        const int cmax = wave_max_count(cnt);
        for (int j0 = 0; j0 < cnt; j0 += kGroup) {
            uint32_t qs[kGroup];
            float4 xs[kGroup], vs[kGroup];
            load_group(nbr, s, j0, K.nstride, qs);
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                xs[u] = slot_at(pos, qs[u]);
                vs[u] = slot_at(vel, qs[u]);
            }
            // all gathers issued ahead of the tail branches (as in for_neighbors)
            asm volatile("" ::"v"(xs[0].x), "v"(xs[1].x), "v"(xs[2].x), "v"(xs[3].x), "v"(xs[4].x),
                         "v"(xs[5].x), "v"(xs[6].x), "v"(xs[7].x));
            asm volatile("" ::"v"(vs[0].x), "v"(vs[1].x), "v"(vs[2].x), "v"(vs[3].x), "v"(vs[4].x),
                         "v"(vs[5].x), "v"(vs[6].x), "v"(vs[7].x));
            run_group(cmax - j0, [&](int u) {
                const float4 xj = xs[u], vj = vs[u];
                float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
                float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
                float r2 = dot3f(rx, ry, rz);
                float t = spiky_scale_t<FAST>(r2, K);
                float gx = rx * t, gy = ry * t, gz = rz * t;
                wx = __builtin_fmaf(uy, gz, __builtin_fmaf(-uz, gy, wx));
                wy = __builtin_fmaf(uz, gx, __builtin_fmaf(-ux, gz, wy));
                wz = __builtin_fmaf(ux, gy, __builtin_fmaf(-uy, gx, wz));
                const float w = poly6(r2, K);
                px = madd(ux, w, px);
                py = madd(uy, w, py);
                pz = madd(uz, w, pz);
                if constexpr (STATS) rho = rho + w;
            });
        }
        const float wlen = sqrtf(wx * wx + wy * wy + wz * wz);
        omega[s] = make_float4(wx, wy, wz, wlen);
        reinterpret_cast<float*>(pos + s)[3] = wlen;
        xsph[s] = make_float4(px, py, pz, 0.0f);
        if constexpr (STATS) {
            float e = rho / K.restRho - 1.0f;
            if (err) err[s] = e;
            sa = sa + fabsf(e);
            sp = sp + fmaxf(e, 0.0f);
            sm = fmaxf(sm, e);
        }
    }
    if (STATS && partial) {
        __shared__ float ra[BS], rp[BS], rm[BS];
        int t = threadIdx.x;
        ra[t] = sa;
        rp[t] = sp;
        rm[t] = sm;
        __syncthreads();
        for (int w = BS / 2; w > 0; w >>= 1) {
            if (t < w) {
                ra[t] = ra[t] + ra[t + w];
                rp[t] = rp[t] + rp[t + w];
                rm[t] = fmaxf(rm[t], rm[t + w]);
            }
            __syncthreads();
        }
        if (t == 0) {
            partial[blockIdx.x * 3 + 0] = ra[0];
            partial[blockIdx.x * 3 + 1] = rp[0];
            partial[blockIdx.x * 3 + 2] = rm[0];
        }
    }
}

// calculateVorticityAndViscosity (cu:217-239) + applyVorticityAndViscosity (cu:24-29): the
// vorticity gradient from the neighbours' |omega| (pos.w, left by the curl pass), the XSPH sum
// from the curl pass.  HASH: also the next step's updateParticles + calculateParticlesGid
// (cu:4-13, 31-40) on the final state, which is in registers here: predicted cell and bucket
// place per slot, counted into the cell counters (the scan of this step has cleared them).  The
// next step's reorder predicts again from the same values, so both agree bit for bit.
template <int BS, bool HASH, bool PF, bool FAST>
__global__ __launch_bounds__(BS) void k_vort_apply(StepConsts K, const SceneGeom* __restrict__ geoms,
                                                   const uint8_t* __restrict__ sid, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float4* __restrict__ omega,
                                                   const float4* __restrict__ xsph, const uint32_t* __restrict__ nbr,
                                                   const uint8_t* __restrict__ ncnt, float4* __restrict__ vel_out,
                                                   uint32_t* __restrict__ cell, uint32_t* __restrict__ local,
                                                   uint32_t* __restrict__ count, int xcd_map) {
    const BlockRange R((K.n + BS - 1) / BS, xcd_map);
    for (int blk = R.first; blk < R.last; ++blk) {
        const int s = blk * BS + threadIdx.x;
        if (s >= K.n) continue;
        const float4 xi = pos[s];
        const float4 vi = vel[s];
        const float4 wi = omega[s];
        const float4 p = xsph[s];
        int cnt = ncnt[s];
        float ex = 0.0f, ey = 0.0f, ez = 0.0f;
        for_neighbors<PF>(nbr, s, K.nstride, cnt, pos, [&](uint32_t, const float4& xj) {
            const float wl = xj.w;  // |omega_j|
            float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
            float r2 = dot3f(rx, ry, rz);
            float t = spiky_scale_t<FAST>(r2, K);
            ex = madd(wl, rx * t, ex);
            ey = madd(wl, ry * t, ey);
            ez = madd(wl, rz * t, ez);
        });
        float el = sqrtf(ex * ex + ey * ey + ez * ez);
        if ((double)fabsf(el) > 1e-3) {
            ex = ex / el;
            ey = ey / el;
            ez = ez / el;
        }
        float cx = ey * wi.z - ez * wi.y;
        float cy = ez * wi.x - ex * wi.z;
        float cz = ex * wi.y - ey * wi.x;
        float fx = K.vort_eps * cx, fy = K.vort_eps * cy, fz = K.vort_eps * cz;
        float vx = K.visc_c * p.x, vy = K.visc_c * p.y, vz = K.visc_c * p.z;
        float4 vn = make_float4(vi.x + (K.dt * fx + vx), vi.y + (K.dt * fy + vy), vi.z + (K.dt * fz + vz), 0.0f);
        vel_out[s] = vn;
        if constexpr (HASH) {
            const SceneGeom G = geom_of(K, geoms, sid, s);
            float4 x = xi;
            predict(x, vn, K, G);
            const int c = cell_of(x.x, x.y, x.z, K, G);
            const uint32_t l = count_in_cell(count, c);
            cell[s] = (uint32_t)c;
            local[s] = l;
        }
    }
}

// Acceptance of a learned first pass, on the device: reduces the per-workgroup (sum max(C,0),
// sum |C|) pairs at the predicted positions (r0) and after the guess pass (rg) in a fixed order
// and keeps the guess iff mean |C| did not grow.  gate[0] = 1 kept / 0 rejected; stats = the
// four means (compression before, after, |C| before, after).
__global__ __launch_bounds__(256) void k_guess_gate(const float* __restrict__ r0, const float* __restrict__ rg,
                                                    int np, float inv_n, int* __restrict__ gate,
                                                    float* __restrict__ stats) {
    __shared__ float sh[4][256];
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = threadIdx.x; i < np; i += 256) {
        a[0] = a[0] + r0[2 * i];
        a[1] = a[1] + rg[2 * i];
        a[2] = a[2] + r0[2 * i + 1];
        a[3] = a[3] + rg[2 * i + 1];
    }
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float c0 = sh[0][0] * inv_n, cg = sh[1][0] * inv_n, d0 = sh[2][0] * inv_n, dg = sh[3][0] * inv_n;
        gate[0] = dg <= d0 ? 1 : 0;
        stats[0] = c0;
        stats[1] = cg;
        stats[2] = d0;
        stats[3] = dg;
    }
}


void launch_guess_gate(hipStream_t st, const float* r0, const float* rg, int np, int n, int* gate, float* stats) {
    hipLaunchKernelGGL(k_guess_gate, dim3(1), dim3(256), 0, st, r0, rg, np, n > 0 ? 1.0f / (float)n : 0.0f, gate, stats);
}


// updatePosition's velocity (cu:21) for a constraint loop that stopped early.
__global__ __launch_bounds__(kBlock) void k_velocity(StepConsts K, const float4* __restrict__ pos,
                                                     const float4* __restrict__ xold, float4* __restrict__ vel) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= K.n) return;
    float4 x = pos[s], o = xold[s];
    vel[s] = make_float4((x.x - o.x) / K.dt, (x.y - o.y) / K.dt, (x.z - o.z) / K.dt, 0.0f);
}

// ---------------------------------------------------------------------------------------
// State in / out in the caller's particle order (AoS xyz).
__global__ __launch_bounds__(kBlock) void k_export(int n, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const uint32_t* __restrict__ ids,
                                                   float* __restrict__ pos_xyz, float* __restrict__ vel_xyz) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    uint32_t i = ids[s];
    if (pos_xyz) {
        float4 p = pos[s];
        pos_xyz[3 * (size_t)i + 0] = p.x;
        pos_xyz[3 * (size_t)i + 1] = p.y;
        pos_xyz[3 * (size_t)i + 2] = p.z;
    }
    if (vel_xyz) {
        float4 v = vel[s];
        vel_xyz[3 * (size_t)i + 0] = v.x;
        vel_xyz[3 * (size_t)i + 1] = v.y;
        vel_xyz[3 * (size_t)i + 2] = v.z;
    }
}

__global__ __launch_bounds__(kBlock) void k_import(int n, const float* __restrict__ pos_xyz,
                                                   const float* __restrict__ vel_xyz, float4* __restrict__ pos,
                                                   float4* __restrict__ vel, uint32_t* __restrict__ ids) {
    int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    pos[s] = make_float4(pos_xyz[3 * (size_t)s], pos_xyz[3 * (size_t)s + 1], pos_xyz[3 * (size_t)s + 2], 0.0f);
    vel[s] = make_float4(vel_xyz[3 * (size_t)s], vel_xyz[3 * (size_t)s + 1], vel_xyz[3 * (size_t)s + 2], 0.0f);
    ids[s] = (uint32_t)s;
}

// Calibration of the launch chain: a kernel with the particle grid's shape (one workgroup of
// kBlock per kBlock particles) that does nothing.  Replayed back to back in a graph, it gives
// the cost of one dependent launch that the stages' short kernels cannot go below.
__global__ __launch_bounds__(kBlock) void k_noop() {}

// Calibration of the VALU issue rate: 8 independent multiply-add chains per lane, each link 2
// wave64 VALU instructions (no contraction, no packed math), kValuProbeLinks links: 65,536 VALU
// instructions per wavefront plus a few for the set-up and the sum.  The sum is stored only if
// it is NaN (never, for the host's constants), which keeps the chains without a store per lane.
__global__ __launch_bounds__(kBlock) void k_valu_probe(float* __restrict__ sink, float m, float k) {
    float x[8];
    for (int i = 0; i < 8; ++i) x[i] = (float)(threadIdx.x + i);
    for (int it = 0; it < kValuProbeLinks; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = x[i] * m + k;
    }
    float s = 0.0f;
    for (int i = 0; i < 8; ++i) s = s + x[i];
    if (s != s) sink[0] = s;
}

// Plain streaming copy (counter calibration: reads and writes exactly `count` uint4).
__global__ __launch_bounds__(kBlock) void k_copy_u128(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      size_t count) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, size_t m, uint32_t v) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) p[i] = v;
}

// ---------------------------------------------------------------------------------------
// launch wrappers
static inline unsigned nblk(size_t n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

int scan_tiles(int m) { return (m + kScanTile - 1) / kScanTile; }

// Workgroups of a neighbour-kernel launch: one per BS particles, or the persistent grid size.
static inline unsigned ngrid(const Tuning& T, int n) {
    unsigned full = nblk(n, T.block);
    return (T.grid > 0 && (unsigned)T.grid < full) ? (unsigned)T.grid : full;
}

int reduce_blocks(const Tuning& T, int n) { return (int)ngrid(T, n > 0 ? n : 1); }

// The search runs kSub lanes per particle: blocks of T.block / kSub particles.
static inline unsigned ngrid_search(const Tuning& T, int n) {
    const unsigned per = (unsigned)(T.block / kSub);
    const unsigned nb = ((unsigned)n + per - 1) / per;
    return T.grid > 0 ? std::min<unsigned>((unsigned)T.grid, nb) : nb;
}

// Dynamic LDS of a search workgroup of bs lanes: one list row per particle.
static inline size_t search_lds_bytes(const StepConsts& K, int bs) {
    return (size_t)(bs / kSub) * (size_t)search_lds_row(K.cap, bs / kSub) * sizeof(uint32_t);
}

void launch_predict_hash(hipStream_t st, const StepConsts& K, const Geometry& GE, const float4* pos, const float4* vel,
                         const StepBuffers& B) {
    if (K.n > 0)
        hipLaunchKernelGGL(k_predict_hash, dim3(nblk(K.n)), dim3(kBlock), 0, st, K, GE.geoms, GE.sid, pos, vel, B.cell,
                           B.local, B.count);
}

void launch_noop(hipStream_t st, int n) {
    if (n > 0) hipLaunchKernelGGL(k_noop, dim3(nblk(n)), dim3(kBlock), 0, st);
}

void launch_valu_probe(hipStream_t st, float* sink) {
    hipLaunchKernelGGL(k_valu_probe, dim3(kValuProbeBlocks), dim3(kBlock), 0, st, sink, 1.0000001f, 1e-7f);
}

void launch_scan(hipStream_t st, const StepConsts& K, const StepBuffers& B) {
    int m = K.nbuckets;
    int nt = scan_tiles(m);
    if (nt <= kScanFusedMax) {
        hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, st, B.count, m, B.tile_sum);
        hipLaunchKernelGGL(k_scan_fused, dim3(nt), dim3(kBlock), 0, st, B.count, m, B.tile_sum, B.cstart);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(kBlock), 0, st, B.count, m, B.tile_sum);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, B.tile_sum, nt, B.cstart, m);
    hipLaunchKernelGGL(k_scan_down, dim3(nt), dim3(kBlock), 0, st, B.count, m, B.tile_sum, B.cstart);
}

void launch_sort_reorder(hipStream_t st, const StepConsts& K, const Geometry& GE, const StepBuffers& B,
                         const float4* pos_in, const float4* vel_in, const uint32_t* ids_in, float4* pos_out,
                         float4* vel_out, uint32_t* ids_out, int32_t* dbg_gid) {
    int n = K.n;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(kBlock), 0, st, n, B.cell, B.local, B.cstart, ids_in, B.bucket,
                       B.bucket_id);
    hipLaunchKernelGGL(k_rank_reorder, dim3(nblk(n)), dim3(kBlock), 0, st, K, GE.geoms, GE.sid, B.bucket, B.bucket_id,
                       B.cell, B.cstart, vel_in, pos_in, pos_out, vel_out, B.xold, ids_out, dbg_gid);
}

// The FAST instances (hardware rsqrt in the spiky gradient) when the step constants ask for them.
#define PBF_DISPATCH_FAST_ONLY(K, CALL)            \
    if ((K).fast) {                                \
        constexpr bool FAST = true;                \
        CALL;                                      \
    } else {                                       \
        constexpr bool FAST = false;               \
        CALL;                                      \
    }

// The neighbour passes' instances by (K).fast and (K).lpf: FAST and PF for CALL.
#define PBF_DISPATCH_FAST(K, CALL)                 \
    if ((K).fast) {                                \
        constexpr bool FAST = true;                \
        if ((K).lpf) {                             \
            constexpr bool PF = true;              \
            CALL;                                  \
        } else {                                   \
            constexpr bool PF = false;             \
            CALL;                                  \
        }                                          \
    } else {                                       \
        constexpr bool FAST = false;               \
        if ((K).lpf) {                             \
            constexpr bool PF = true;              \
            CALL;                                  \
        } else {                                   \
            constexpr bool PF = false;             \
            CALL;                                  \
        }                                          \
    }

#define PBF_DISPATCH_BS(T, CALL)            \
    switch ((T).block) {                    \
        case 64: { constexpr int BS = 64; CALL; } break;   \
        case 128: { constexpr int BS = 128; CALL; } break; \
        case 512: { constexpr int BS = 512; CALL; } break; \
        default: { constexpr int BS = 256; CALL; } break;  \
    }

void launch_neighbors_lambda(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, float4* pos,
                             const StepBuffers& B, float* dbg_lam, const LearnArgs& L) {
    if (K.n <= 0) return;
    const bool wide = (int64_t)K.nstride * 8 >= (int64_t)1 << 24;
    const bool slack = search_lds_slack(K.cap, T.block / kSub);
#define PBF_SEARCH(W, SL)                                                                                              \
    PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_search<BS, W, SL>), dim3(ngrid_search(T, K.n)), dim3(BS),                 \
                                          search_lds_bytes(K, BS), st, K, GE.geoms, GE.sid, pos, B.cstart, B.nbr, B.ncnt, \
                                          T.xcd_map))
    if (wide) {
        if (slack) {
            PBF_SEARCH(true, true);
        } else {
            PBF_SEARCH(true, false);
        }
    } else {
        if (slack) {
            PBF_SEARCH(false, true);
        } else {
            PBF_SEARCH(false, false);
        }
    }
#undef PBF_SEARCH
    PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_lambda_first<BS, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, pos, B.nbr,
                                          B.ncnt, B.lam1, dbg_lam, L, T.xcd_map)));
}

void launch_lambda(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, float4* pos,
                   const StepBuffers& B, const LearnArgs& L) {
    if (K.n <= 0) return;
    if (L.feat) {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_lambda<BS, true, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, GE.geoms,
                                              GE.sid, pos, B.nbr, B.ncnt, L, T.xcd_map)));
    } else {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_lambda<BS, false, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, GE.geoms,
                                              GE.sid, pos, B.nbr, B.ncnt, L, T.xcd_map)));
    }
}

void launch_delta_apply(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, const float4* pos,
                        float4* pos_next, float4* vel, const StepBuffers& B, int last, int first, const LearnArgs& L) {
    if (K.n <= 0) return;
    if (L.lam_src) {  // the fallback after a rejected learned guess (rare): the generic s_corr kernel
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_delta_apply<BS, 0, true, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K,
                                              GE.geoms, GE.sid, pos, pos_next, B.nbr, B.ncnt, B.xold, vel, last, first,
                                              L, T.xcd_map)));
    } else if (K.n_int == 4) {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_delta_apply<BS, 4, false, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K,
                                              GE.geoms, GE.sid, pos, pos_next, B.nbr, B.ncnt, B.xold, vel, last, first,
                                              L, T.xcd_map)));
    } else {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_delta_apply<BS, 0, false, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K,
                                              GE.geoms, GE.sid, pos, pos_next, B.nbr, B.ncnt, B.xold, vel, last, first,
                                              L, T.xcd_map)));
    }
}

void launch_curl(hipStream_t st, const StepConsts& K, const Tuning& T, float4* pos, const float4* vel,
                 const StepBuffers& B, float* err, float* partial) {
    if (K.n <= 0) return;
    if (err || partial) {
        PBF_DISPATCH_FAST_ONLY(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_curl<BS, true, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, pos, vel,
                                              B.nbr, B.ncnt, B.omega, B.xold, err, partial, T.xcd_map)));
    } else {
        PBF_DISPATCH_FAST_ONLY(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_curl<BS, false, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, pos, vel,
                                              B.nbr, B.ncnt, B.omega, B.xold, err, partial, T.xcd_map)));
    }
}

void launch_vort_apply(hipStream_t st, const StepConsts& K, const Geometry& GE, const Tuning& T, const float4* pos,
                       const float4* vel, float4* vel_out, const StepBuffers& B, bool hash_next) {
    if (K.n <= 0) return;
    if (hash_next) {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_vort_apply<BS, true, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K, GE.geoms,
                                              GE.sid, pos, vel, B.omega, B.xold, B.nbr, B.ncnt, vel_out, B.cell, B.local,
                                              B.count, T.xcd_map)));
    } else {
        PBF_DISPATCH_FAST(K, PBF_DISPATCH_BS(T, hipLaunchKernelGGL((k_vort_apply<BS, false, PF, FAST>), dim3(ngrid(T, K.n)), dim3(BS), 0, st, K,
                                              GE.geoms, GE.sid, pos, vel, B.omega, B.xold, B.nbr, B.ncnt, vel_out, B.cell,
                                              B.local, B.count, T.xcd_map)));
    }
}

void launch_velocity(hipStream_t st, const StepConsts& K, const float4* pos, float4* vel, const StepBuffers& B) {
    if (K.n > 0) hipLaunchKernelGGL(k_velocity, dim3(nblk(K.n)), dim3(kBlock), 0, st, K, pos, B.xold, vel);
}

void launch_export(hipStream_t st, int n, const float4* pos, const float4* vel, const uint32_t* ids, float* pos_xyz,
                   float* vel_xyz) {
    if (n > 0) hipLaunchKernelGGL(k_export, dim3(nblk(n)), dim3(kBlock), 0, st, n, pos, vel, ids, pos_xyz, vel_xyz);
}

void launch_import(hipStream_t st, int n, const float* pos_xyz, const float* vel_xyz, float4* pos, float4* vel,
                   uint32_t* ids) {
    if (n > 0) hipLaunchKernelGGL(k_import, dim3(nblk(n)), dim3(kBlock), 0, st, n, pos_xyz, vel_xyz, pos, vel, ids);
}

void launch_copy_u128(hipStream_t st, const void* src, void* dst, size_t bytes) {
    size_t count = bytes / 16;
    if (count == 0) return;
    unsigned grid = (unsigned)std::min<size_t>((count + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_copy_u128, dim3(grid), dim3(kBlock), 0, st, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(dst), count);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, size_t m, uint32_t v) {
    if (m > 0) hipLaunchKernelGGL(k_fill_u32, dim3(nblk(m)), dim3(kBlock), 0, st, p, m, v);
}

}  // namespace pbf
