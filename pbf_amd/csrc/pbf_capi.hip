// Host runtime and C ABI (include/pbf.h) of the gfx950 PBF solver.
//
// One pbf_solver owns its device buffers and one HIP stream.  A step is the launch sequence
// of the reference's cudaPBFSolve (kernels/cudaPBF.cu:50-114) without its per-launch device
// synchronisations: predict+hash+count, scan, scatter, rank+reorder, neighbours+lambda,
// I x (delta-p+apply [, lambda]), curl, vorticity+viscosity+apply.  Optionally the whole step
// is captured once into a hipGraph and replayed.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pbf.h"
#include "pbf_kernels.h"
#include "pbf_learned.h"

using namespace pbf;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return set_error(PBF_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                               __FILE__, __LINE__);                                       \
    } while (0)

const double kInvPi = 0.31830988618;  // INVPI, include/solverBase.h:4

}  // namespace

struct pbf_solver {
    pbf_params P;
    float gs[6];
    float bd[6];
    int n = 0;
    int ncells = 0;
    int nstride = 0;
    int cap = 0;
    int device = 0;
    int cus = 256;  // compute units of the device
    hipStream_t stream = nullptr;
    StepBuffers B{};
    int nbuckets = 0;            // cell counters: ncells + 1, or the sum over a batch
    Geometry geo{nullptr, nullptr};
    int nscenes = 1;
    Tuning tune{256, 0, 1};
    int pc = 0, vc = 0, ic = 0;  // which of pos/vel/ids[2] hold the state
    int lam_last = -1;           // pos buffer whose .w holds the last iteration's lambda
    int steps_done = 0;
    // diagnostics
    int32_t* dbg_gid = nullptr;
    float* dbg_lam = nullptr;
    float* dbg_err = nullptr;
    float* partial = nullptr;
    int npartial = 0;
    bool stats_valid = false;
    bool diag_valid = false;     // the step diagnostics describe the current state's last step
    // The last step already predicted, hashed and counted the current state for the next step
    // (cell, local and the cell counters hold it), so the next step starts at the scan.  Cleared,
    // with the counters zeroed, by anything that changes the state, the step parameters or the
    // cell order (invalidate_hash).
    bool hash_ready = false;
    // consecutive pbf_solve / pbf_step calls that changed the parameters: from the second on, steps
    // launch eagerly (no capture, no drain of the stream) until a call leaves them unchanged
    int param_streak = 0;
    // warm start: features, accumulated multiplier, residual partials, attached network
    float* feat = nullptr;
    float* lam_sum = nullptr;
    float* lam_keep = nullptr;
    float* resid = nullptr;
    float* resid0 = nullptr;
    float* resid_host = nullptr;
    int* gate = nullptr;          // device: 1 if the last learned first pass was kept
    float* gate_stats = nullptr;  // device: compression / mean |C| before and after that pass
    int learned_passes = 0;       // > 0: steps run the learned first pass + (passes - 1) plain
    const pbf::MlpModel* mlp = nullptr;
    // one device-side copy of the state (pbf_snapshot)
    float4* snap_pos = nullptr;
    float4* snap_vel = nullptr;
    uint32_t* snap_ids = nullptr;
    bool snap_valid = false;
    // per-launch profiling: when set, enqueue_step records an event after every launch
    std::vector<hipEvent_t>* prof_events = nullptr;
    std::vector<const char*>* prof_names = nullptr;
    // graph replay, one graph per (step kind, state parity)
    std::map<int, hipGraphExec_t> graphs;
    std::vector<void*> allocs;
    // cell-row order (pbf_set_ordering): 0 = z, then y; 1 = a Morton curve over (y, z) per scene
    std::vector<SceneGeom> scenes;  // host copy of each scene's grid (rowoff, dims)
    int order = 0;
    int* rowmap = nullptr;          // device: the Morton rows of every scene (built on first use)
    std::vector<int> rowmap_host;   // the same on the host (pbf_get_grid)
};

namespace {

StepConsts make_consts(const pbf_solver* s) {
    const pbf_params& P = s->P;
    StepConsts K;
    std::memset(&K, 0, sizeof(K));
    K.h = P.h;
    K.h2 = P.h * P.h;
    float h9 = (float)std::pow((double)P.h, 9.0);
    float h6 = (float)std::pow((double)P.h, 6.0);
    K.c6 = (float)((315.0 * kInvPi) / (double)(64.0f * h9));
    K.cs = (float)((-45.0 * kInvPi) / (double)h6);
    float dq2 = P.deltaq * P.deltaq + 0.0f * 0.0f + 0.0f * 0.0f;
    if (dq2 > K.h2) {
        K.wdq = 0.0f;
    } else {
        float d = K.h2 - dq2;
        K.wdq = K.c6 * (d * d * d);
    }
    K.inv_wdq = 1.0f / K.wdq;
    K.k = P.k;
    K.scorr_n = P.n;
    K.n_int = (P.n >= 0.0f && P.n <= 64.0f && P.n == (float)(int)P.n) ? (int)P.n : -1;
    K.restRho = P.restRho;
    K.eps = P.epsilon;
    K.dt = P.dt;
    K.visc_c = P.c;
    K.vort_eps = P.vorticityConfinementEpilon;
    K.jitter = P.jitter_coeff;
    K.rmin = P.grad_rmin;
    K.rmin2 = P.grad_rmin * P.grad_rmin;
    K.fast = (P.flags & PBF_FLAG_FAST_MATH) ? 1 : 0;
    // at most one wavefront per SIMD: the neighbour passes load list entries a group ahead
    K.lpf = (int64_t)s->n <= (int64_t)64 * 4 * s->cus ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        K.g[a] = P.gravity[a];
        K.gmin[a] = s->gs[a];
        K.dims[a] = P.grid_dims[a];
        K.bmin[a] = s->bd[a];
        K.bmax[a] = s->bd[3 + a];
    }
    K.invH = 1.0f / P.h;
    K.ncells = s->ncells;
    K.cap = s->cap;
    K.n = s->n;
    K.nstride = s->nstride;
    K.nbuckets = s->nbuckets;
    K.rowmap = s->order ? s->rowmap : nullptr;
    return K;
}

int check_params(const pbf_params* p) {
    if (!p) return set_error(PBF_ERR_ARG, "params is NULL");
    if (!(p->h > 0.0f)) return set_error(PBF_ERR_ARG, "h must be > 0");
    if (p->iterations < 0 || p->iterations > 1000) return set_error(PBF_ERR_ARG, "iterations out of range");
    if (p->max_neighbors < 1 || p->max_neighbors > 255)
        return set_error(PBF_ERR_ARG, "max_neighbors must be in 1..255");
    for (int a = 0; a < 3; ++a)
        if (p->grid_dims[a] < 1 || p->grid_dims[a] > 4096) return set_error(PBF_ERR_ARG, "grid_dims out of range");
    int64_t c = (int64_t)p->grid_dims[0] * p->grid_dims[1] * p->grid_dims[2];
    if (c > (int64_t)1 << 28) return set_error(PBF_ERR_ARG, "too many grid cells");
    return PBF_OK;
}

template <class T>
int dalloc(pbf_solver* s, T** p, size_t count) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, bytes));
    s->allocs.push_back(q);
    *p = (T*)q;
    return PBF_OK;
}

void free_all(pbf_solver* s) {
    std::free(s->resid_host);
    s->resid_host = nullptr;
    for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second);
    s->graphs.clear();
    for (void* p : s->allocs) (void)hipFree(p);
    s->allocs.clear();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

// A captured step may still be running (steps only enqueue): the stream drains before the
// executable graphs go.  Graphs are dropped on a parameter change, which recaptures anyway.
void drop_graphs(pbf_solver* s) {
    if (s->graphs.empty()) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second);
    s->graphs.clear();
}

// The next step hashes the state itself: the counters the last step filled for it are zeroed
// (a fill kernel on the solver's stream, ordered before the next step).
void invalidate_hash(pbf_solver* s) {
    if (!s->hash_ready) return;
    (void)hipSetDevice(s->device);
    launch_fill_u32(s->stream, s->B.count, (size_t)scan_tiles(s->nbuckets) * 1024, 0u);
    s->hash_ready = false;
}

// Enqueue one step.  Stage events (7 stage boundaries + start) are recorded when ev != NULL.
int prof_mark(pbf_solver* s, const char* name) {
    if (!s->prof_events) return PBF_OK;
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(e, s->stream));
    s->prof_events->push_back(e);
    s->prof_names->push_back(name);
    return PBF_OK;
}

#define PROF(name)                          \
    do {                                    \
        int rc_ = prof_mark(s, name);       \
        if (rc_) return rc_;                \
    } while (0)

// The learned pass (pass 2 of the step).  pos[a] holds the positions after the plain pass 1 and,
// from the feature lambda pass, lambda 2 in .w (a copy in lam_keep) and the residual partials
// resid0.  The network's multiplier (from the features) replaces lambda 2 and one delta-p pass
// runs on it.  The gate keeps the result only if mean |C| did not grow over the state after pass
// 1; otherwise the pass reruns plain, its delta-p reading lambda 2 from lam_keep (both fallback
// launches return at once when the guess was kept).  Leaves the positions after the pass in
// pos[a ^ 1] and, with need_lambda, their lambda in .w and residual partials in resid.
int enqueue_guess_pass(pbf_solver* s, const StepConsts& K, int a, int v, bool need_lambda) {
    StepBuffers& B = s->B;
    hipStream_t st = s->stream;
    const LearnArgs L0{};
    launch_mlp_infer(st, s->mlp, s->feat, s->n, B.pos[a], nullptr, s->P.h * s->P.h);
    PROF("mlp_infer");
    launch_delta_apply(st, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[v], B, 0, 0, L0);
    PROF("delta_apply");
    LearnArgs LR{};
    LR.resid = s->resid;
    launch_lambda(st, K, s->geo, s->tune, B.pos[a ^ 1], B, LR);
    PROF("lambda");
    launch_guess_gate(st, s->resid0, s->resid, reduce_blocks(s->tune, s->n), s->n, s->gate, s->gate_stats);
    LearnArgs LS{};
    LS.skip = s->gate;
    LS.lam_src = s->lam_keep;
    launch_delta_apply(st, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[v], B, 0, 0, LS);
    if (need_lambda) {
        LearnArgs LL{};
        LL.skip = s->gate;
        LL.resid = s->resid;
        launch_lambda(st, K, s->geo, s->tune, B.pos[a ^ 1], B, LL);
    }
    PROF("guess_gate_fallback");
    HIP_TRY(hipGetLastError());
    return PBF_OK;
}

// Arguments of the feature lambda pass (pass 2's lambda with the warm-start features).
LearnArgs feature_args(pbf_solver* s, int v, float* resid) {
    LearnArgs L{};
    L.feat = s->feat;
    L.vel = s->B.vel[v];
    L.lam1 = s->B.lam1;
    L.lam_keep = s->lam_keep;
    L.resid = resid;
    return L;
}

// Enqueue one step: plain (P.iterations Jacobi iterations) or, with learned_passes >= 2, a plain
// pass, the learned pass and learned_passes - 2 plain iterations.  Stage events (7 stage
// boundaries + start) are recorded when ev != NULL.
int enqueue_step(pbf_solver* s, hipEvent_t* ev) {
    const StepConsts K = make_consts(s);
    StepBuffers& B = s->B;
    hipStream_t st = s->stream;
    const bool dbg = (s->P.flags & PBF_FLAG_DEBUG) != 0;
    const bool stats = dbg || (s->P.flags & PBF_FLAG_DENSITY_STATS);
    const bool learned = s->learned_passes > 0 && s->mlp;
    if (ev) HIP_TRY(hipEventRecord(ev[0], st));
    PROF("start");
    if (!s->hash_ready) {
        launch_predict_hash(st, K, s->geo, B.pos[s->pc], B.vel[s->vc], B);
        PROF("predict_hash");
    }
    if (ev) HIP_TRY(hipEventRecord(ev[1], st));
    launch_scan(st, K, B);
    PROF("scan");
    if (ev) HIP_TRY(hipEventRecord(ev[2], st));
    launch_sort_reorder(st, K, s->geo, B, B.pos[s->pc], B.vel[s->vc], B.ids[s->ic], B.pos[s->pc ^ 1], B.vel[s->vc ^ 1],
                        B.ids[s->ic ^ 1],
                        dbg ? s->dbg_gid : nullptr);
    PROF("scatter_rank_reorder");
    if (ev) HIP_TRY(hipEventRecord(ev[3], st));
    int a = s->pc ^ 1;
    int v = s->vc ^ 1;
    const LearnArgs L0{};
    LearnArgs LF{};
    if (learned) LF.feat = s->feat;  // C at the predicted positions
    launch_neighbors_lambda(st, K, s->geo, s->tune, B.pos[a], B, dbg ? s->dbg_lam : nullptr, LF);
    PROF("neighbors_lambda");
    if (ev) HIP_TRY(hipEventRecord(ev[4], st));
    const int I = learned ? s->learned_passes : s->P.iterations;
    int it = 0;
    if (learned) {
        launch_delta_apply(st, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[v], B, 0, 1, L0);
        PROF("delta_apply");
        a ^= 1;
        launch_lambda(st, K, s->geo, s->tune, B.pos[a], B, feature_args(s, v, s->resid0));
        PROF("lambda_features");
        int rc = enqueue_guess_pass(s, K, a, v, I > 2);
        if (rc) return rc;
        a ^= 1;
        it = 2;
    }
    for (; it < I; ++it) {
        if (it > (learned ? 2 : 0)) {
            launch_lambda(st, K, s->geo, s->tune, B.pos[a], B, L0);
            PROF("lambda");
        }
        launch_delta_apply(st, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[v], B, it == I - 1, it == 0, L0);
        PROF("delta_apply");
        a ^= 1;
    }
    if (learned && I == 2) {
        launch_velocity(st, K, B.pos[a], B.vel[v], B);
        PROF("velocity");
    }
    if (ev) HIP_TRY(hipEventRecord(ev[5], st));
    launch_curl(st, K, s->tune, B.pos[a], B.vel[v], B, dbg ? s->dbg_err : nullptr, stats ? s->partial : nullptr);
    PROF("curl");
    if (ev) HIP_TRY(hipEventRecord(ev[6], st));
    launch_vort_apply(st, K, s->geo, s->tune, B.pos[a], B.vel[v], B.vel[v ^ 1], B, true);
    PROF("vort_apply_hash");
    if (ev) HIP_TRY(hipEventRecord(ev[7], st));
    HIP_TRY(hipGetLastError());
    s->lam_last = I > 0 ? (a ^ 1) : -1;
    s->pc = a;
    s->vc = v ^ 1;
    s->ic ^= 1;
    s->stats_valid = stats;
    s->diag_valid = true;
    s->hash_ready = true;
    s->steps_done++;
    return PBF_OK;
}

// Steps per captured multi-step graph: after an even number of steps every double-buffer
// parity is back where it started, so one such graph serves any number of launches.  Steps
// inside a graph run back to back; between graph launches the device idles for a few µs.
constexpr int kGraphSteps = 8;

// The executable graph of k steps (1 or kGraphSteps) from the current parity, captured (not
// launched) on first use.  A captured step bakes in which buffer of each pair it reads; vc is
// unchanged by a step, so (pc, ic), the kind of step and the number of steps name the variant.
int graph_for(pbf_solver* s, int k, hipGraphExec_t* out) {
    // the first step of a graph starts at the hash or, when the last step hashed for it, at the
    // scan: two variants
    const bool learned = s->learned_passes > 0 && s->mlp;
    const int key = ((learned ? s->learned_passes : 0) << 4) | (s->hash_ready ? 8 : 0) | (k > 1 ? 4 : 0) |
                    (s->pc * 2 + s->ic);
    auto found = s->graphs.find(key);
    if (found == s->graphs.end()) {
        int pc0 = s->pc, vc0 = s->vc, ic0 = s->ic, ll0 = s->lam_last, sd0 = s->steps_done;
        const bool st0 = s->stats_valid, dv0 = s->diag_valid, hr0 = s->hash_ready;
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = PBF_OK;
        for (int j = 0; j < k && rc == PBF_OK; ++j) rc = enqueue_step(s, nullptr);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        // capture only recorded the steps: restore the parity (also on the error paths below, so
        // a failed capture leaves the solver as it was)
        s->pc = pc0;
        s->vc = vc0;
        s->ic = ic0;
        s->lam_last = ll0;
        s->steps_done = sd0;
        s->stats_valid = st0;
        s->diag_valid = dv0;
        s->hash_ready = hr0;
        if (rc != PBF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(PBF_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_error(PBF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        found = s->graphs.emplace(key, ge).first;
    }
    *out = found->second;
    return PBF_OK;
}

// The parity and step bookkeeping of k launched steps, exactly as enqueue_step leaves it.
void advance_parity(pbf_solver* s, int k) {
    const bool learned = s->learned_passes > 0 && s->mlp;
    const int I = learned ? s->learned_passes : s->P.iterations;
    for (int j = 0; j < k; ++j) {
        int a = s->pc ^ 1;
        for (int it = 0; it < I; ++it) a ^= 1;
        s->lam_last = I > 0 ? (a ^ 1) : -1;
        s->pc = a;
        s->ic ^= 1;
        s->stats_valid = (s->P.flags & PBF_FLAG_DENSITY_STATS) != 0;
        s->diag_valid = true;
        s->hash_ready = true;
        s->steps_done++;
    }
}

bool uses_graphs(const pbf_solver* s) { return (s->P.flags & PBF_FLAG_GRAPH) && !(s->P.flags & PBF_FLAG_DEBUG); }

// k steps (1 or kGraphSteps): eagerly, or from the graph captured for (kind of step, k, parity).
// Parameters that change on every call (a caller adapting dt each step) would recapture a graph
// per step; those steps launch eagerly instead.
int run_steps(pbf_solver* s, int k) {
    HIP_TRY(hipSetDevice(s->device));
    if (!uses_graphs(s) || s->param_streak >= 2) {
        for (int j = 0; j < k; ++j)
            if (int rc = enqueue_step(s, nullptr)) return rc;
        return PBF_OK;
    }
    hipGraphExec_t ge = nullptr;
    if (int rc = graph_for(s, k, &ge)) return rc;
    HIP_TRY(hipGraphLaunch(ge, s->stream));
    advance_parity(s, k);
    return PBF_OK;
}

int run_step(pbf_solver* s) { return run_steps(s, 1); }

// nsteps steps: graphs of kGraphSteps steps while that many remain, then single steps.  The
// parameters stay as they are for all of them.
int run_many(pbf_solver* s, int nsteps) {
    s->param_streak = 0;
    int i = 0;
    for (; i + kGraphSteps <= nsteps; i += kGraphSteps)
        if (int rc = run_steps(s, kGraphSteps)) return rc;
    for (; i < nsteps; ++i)
        if (int rc = run_steps(s, 1)) return rc;
    return PBF_OK;
}

// Captures every graph run_many(nsteps) will launch, and launches none, so that a timed
// run_many does not include host capture and instantiation.
int prepare_graphs(pbf_solver* s, int nsteps) {
    if (!uses_graphs(s)) return PBF_OK;
    HIP_TRY(hipSetDevice(s->device));
    const int pc0 = s->pc, ic0 = s->ic, ll0 = s->lam_last, sd0 = s->steps_done;
    const bool st0 = s->stats_valid, dv0 = s->diag_valid, hr0 = s->hash_ready;
    int rc = PBF_OK;
    hipGraphExec_t ge = nullptr;
    int i = 0;
    for (; i + kGraphSteps <= nsteps && rc == PBF_OK; i += kGraphSteps) {
        rc = graph_for(s, kGraphSteps, &ge);
        advance_parity(s, kGraphSteps);
    }
    for (; i < nsteps && rc == PBF_OK; ++i) {
        rc = graph_for(s, 1, &ge);
        advance_parity(s, 1);
    }
    s->pc = pc0;
    s->ic = ic0;
    s->lam_last = ll0;
    s->steps_done = sd0;
    s->stats_valid = st0;
    s->diag_valid = dv0;
    s->hash_ready = hr0;
    return rc;
}

}  // namespace

extern "C" {

const char* pbf_last_error(void) { return g_last_error.c_str(); }

const char* pbf_version(void) { return "pbf_hip 0.1 gfx950"; }

}  // extern "C"

namespace {

// Shared by pbf_create and pbf_create_batch: allocates every buffer for n particles and
// nbuckets cell counters and uploads the state.  geoms / sid (host) describe a batch.
int create_solver(const pbf_params* para, const float* grid_space, const float* boundary, const float* pos_xyz,
                  const float* vel_xyz, int32_t n, int nbuckets, const std::vector<SceneGeom>* geoms,
                  const std::vector<uint8_t>* sid, pbf_solver** out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(PBF_ERR_NOGPU, "no HIP device");
    pbf_solver* s = new pbf_solver();
    s->P = *para;
    s->P.nParticles = n;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
            prop.multiProcessorCount > 0)
            s->cus = prop.multiProcessorCount;
    }
    std::memcpy(s->gs, grid_space, sizeof(s->gs));
    std::memcpy(s->bd, boundary, sizeof(s->bd));
    s->n = n;
    s->ncells = para->grid_dims[0] * para->grid_dims[1] * para->grid_dims[2];
    s->nbuckets = nbuckets;
    s->nstride = ((n + 63) / 64) * 64;
    if (s->nstride == 0) s->nstride = 64;
    s->cap = para->max_neighbors;
    if (geoms) {
        s->scenes = *geoms;
    } else {
        SceneGeom G{};
        for (int a = 0; a < 3; ++a) G.dims[a] = para->grid_dims[a];
        G.ncells = s->ncells;
        s->scenes.push_back(G);
    }
    (void)hipGetDevice(&s->device);
    auto fail = [&](int code) {
        free_all(s);
        delete s;
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(PBF_ERR_HIP, "hipStreamCreate failed"));
    StepBuffers& B = s->B;
    size_t np = (size_t)std::max(n, 1);
    int m = nbuckets;
    size_t mpad = (size_t)scan_tiles(m) * 1024;
    int rc;
    // pos / vel hold the padding slot n after the particles (see StepBuffers)
    const float4 far_pos = make_float4(kFarPos, kFarPos, kFarPos, 0.0f), still = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const std::vector<float4> far_tail((size_t)kSearchTail, far_pos);
    for (int i = 0; i < 2; ++i) {
        if ((rc = dalloc(s, &B.pos[i], np + kSearchTail))) return fail(rc);
        if ((rc = dalloc(s, &B.vel[i], np + 1))) return fail(rc);
        if ((rc = dalloc(s, &B.ids[i], np))) return fail(rc);
        if (hipMemcpy(B.pos[i] + n, far_tail.data(), sizeof(float4) * kSearchTail, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(B.vel[i] + n, &still, sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(PBF_ERR_HIP, "padding slot upload failed"));
    }
    if ((rc = dalloc(s, &B.xold, np)) || (rc = dalloc(s, &B.xp, np)) || (rc = dalloc(s, &B.vp, np)) ||
        (rc = dalloc(s, &B.omega, np)) || (rc = dalloc(s, &B.cell, np)) || (rc = dalloc(s, &B.local, np)) ||
        (rc = dalloc(s, &B.count, mpad)) || (rc = dalloc(s, &B.cstart, mpad + 4)) ||
        (rc = dalloc(s, &B.tile_sum, (size_t)scan_tiles(m) + 4)) || (rc = dalloc(s, &B.bucket, np)) || (rc = dalloc(s, &B.bucket_id, np)) ||
        (rc = dalloc(s, &B.nbr, (size_t)((s->cap + 7) / 8) * 8 * s->nstride)) || (rc = dalloc(s, &B.ncnt, np + 16)) ||
        (rc = dalloc(s, &B.lam1, np)))
        return fail(rc);
    s->npartial = (int)((np + 63) / 64) + 1;  // enough for the smallest workgroup size
    if ((rc = dalloc(s, &s->feat, np * kNumFeatures)) || (rc = dalloc(s, &s->lam_sum, np)) || (rc = dalloc(s, &s->lam_keep, np + 1)) ||
        (rc = dalloc(s, &s->resid, (size_t)s->npartial * 2)) ||
        (rc = dalloc(s, &s->resid0, (size_t)s->npartial * 2)) || (rc = dalloc(s, &s->gate, 4)) ||
        (rc = dalloc(s, &s->gate_stats, 4)))
        return fail(rc);
    s->resid_host = (float*)std::malloc(sizeof(float) * s->npartial * 2);
    if (!s->resid_host) return fail(set_error(PBF_ERR_ARG, "host allocation failed"));
    if ((rc = dalloc(s, &s->dbg_gid, np)) || (rc = dalloc(s, &s->dbg_lam, np)) || (rc = dalloc(s, &s->dbg_err, np)) ||
        (rc = dalloc(s, &s->partial, (size_t)s->npartial * 3)))
        return fail(rc);
    if (geoms) {
        SceneGeom* dg = nullptr;
        uint8_t* ds = nullptr;
        if ((rc = dalloc(s, &dg, geoms->size())) || (rc = dalloc(s, &ds, np))) return fail(rc);
        if (hipMemcpy(dg, geoms->data(), sizeof(SceneGeom) * geoms->size(), hipMemcpyHostToDevice) != hipSuccess ||
            (n > 0 && hipMemcpy(ds, sid->data(), (size_t)n, hipMemcpyHostToDevice) != hipSuccess))
            return fail(set_error(PBF_ERR_HIP, "scene table upload failed"));
        s->geo.geoms = dg;
        s->geo.sid = ds;
        s->nscenes = (int)geoms->size();
    }
    launch_fill_u32(s->stream, B.count, mpad, 0u);
    launch_fill_u32(s->stream, reinterpret_cast<uint32_t*>(s->gate), 4, 0u);
    launch_fill_u32(s->stream, reinterpret_cast<uint32_t*>(s->lam_keep), np + 1, 0u);  // entry n: the padding slot
    launch_fill_u32(s->stream, reinterpret_cast<uint32_t*>(s->gate_stats), 4, 0u);
    if (hipGetLastError() != hipSuccess) return fail(set_error(PBF_ERR_HIP, "fill launch failed"));
    *out = s;
    rc = pbf_set_state(s, pos_xyz, vel_xyz);
    if (rc) {
        *out = nullptr;
        return fail(rc);
    }
    return PBF_OK;
}

}  // namespace

extern "C" {

int pbf_create(const pbf_params* para, const float grid_space[6], const float boundary[6], const float* pos_xyz,
               const float* vel_xyz, int32_t n, pbf_solver** out) {
    if (!out) return set_error(PBF_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(para);
    if (rc) return rc;
    if (n < 0) return set_error(PBF_ERR_ARG, "n < 0");
    if (n >= kMaxSlots) return set_error(PBF_ERR_ARG, "n must be below 2^28 - 32");
    if (n > 0 && !pos_xyz) return set_error(PBF_ERR_ARG, "pos_xyz is NULL");
    if (!grid_space || !boundary) return set_error(PBF_ERR_ARG, "grid_space / boundary is NULL");
    int ncells = para->grid_dims[0] * para->grid_dims[1] * para->grid_dims[2];
    return create_solver(para, grid_space, boundary, pos_xyz, vel_xyz, n, ncells + 1, nullptr, nullptr, out);
}

int pbf_create_batch(const pbf_params* para, int32_t nscenes, const int32_t* counts, const float* grid_space,
                     const float* boundary, const int32_t* grid_dims, const float* pos_xyz, const float* vel_xyz,
                     pbf_solver** out) {
    if (!out) return set_error(PBF_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_params(para);
    if (rc) return rc;
    if (nscenes < 1 || nscenes > 255 || !counts || !grid_space || !boundary || !grid_dims)
        return set_error(PBF_ERR_ARG, "bad scene description (1..255 scenes)");
    std::vector<SceneGeom> geoms(nscenes);
    int64_t n = 0, nb = 0, rows = 0;
    for (int k = 0; k < nscenes; ++k) {
        if (counts[k] < 0) return set_error(PBF_ERR_ARG, "negative scene count");
        SceneGeom& G = geoms[k];
        int64_t cells = 1;
        for (int a = 0; a < 3; ++a) {
            int d = grid_dims[3 * k + a];
            if (d < 1 || d > 4096) return set_error(PBF_ERR_ARG, "scene grid dims out of range");
            G.dims[a] = d;
            G.gmin[a] = grid_space[6 * k + a];
            G.bmin[a] = boundary[6 * k + a];
            G.bmax[a] = boundary[6 * k + 3 + a];
            cells *= d;
        }
        G.coff = (int)nb;
        G.ncells = (int)cells;
        G.rowoff = (int)rows;
        rows += (int64_t)G.dims[1] * G.dims[2];
        nb += cells + 1;
        n += counts[k];
        if (nb > ((int64_t)1 << 28) || n >= kMaxSlots) return set_error(PBF_ERR_ARG, "batch too large");
    }
    if (n > 0 && !pos_xyz) return set_error(PBF_ERR_ARG, "pos_xyz is NULL");
    std::vector<uint8_t> sid((size_t)n);
    for (int64_t k = 0, i = 0; k < nscenes; ++k)
        for (int j = 0; j < counts[k]; ++j) sid[i++] = (uint8_t)k;
    pbf_params p = *para;
    for (int a = 0; a < 3; ++a) p.grid_dims[a] = geoms[0].dims[a];
    return create_solver(&p, grid_space, boundary, pos_xyz, vel_xyz, (int32_t)n, (int)nb, &geoms, &sid, out);
}

int pbf_set_tuning(pbf_solver* s, int32_t block, int32_t grid, int32_t xcd_map) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (block != 64 && block != 128 && block != 256 && block != 512) return set_error(PBF_ERR_ARG, "block must be 64/128/256/512");
    if (grid < 0) return set_error(PBF_ERR_ARG, "grid < 0");
    s->tune.block = block;
    s->tune.grid = grid;
    s->tune.xcd_map = xcd_map ? 1 : 0;
    drop_graphs(s);
    return PBF_OK;
}

namespace {
// Morton key of row (y, z): the bits of y and z interleaved (y in the even bits).
uint32_t morton2(uint32_t y, uint32_t z) {
    uint32_t k = 0;
    for (int b = 0; b < 16; ++b) k |= ((y >> b) & 1u) << (2 * b) | ((z >> b) & 1u) << (2 * b + 1);
    return k;
}
}  // namespace

int pbf_set_ordering(pbf_solver* s, int32_t order) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (order != 0 && order != 1) return set_error(PBF_ERR_ARG, "order must be 0 (z, y rows) or 1 (Morton rows)");
    if (order == 1 && !s->rowmap) {
        size_t total = 0;
        for (const SceneGeom& G : s->scenes) total += (size_t)G.dims[1] * G.dims[2];
        std::vector<int> map(total);
        for (const SceneGeom& G : s->scenes) {
            const int ny = G.dims[1], nz = G.dims[2];
            std::vector<int> r(ny * nz);
            for (int i = 0; i < ny * nz; ++i) r[i] = i;
            std::sort(r.begin(), r.end(), [&](int a, int b) {
                return morton2((uint32_t)(a % ny), (uint32_t)(a / ny)) < morton2((uint32_t)(b % ny), (uint32_t)(b / ny));
            });
            for (int rank = 0; rank < ny * nz; ++rank) map[(size_t)G.rowoff + r[rank]] = rank;
        }
        HIP_TRY(hipSetDevice(s->device));
        int* d = nullptr;
        int rc = dalloc(s, &d, total);
        if (rc) return rc;
        HIP_TRY(hipMemcpy(d, map.data(), sizeof(int) * total, hipMemcpyHostToDevice));
        s->rowmap = d;
        s->rowmap_host = std::move(map);
    }
    if (order != s->order) {
        invalidate_hash(s);  // cell ids depend on the row order
        s->order = order;
        drop_graphs(s);
    }
    return PBF_OK;
}

int pbf_get_ordering(pbf_solver* s, int32_t* out) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    *out = s->order;
    return PBF_OK;
}

int pbf_get_tuning(pbf_solver* s, int32_t out[3]) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    out[0] = s->tune.block;
    out[1] = s->tune.grid;
    out[2] = s->tune.xcd_map;
    return PBF_OK;
}

int pbf_set_state(pbf_solver* s, const float* pos_xyz, const float* vel_xyz) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (s->n == 0) return PBF_OK;
    if (!pos_xyz) return set_error(PBF_ERR_ARG, "pos_xyz is NULL");
    HIP_TRY(hipSetDevice(s->device));
    size_t bytes = (size_t)s->n * 3 * sizeof(float);
    // stage through the (not yet used) xp / vp buffers
    float* dp = reinterpret_cast<float*>(s->B.xp);
    float* dv = reinterpret_cast<float*>(s->B.vp);
    HIP_TRY(hipMemcpyAsync(dp, pos_xyz, bytes, hipMemcpyHostToDevice, s->stream));
    if (vel_xyz) {
        HIP_TRY(hipMemcpyAsync(dv, vel_xyz, bytes, hipMemcpyHostToDevice, s->stream));
    } else {
        launch_fill_u32(s->stream, reinterpret_cast<uint32_t*>(dv), (size_t)s->n * 3, 0u);
    }
    invalidate_hash(s);
    s->pc = s->vc = s->ic = 0;
    launch_import(s->stream, s->n, dp, dv, s->B.pos[0], s->B.vel[0], s->B.ids[0]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    // captured steps name buffers and parity only, and the parity is back at (0, 0): they stay
    s->lam_last = -1;
    s->stats_valid = false;
    s->diag_valid = false;
    return PBF_OK;
}

int pbf_destroy(pbf_solver* s) {
    if (!s) return PBF_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return PBF_OK;
}

int pbf_get_params(pbf_solver* s, pbf_params* out) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    *out = s->P;
    return PBF_OK;
}

int pbf_solve(pbf_solver* s, const pbf_params* para) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    int rc = check_params(para);
    if (rc) return rc;
    if (para->nParticles != s->n) return set_error(PBF_ERR_CAPACITY, "nParticles differs from the solver's");
    if (para->grid_dims[0] != s->P.grid_dims[0] || para->grid_dims[1] != s->P.grid_dims[1] ||
        para->grid_dims[2] != s->P.grid_dims[2])
        return set_error(PBF_ERR_CAPACITY, "grid_dims differ from the solver's");
    if (para->max_neighbors != s->cap) return set_error(PBF_ERR_CAPACITY, "max_neighbors differs from the solver's");
    if (std::memcmp(para, &s->P, sizeof(pbf_params)) != 0) {
        invalidate_hash(s);  // the last step predicted with the old parameters
        s->P = *para;
        drop_graphs(s);
        s->param_streak++;
    } else {
        s->param_streak = 0;
    }
    return run_step(s);
}

int pbf_step(pbf_solver* s, float dt) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (!(dt > 0.0f)) return set_error(PBF_ERR_ARG, "dt must be > 0");
    if (s->P.dt != dt) {
        invalidate_hash(s);  // the last step predicted with the old time step
        s->P.dt = dt;
        s->P.invDt = 1.0f / dt;
        drop_graphs(s);
        s->param_streak++;
    } else {
        s->param_streak = 0;
    }
    return run_step(s);
}

int pbf_advance(pbf_solver* s, int32_t nsteps) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    return run_many(s, nsteps);
}

int pbf_synchronize(pbf_solver* s) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

int pbf_export_device(pbf_solver* s, float* dev_pos_xyz, float* dev_vel_xyz) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    HIP_TRY(hipSetDevice(s->device));
    launch_export(s->stream, s->n, s->B.pos[s->pc], s->B.vel[s->vc], s->B.ids[s->ic], dev_pos_xyz, dev_vel_xyz);
    HIP_TRY(hipGetLastError());
    return PBF_OK;
}

int pbf_get_state(pbf_solver* s, float* pos_xyz, float* vel_xyz) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (s->n == 0) return PBF_OK;
    HIP_TRY(hipSetDevice(s->device));
    // xp / vp are free between steps: use them as staging
    float* dp = reinterpret_cast<float*>(s->B.xp);
    float* dv = reinterpret_cast<float*>(s->B.vp);
    launch_export(s->stream, s->n, s->B.pos[s->pc], s->B.vel[s->vc], s->B.ids[s->ic], dp, dv);
    HIP_TRY(hipGetLastError());
    size_t bytes = (size_t)s->n * 3 * sizeof(float);
    if (pos_xyz) HIP_TRY(hipMemcpyAsync(pos_xyz, dp, bytes, hipMemcpyDeviceToHost, s->stream));
    if (vel_xyz) HIP_TRY(hipMemcpyAsync(vel_xyz, dv, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return PBF_OK;
}

int pbf_device_state(pbf_solver* s, pbf_device_view* out) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    out->pos4 = s->B.pos[s->pc];
    out->vel4 = s->B.vel[s->vc];
    out->ids = s->B.ids[s->ic];
    out->n = s->n;
    out->ncells = s->ncells;
    return PBF_OK;
}

int pbf_get_grid(pbf_solver* s, int32_t* gid, int32_t* count, int32_t* cell_start, int32_t* sorted_ids) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (!(s->P.flags & PBF_FLAG_DEBUG) || !s->diag_valid)
        return set_error(PBF_ERR_STATE, "needs a step taken with PBF_FLAG_DEBUG on the current state");
    if (s->nscenes > 1) return set_error(PBF_ERR_STATE, "pbf_get_grid describes single-scene solvers");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> cs(s->ncells + 1);
    HIP_TRY(hipMemcpy(cs.data(), s->B.cstart, sizeof(uint32_t) * cs.size(), hipMemcpyDeviceToHost));
    if (gid && s->n) HIP_TRY(hipMemcpy(gid, s->dbg_gid, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost));
    if (!s->order) {
        if (cell_start)
            for (int c = 0; c <= s->ncells; ++c) cell_start[c] = (int32_t)cs[c];
        if (count)
            for (int c = 0; c < s->ncells; ++c) count[c] = (int32_t)(cs[c + 1] - cs[c]);
        if (sorted_ids && s->n)
            HIP_TRY(hipMemcpy(sorted_ids, s->B.ids[s->ic], sizeof(int32_t) * s->n, hipMemcpyDeviceToHost));
        return (int)cs[s->ncells];
    }
    // Morton rows: the table in the reference's cell order (x-fastest, then y, then z), the
    // members of each cell as they lie in its bucket, then the out-of-grid particles
    std::vector<uint32_t> ids(s->n);
    if (s->n) HIP_TRY(hipMemcpy(ids.data(), s->B.ids[s->ic], sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost));
    const int dx = s->P.grid_dims[0];
    uint32_t run = 0;
    size_t k = 0;
    for (int c = 0; c < s->ncells; ++c) {
        const int b = s->rowmap_host[c / dx] * dx + c % dx;
        const uint32_t len = cs[b + 1] - cs[b];
        if (cell_start) cell_start[c] = (int32_t)run;
        if (count) count[c] = (int32_t)len;
        if (sorted_ids)
            for (uint32_t j = cs[b]; j < cs[b + 1]; ++j) sorted_ids[k++] = (int32_t)ids[j];
        run += len;
    }
    if (cell_start) cell_start[s->ncells] = (int32_t)run;
    if (sorted_ids)
        for (uint32_t j = cs[s->ncells]; j < (uint32_t)s->n; ++j) sorted_ids[k++] = (int32_t)ids[j];
    return (int)cs[s->ncells];
}

int pbf_get_neighbors(pbf_solver* s, int32_t* nbr, int32_t* cnt) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (!(s->P.flags & PBF_FLAG_DEBUG) || !s->diag_valid)
        return set_error(PBF_ERR_STATE, "needs a step taken with PBF_FLAG_DEBUG on the current state");
    if (s->n == 0) return PBF_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> tab((size_t)((s->cap + 7) / 8) * 8 * s->nstride), ids(s->n);
    std::vector<uint8_t> c(s->n);
    HIP_TRY(hipMemcpy(tab.data(), s->B.nbr, sizeof(uint32_t) * tab.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ids.data(), s->B.ids[s->ic], sizeof(uint32_t) * s->n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(c.data(), s->B.ncnt, s->n, hipMemcpyDeviceToHost));
    for (int sl = 0; sl < s->n; ++sl) {
        uint32_t id = ids[sl];
        if (cnt) cnt[id] = c[sl];
        if (nbr)
            for (int j = 0; j < s->cap; ++j)
                nbr[(size_t)id * s->cap + j] =
                    j < c[sl] ? (int32_t)ids[tab[((size_t)(j / 8) * s->nstride + sl) * 8 + (j % 8)] >> 4] : -1;
    }
    return PBF_OK;
}

int pbf_get_fields(pbf_solver* s, float* lam_first, float* lam_last, float* err, float* curl_xyz) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (!(s->P.flags & PBF_FLAG_DEBUG) || !s->diag_valid)
        return set_error(PBF_ERR_STATE, "needs a step taken with PBF_FLAG_DEBUG on the current state");
    if (s->n == 0) return PBF_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int n = s->n;
    std::vector<uint32_t> ids(n);
    std::vector<float> lf(n), e(n);
    std::vector<float4> pl(n), om(n);
    HIP_TRY(hipMemcpy(ids.data(), s->B.ids[s->ic], sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lf.data(), s->dbg_lam, sizeof(float) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(e.data(), s->dbg_err, sizeof(float) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(om.data(), s->B.omega, sizeof(float4) * n, hipMemcpyDeviceToHost));
    if (s->lam_last >= 0)
        HIP_TRY(hipMemcpy(pl.data(), s->B.pos[s->lam_last], sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int sl = 0; sl < n; ++sl) {
        uint32_t id = ids[sl];
        if (lam_first) lam_first[id] = lf[sl];
        if (lam_last) lam_last[id] = s->lam_last >= 0 ? pl[sl].w : lf[sl];
        if (err) err[id] = e[sl];
        if (curl_xyz) {
            curl_xyz[3 * (size_t)id] = om[sl].x;
            curl_xyz[3 * (size_t)id + 1] = om[sl].y;
            curl_xyz[3 * (size_t)id + 2] = om[sl].z;
        }
    }
    return PBF_OK;
}

int pbf_density_stats(pbf_solver* s, float out[3]) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    if (!s->stats_valid) return set_error(PBF_ERR_STATE, "no step with density statistics taken");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int nb = reduce_blocks(s->tune, s->n);
    std::vector<float> p((size_t)nb * 3);
    HIP_TRY(hipMemcpy(p.data(), s->partial, sizeof(float) * p.size(), hipMemcpyDeviceToHost));
    double sa = 0.0, sp = 0.0;
    float mx = -INFINITY;
    for (int b = 0; b < nb; ++b) {
        sa += p[3 * b];
        sp += p[3 * b + 1];
        mx = std::max(mx, p[3 * b + 2]);
    }
    double n = std::max(s->n, 1);
    out[0] = (float)(sa / n);
    out[1] = (float)(sp / n);
    out[2] = mx;
    return PBF_OK;
}

int pbf_time_steps(pbf_solver* s, int32_t nsteps, float* ms_total) {
    if (!s || !ms_total) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(s->device));
    // every graph the run launches is captured before the first event: the window holds device
    // work and launches only
    if (int rc = prepare_graphs(s, nsteps)) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    int rc = PBF_OK;
    rc = run_many(s, nsteps);
    HIP_TRY(hipEventRecord(e1, s->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_total = ms;
    return rc;
}

int pbf_stage_times(pbf_solver* s, float out[7]) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(s->device));
    hipEvent_t ev[8];
    for (int i = 0; i < 8; ++i) HIP_TRY(hipEventCreate(&ev[i]));
    int rc = enqueue_step(s, ev);
    for (int i = 0; i < 7; ++i) out[i] = 0.0f;
    if (rc == PBF_OK) {
        HIP_TRY(hipEventSynchronize(ev[7]));
        for (int i = 0; i < 7; ++i) HIP_TRY(hipEventElapsedTime(&out[i], ev[i], ev[i + 1]));
    }
    for (int i = 0; i < 8; ++i) (void)hipEventDestroy(ev[i]);
    return rc;
}

// One step, eager, with an event after every stage launch group; writes up to max entries of
// (ms) into ms[] and the stage names (static strings) into names[]; returns the count.
int pbf_profile_step(pbf_solver* s, float* ms, const char** names, int32_t max) {
    if (!s || !ms || !names || max < 1) return set_error(PBF_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(s->device));
    std::vector<hipEvent_t> evs;
    std::vector<const char*> nm;
    s->prof_events = &evs;
    s->prof_names = &nm;
    int rc = enqueue_step(s, nullptr);
    s->prof_events = nullptr;
    s->prof_names = nullptr;
    int cnt = 0;
    if (rc == PBF_OK && !evs.empty()) {
        HIP_TRY(hipEventSynchronize(evs.back()));
        for (size_t i = 1; i < evs.size() && cnt < max; ++i, ++cnt) {
            HIP_TRY(hipEventElapsedTime(&ms[cnt], evs[i - 1], evs[i]));
            names[cnt] = nm[i];
        }
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    return rc == PBF_OK ? cnt : rc;
}


// Average duration of one launch of a step stage (0: neighbour build + lambda 1, 1: lambda,
// 2: delta-p + apply, 3: predict + hash, scan, scatter + rank + reorder, 4: an empty kernel of
// the particle grid's shape, the floor of one dependent launch, 5: the VALU issue-rate probe,
// 6: the vorticity pass with the next step's hash, then the sort stage, 7: the vorticity pass,
// 8: the curl pass with the XSPH sum), replayed reps times
// back to back on the state of the last step between two HIP events on the solver's stream.
// The replays write only buffers that the next step rebuilds (cell tables, lists, lambda in
// pos.w, the free position / velocity / id buffers, x_old; the scan clears the cell counters
// again; the curl pass rewrites omega, |omega| in pos.w and the XSPH sums in x_old with the
// values the step left there), so the state is unchanged.
int pbf_time_stage(pbf_solver* s, int32_t stage, int32_t reps, float* ms) {
    if (!s || !ms || reps < 1 || stage < 0 || stage > 8) return set_error(PBF_ERR_ARG, "bad argument");
    if (s->steps_done < 1 || s->n == 0) return set_error(PBF_ERR_STATE, "run a step first");
    HIP_TRY(hipSetDevice(s->device));
    const StepConsts K = make_consts(s);
    StepBuffers& B = s->B;
    const int a = s->pc;
    const LearnArgs L0{};
    // the sort-stage replays hash from zeroed counters (and leave them zeroed): the next step
    // hashes for itself
    if (stage == 3 || stage == 6) invalidate_hash(s);
    // the reps replays are captured in one graph, as a step runs, so host launch overhead does
    // not enter short stages; one untimed launch of the graph warms up
    HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < reps; ++r) {
        if (stage == 0) launch_neighbors_lambda(s->stream, K, s->geo, s->tune, B.pos[a], B, nullptr, L0);
        if (stage == 1) launch_lambda(s->stream, K, s->geo, s->tune, B.pos[a], B, L0);
        if (stage == 2) launch_delta_apply(s->stream, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[s->vc], B, 0, 1, L0);
        if (stage == 3) {
            launch_predict_hash(s->stream, K, s->geo, B.pos[s->pc], B.vel[s->vc], B);
            launch_scan(s->stream, K, B);
            launch_sort_reorder(s->stream, K, s->geo, B, B.pos[s->pc], B.vel[s->vc], B.ids[s->ic], B.pos[s->pc ^ 1],
                                B.vel[s->vc ^ 1], B.ids[s->ic ^ 1], nullptr);
        }
        if (stage == 4) launch_noop(s->stream, s->n);
        if (stage == 5) launch_valu_probe(s->stream, reinterpret_cast<float*>(B.xp));
        // the vorticity pass writes into the spare staging buffer vp, never the state's velocity
        if (stage == 6) {
            launch_vort_apply(s->stream, K, s->geo, s->tune, B.pos[a], B.vel[s->vc ^ 1], B.vp, B, true);
            launch_scan(s->stream, K, B);
            launch_sort_reorder(s->stream, K, s->geo, B, B.pos[s->pc], B.vel[s->vc], B.ids[s->ic], B.pos[s->pc ^ 1],
                                B.vel[s->vc ^ 1], B.ids[s->ic ^ 1], nullptr);
        }
        if (stage == 7) launch_vort_apply(s->stream, K, s->geo, s->tune, B.pos[a], B.vel[s->vc ^ 1], B.vp, B, false);
        // the curl pass reads the velocity before the vorticity pass, as the step does
        if (stage == 8) launch_curl(s->stream, K, s->tune, B.pos[a], B.vel[s->vc ^ 1], B, nullptr, nullptr);
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(s->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float t = 0.0f;
    e = hipGraphLaunch(ge, s->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
    if (e == hipSuccess) e = hipGraphLaunch(ge, s->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(ge);
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "stage timing: %s", hipGetErrorString(e));
    HIP_TRY(hipGetLastError());
    // the replays rewrote the cell tables, lists and lambda: the step diagnostics are stale
    s->lam_last = -1;
    s->stats_valid = false;
    s->diag_valid = false;
    *ms = t / (float)reps;
    return PBF_OK;
}

// Times every launch shape (workgroup size x persistent grid x XCD mapping) over `steps`
// graph-replayed steps each and keeps the fastest.  Results do not depend on the launch shape,
// and the state is restored afterwards, so tuning does not advance the simulation.  best[3]
// (may be NULL) receives block, grid, xcd_map; best_ms the fastest time per step.
int pbf_autotune(pbf_solver* s, int32_t steps, int32_t* best, float* best_ms) {
    if (!s || steps < 1) return set_error(PBF_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int n = s->n;
    size_t np = (size_t)std::max(n, 1);
    float4 *sp = nullptr, *sv = nullptr;
    uint32_t* si = nullptr;
    HIP_TRY(hipMalloc(&sp, sizeof(float4) * np));
    HIP_TRY(hipMalloc(&sv, sizeof(float4) * np));
    HIP_TRY(hipMalloc(&si, sizeof(uint32_t) * np));
    HIP_TRY(hipMemcpyAsync(sp, s->B.pos[s->pc], sizeof(float4) * np, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(sv, s->B.vel[s->vc], sizeof(float4) * np, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(si, s->B.ids[s->ic], sizeof(uint32_t) * np, hipMemcpyDeviceToDevice, s->stream));
    const int pc0 = s->pc, vc0 = s->vc, ic0 = s->ic, sd0 = s->steps_done;
    const Tuning keep = s->tune;
    const uint32_t flags0 = s->P.flags;
    s->P.flags = (flags0 | PBF_FLAG_GRAPH) & ~(PBF_FLAG_DEBUG | PBF_FLAG_DENSITY_STATS);
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->device) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    const int blocks[4] = {64, 128, 256, 512};
    const int grids[4] = {0, 2 * cus, 4 * cus, 8 * cus};
    Tuning bestT = keep;
    float bestMs = INFINITY;
    int rc = PBF_OK;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // puts the solver back on the state the tuning started from (every candidate is timed from it:
    // the work per step depends on the state, and a candidate timed on later steps would be
    // charged for them)
    auto restore = [&]() -> int {
        s->pc = pc0;
        s->vc = vc0;
        s->ic = ic0;
        s->steps_done = sd0;
        drop_graphs(s);
        invalidate_hash(s);  // the tuning steps hashed their own successors
        hipError_t e = hipMemcpyAsync(s->B.pos[s->pc], sp, sizeof(float4) * np, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(s->B.vel[s->vc], sv, sizeof(float4) * np, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(s->B.ids[s->ic], si, sizeof(uint32_t) * np, hipMemcpyDeviceToDevice, s->stream);
        if (e != hipSuccess) return set_error(PBF_ERR_HIP, "state restore failed: %s", hipGetErrorString(e));
        return PBF_OK;
    };
    // ms per step of the current shape and order from the starting state, stepped as pbf_advance /
    // pbf_time_steps step: `steps` rounded up to whole multi-step graphs, captured first, after one
    // warm-up graph
    const int rounds = ((steps + kGraphSteps - 1) / kGraphSteps) * kGraphSteps;
    auto time_config = [&](float& ms) -> int {
        if (int r = restore()) return r;
        if (int r = prepare_graphs(s, rounds)) return r;
        if (int r = run_many(s, kGraphSteps)) return r;
        if (hipEventRecord(e0, s->stream) != hipSuccess) return set_error(PBF_ERR_HIP, "event");
        if (int r = run_many(s, rounds)) return r;
        if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return set_error(PBF_ERR_HIP, "event");
        ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)rounds;
        return PBF_OK;
    };
    for (int bi = 0; bi < 4 && rc == PBF_OK; ++bi)
        for (int gi = 0; gi < 4 && rc == PBF_OK; ++gi)
            for (int xm = 0; xm < 2 && rc == PBF_OK; ++xm) {
                if (grids[gi] > 0 && (int)((n + blocks[bi] - 1) / blocks[bi]) <= grids[gi]) continue;
                s->tune = Tuning{blocks[bi], grids[gi], xm};
                float ms = 0.0f;
                rc = time_config(ms);
                if (rc == PBF_OK && ms < bestMs) {
                    bestMs = ms;
                    bestT = s->tune;
                }
            }
    // the best shape with the other row order
    const int order0 = s->order;
    int bestOrder = order0;
    if (rc == PBF_OK) {
        s->tune = bestT;
        rc = pbf_set_ordering(s, 1 - order0);
        float ms = 0.0f;
        if (rc == PBF_OK) rc = time_config(ms);
        if (rc == PBF_OK && ms < bestMs) {
            bestMs = ms;
            bestOrder = 1 - order0;
        }
    }
    s->order = rc == PBF_OK ? bestOrder : order0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // restore the state the tuning started from
    s->P.flags = flags0;
    // the tuning steps overwrote the spare buffers, lambda and the cell tables: the last step's
    // diagnostics (lambda, density statistics, grid, lists) are gone until the next step
    s->lam_last = -1;
    s->stats_valid = false;
    s->diag_valid = false;
    const int rr = restore();
    hipError_t e = hipStreamSynchronize(s->stream);
    (void)hipFree(sp);
    (void)hipFree(sv);
    (void)hipFree(si);
    if (rr) return rr;
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "state restore failed: %s", hipGetErrorString(e));
    s->tune = rc == PBF_OK ? bestT : keep;
    if (rc) return rc;
    if (best) {
        best[0] = bestT.block;
        best[1] = bestT.grid;
        best[2] = bestT.xcd_map;
    }
    if (best_ms) *best_ms = bestMs;
    return PBF_OK;
}

}  // extern "C"

// ============================================================================================
// Learned warm start and adaptive stepping (no counterpart in the reference).

struct pbf_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* p = nullptr;        // fp32 master parameters
    float* m = nullptr;        // Adam first moment
    float* v = nullptr;        // Adam second moment
    float* g = nullptr;        // gradient (+ sum of squared errors at [kMlpParams])
    pbf::MlpModel* M = nullptr;
    float* partial = nullptr;  // per-workgroup gradient sums
    uint32_t* idx = nullptr;
    int idx_cap = 0;
    int adam_t = 0;
    std::vector<void*> allocs;
};

namespace {

template <class T>
int malloc_mlp(pbf_mlp* m, T** p, size_t count) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256)));
    m->allocs.push_back(q);
    *p = (T*)q;
    return PBF_OK;
}

void free_mlp(pbf_mlp* m) {
    for (void* q : m->allocs) (void)hipFree(q);
    m->allocs.clear();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    m->stream = nullptr;
}

int mlp_ensure_idx(pbf_mlp* m, int n) {
    if (n <= m->idx_cap) return PBF_OK;
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, sizeof(uint32_t) * (size_t)n));
    if (m->idx) {
        for (auto& a : m->allocs)
            if (a == m->idx) a = q;
        (void)hipFree(m->idx);
    } else {
        m->allocs.push_back(q);
    }
    m->idx = (uint32_t*)q;
    m->idx_cap = n;
    return PBF_OK;
}

uint64_t sm64(uint64_t& st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Means over particles of max(C, 0) (compression residual) and |C| from the workgroup partials.
int read_residual(pbf_solver* s, const float* partials, float* comp, float* dev = nullptr) {
    int nb = reduce_blocks(s->tune, s->n);
    HIP_TRY(hipMemcpyAsync(s->resid_host, partials, sizeof(float) * 2 * nb, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nb; ++k) {
        a += s->resid_host[2 * k];
        b += s->resid_host[2 * k + 1];
    }
    const double n = std::max(s->n, 1);
    *comp = (float)(a / n);
    if (dev) *dev = (float)(b / n);
    return PBF_OK;
}

}  // namespace

extern "C" {

int pbf_mlp_param_count(void) { return pbf::kMlpParams; }

int pbf_mlp_create(uint64_t seed, pbf_mlp** out) {
    if (!out) return set_error(PBF_ERR_ARG, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(PBF_ERR_NOGPU, "no HIP device");
    pbf_mlp* m = new pbf_mlp();
    (void)hipGetDevice(&m->device);
    auto fail = [&](int rc) {
        free_mlp(m);
        delete m;
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(PBF_ERR_HIP, "hipStreamCreate failed"));
    int rc;
    if ((rc = malloc_mlp(m, &m->p, kMlpParams)) || (rc = malloc_mlp(m, &m->m, kMlpParams)) ||
        (rc = malloc_mlp(m, &m->v, kMlpParams)) || (rc = malloc_mlp(m, &m->g, kMlpParams + 1)) ||
        (rc = malloc_mlp(m, &m->M, 1)) || (rc = malloc_mlp(m, &m->partial, (size_t)kGradMaxBlocks * (kMlpParams + 1))))
        return fail(rc);
    // He-uniform weights (fan-in), zero biases, small output layer.
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t st = seed;
    auto uni = [&](float a) { return (float)(((sm64(st) >> 40) * (1.0 / 16777216.0)) * 2.0 - 1.0) * a; };
    for (int i = 0; i < kHid * kIn; ++i) p[kOffW1 + i] = uni(std::sqrt(6.0f / kIn));
    for (int i = 0; i < kHid * kHid; ++i) p[kOffW2 + i] = uni(std::sqrt(6.0f / kHid));
    for (int i = 0; i < kHid * kHid; ++i) p[kOffW3 + i] = uni(std::sqrt(6.0f / kHid));
    for (int i = 0; i < kHid; ++i) p[kOffW4 + i] = uni(0.1f * std::sqrt(6.0f / kHid));
    std::vector<float> zeros(kMlpParams, 0.0f);
    if (hipMemcpy(m->p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->v, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(PBF_ERR_HIP, "parameter upload failed"));
    // identity normalisation until set
    float fm[kIn], fi[kIn];
    for (int i = 0; i < kIn; ++i) {
        fm[i] = 0.0f;
        fi[i] = 1.0f;
    }
    float ys[2] = {1.0f, 0.0f};
    if (hipMemcpy(m->M->fmean, fm, sizeof(fm), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->M->finv, fi, sizeof(fi), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(&m->M->yscale, ys, sizeof(ys), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(PBF_ERR_HIP, "normalisation upload failed"));
    launch_mlp_pack(m->stream, m->p, m->M);
    if (hipStreamSynchronize(m->stream) != hipSuccess) return fail(set_error(PBF_ERR_HIP, "pack failed"));
    *out = m;
    return PBF_OK;
}

int pbf_mlp_destroy(pbf_mlp* m) {
    if (!m) return PBF_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_mlp(m);
    delete m;
    return PBF_OK;
}

int pbf_mlp_get_params(pbf_mlp* m, float* params) {
    if (!m || !params) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(params, m->p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return PBF_OK;
}

int pbf_mlp_set_params(pbf_mlp* m, const float* params) {
    if (!m || !params) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpyAsync(m->p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    launch_mlp_pack(m->stream, m->p, m->M);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    return PBF_OK;
}

int pbf_mlp_set_normalization(pbf_mlp* m, const float* feat_mean, const float* feat_inv_std, float yscale,
                              float yshift) {
    if (!m || !feat_mean || !feat_inv_std) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(m->device));
    float ys[2] = {yscale, yshift};
    HIP_TRY(hipMemcpy(m->M->fmean, feat_mean, sizeof(float) * kIn, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->M->finv, feat_inv_std, sizeof(float) * kIn, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(&m->M->yscale, ys, sizeof(ys), hipMemcpyHostToDevice));
    return PBF_OK;
}

int pbf_mlp_forward(pbf_mlp* m, const float* x, int32_t n, float* y) {
    if (!m || (n > 0 && (!x || !y)) || n < 0) return set_error(PBF_ERR_ARG, "bad argument");
    if (n == 0) return PBF_OK;
    HIP_TRY(hipSetDevice(m->device));
    float *dx = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc(&dx, sizeof(float) * (size_t)n * kIn));
    hipError_t e = hipMalloc(&dy, sizeof(float) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        launch_mlp_infer(m->stream, m->M, dx, n, nullptr, dy, 1.0f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "mlp forward: %s", hipGetErrorString(e));
    return PBF_OK;
}

// Gradient of the mean squared error over (x, t) with respect to the packed parameters.
int pbf_mlp_loss_grad(pbf_mlp* m, const float* x, const float* t, int32_t n, float* grad, float* loss) {
    if (!m || n <= 0 || !x || !t) return set_error(PBF_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(m->device));
    int rc = mlp_ensure_idx(m, n);
    if (rc) return rc;
    float *dx = nullptr, *dt = nullptr;
    HIP_TRY(hipMalloc(&dx, sizeof(float) * (size_t)n * kIn));
    hipError_t e = hipMalloc(&dt, sizeof(float) * (size_t)n);
    std::vector<uint32_t> id(n);
    for (int i = 0; i < n; ++i) id[i] = (uint32_t)i;
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * kIn, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(m->idx, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        launch_mlp_grad(m->stream, m->M, dx, dt, m->idx, n, m->partial, m->g);
        e = hipGetLastError();
    }
    std::vector<float> g(kMlpParams + 1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g.data(), m->g, sizeof(float) * (kMlpParams + 1), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "mlp grad: %s", hipGetErrorString(e));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * kMlpParams);
    if (loss) *loss = g[kMlpParams] / (float)n;
    return PBF_OK;
}

// pbf_mlp_loss_grad_noalloc's scratch: the per-workgroup partial gradient rows of one gradient
// launch (at most kGradMaxBlocks rows of kMlpParams + 1 floats, 9.25 MiB), one copy per device,
// held until the process ends.  Host code never uses this name as an address: it uses the
// per-device pointer hipGetSymbolAddress returns (noalloc_scratch).
static __device__ __align__(256) float scratch[kGradMaxBlocks * (kMlpParams + 1)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once per device.
int noalloc_scratch(int device, float** out) {
    if (device < 0 || device >= kMaxDevices) return set_error(PBF_ERR_ARG, "device index %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(PBF_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return PBF_OK;
}
}  // namespace

// pbf_mlp_loss_grad on device arrays, with no allocation: the partial rows go to the file-scope
// scratch (every row the reduction reads is fully written by its workgroup, so nothing needs
// clearing), samples are taken in order (no index array), and the result stays on the device.
int pbf_mlp_loss_grad_noalloc(pbf_mlp* m, const float* dev_x, const float* dev_t, int32_t n, float* dev_grad) {
    if (!m || !dev_x || !dev_t || !dev_grad) return set_error(PBF_ERR_ARG, "NULL argument");
    if (n < 1) return set_error(PBF_ERR_ARG, "n must be >= 1");
    const int blocks = mlp_grad_blocks(n);
    if (blocks > kGradMaxBlocks)
        return set_error(PBF_ERR_CAPACITY, "%d partial rows exceed the scratch's %d", blocks, kGradMaxBlocks);
    HIP_TRY(hipSetDevice(m->device));
    float* part = nullptr;
    if (int rc = noalloc_scratch(m->device, &part)) return rc;
    launch_mlp_grad(m->stream, m->M, dev_x, dev_t, nullptr, n, part, dev_grad);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    return PBF_OK;
}

// One Adam step with the gradient of the last pbf_mlp_loss_grad (grad == NULL) or the given one.
int pbf_mlp_adam_step(pbf_mlp* m, const float* grad, float lr, float beta1, float beta2, float eps) {
    if (!m) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(m->device));
    if (grad) HIP_TRY(hipMemcpyAsync(m->g, grad, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    m->adam_t++;
    float c1 = (float)(1.0 - std::pow((double)beta1, m->adam_t));
    float c2 = (float)(1.0 - std::pow((double)beta2, m->adam_t));
    launch_adam(m->stream, m->p, m->m, m->v, m->g, lr, beta1, beta2, eps, c1, c2);
    launch_mlp_pack(m->stream, m->p, m->M);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    return PBF_OK;
}

// Trains on a dataset already in device memory (X [nsamples][16] raw features, T [nsamples]
// normalised targets): `steps` Adam steps on batches drawn with SplitMix64(seed).  losses[i]
// (host, may be NULL) receives the batch loss of step i.  Learning rate decays by cosine to
// lr_final.
int pbf_mlp_train(pbf_mlp* m, const float* dev_x, const float* dev_t, int64_t nsamples, int32_t batch, int32_t steps,
                  float lr, float lr_final, uint64_t seed, float* losses) {
    if (!m || !dev_x || !dev_t || nsamples <= 0 || batch <= 0 || steps < 0)
        return set_error(PBF_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(m->device));
    int rc = mlp_ensure_idx(m, batch);
    if (rc) return rc;
    float* dloss = nullptr;
    HIP_TRY(hipMalloc(&dloss, sizeof(float) * (size_t)std::max(steps, 1)));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    hipError_t e = hipSuccess;
    for (int it = 0; it < steps && e == hipSuccess; ++it) {
        launch_batch_indices(m->stream, seed, (uint64_t)it, batch, (uint64_t)nsamples, m->idx);
        launch_mlp_grad(m->stream, m->M, dev_x, dev_t, m->idx, batch, m->partial, m->g);
        e = hipMemcpyAsync(dloss + it, m->g + kMlpParams, sizeof(float), hipMemcpyDeviceToDevice, m->stream);
        m->adam_t++;
        float c1 = (float)(1.0 - std::pow((double)b1, m->adam_t));
        float c2 = (float)(1.0 - std::pow((double)b2, m->adam_t));
        double frac = steps > 1 ? (double)it / (double)(steps - 1) : 1.0;
        float lr_t = (float)(lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(3.14159265358979 * frac)));
        launch_adam(m->stream, m->p, m->m, m->v, m->g, lr_t, b1, b2, eps, c1, c2);
        launch_mlp_pack(m->stream, m->p, m->M);
        if (e == hipSuccess) e = hipGetLastError();
    }
    if (e == hipSuccess && losses && steps > 0) {
        std::vector<float> l(steps);
        e = hipMemcpyAsync(l.data(), dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        for (int i = 0; i < steps; ++i) losses[i] = l[i] / (float)batch;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dloss);
    if (e != hipSuccess) return set_error(PBF_ERR_HIP, "mlp train: %s", hipGetErrorString(e));
    return PBF_OK;
}

int pbf_attach_mlp(pbf_solver* s, pbf_mlp* m) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    if (m && m->device != s->device) return set_error(PBF_ERR_ARG, "network and solver are on different devices");
    if (m) HIP_TRY(hipStreamSynchronize(m->stream));
    s->mlp = m ? m->M : nullptr;
    if (!m) s->learned_passes = 0;
    drop_graphs(s);  // captured learned steps name the previous network
    return PBF_OK;
}

// Device memory helpers for callers without an allocator of their own (datasets, exports).
int pbf_device_alloc(uint64_t bytes, void** out) {
    if (!out) return set_error(PBF_ERR_ARG, "out is NULL");
    HIP_TRY(hipMalloc(out, std::max<uint64_t>(bytes, 256)));
    return PBF_OK;
}

int pbf_copy_kernel(void* dst, const void* src, uint64_t bytes, float* ms) {
    if (!dst || !src || (bytes % 16) != 0) return set_error(PBF_ERR_ARG, "bad argument (bytes must be a multiple of 16)");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, nullptr));
    launch_copy_u128(nullptr, src, dst, (size_t)bytes);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return PBF_OK;
}

int pbf_device_free(void* p) {
    if (p) HIP_TRY(hipFree(p));
    return PBF_OK;
}

int pbf_memcpy_to_host(void* dst, const void* src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return PBF_OK;
}

int pbf_memcpy_to_device(void* dst, const void* src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return PBF_OK;
}

}  // extern "C"

namespace {

// One step whose constraint loop stops at residual <= tol (see pbf_step_adaptive).  record: the
// pass-2 lambda writes the warm-start features and lam_sum accumulates the multipliers of passes
// 2..I (the training target); the residual is not read.
int step_adaptive(pbf_solver* s, int32_t use_mlp, int32_t metric, float tol, int32_t min_iters, int32_t max_iters,
                  int32_t* iters_used, float* resid, bool record) {
    if (!s || max_iters < 1 || min_iters < 0 || metric < 0 || metric > 1) return set_error(PBF_ERR_ARG, "bad argument");
    if (use_mlp && !s->mlp) return set_error(PBF_ERR_STATE, "no network attached");
    HIP_TRY(hipSetDevice(s->device));
    const StepConsts K = make_consts(s);
    StepBuffers& B = s->B;
    hipStream_t st = s->stream;
    const bool dbg = (s->P.flags & PBF_FLAG_DEBUG) != 0;
    const bool stats = dbg || (s->P.flags & PBF_FLAG_DENSITY_STATS);
    const bool features = use_mlp || record;
    if (!s->hash_ready) launch_predict_hash(st, K, s->geo, B.pos[s->pc], B.vel[s->vc], B);
    launch_scan(st, K, B);
    launch_sort_reorder(st, K, s->geo, B, B.pos[s->pc], B.vel[s->vc], B.ids[s->ic], B.pos[s->pc ^ 1], B.vel[s->vc ^ 1],
                        B.ids[s->ic ^ 1], dbg ? s->dbg_gid : nullptr);
    int a = s->pc ^ 1;
    const int v = s->vc ^ 1;
    LearnArgs LF{};
    if (features) LF.feat = s->feat;
    launch_neighbors_lambda(st, K, s->geo, s->tune, B.pos[a], B, dbg ? s->dbg_lam : nullptr, LF);
    const LearnArgs L0{};
    LearnArgs LA{};
    if (record) LA.lam_sum = s->lam_sum;
    LearnArgs LR{};
    LR.resid = s->resid;
    const bool check = tol >= 0.0f && !record;  // read residuals back
    int it = 0;
    float r = -1.0f;
    bool done = false;
    while (!done) {
        // pass it + 1 on the lambda in pos[a].w
        launch_delta_apply(st, K, s->geo, s->tune, B.pos[a], B.pos[a ^ 1], B.vel[v], B, 0, it == 1,
                           it >= 1 ? LA : L0);
        a ^= 1;
        ++it;
        if (it >= max_iters) break;
        const bool feat_pass = features && it == 1;
        launch_lambda(st, K, s->geo, s->tune, B.pos[a], B, feat_pass ? feature_args(s, v, s->resid0) : LR);
        HIP_TRY(hipGetLastError());
        if (check && it >= min_iters) {
            float comp = 0.0f, dev = 0.0f;
            int rc = read_residual(s, feat_pass ? s->resid0 : s->resid, &comp, &dev);
            if (rc) return rc;
            r = metric ? dev : comp;
            if (r <= tol) break;
        }
        if (use_mlp && it == 1) {
            // the learned pass (pass 2)
            int rc = enqueue_guess_pass(s, K, a, v, max_iters > 2);
            if (rc) return rc;
            a ^= 1;
            ++it;
            if (it >= max_iters) break;
            if (check && it >= min_iters) {
                float comp = 0.0f, dev = 0.0f;
                rc = read_residual(s, s->resid, &comp, &dev);
                if (rc) return rc;
                r = metric ? dev : comp;
                if (r <= tol) break;
            }
        }
    }
    launch_velocity(st, K, B.pos[a], B.vel[v], B);
    launch_curl(st, K, s->tune, B.pos[a], B.vel[v], B, dbg ? s->dbg_err : nullptr, stats ? s->partial : nullptr);
    launch_vort_apply(st, K, s->geo, s->tune, B.pos[a], B.vel[v], B.vel[v ^ 1], B, true);
    HIP_TRY(hipGetLastError());
    s->lam_last = a ^ 1;
    s->pc = a;
    s->vc = v ^ 1;
    s->ic ^= 1;
    s->stats_valid = stats;
    s->diag_valid = true;
    s->hash_ready = true;
    s->steps_done++;
    if (iters_used) *iters_used = it;
    if (resid) *resid = r;
    return PBF_OK;
}

}  // namespace

extern "C" {

// One plain step with `iterations` Jacobi iterations that also records, for every particle
// with id % stride == phase, its raw features (at the predicted positions) and the multiplier
// the step accumulated over its iterations divided by h^2, at row base + id / stride of
// (dev_x [cap][16], dev_t [cap]).
int pbf_record_step(pbf_solver* s, int32_t iterations, float* dev_x, float* dev_t, int64_t base, int64_t cap,
                    int32_t stride, int32_t phase) {
    if (!s || !dev_x || !dev_t || iterations < 2 || stride < 1 || phase < 0 || phase >= stride || base < 0)
        return set_error(PBF_ERR_ARG, "bad argument");
    HIP_TRY(hipSetDevice(s->device));
    int32_t used = 0;
    int rc = step_adaptive(s, 0, 0, -1.0f, iterations, iterations, &used, nullptr, true);
    if (rc) return rc;
    const float inv_h2 = 1.0f / (s->P.h * s->P.h);
    launch_record(s->stream, s->n, s->B.ids[s->ic], s->feat, s->lam_sum, inv_h2, stride, phase, (size_t)base,
                  (size_t)cap, dev_x, dev_t);
    HIP_TRY(hipGetLastError());
    return PBF_OK;
}

// One step whose constraint loop stops as soon as the residual (metric 0: compression, the mean
// of max(C_i, 0); metric 1: mean |C_i|; over the frozen lists at the current positions) is
// <= tol after at least min_iters iterations, or after max_iters.  Each residual is read back to
// the host, so this is the measuring variant; fixed-pass stepping is pbf_set_learned + pbf_step.
//
// With use_mlp (a network attached) iteration 1 is the learned first pass: the network's
// multiplier replaces lambda, and the pass is kept only if mean |C| does not grow (else it reruns
// plain from the predicted state, exactly as without the network).  Plain iterations then continue
// until the tolerance holds: they are the fallback that judges the guess.  iters_used / resid (may
// be NULL) report the passes run (the learned pass counts as one) and the last residual measured.
int pbf_step_adaptive(pbf_solver* s, int32_t use_mlp, int32_t metric, float tol, int32_t min_iters, int32_t max_iters,
                      int32_t* iters_used, float* resid) {
    return step_adaptive(s, use_mlp, metric, tol, min_iters, max_iters, iters_used, resid, false);
}

// The last learned pass: out[0] / out[1] = compression after the plain pass 1 / after the
// network's pass, out[2] = 1 if that pass was kept, out[3] / out[4] = mean |C| before / after.
int pbf_last_guess_info(pbf_solver* s, float out[5]) {
    if (!s || !out) return set_error(PBF_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(s->device));
    float st4[4];
    int g = 0;
    HIP_TRY(hipMemcpyAsync(st4, s->gate_stats, sizeof(st4), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(&g, s->gate, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    out[0] = st4[0];
    out[1] = st4[1];
    out[2] = (float)g;
    out[3] = st4[2];
    out[4] = st4[3];
    return PBF_OK;
}

// passes >= 2: every following pbf_step / pbf_advance / pbf_time_steps runs a plain pass, the
// learned pass (network attached) and passes - 2 plain iterations; 0: back to P.iterations plain
// iterations.
int pbf_set_learned(pbf_solver* s, int32_t passes) {
    if (!s || passes < 0 || passes == 1 || passes > 1000)
        return set_error(PBF_ERR_ARG, "passes: 0 (plain) or 2..1000 (a plain pass, the learned pass, the rest plain)");
    if (passes > 0 && !s->mlp) return set_error(PBF_ERR_STATE, "no network attached");
    s->learned_passes = passes;
    return PBF_OK;
}

// save != 0: copy the state (positions, velocities with their .w, particle order) into the
// solver's snapshot buffers; save == 0: restore it.  Both are stream-ordered device copies.
int pbf_snapshot(pbf_solver* s, int32_t save) {
    if (!s) return set_error(PBF_ERR_ARG, "solver is NULL");
    HIP_TRY(hipSetDevice(s->device));
    const size_t np = (size_t)std::max(s->n, 1);
    if (!s->snap_pos) {
        int rc;
        if ((rc = dalloc(s, &s->snap_pos, np)) || (rc = dalloc(s, &s->snap_vel, np)) || (rc = dalloc(s, &s->snap_ids, np)))
            return rc;
    }
    if (!save && !s->snap_valid) return set_error(PBF_ERR_STATE, "no snapshot saved");
    float4* p = s->B.pos[s->pc];
    float4* v = s->B.vel[s->vc];
    uint32_t* id = s->B.ids[s->ic];
    if (save) {
        launch_copy_u128(s->stream, p, s->snap_pos, sizeof(float4) * np);
        launch_copy_u128(s->stream, v, s->snap_vel, sizeof(float4) * np);
        HIP_TRY(hipMemcpyAsync(s->snap_ids, id, sizeof(uint32_t) * np, hipMemcpyDeviceToDevice, s->stream));
        s->snap_valid = true;
    } else {
        invalidate_hash(s);
        launch_copy_u128(s->stream, s->snap_pos, p, sizeof(float4) * np);
        launch_copy_u128(s->stream, s->snap_vel, v, sizeof(float4) * np);
        HIP_TRY(hipMemcpyAsync(id, s->snap_ids, sizeof(uint32_t) * np, hipMemcpyDeviceToDevice, s->stream));
        s->lam_last = -1;
        s->stats_valid = false;
        s->diag_valid = false;
    }
    HIP_TRY(hipGetLastError());
    return PBF_OK;
}

}  // extern "C"
