/* pbf.h -- C ABI of the MI355X (gfx950) Position Based Fluids solver (libpbf_hip.so).
 *
 * Drop-in boundary for the reference's solver host API, kernels/cudaPBF.h:8-12
 * (initCUDASolverDataBuffer / cudaPBFSolve / freeCUDASolverDataBuffer, a C++ ABI taking
 * std::vector<glm::vec3>, SolverData& and SolverParameter&).  Here every entry point is
 * extern "C" with plain pointers and sizes.  A header-only C++ facade that re-creates the
 * reference's function names and its PBFSolver class on top of this ABI is in
 * include/pbf_solver.hpp.
 *
 * Conventions
 *   - Every function returns PBF_OK (0) or a negative PBF_ERR_* code; pbf_last_error()
 *     describes the last failure of the calling thread.  Nothing calls exit() (the reference's
 *     CHECK_CUDA does, kernels/cudaPBFSolver.cuh:5-10).
 *   - Host arrays of positions / velocities are float xyz triples (AoS, 12 bytes per particle)
 *     in the caller's particle order, as the reference's SolverData.position / .velocity
 *     (include/solverBase.h:45-47).  Inside, the solver keeps cell-sorted SoA-of-float4.
 *   - One solver = one HIP stream on the current device.  pbf_step / pbf_solve only enqueue
 *     work; functions returning host data synchronise that stream.  Solvers are independent.
 */
#ifndef PBF_H_
#define PBF_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_OK 0
#define PBF_ERR_ARG (-1)      /* bad argument */
#define PBF_ERR_HIP (-2)      /* a HIP runtime call failed */
#define PBF_ERR_CAPACITY (-3) /* size beyond what the solver was created for */
#define PBF_ERR_STATE (-4)    /* call not valid in the solver's current state */
#define PBF_ERR_NOGPU (-5)    /* no HIP device */

/* flags */
#define PBF_FLAG_DEBUG 1u         /* keep per-step intermediates for pbf_get_grid / _neighbors / _fields */
#define PBF_FLAG_DENSITY_STATS 2u /* compute density-error statistics every step */
#define PBF_FLAG_GRAPH 4u         /* replay each step from a captured hipGraph */
#define PBF_FLAG_FAST_MATH 8u     /* fast arithmetic in the constraint passes' neighbour terms: the
                                     hardware reciprocal square root for 1/|r|.  Results then match
                                     the bit-exact default to tolerances, not bits. */

/* Solver parameters.  The first 13 fields are the reference's SolverParameter
 * (include/solverBase.h:16-40) in the same order, size and meaning.  restRho is the reference's
 * mass-free rest density (rho = sum of W, kernels/cudaPBFSolver.cu:170), i.e. rest density
 * divided by particle mass.  The fields after `radius` are constants the reference hard-codes:
 * gravity (cu:9: (0,-49,0)), the clamp jitter coefficient (solverBase.h:5: 0.01), the
 * spiky-gradient cut-off below which the gradient is zero (cuh:107: 0.01), the grid
 * dimensions (solverBase.h:11-13: 64^3) and the neighbour cap (solverBase.h:6: 63). */
typedef struct pbf_params {
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
    float grad_rmin;       /* spiky gradient is 0 below this |r| (cuh:107: 0.01) */
    int32_t grid_dims[3];
    int32_t max_neighbors; /* 1..255 */
    uint32_t flags;        /* PBF_FLAG_* */
} pbf_params;

typedef struct pbf_solver pbf_solver;

/* Replaces initCUDASolverDataBuffer (kernels/cudaPBF.h:8-10, kernels/cudaPBF.cu:4-30):
 * allocates every device buffer for n particles (0 <= n < 2^28 - 32) and uploads positions and
 * velocities.  grid_space = {pmin xyz, pmax xyz}: only pmin anchors the grid (cuh:156-160), as in the
 * reference where the app passes the box (src/renderapp.cpp:251).  boundary = the clamp box
 * (box shrunk by the particle radius, renderapp.cpp:131).  vel_xyz may be NULL (zero). */
int pbf_create(const pbf_params* para, const float grid_space[6], const float boundary[6], const float* pos_xyz,
               const float* vel_xyz, int32_t n, pbf_solver** out);

/* A batch of nscenes independent scenes (1..255) advanced by one solver: every step is one
 * launch per kernel over all scenes.  counts[k] particles of scene k follow one another in
 * pos_xyz / vel_xyz (the batch's particle order); grid_space[6k..6k+5], boundary[6k..6k+5] and
 * grid_dims[3k..3k+2] are scene k's.  para is shared; its grid_dims are ignored.  Each scene
 * evolves exactly as a pbf_create solver of that scene alone would.  No counterpart in the
 * reference, which runs one fixed 64^3 scene. */
int pbf_create_batch(const pbf_params* para, int32_t nscenes, const int32_t* counts, const float* grid_space,
                     const float* boundary, const int32_t* grid_dims, const float* pos_xyz, const float* vel_xyz,
                     pbf_solver** out);

/* Launch shape of the neighbour kernels: block 64/128/256/512, grid > 0 = persistent grid of
 * that many workgroups (0 = one per block), xcd_map = contiguous particle chunks per XCD.
 * Results do not depend on it.  No counterpart in the reference, which launches every kernel
 * with one fixed shape (kernels/cudaPBF.cu:54-56). */
int pbf_set_tuning(pbf_solver* s, int32_t block, int32_t grid, int32_t xcd_map);
int pbf_get_tuning(pbf_solver* s, int32_t out[3]);

/* Order of the cell rows (the grid_dims[0] cells of one (y, z)) in the cell table, which is also
 * the particles' order in memory: 0 = z, then y (the reference's computeGid order,
 * kernels/cudaPBFSolver.cuh:156-160); 1 = a Morton curve over (y, z), per scene.  Cells stay
 * x-fastest inside a row, and the neighbour lists keep the reference's z / y / x order either
 * way, so results do not depend on it (pbf_get_grid always reports the reference's cell
 * order).  No counterpart in the reference. */
int pbf_set_ordering(pbf_solver* s, int32_t order);
int pbf_get_ordering(pbf_solver* s, int32_t* out);

/* Autotuner: times every launch shape over `steps` steps (rounded up to whole multi-step graphs)
 * replayed as pbf_advance replays them, each from the state it was called on, keeps the fastest,
 * then times that shape with the other row order (pbf_set_ordering) and keeps the faster; it
 * restores the state (tuning does not advance the simulation).  The last step's diagnostics (pbf_get_grid, _neighbors, _fields,
 * pbf_density_stats) are invalid until the next step.  best[3] = block, grid, xcd_map; best_ms =
 * time per step; the order chosen is pbf_get_ordering's.  Either may be NULL.  No counterpart in
 * the reference. */
int pbf_autotune(pbf_solver* s, int32_t steps, int32_t* best, float* best_ms);

/* Replaces cudaPBFSolve (kernels/cudaPBF.h:12, kernels/cudaPBF.cu:50-114): one solver step
 * with para's values (dt, iterations, ...).  para->nParticles and para->grid_dims must equal
 * the solver's. */
int pbf_solve(pbf_solver* s, const pbf_params* para);

/* One step with the solver's current parameters and time step dt (para.dt = dt,
 * para.invDt = 1/dt), as cudaPBFSolve (kernels/cudaPBF.cu:50) called with an updated para.dt.
 * With PBF_FLAG_GRAPH, a pbf_solve / pbf_step call whose parameters differ from the last ones
 * recaptures the step's graph; from the second such call in a row on, steps launch eagerly (no
 * capture) until a call leaves the parameters unchanged, so a caller that adapts dt every step
 * pays no capture per step.  Results are the same either way. */
int pbf_step(pbf_solver* s, float dt);

/* nsteps steps with the current parameters: the app's frame loop of cudaPBFSolve calls
 * (src/renderapp.cpp:497-502) without its per-frame host timing. */
int pbf_advance(pbf_solver* s, int32_t nsteps);

/* Host copy of the state in the caller's particle order (either pointer may be NULL): what the
 * app reads from SolverData.position / .velocity (include/solverBase.h:45-47). */
int pbf_get_state(pbf_solver* s, float* pos_xyz, float* vel_xyz);

/* Replace the state (caller's particle order; vel_xyz NULL = zero): the upload of
 * initCUDASolverDataBuffer (kernels/cudaPBF.cu:4-30) into an existing solver.  Captured step
 * graphs stay valid (a caller that uploads every frame keeps replaying them). */
int pbf_set_state(pbf_solver* s, const float* pos_xyz, const float* vel_xyz);

/* Write positions (and velocities if non-NULL) in the caller's particle order into DEVICE
 * buffers the caller owns, on the solver's stream -- the reference's model where the position
 * buffer belongs to the app (a mapped GL VBO, src/renderapp.cpp:236-249, cudaPBF.cu:7).  Only
 * enqueued: pbf_synchronize before another stream or the host reads the buffers (the C++ facade's
 * exportPositions does). */
int pbf_export_device(pbf_solver* s, float* dev_pos_xyz, float* dev_vel_xyz);

/* Replaces freeCUDASolverDataBuffer (kernels/cudaPBF.h:11, kernels/cudaPBF.cu:32-48). */
int pbf_destroy(pbf_solver* s);

/* Text of the calling thread's last error (the reference's CHECK_CUDA prints and exits instead,
 * kernels/cudaPBFSolver.cuh:5-10). */
const char* pbf_last_error(void);

/* Current parameters of the solver.  No counterpart in the reference (the app keeps para). */
int pbf_get_params(pbf_solver* s, pbf_params* out);

/* Block until the solver's queued work is done (the reference synchronises the device after
 * every launch, kernels/cudaPBF.cu:62-113). */
int pbf_synchronize(pbf_solver* s);

/* Device view of the cell-sorted state after the last step: n float4 positions (x, y, z, |curl|
 * of the step), n float4 velocities (w = 0), uint32 particle ids per slot -- the device pointers
 * the reference's SolverData exposes (include/solverBase.h:45-47), in slot order. */
typedef struct pbf_device_view {
    const void* pos4;
    const void* vel4;
    const uint32_t* ids;
    int32_t n;
    int32_t ncells;
} pbf_device_view;
int pbf_device_state(pbf_solver* s, pbf_device_view* out);

/* ---- diagnostics of the last step (need PBF_FLAG_DEBUG at create / solve time) -----------
 * The intermediates the reference keeps in SolverData (include/solverBase.h:45-63: grid,
 * gridParticlesCnt, gridParticlesID, neighbors, lambda, particlesCurl), read back by particle
 * id; the reference has no read-back functions of its own. */

/* gid[n] by particle id (-1 = outside the grid), count[ncells], cell_start[ncells+1]
 * (cell_start[ncells] = particles inside the grid), sorted_ids[n] (particle ids in cell
 * order, inside-grid particles first).  Any pointer may be NULL.  Returns the number of
 * particles inside the grid (>= 0) or an error. */
int pbf_get_grid(pbf_solver* s, int32_t* gid, int32_t* count, int32_t* cell_start, int32_t* sorted_ids);

/* Frozen neighbour lists by particle id: nbr[n*max_neighbors] (particle ids, list order),
 * cnt[n]. */
int pbf_get_neighbors(pbf_solver* s, int32_t* nbr, int32_t* cnt);

/* lambda of the first and last Jacobi iteration, density error rho/restRho-1 at the end of the
 * constraint loop, curl (xyz) -- all by particle id; any pointer may be NULL.  After a learned
 * step of 2 passes whose guess was rejected, the last lambda is the network's multiplier (the
 * positions are those of the plain pass). */
int pbf_get_fields(pbf_solver* s, float* lam_first, float* lam_last, float* err, float* curl_xyz);

/* Density-error statistics of the last step: out[0] = mean |err|, out[1] = mean max(err,0),
 * out[2] = max err.  Needs PBF_FLAG_DENSITY_STATS or PBF_FLAG_DEBUG. */
int pbf_density_stats(pbf_solver* s, float out[3]);

/* ---- measurement (no counterpart in the reference, which times cudaPBFSolve with the host
 * clock, src/renderapp.cpp:498-501) ---------------------------------------------------------- */

/* Runs nsteps steps and returns the elapsed time measured with HIP events recorded on the
 * solver's stream around them.  The step graphs the run replays are captured before the first
 * event, so the window holds no host capture or instantiation. */
int pbf_time_steps(pbf_solver* s, int32_t nsteps, float* ms_total);

/* Per-stage times of one step (HIP events between stages, one step, no graph):
 * out[0] predict+hash, [1] scan, [2] scatter+rank+reorder, [3] neighbours+lambda1,
 * [4] Jacobi iterations, [5] curl, [6] vorticity+viscosity.  ms.  A step's vorticity pass also
 * predicts, hashes and counts for the next step, so out[0] is about 0 and out[6] includes that
 * hash, unless the state changed since the last step (pbf_set_state, dt, parameters, row order),
 * which makes the step hash for itself in out[0]. */
int pbf_stage_times(pbf_solver* s, float out[7]);

/* One step launched eagerly with a HIP event recorded on the solver's stream after every
 * launch group; ms[i] is the time of group i and names[i] its name (static strings:
 * predict_hash, scan, scatter_rank_reorder, neighbors_lambda, lambda, delta_apply, curl,
 * vort_apply).  Returns the number of groups written (<= max) or an error. */
int pbf_profile_step(pbf_solver* s, float* ms, const char** names, int32_t max);

/* ---- learned warm start and adaptive iteration (no counterpart in the reference) ----------
 * A per-particle MLP 16 -> 64 -> 64 -> 64 -> 1 (hard-swish), BF16 weights and activations, FP32
 * accumulation on MFMA, predicts the multiplier a plain step accumulates over its Jacobi
 * iterations 2..I (divided by h^2) from 16 features of the particle after the plain pass 1:
 *   0 C = rho/restRho - 1   1 neighbour count   2 sum_j |grad_j C|^2 / restRho^2
 *   3 |grad_i C| / restRho   4-6 v dt / h (predicted velocity)   7 distance to the nearest wall
 *   / h (<= 4)   8 lambda of pass 2 / h^2   9 lambda of pass 1 / h^2   10 C at the predicted
 *   positions   11-15 distance to the walls x-, x+, y-, y+, z- / h (<= 4)
 * Parameters are packed FP32: W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4. */
typedef struct pbf_mlp pbf_mlp;
int pbf_mlp_param_count(void);
/* New network with seeded He-uniform weights on the current device. */
int pbf_mlp_create(uint64_t seed, pbf_mlp** out);
int pbf_mlp_destroy(pbf_mlp* m);
int pbf_mlp_get_params(pbf_mlp* m, float* params);
int pbf_mlp_set_params(pbf_mlp* m, const float* params);
/* Input standardisation x' = clip((x - mean) * inv_std, -8, 8); output multiplier / h^2 =
 * y * yscale + yshift. */
int pbf_mlp_set_normalization(pbf_mlp* m, const float* feat_mean, const float* feat_inv_std, float yscale,
                              float yshift);
/* y[n] (the network output before yscale / yshift) for raw features x[n][16] (host arrays). */
int pbf_mlp_forward(pbf_mlp* m, const float* x, int32_t n, float* y);
/* Mean squared error of y against t over n samples and its gradient (host arrays; grad may be
 * NULL). */
int pbf_mlp_loss_grad(pbf_mlp* m, const float* x, const float* t, int32_t n, float* grad, float* loss);
/* pbf_mlp_loss_grad for callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc): no counterpart in the reference.  dev_x [n][16] raw features and
 * dev_t [n] targets are DEVICE arrays; dev_grad (device) receives the kMlpParams = 9473 gradient
 * values followed by the sum of squared errors (the loss times n).  Allocates nothing: the
 * kernels' per-workgroup partial sums go to a fixed library scratch of 256 rows (9.25 MiB per
 * device, held until the process ends), which bounds a launch to 256 workgroups for any n >= 1.
 * Results equal pbf_mlp_loss_grad's bit for bit.  Calls that overlap on one device (other
 * threads, other networks) share that scratch and race: serialise them. */
int pbf_mlp_loss_grad_noalloc(pbf_mlp* m, const float* dev_x, const float* dev_t, int32_t n, float* dev_grad);
/* One Adam step with grad (host) or, when NULL, the gradient of the last pbf_mlp_loss_grad. */
int pbf_mlp_adam_step(pbf_mlp* m, const float* grad, float lr, float beta1, float beta2, float eps);
/* steps Adam steps on random batches of a device-resident dataset (dev_x [nsamples][16] raw
 * features, dev_t [nsamples] normalised targets), learning rate cosine-decayed from lr to
 * lr_final; losses[steps] (host, may be NULL) receives the batch losses. */
int pbf_mlp_train(pbf_mlp* m, const float* dev_x, const float* dev_t, int64_t nsamples, int32_t batch, int32_t steps,
                  float lr, float lr_final, uint64_t seed, float* losses);
/* Use m for the warm start of s (NULL detaches); m must outlive its use. */
int pbf_attach_mlp(pbf_solver* s, pbf_mlp* m);

/* One plain step with `iterations` (>= 2) iterations that also records, for each particle whose
 * id satisfies id % stride == phase, its 16 raw features of pass 2 and the multiplier accumulated
 * over passes 2..iterations, divided by h^2, at row base + id / stride of the device arrays
 * dev_x [cap][16] and dev_t [cap]. */
int pbf_record_step(pbf_solver* s, int32_t iterations, float* dev_x, float* dev_t, int64_t base, int64_t cap,
                    int32_t stride, int32_t phase);

/* One step whose constraint loop stops once the residual is <= tol after at least min_iters
 * iterations, or after max_iters (tol < 0: exactly max_iters, no read-back).  metric 0:
 * compression residual, the mean over particles of max(C_i, 0), C_i = rho_i/restRho - 1 over the
 * frozen lists; metric 1: the mean of |C_i|.  use_mlp: pass 2 is the learned pass.  After the
 * plain pass 1, the lambda pass writes the warm-start features, the attached network's multiplier
 * replaces lambda 2 for one delta-p pass, and the result is kept only if mean |C| does not grow
 * (otherwise the pass reruns plain, exactly as with use_mlp = 0).  Plain iterations then continue
 * until the tolerance holds (the fallback that judges the guess).  Writes the passes run (the
 * learned pass counts as one) and the last residual measured (either pointer may be NULL). */
int pbf_step_adaptive(pbf_solver* s, int32_t use_mlp, int32_t metric, float tol, int32_t min_iters, int32_t max_iters,
                      int32_t* iters_used, float* resid);

/* Of the last learned pass: out[0] / out[1] = compression after the plain pass 1 / after the
 * network's pass, out[2] = 1 if it was kept (0 if rejected), out[3] / out[4] = mean |C| after the
 * plain pass 1 / after the network's pass.  Synchronises the solver's stream. */
int pbf_last_guess_info(pbf_solver* s, float out[5]);

/* passes >= 2 (a network attached): every following pbf_step / pbf_advance / pbf_time_steps runs
 * a plain pass, the learned pass and passes - 2 plain iterations, with the keep / reject decision
 * made on the device (no read-back, graph-captured with PBF_FLAG_GRAPH).  passes = 0: plain
 * steps with params.iterations iterations again. */
int pbf_set_learned(pbf_solver* s, int32_t passes);

/* Measurement: the average duration of one launch of a step stage (stage 0: neighbour build +
 * iteration-1 lambda, 1: lambda, 2: delta-p + apply, 3: predict + hash, scan, scatter + rank +
 * reorder (the sort stage of a step that hashes for itself), 4: an empty kernel of the particle
 * grid's shape, i.e. the cost of one dependent launch, 5: a VALU issue-rate probe of 2048 x 256
 * threads issuing 65,536 wave64 VALU instructions per wavefront, 6: the vorticity pass with the
 * next step's hash, scan, scatter + rank + reorder (the sort stage of a step whose predecessor
 * hashed for it, with the vorticity pass), 7: the vorticity pass alone; the vorticity replays
 * write a spare buffer; 8: the curl pass with the XSPH sum), replayed reps times back to back on
 * the state of the last step,
 * between HIP events on the solver's stream.  The state is unchanged (the replays write only
 * what the next step rebuilds); the diagnostics of the last step (pbf_get_grid,
 * pbf_get_neighbors, pbf_get_fields) no longer describe it afterwards.  No counterpart in the
 * reference. */
int pbf_time_stage(pbf_solver* s, int32_t stage, int32_t reps, float* ms);

/* save != 0: copy the current state (positions, velocities, particle order) into one device-side
 * snapshot of the solver; save == 0: restore the state from it.  Stream-ordered, no host copies.
 * No counterpart in the reference. */
int pbf_snapshot(pbf_solver* s, int32_t save);

/* Device memory helpers for callers without an allocator (datasets, export targets).  No
 * counterpart in the reference. */
int pbf_device_alloc(uint64_t bytes, void** out);
int pbf_device_free(void* p);
int pbf_memcpy_to_host(void* dst, const void* src, uint64_t bytes);
int pbf_memcpy_to_device(void* dst, const void* src, uint64_t bytes);
/* Streaming copy kernel of `bytes` (multiple of 16) on the null stream, timed with HIP events:
 * a known-traffic calibration for hardware counters.  No counterpart in the reference. */
int pbf_copy_kernel(void* dst, const void* src, uint64_t bytes, float* ms);

/* Library build string.  No counterpart in the reference. */
const char* pbf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PBF_H_ */
