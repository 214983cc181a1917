// Test-only driver around the reference's CPU solver and scan test, compiled together with
// their sources where they lie (src/solver.cpp, src/AABB.cpp, test/testPrefixSum.cpp) by
// oracle/Makefile into oracle/_ref/libpbfref.so.  Nothing of the reference is copied here:
// this file only calls into the reference's own functions and reads the members that
// PBFSolver::solve (src/solver.cpp:60-157) leaves behind.
//
// Used by tests/test_oracle.py to pin the oracle's grid hash, cell table, sorted cell
// members, neighbour sets, lambda and delta-p, and (ref_step_probe) its multi-iteration
// positions, loop velocity, curl and XSPH viscosity against the reference implementation.

// Standard and glm headers first, so their own access specifiers are untouched.
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <atomic>
#include <thread>
#include <chrono>
#include <iostream>
#include <vector>
#include <unistd.h>
#include <fcntl.h>
#include "config.h"

// PBFSolver keeps the grid, neighbour lists, lambda and offsets private; expose them to
// this driver only.
#define private public
#include "solver.h"
#undef private

void computePrefixSum(uint32_t* data, uint32_t* prefixSum, int n);                         // test/testPrefixSum.cpp:3
void computeParllelPrefixSum(uint32_t* data, uint32_t* prefixSum, int n, int pow2GreaterN);  // test/testPrefixSum.cpp:26

namespace {
struct QuietStdout {
    int saved;
    QuietStdout() {
        fflush(stdout);
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        dup2(devnull, 1);
        close(devnull);
    }
    ~QuietStdout() {
        fflush(stdout);
        dup2(saved, 1);
        close(saved);
    }
};
}  // namespace

extern "C" {

// Serial and Blelloch exclusive scans of the reference's own scan test.  out must hold the
// next power of two >= n entries for the Blelloch variant.
void ref_scan_serial(const uint32_t* data, uint32_t* out, int n) {
    QuietStdout q;
    std::vector<uint32_t> d(data, data + n);
    computePrefixSum(d.data(), out, n);
}

void ref_scan_blelloch(const uint32_t* data, uint32_t* out, int n) {
    QuietStdout q;
    uint32_t p2 = 1;
    while ((int)p2 < n) p2 <<= 1;
    std::vector<uint32_t> d(p2, 0u);
    std::memcpy(d.data(), data, sizeof(uint32_t) * n);
    std::vector<uint32_t> o(p2, 0u);
    computeParllelPrefixSum(d.data(), o.data(), n, (int)p2);
    std::memcpy(out, o.data(), sizeof(uint32_t) * n);
}

float ref_wpoly6(const float* r, float h) { return Wpoly6(glm::vec3(r[0], r[1], r[2]), h); }

void ref_spiky_grad(const float* r, float h, float* g) {
    glm::vec3 o = delWspiky(glm::vec3(r[0], r[1], r[2]), h);
    g[0] = o.x;
    g[1] = o.y;
    g[2] = o.z;
}

// Runs PBFSolver::solve once with dt = 0 and one iteration, so that the predicted positions
// equal the given ones (all of which must lie inside the boundary), and returns what the
// reference computed on them: per-particle cell id, the 64^3 cell table, the sorted cell
// members, neighbour lists (x-outer order, src/solver.cpp:31-57), lambda (:78-95) and the
// position offset (:96-108) of that iteration.
//
// params: h, restRho, epsilon, k, n, deltaq, radius (SolverParameter fields).
int ref_probe(const float* params, const float box[6], const float* pos, int n, int32_t* gid, int32_t* cell_start,
              int32_t* cell_end, int32_t* sorted_ids, int32_t* nbr /* n*64 */, float* lambda, float* offset) {
    SolverParameter para;
    std::memset(&para, 0, sizeof(para));
    para.h = params[0];
    para.nParticles = n;
    para.restRho = params[1];
    para.epsilon = params[2];
    para.k = params[3];
    para.n = params[4];
    para.deltaq = params[5];
    para.radius = params[6];
    para.c = 0.0f;
    para.dt = 0.0f;
    para.invDt = 0.0f;
    para.iterations = 1;
    para.vorticityConfinementEpilon = 0.0f;
    AABB aabb(glm::vec3(box[0], box[1], box[2]), glm::vec3(box[3], box[4], box[5]));
    PBFSolver* solver = new PBFSolver(para, aabb);
    // The constructor's memset clears only sizeof(int*) bytes of the counters (solver.h:102);
    // clear all of them, as intended, so a reused heap block cannot leave stale counts.
    std::memset(solver->gridParticlesCnt, 0, sizeof(int) * solver->gridsCnt);
    std::vector<Particle> particles(n);
    for (int i = 0; i < n; ++i) {
        particles[i].position = glm::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        particles[i].velocity = glm::vec3(0.0f);
    }
    {
        QuietStdout q;
        solver->solve(particles.data());
    }
    for (int i = 0; i < n; ++i) gid[i] = solver->particlesGid[i];
    for (int c = 0; c < solver->gridsCnt; ++c) {
        cell_start[c] = (int32_t)solver->grids[c].startIdx;
        cell_end[c] = (int32_t)solver->grids[c].endIdx;
    }
    int in_grid = solver->gridsCnt ? (int)solver->grids[solver->gridsCnt - 1].endIdx : 0;
    for (int i = 0; i < n; ++i) sorted_ids[i] = i < in_grid ? solver->gridParticlesID[i] : -1;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < 64; ++j) nbr[(size_t)i * 64 + j] = solver->neighbors[i][j];
        lambda[i] = solver->lambda[i];
        offset[3 * i] = solver->particlesOffset[i].x;
        offset[3 * i + 1] = solver->particlesOffset[i].y;
        offset[3 * i + 2] = solver->particlesOffset[i].z;
    }
    // PBFSolver has no destructor; release what its constructor allocated.
    for (int i = 0; i < n; ++i) delete[] solver->neighbors[i];
    delete[] solver->neighbors;
    delete[] solver->grids;
    delete[] solver->gridParticlesCnt;
    delete[] solver->particlesGid;
    delete[] solver->gridParticlesID;
    delete solver;
    return in_grid;
}

// Runs the reference's PBFSolver::solve for nsteps steps on AoS particles in place.
// params: h, restRho, epsilon, k, n, deltaq, radius, c, dt, iterations, vorticity epsilon.
int ref_run(const float* params, const float box[6], float* pos, float* vel, int n, int nsteps) {
    SolverParameter para;
    std::memset(&para, 0, sizeof(para));
    para.h = params[0];
    para.nParticles = n;
    para.restRho = params[1];
    para.epsilon = params[2];
    para.k = params[3];
    para.n = params[4];
    para.deltaq = params[5];
    para.radius = params[6];
    para.c = params[7];
    para.dt = params[8];
    para.invDt = 1.0f / params[8];
    para.iterations = (int)params[9];
    para.vorticityConfinementEpilon = params[10];
    AABB aabb(glm::vec3(box[0], box[1], box[2]), glm::vec3(box[3], box[4], box[5]));
    PBFSolver* solver = new PBFSolver(para, aabb);
    // The constructor's memset clears only sizeof(int*) bytes of the counters (solver.h:102);
    // clear all of them, as intended, so a reused heap block cannot leave stale counts.
    std::memset(solver->gridParticlesCnt, 0, sizeof(int) * solver->gridsCnt);
    std::vector<Particle> particles(n);
    for (int i = 0; i < n; ++i) {
        particles[i].position = glm::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        particles[i].velocity = glm::vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
    {
        QuietStdout q;
        for (int s = 0; s < nsteps; ++s) solver->solve(particles.data());
    }
    for (int i = 0; i < n; ++i) {
        pos[3 * i] = particles[i].position.x;
        pos[3 * i + 1] = particles[i].position.y;
        pos[3 * i + 2] = particles[i].position.z;
        vel[3 * i] = particles[i].velocity.x;
        vel[3 * i + 1] = particles[i].velocity.y;
        vel[3 * i + 2] = particles[i].velocity.z;
    }
    for (int i = 0; i < n; ++i) delete[] solver->neighbors[i];
    delete[] solver->neighbors;
    delete[] solver->grids;
    delete[] solver->gridParticlesCnt;
    delete[] solver->particlesGid;
    delete[] solver->gridParticlesID;
    delete solver;
    return 0;
}

// Runs PBFSolver::solve once on AoS particles (positions and velocities given) and returns what
// the step leaves behind, for the parts the CPU solver shares with the GPU method:
//   pos_out    positions after the constraint loop (src/solver.cpp:74-116; the vorticity and
//              viscosity only change velocities);
//   vel_loop   the velocity the loop hands to the curl and viscosity sums, (x - x_old) * invDt
//              (:117-120), recomputed here from the same float operands, so it is bit-identical;
//   curl       particlesCurl (:122-132, with v_ij = v_i - v_j);
//   visc       particlesViscosity = c * sum_j -(v_i - v_j) W(r) (:138-150);
//   vel_out    the final velocities (:152-155);
//   lambda     the multiplier of the last iteration (:78-95).
// params: h, restRho, epsilon, k, n, deltaq, radius, c, dt, iterations, vorticity epsilon.
int ref_step_probe(const float* params, const float box[6], const float* pos, const float* vel, int n, float* pos_out,
                   float* vel_loop, float* curl, float* visc, float* vel_out, float* lambda) {
    SolverParameter para;
    std::memset(&para, 0, sizeof(para));
    para.h = params[0];
    para.nParticles = n;
    para.restRho = params[1];
    para.epsilon = params[2];
    para.k = params[3];
    para.n = params[4];
    para.deltaq = params[5];
    para.radius = params[6];
    para.c = params[7];
    para.dt = params[8];
    para.invDt = 1.0f / params[8];
    para.iterations = (int)params[9];
    para.vorticityConfinementEpilon = params[10];
    AABB aabb(glm::vec3(box[0], box[1], box[2]), glm::vec3(box[3], box[4], box[5]));
    PBFSolver* solver = new PBFSolver(para, aabb);
    std::memset(solver->gridParticlesCnt, 0, sizeof(int) * solver->gridsCnt);
    std::vector<Particle> particles(n);
    for (int i = 0; i < n; ++i) {
        particles[i].position = glm::vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        particles[i].velocity = glm::vec3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
    {
        QuietStdout q;
        solver->solve(particles.data());
    }
    for (int i = 0; i < n; ++i) {
        const glm::vec3 x = particles[i].position;
        const glm::vec3 vl = (x - solver->oldPosition[i]) * para.invDt;
        const glm::vec3 c = solver->particlesCurl[i];
        const glm::vec3 vs = solver->particlesViscosity[i];
        const glm::vec3 v = particles[i].velocity;
        for (int a = 0; a < 3; ++a) {
            pos_out[3 * i + a] = x[a];
            vel_loop[3 * i + a] = vl[a];
            curl[3 * i + a] = c[a];
            visc[3 * i + a] = vs[a];
            vel_out[3 * i + a] = v[a];
        }
        lambda[i] = solver->lambda[i];
    }
    for (int i = 0; i < n; ++i) delete[] solver->neighbors[i];
    delete[] solver->neighbors;
    delete[] solver->grids;
    delete[] solver->gridParticlesCnt;
    delete[] solver->particlesGid;
    delete[] solver->gridParticlesID;
    delete solver;
    return 0;
}

int ref_grid_cells(void) { return UNIFORM_GRID_DIM_SIZE_X * UNIFORM_GRID_DIM_SIZE_Y * UNIFORM_GRID_DIM_SIZE_Z; }

}  // extern "C"
