// pbf_solver.hpp -- header-only C++ facade over the C ABI (pbf.h) with the reference's host API
// shape, so an app written against the reference's kernels/cudaPBF.h and include/solver.h
// compiles against the MI355X solver by swapping the include.
//
//   reference                                      here (namespace pbf_ref)
//   SolverParameter        include/solverBase.h:16  SolverParameter (pbf_params, same leading fields;
//                                                   the appended ones default to the reference's constants)
//   SolverData             include/solverBase.h:45  SolverData (holds the pbf_solver handle)
//   AABB                   include/AABB.h:6         AABB {vec3 pMin, pMax}
//   Particle               include/particle.h:9     Particle {vec3 position, velocity}
//   initCUDASolverDataBuffer  kernels/cudaPBF.h:8   initCUDASolverDataBuffer (template on the vec3 type)
//   cudaPBFSolve           kernels/cudaPBF.h:12     cudaPBFSolve
//   freeCUDASolverDataBuffer  kernels/cudaPBF.h:11  freeCUDASolverDataBuffer
//   PBFSolver              include/solver.h:83      PBFSolver (ctor(para, box, fext), solve(Particle*))
//
// Differences, by design: errors throw pbf_ref::Error (the reference's CHECK_CUDA exits the
// process); the positions the app reads are written into its own device buffer with
// exportPositions() instead of the solver stepping inside a mapped GL buffer.
#pragma once

#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pbf.h"

namespace pbf_ref {

struct vec3 {
    float x, y, z;
};
static_assert(sizeof(vec3) == 12, "vec3 must be three packed floats (glm::vec3 layout)");

struct AABB {
    vec3 pMin, pMax;
};

struct Particle {
    vec3 position;
    vec3 velocity;
};

// The reference's 13 fields (include/solverBase.h:16-40), zero until the caller sets them, and the
// constants the reference hard-codes, preset to its values so that a caller that fills only the
// 13 fields (as src/renderapp.cpp:116-129 does) runs the reference's configuration.
struct SolverParameter : pbf_params {
    SolverParameter() : pbf_params{} {
        gravity[0] = 0.0f;  // kernels/cudaPBFSolver.cu:9
        gravity[1] = -49.0f;
        gravity[2] = 0.0f;
        jitter_coeff = 0.01f;  // RANDOM_OFFSET_COEFF, include/solverBase.h:5
        grad_rmin = 0.01f;     // kernels/cudaPBFSolver.cuh:107
        grid_dims[0] = grid_dims[1] = grid_dims[2] = 64;  // UNIFORM_GRID_DIM_SIZE_*, solverBase.h:11-13
        max_neighbors = 63;    // MAX_NEIGHBOR, solverBase.h:6
        flags = 0;
    }
};
static_assert(sizeof(SolverParameter) == sizeof(pbf_params), "the facade adds no fields");

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc < 0) throw Error(rc, pbf_last_error());
}

struct SolverData {
    pbf_solver* solver = nullptr;
};

inline void aabb_to_array(const AABB& b, float out[6]) {
    out[0] = b.pMin.x;
    out[1] = b.pMin.y;
    out[2] = b.pMin.z;
    out[3] = b.pMax.x;
    out[4] = b.pMax.y;
    out[5] = b.pMax.z;
}

// kernels/cudaPBF.cu:4-30.  Vec is any 12-byte xyz type (glm::vec3, pbf_ref::vec3, ...).
template <class Vec>
void initCUDASolverDataBuffer(const std::vector<Vec>& position, const std::vector<Vec>& velocity, SolverData& data,
                              const SolverParameter& para, const AABB& gridSpace, const AABB& boundary) {
    static_assert(sizeof(Vec) == 3 * sizeof(float), "position type must be three floats");
    if ((int)position.size() != para.nParticles) throw Error(PBF_ERR_ARG, "position.size() != para.nParticles");
    if (!velocity.empty() && velocity.size() != position.size()) throw Error(PBF_ERR_ARG, "velocity size");
    float gs[6], bd[6];
    aabb_to_array(gridSpace, gs);
    aabb_to_array(boundary, bd);
    if (data.solver) {
        pbf_destroy(data.solver);
        data.solver = nullptr;
    }
    check(pbf_create(&para, gs, bd, reinterpret_cast<const float*>(position.data()),
                     velocity.empty() ? nullptr : reinterpret_cast<const float*>(velocity.data()), para.nParticles,
                     &data.solver));
}

// kernels/cudaPBF.cu:50-114: one step with para (dt, iterations, ...).
inline void cudaPBFSolve(SolverData& data, const SolverParameter& para) {
    if (!data.solver) throw Error(PBF_ERR_STATE, "solver data not initialised");
    check(pbf_solve(data.solver, &para));
}

// kernels/cudaPBF.cu:32-48.
inline void freeCUDASolverDataBuffer(SolverData& data) {
    if (data.solver) check(pbf_destroy(data.solver));
    data.solver = nullptr;
}

// Positions (and velocities) in the app's particle order into device buffers it owns -- the
// reference steps directly in a mapped GL buffer (src/renderapp.cpp:236-249).  Returns once the
// buffers are written, as the reference's synchronous cudaPBFSolve does (cudaPBF.cu:62-113), so
// the app may draw from them right away.
inline void exportPositions(SolverData& data, float* dev_pos_xyz, float* dev_vel_xyz = nullptr) {
    check(pbf_export_device(data.solver, dev_pos_xyz, dev_vel_xyz));
    check(pbf_synchronize(data.solver));
}

inline void readState(SolverData& data, std::vector<vec3>& position, std::vector<vec3>& velocity) {
    pbf_params p;
    check(pbf_get_params(data.solver, &p));
    position.resize(p.nParticles);
    velocity.resize(p.nParticles);
    check(pbf_get_state(data.solver, reinterpret_cast<float*>(position.data()),
                        reinterpret_cast<float*>(velocity.data())));
}

// include/solver.h:83-111: in-place step of an AoS particle array.  The clamp boundary is the
// box shrunk by para.radius (solver.h:97) and the grid is anchored at box.pMin (solver.h:109).
// fext is kept as in the reference, which never calls it (solver.h:201).
class PBFSolver {
   public:
    PBFSolver() = default;
    PBFSolver(const SolverParameter& para, const AABB& box,
              const std::function<vec3(const vec3&)>& fext = [](const vec3&) { return vec3{0.0f, 0.0f, 0.0f}; })
        : para_(para), box_(box), fext_(fext) {}
    PBFSolver(const PBFSolver&) = delete;
    PBFSolver& operator=(const PBFSolver&) = delete;
    ~PBFSolver() {
        if (data_.solver) pbf_destroy(data_.solver);
    }

    void solve(Particle* particles) {
        const int n = para_.nParticles;
        pos_.resize(n);
        vel_.resize(n);
        for (int i = 0; i < n; ++i) {
            pos_[i] = particles[i].position;
            vel_[i] = particles[i].velocity;
        }
        if (!data_.solver) {
            AABB boundary{{box_.pMin.x + para_.radius, box_.pMin.y + para_.radius, box_.pMin.z + para_.radius},
                          {box_.pMax.x - para_.radius, box_.pMax.y - para_.radius, box_.pMax.z - para_.radius}};
            initCUDASolverDataBuffer(pos_, vel_, data_, para_, box_, boundary);
        } else {
            check(pbf_set_state(data_.solver, reinterpret_cast<const float*>(pos_.data()),
                                reinterpret_cast<const float*>(vel_.data())));
        }
        cudaPBFSolve(data_, para_);
        check(pbf_get_state(data_.solver, reinterpret_cast<float*>(pos_.data()), reinterpret_cast<float*>(vel_.data())));
        for (int i = 0; i < n; ++i) {
            particles[i].position = pos_[i];
            particles[i].velocity = vel_[i];
        }
    }

    SolverData& data() { return data_; }

   private:
    SolverParameter para_{};
    AABB box_{};
    std::function<vec3(const vec3&)> fext_;
    SolverData data_;
    std::vector<vec3> pos_, vel_;
};

}  // namespace pbf_ref
