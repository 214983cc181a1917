"""The C ABI library without a GPU: it loads, exports every function include/pbf.h declares,
its parameter struct matches the ctypes mirror, the C++ facade compiles and links against it,
and creation fails cleanly (an error code, no exit) when no device is present."""
import ctypes as C
import os
import re
import subprocess

import pytest

import pbf_amd
from pbf_amd import SolverParameter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            text = open(os.path.join(INCLUDE, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(pbf_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = C.CDLL(pbf_amd.LIB_PATH)
    names = declared_functions()
    assert "pbf_create" in names and "pbf_solve" in names and "pbf_destroy" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_params_struct_layout():
    # 13 reference fields + gravity[3] + jitter + grad_rmin + dims[3] + cap + flags = 23 words
    assert C.sizeof(SolverParameter) == 23 * 4
    assert SolverParameter.radius.offset == 12 * 4  # the reference's SolverParameter ends here (52 B)
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "pbf.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu\n", sizeof(pbf_params), offsetof(pbf_params, radius),
           offsetof(pbf_params, grid_dims), offsetof(pbf_params, max_neighbors), offsetof(pbf_params, flags));
    return 0;
}
"""
    out = compile_and_run(src, "c")
    vals = [int(x) for x in out.split()]
    assert vals == [
        C.sizeof(SolverParameter),
        SolverParameter.radius.offset,
        SolverParameter.grid_dims.offset,
        SolverParameter.max_neighbors.offset,
        SolverParameter.flags.offset,
    ]


def compile_and_run(src, lang, link=False, tmp=None):
    import tempfile

    d = tmp or tempfile.mkdtemp()
    ext = "c" if lang == "c" else "cpp"
    path = os.path.join(d, f"t.{ext}")
    open(path, "w").write(src)
    exe = os.path.join(d, "t")
    cc = "gcc" if lang == "c" else "g++"
    cmd = [cc, "-O1", "-I", INCLUDE, path, "-o", exe]
    if lang != "c":
        cmd.insert(1, "-std=c++17")
    if link:
        cmd += [pbf_amd.LIB_PATH, f"-Wl,-rpath,{os.path.dirname(pbf_amd.LIB_PATH)}"]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    return r.stdout + r.stderr


def test_cpp_facade_compiles_links_and_reports_errors():
    """A reference-shaped caller (cudaPBF.h / solver.h API) builds against the facade.  Without a
    GPU, creation returns PBF_ERR_NOGPU as an exception instead of exiting."""
    src = r"""
#include <cstdio>
#include <vector>
#include "pbf_solver.hpp"
using namespace pbf_ref;
int main() {
    SolverParameter para{};
    para.h = 0.1f; para.nParticles = 2; para.restRho = 8000.f; para.epsilon = 100.f; para.k = 1e-3f;
    para.n = 4.f; para.deltaq = 0.02f; para.c = 0.f; para.dt = 1.f / 120.f; para.invDt = 120.f;
    para.iterations = 4; para.radius = 0.03f; para.grid_dims[0] = para.grid_dims[1] = para.grid_dims[2] = 16;
    para.max_neighbors = 63;
    std::vector<vec3> pos{{0.5f, 0.5f, 0.5f}, {0.55f, 0.5f, 0.5f}}, vel(2, vec3{0, 0, 0});
    AABB box{{0, 0, 0}, {1.6f, 1.6f, 1.6f}}, bd{{0.03f, 0.03f, 0.03f}, {1.57f, 1.57f, 1.57f}};
    SolverData data;
    try {
        initCUDASolverDataBuffer(pos, vel, data, para, box, bd);
        cudaPBFSolve(data, para);
        freeCUDASolverDataBuffer(data);
        std::printf("ran\n");
    } catch (const Error& e) {
        std::printf("error %d\n", e.code);
    }
    PBFSolver solver(para, box);
    (void)solver;
    return 0;
}
"""
    out = compile_and_run(src, "c++", link=True)
    from conftest import gpu_available

    if gpu_available():
        assert "ran" in out
    else:
        assert f"error {pbf_amd.PBF_ERR_NOGPU}" in out


def test_create_without_gpu_is_an_error_not_a_crash():
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    from pbf_amd import PBFError, Solver, scenes

    sc = scenes.dam_break((4, 4, 4))
    with pytest.raises(PBFError) as ei:
        Solver.from_scene(sc)
    assert ei.value.code == pbf_amd.PBF_ERR_NOGPU


@pytest.mark.parametrize("n", [(1 << 28) - 32, 1 << 28])
def test_create_rejects_too_many_particles(n):
    """Particle counts of 2^28 - 32 and more are refused before any device work: float4 slots,
    the search's 32 padding slots included, are addressed with 32-bit byte offsets."""
    from pbf_amd import scenes

    sc = scenes.dam_break((4, 4, 4))
    para = SolverParameter.from_dict(dict(sc.params), n, sc.grid_dims, 0)
    box = (C.c_float * 6)(*sc.box)
    bnd = (C.c_float * 6)(*sc.boundary)
    out = C.c_void_p()
    lib = pbf_amd.lib()
    rc = lib.pbf_create(C.byref(para), box, bnd, None, None, n, C.byref(out))
    assert rc == pbf_amd.PBF_ERR_ARG
    assert out.value is None
    assert b"2^28" in lib.pbf_last_error()


def test_facade_parameters_default_to_the_reference_constants():
    """A caller that sets only the reference's 13 SolverParameter fields (src/renderapp.cpp:116-129)
    gets the constants the reference hard-codes for the appended fields: gravity (0, -49, 0)
    (kernels/cudaPBFSolver.cu:9), jitter 0.01 (solverBase.h:5), gradient cut-off 0.01
    (cudaPBFSolver.cuh:107), a 64^3 grid (solverBase.h:11-13) and 63 neighbours (solverBase.h:6)."""
    src = r"""
#include <cstdio>
#include "pbf_solver.hpp"
using namespace pbf_ref;
int main() {
    SolverParameter para;
    para.h = 1.0; para.nParticles = 8; para.restRho = 6; para.epsilon = 1.0f; para.k = 0.01;
    para.deltaq = 0.1; para.n = 4; para.c = 0.1f; para.dt = 0.016f; para.invDt = 1.0f / para.dt;
    para.iterations = 4; para.vorticityConfinementEpilon = 0.1; para.radius = 0.3;
    std::printf("%g %g %g %g %g %d %d %d %d %u\n", para.gravity[0], para.gravity[1], para.gravity[2],
                para.jitter_coeff, para.grad_rmin, para.grid_dims[0], para.grid_dims[1], para.grid_dims[2],
                para.max_neighbors, para.flags);
    SolverParameter z{};
    std::printf("%g %d\n", z.h, z.iterations);
    return 0;
}
"""
    out = compile_and_run(src, "c++").split()
    assert out[:10] == ["0", "-49", "0", "0.01", "0.01", "64", "64", "64", "63", "0"]
    assert out[10:] == ["0", "0"]
