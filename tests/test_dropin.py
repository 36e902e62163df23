"""The C++ drop-in header (include/spock_amd.hpp): its file readers on the CPU, and the
reference's timing driver (examples/random/main.cu:24-85, as spock.cu_amd/examples/random_main.cpp)
built against the header and run on the GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from spock_amd import runtime as rt
from spock_amd.api import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.dirname(rt.LIB_PATH)
ROCM = "/opt/rocm"
HAVE_GXX = shutil.which("g++") is not None and os.path.isdir(os.path.join(ROCM, "include", "hip"))
needs_lib = pytest.mark.skipif(not (HAVE_GXX and os.path.exists(rt.LIB_PATH)),
                               reason="needs g++, the HIP headers and the built library")

READER = r'''
#include <cstdio>
#include "spock_amd.hpp"
using namespace spock_amd;
int main(int argc, char **argv) {
    ScenarioTree<double> tree(argv[1]);
    std::vector<double> x0 = parseFromFile<double>(tree.path() + "initialState" + tree.fpFileExt());
    std::vector<double> cp = parseFromFile<double>(tree.path() + "conditionalProbabilities" + tree.fpFileExt());
    std::vector<float> anc = parseFromFile<float>(tree.path() + "ancestors_u.bt");
    std::printf("%zu %zu %zu %zu %zu %zu\n", tree.numStates(), tree.numInputs(), tree.numNodes(),
                tree.numNonleafNodes(), tree.numStages(), tree.numEvents());
    for (double v : x0) std::printf("%.17g ", v);
    std::printf("\n%.17g %g\n", cp[5], (double) anc[5]);
    // the reference drivers' way (examples/random/main.cu:48-50)
    DTensor<double> d_initState = DTensor<double>::parseFromFile(tree.path() + "initialState" + tree.fpFileExt());
    std::vector<double> initState(tree.numStates());
    d_initState.download(initState);
    std::printf("%zu %zu", tree.numStatesAndInputs(), d_initState.numEl());
    for (double v : initState) std::printf(" %.17g", v);
    std::printf("\n");
    return 0;
}
'''


ACCESSORS = r'''
#include <cmath>
#include <cstdio>
#include "spock_amd.hpp"
using namespace spock_amd;
int main(int argc, char **argv) {
    ScenarioTree<double> tree(argv[1]);
    ProblemData<double> problem(tree);
    CacheBuilder<double> builder(tree, problem);
    std::unique_ptr<Cache<double>> cache = builder.tol(1e-3).maxIters(3000).make_unique();
    std::vector<double> x0 = parseFromFile<double>(tree.path() + "initialState" + tree.fpFileExt());
    const int status = cache->runSpock(x0);
    // the reference's way: device tensors returned by reference, downloaded (src/cache.cuh:371-382)
    DTensor<double> &sol = cache->solution();
    std::vector<double> z;
    sol.download(z);
    std::vector<double> p = cache->solutionPrim(), d = cache->solutionDual();
    DTensor<double> prim(p);
    DTensor<double> dual = cache->getDual(prim);
    std::vector<double> g;
    dual.download(g);
    double diff = 0., scale = 0.;
    for (size_t k = 0; k < d.size(); k++) {
        diff = std::fmax(diff, std::fabs(g[k] - d[k]));
        scale = std::fmax(scale, std::fabs(d[k]));
    }
    std::vector<double> xs = cache->states(), us = cache->inputs();
    std::printf("%d %zu %zu %zu %zu %.3e %zu\n%s\n", status, p.size(), d.size(), z.size(), xs.size(), diff / scale,
                sol.numEl(), toString(status));
    FILE *f = std::fopen(argv[2], "wb");
    std::fwrite(z.data(), sizeof(double), z.size(), f);
    std::fwrite(xs.data(), sizeof(double), xs.size(), f);
    std::fwrite(us.data(), sizeof(double), us.size(), f);
    std::fclose(f);
    return 0;
}
'''


def _compile(src, out, tmp_path):
    path = tmp_path / "prog.cpp"
    path.write_text(src)
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{INC}", f"-I{ROCM}/include", str(path),
           f"-L{LIBDIR}", "-lspock_amd", f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{LIBDIR}",
           f"-Wl,-rpath,{ROCM}/lib", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)


def _problem_with_initial_state(folder):
    prob = configs.generate("small", folder)
    x0 = configs.initial_states(1, 3, 1, "uniform")[0]
    prob.tree.write_to_file_fp(folder, "initialState", x0)  # as examples/random/main.py:128
    return x0


@needs_lib
def test_header_reads_the_problem_folder(tmp_path):
    folder = str(tmp_path / "p")
    x0 = _problem_with_initial_state(folder)
    exe = tmp_path / "reader"
    _compile(READER, exe, tmp_path)
    out = subprocess.run([str(exe), folder], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    assert [int(v) for v in out[0].split()] == [3, 2, 31, 15, 5, 2]
    assert np.array_equal(np.array([float(v) for v in out[1].split()]), x0)
    from spock_amd.api import btio
    cp = btio.read_flat(os.path.join(folder, "conditionalProbabilities_d.bt"))
    anc = btio.read_flat(os.path.join(folder, "ancestors_u.bt"))
    assert float(out[2].split()[0]) == cp[5] and float(out[2].split()[1]) == anc[5]
    ref_way = out[3].split()
    assert [int(v) for v in ref_way[:2]] == [5, 3]
    assert np.array_equal(np.array([float(v) for v in ref_way[2:]]), x0)


@needs_lib
@pytest.mark.gpu
def test_reference_driver_runs_on_the_drop_in(tmp_path):
    folder = str(tmp_path / "data")
    _problem_with_initial_state(folder)
    exe = tmp_path / "random_main"
    src = open(os.path.join(ROOT, "spock.cu_amd", "examples", "random_main.cpp")).read()
    _compile(src, exe, tmp_path)
    run = subprocess.run([str(exe), folder + "/"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(run.stdout)
    assert run.returncode == 0 and "Saved" in run.stdout
    row = (tmp_path / "time.csv").read_text().strip().split(",")
    assert len(row) == 3 and float(row[0]) > 0.0


@needs_lib
@pytest.mark.gpu
def test_drop_in_accessors(tmp_path):
    """make_unique, solutionPrim/solutionDual, states and getDual: at a converged solution the dual
    step from the solution's own primal part returns (nearly) the solution's dual part."""
    folder = str(tmp_path / "data")
    _problem_with_initial_state(folder)
    exe = tmp_path / "accessors"
    _compile(ACCESSORS, exe, tmp_path)
    dump = tmp_path / "solution.bin"
    out = subprocess.run([str(exe), folder + "/", str(dump)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.split("\n")
    head = out[0].split()
    status, n_prim, n_dual, n_iter, n_states, rel, n_el = (int(head[0]), int(head[1]), int(head[2]), int(head[3]),
                                                            int(head[4]), float(head[5]), int(head[6]))
    assert status == 0 and n_prim + n_dual == n_iter == n_el and n_states == 31 * 3
    assert out[1] == "Converged!"  # the reference's string (src/cache.cuh:75-88)
    assert rel < 1e-2
    # states()/inputs(): the solution's x and u divided by the preconditioning scaling
    # (src/cache.cuh:386-412)
    from oracle.oracle import Oracle
    o = Oracle(folder)
    vals = np.fromfile(str(dump), dtype=np.float64)
    z, xs, us = vals[:n_iter], vals[n_iter:n_iter + o.sX], vals[n_iter + o.sX:]
    nx, nu = o.nx, o.info["numInputs"]
    assert us.size == o.sU and o.scaling is not None
    assert np.array_equal(xs, z[o.sU:o.sU + o.sX] / np.tile(o.scaling[:nx], o.sX // nx))
    assert np.array_equal(us, z[:o.sU] / np.tile(o.scaling[nx:nx + nu], o.sU // nu))


@needs_lib
@pytest.mark.gpu
def test_reference_driver_on_the_229_node_problem(tmp_path):
    """The timing driver on the 229-node problem (52,360-element iterate): the drop-in's Cache
    spreads the single instance over 64 cooperating workgroups (grid mode) and converges."""
    folder = str(tmp_path / "data")
    prob = configs.generate("mid", folder)
    x0 = configs.initial_states(2, 50, 1)[0]
    prob.tree.write_to_file_fp(folder, "initialState", x0)
    exe = tmp_path / "random_main"
    src = open(os.path.join(ROOT, "spock.cu_amd", "examples", "random_main.cpp")).read()
    _compile(src, exe, tmp_path)
    run = subprocess.run([str(exe), folder + "/"], cwd=str(tmp_path), capture_output=True, text=True, timeout=170)
    print(run.stdout)
    assert run.returncode == 0 and "Saved" in run.stdout
    # which launch shape the Cache used: the automatic 64-workgroup grid mode, with no fallback
    exe2 = tmp_path / "gridmode"
    _compile(GRIDMODE, exe2, tmp_path)
    out = subprocess.run([str(exe2), folder + "/"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0
    assert out.stdout.split() == ["0", "64", "0"]


GRIDMODE = r'''
#include <cstdio>
#include "spock_amd.hpp"
using namespace spock_amd;
int main(int argc, char **argv) {
    ScenarioTree<double> tree(argv[1]);
    ProblemData<double> problem(tree);
    CacheBuilder<double> builder(tree, problem);
    Cache<double> cache = builder.tol(1e-3).maxIters(400).build();
    std::vector<double> x0 = parseFromFile<double>(tree.path() + "initialState" + tree.fpFileExt());
    const int status = cache.runSpock(x0);
    std::printf("%d %d %d\n", status == converged || status == outOfIter ? 0 : 1, cache.workgroups(),
                (int) cache.gridFallback());
    return 0;
}
'''


@needs_lib
def test_anderson_buffer_driver_builds(tmp_path):
    """spock.cu_amd/examples/anderson_buffer_main.cpp (the reference's examples/andersonBuffer
    study on the drop-in header) compiles and links against the library; running it needs a GPU."""
    src = open(os.path.join(ROOT, "spock.cu_amd", "examples", "anderson_buffer_main.cpp")).read()
    exe = tmp_path / "anderson_buffer_main"
    _compile(src, exe, tmp_path)
    assert exe.exists()
