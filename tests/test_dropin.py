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


DEBUG = r'''
#include <cstdio>
#include "spock_amd.hpp"
using namespace spock_amd;
int main(int argc, char **argv) {
    ScenarioTree<double> tree(argv[1]);
    ProblemData<double> problem(tree);
    CacheBuilder<double> builder(tree, problem);
    Cache<double> cache = builder.tol(1e-30).enableDebug().build();  // maxIters defaults to 1000
    std::vector<double> x0 = parseFromFile<double>(tree.path() + "initialState" + tree.fpFileExt());
    const int status = cache.runSpock(x0);
    std::printf("status %d\n", status);
    std::vector<double> bad(x0);
    bad[1] = 0.0 / 0.0;
    try {
        cache.runSpock(bad);
        std::printf("no throw\n");
    } catch (const std::invalid_argument &e) {
        std::printf("threw %s", e.what());
    }
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
    spreads the single instance over 32 cooperating workgroups (grid mode) and converges."""
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
    # which launch shape the Cache used: the automatic 32-workgroup grid mode, with no fallback
    exe2 = tmp_path / "gridmode"
    _compile(GRIDMODE, exe2, tmp_path)
    out = subprocess.run([str(exe2), folder + "/"], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0
    assert out.stdout.split() == ["0", "32", "0"]


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


@needs_lib
@pytest.mark.gpu
def test_anderson_buffer_driver_runs(tmp_path):
    """The Anderson-buffer study on a problem of the reference's random generator (seeded; its
    first size interval, 1,000 to 10,000 variables): buffers 3, 5 and 7 all converge and the
    time.csv line carries the sizes and three positive times."""
    folder = str(tmp_path / "data")
    configs.generate("random", folder, seed=1, lo=1000, hi=10000)
    src = open(os.path.join(ROOT, "spock.cu_amd", "examples", "anderson_buffer_main.cpp")).read()
    exe = tmp_path / "anderson_buffer_main"
    _compile(src, exe, tmp_path)
    run = subprocess.run([str(exe), folder + "/"], cwd=str(tmp_path), capture_output=True, text=True, timeout=170)
    print(run.stdout)
    assert run.returncode == 0 and "SPOCK failed" not in run.stdout
    assert run.stdout.count("Run (1), buff (") == 3
    row = [float(v) for v in (tmp_path / "time.csv").read_text().strip().split(",")]
    from oracle.oracle import read_info
    info = read_info(folder)
    assert row[:3] == [info["numNodes"] * (info["numStates"] + info["numInputs"]), info["numNodes"],
                       info["numStates"] + info["numInputs"]]
    assert all(t > 0.0 for t in row[3:])


@needs_lib
@pytest.mark.gpu
def test_server_air_condition_driver_runs(tmp_path):
    """The server air-conditioning example (examples/serverAirCondition/main.cu and main.py,
    horizon 5): three warm-up and two timed runs, each converged, and the average appended."""
    folder = str(tmp_path / "data")
    configs.generate("server", folder, horizon=5)
    src = open(os.path.join(ROOT, "spock.cu_amd", "examples", "server_air_main.cpp")).read()
    exe = tmp_path / "server_air_main"
    _compile(src, exe, tmp_path)
    run = subprocess.run([str(exe), folder + "/"], cwd=str(tmp_path), capture_output=True, text=True, timeout=170)
    print(run.stdout)
    assert run.returncode == 0 and "SPOCK failed" not in run.stdout
    assert all(f"Run ({i}) : " in run.stdout for i in range(5))
    assert float((tmp_path / "time.csv").read_text().strip()) > 0.0


def test_server_air_condition_family(tmp_path):
    """The generator restates the example's model: the oracle solves it at tol(1e-3) from the
    example's initial state 0.1 * ones, and the solution keeps |x| <= 1 and |u| <= 1.5."""
    from oracle.oracle import Oracle
    from spock_amd.api import btio
    folder = str(tmp_path / "server")
    configs.generate("server", folder, horizon=4)
    o = Oracle(folder)
    x0 = btio.read_flat(os.path.join(folder, "initialState_d.bt"))
    assert np.array_equal(x0, 0.1 * np.ones(50))
    r = o.solve(x0, tol_abs=1e-3, tol_rel=1e-3, max_iters=2000)
    assert r["status"] == 0
    nx, nu = o.nx, o.info["numInputs"]
    x = r["z"][o.sU:o.sU + o.sX] / np.tile(o.scaling[:nx], o.sX // nx)
    u = r["z"][:o.sU] / np.tile(o.scaling[nx:nx + nu], o.sU // nu)
    assert np.abs(x).max() <= 1.0 + 1e-2 and np.abs(u).max() <= 1.5 + 1e-2
    assert o.info["numNodes"] == 31 and o.info["numEvents"] == 2


@needs_lib
@pytest.mark.gpu
def test_enable_debug_writes_the_reference_cache_json(tmp_path):
    """enableDebug (src/cache.cuh:1204-1209): maxIters 1,000 by default, the console summary with
    the K counts, log/cacheSp.json with the reference's fields (src/cache.cuh:1274-1311) whose
    callsL and err0..err3 are the oracle's debug caches, and isFinite's exception on a NaN state."""
    import json
    from oracle.oracle import Oracle
    folder = str(tmp_path / "data")
    x0 = _problem_with_initial_state(folder)
    exe = tmp_path / "debug"
    _compile(DEBUG, exe, tmp_path)
    run = subprocess.run([str(exe), folder + "/"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr)
    assert run.returncode == 0
    assert "status 1" in run.stdout and "Max iterations (1000) reached [K0: " in run.stdout
    assert "threw [isFinite] DTensor has a non-finite entry at (" in run.stdout
    doc = json.loads((tmp_path / "log" / "cacheSp.json").read_text())
    assert doc["maxIters"] == doc["sizeCache"] == 1000
    assert (doc["horizon"], doc["branchFactor"], doc["numEvents"], doc["numStates"], doc["numInputs"]) == \
        (4, 5, 2, 3, 2)  # branching never stops: find() returns end(), numStages
    # entries at the error checks (every 25 iterations; the check at iteration 1000 itself falls
    # outside the caches), zero elsewhere
    checks = np.arange(0, 1000, 25)
    mask = np.zeros(1000, dtype=bool)
    mask[checks] = True
    for key in ("callsL", "err0", "err1", "err2", "err3"):
        arr = np.array(doc[key], dtype=np.float64)
        assert arr.size == 1000 and np.all(arr[~mask] == 0.0) and np.all(arr[mask] > 0.0), key
    # values: the oracle's caches over the first 150 iterations (at tol 1e-30 the two arithmetics
    # later take different SuperMann branches in the noise floor, and the traces part)
    o = Oracle(folder)
    rec = o.solve_debug(x0, np.zeros(o.n_prim), tol_abs=1e-30, max_iters=150)["records"]
    idx = rec[:, 0].astype(int)
    assert np.array_equal(np.array(doc["callsL"], dtype=np.float64)[idx], rec[:, 1])
    for col, key in ((2, "err0"), (3, "err1"), (4, "err2"), (5, "err3")):
        v = np.array(doc[key], dtype=np.float64)[idx]
        assert np.max(np.abs(v - rec[:, col])) <= 1e-8 * np.max(np.abs(rec[:, col])), key
    assert len(doc["states"]) == o.sX and len(doc["inputs"]) == o.sU
