"""The problem-folder loader (spock.cu_amd/csrc/loader.cpp, host C++) built on its own with
AddressSanitizer and UndefinedBehaviorSanitizer: it reads every generated problem family, and
truncated or inconsistent files end in an error message, never a crash."""
import os
import json
import shutil
import subprocess

import numpy as np
import pytest

from spock_amd.api import btio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "spock.cu_amd", "csrc")

MAIN = r'''
#include <cstdio>
#include <exception>
#include "loader.hpp"
int main(int argc, char **argv) {
    try {
        spk::HostProblem p = spk::loadProblem(argv[1]);
        std::printf("%d %d %d %d %d %ld %ld %d\n", p.N, p.nl, p.nx, p.nu, p.E, p.sPrim, p.sDual, p.affine ? 1 : 0);
        return 0;
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
}
'''


@pytest.fixture(scope="module")
def loader_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("loader")
    (d / "main.cpp").write_text(MAIN)
    exe = d / "load"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{CSRC}", str(d / "main.cpp"), os.path.join(CSRC, "loader.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    return str(exe)


def run(exe, folder):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe, folder], capture_output=True, text=True, timeout=120, env=env)


@pytest.mark.parametrize("name", ["fixture", "small", "mid", "poly", "nocon", "affine", "pwi", "linear"])
def test_loader_reads_every_family(loader_exe, problem_dirs, name):
    r = run(loader_exe, problem_dirs[name])
    assert r.returncode == 0, r.stdout + r.stderr
    N, nl, nx, nu, E, sp, sd, aff = (int(v) for v in r.stdout.split())

    info = json.load(open(os.path.join(problem_dirs[name], "data.json")))
    assert (N, nl, nx, nu, E) == (info["numNodes"], info["numNonleafNodes"], info["numStates"],
                                  info["numInputs"], info["numEvents"])
    assert aff == int(info["dynamics"]["type"] == "affine")


def _copy(problem_dirs, tmp_path, name="small"):
    dst = tmp_path / "p"
    shutil.copytree(problem_dirs[name], dst)
    return dst


def test_truncated_tensor_is_an_error(loader_exe, problem_dirs, tmp_path):
    d = _copy(problem_dirs, tmp_path)
    f = d / "dynamics_AB_d.bt"
    raw = f.read_bytes()
    f.write_bytes(raw[:len(raw) // 2])
    r = run(loader_exe, str(d))
    assert r.returncode == 2 and "error" in r.stdout, r.stdout + r.stderr


def test_header_only_tensor_is_an_error(loader_exe, problem_dirs, tmp_path):
    d = _copy(problem_dirs, tmp_path)
    (d / "risk_b_d.bt").write_bytes((d / "risk_b_d.bt").read_bytes()[:10])
    r = run(loader_exe, str(d))
    assert r.returncode == 2 and "error" in r.stdout, r.stdout + r.stderr


def test_wrong_sizes_are_an_error(loader_exe, problem_dirs, tmp_path):
    d = _copy(problem_dirs, tmp_path)
    btio.write_bt(str(d / "dynamics_K_d.bt"), np.zeros((2, 2, 3)), "d")
    r = run(loader_exe, str(d))
    assert r.returncode == 2 and "error" in r.stdout, r.stdout + r.stderr


def test_out_of_range_tree_indices_are_an_error(loader_exe, problem_dirs, tmp_path):
    d = _copy(problem_dirs, tmp_path)
    anc = btio.read_flat(str(d / "ancestors_u.bt")).astype(np.int64)
    anc[5] = 10 ** 6
    btio.write_bt(str(d / "ancestors_u.bt"), anc, "u")
    r = run(loader_exe, str(d))
    assert r.returncode == 2 and "error" in r.stdout, r.stdout + r.stderr


def test_missing_json_is_an_error(loader_exe, tmp_path):
    r = run(loader_exe, str(tmp_path))
    assert r.returncode == 2 and "error" in r.stdout, r.stdout + r.stderr


def test_unsupported_constraint_kind_is_named(loader_exe, problem_dirs, tmp_path):
    """A constraint type the reference does not have (src/problem.cuh:74-92) is refused by name."""
    d = _copy(problem_dirs, tmp_path)
    info = json.load(open(d / "data.json"))
    info["constraint"]["nonleaf"] = "ellipsoid"
    json.dump(info, open(d / "data.json", "w"))
    r = run(loader_exe, str(d))
    assert r.returncode == 2 and "ellipsoid" in r.stdout and "not supported" in r.stdout, r.stdout


REORDER = r'''
#include <cstdio>
#include <vector>
#include "loader.hpp"
int main(int argc, char **argv) {
    spk::HostProblem p = spk::loadProblem(argv[1]);
    std::vector<double> z(p.sIter);
    FILE *f = std::fopen(argv[2], "rb");
    if (std::fread(z.data(), sizeof(double), z.size(), f) != z.size()) return 3;
    std::fclose(f);
    spk::reorderIdentityBlocks(p, z.data(), true);
    f = std::fopen(argv[3], "wb");
    std::fwrite(z.data(), sizeof(double), z.size(), f);
    std::fclose(f);
    spk::reorderIdentityBlocks(p, z.data(), false);
    f = std::fopen(argv[4], "wb");
    std::fwrite(z.data(), sizeof(double), z.size(), f);
    std::fclose(f);
    std::printf("%d %d %d %d\n", p.con3, p.id3, p.con5, p.id5);
    return 0;
}
'''


def test_polyhedron_with_identity_dual_order(problem_dirs, tmp_path):
    """polyhedronWithIdentity loads as the polyhedron [I; Gamma] (con3 = nx + nu + rows of Gamma),
    and an iterate's constraint blocks move from the reference's order (identity parts of every
    node, then the Gamma parts; src/constraints.cuh:340-373) to one block per node and back.  The
    expected engine order is written out here index by index from the reference's layout."""
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    (tmp_path / "main.cpp").write_text(REORDER)
    exe = tmp_path / "reorder"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{CSRC}", str(tmp_path / "main.cpp"), os.path.join(CSRC, "loader.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    from oracle.oracle import Oracle
    d = problem_dirs["pwi"]
    o = Oracle(d)
    info = json.load(open(os.path.join(d, "data.json")))
    nx, nu, nl, N = info["numStates"], info["numInputs"], info["numNonleafNodes"], info["numNodes"]
    nlf = N - nl
    pg3 = btio.read_flat(os.path.join(d, "nonleafConstraintGLB_d.bt")).size
    pg5 = btio.read_flat(os.path.join(d, "leafConstraintGLB_d.bt")).size
    z = np.random.default_rng(3).standard_normal(o.n_iter)
    z.tofile(tmp_path / "in.bin")
    r = subprocess.run([str(exe), d, str(tmp_path / "in.bin"), str(tmp_path / "eng.bin"), str(tmp_path / "back.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    con3, id3, con5, id5 = (int(v) for v in r.stdout.split())
    assert (con3, id3, con5, id5) == (nx + nu + pg3, nx + nu, nx + pg5, nx)
    eng = np.fromfile(tmp_path / "eng.bin")
    back = np.fromfile(tmp_path / "back.bin")
    assert np.array_equal(back, z)
    expect = z.copy()
    o3 = o.n_prim + o.sI + o.sII
    for i in range(nl):
        expect[o3 + i * con3:o3 + i * con3 + id3] = z[o3 + i * id3:o3 + (i + 1) * id3]
        g0 = o3 + nl * id3 + i * pg3
        expect[o3 + i * con3 + id3:o3 + (i + 1) * con3] = z[g0:g0 + pg3]
    o5 = o3 + o.sIII + o.sIV
    for l in range(nlf):
        expect[o5 + l * con5:o5 + l * con5 + id5] = z[o5 + l * id5:o5 + (l + 1) * id5]
        g0 = o5 + nlf * id5 + l * pg5
        expect[o5 + l * con5 + id5:o5 + (l + 1) * con5] = z[g0:g0 + pg5]
    assert np.array_equal(eng, expect)
