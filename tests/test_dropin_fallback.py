"""The drop-in Cache's grid-mode fallback (include/spock_amd.hpp), without a GPU: the header is
compiled against a stub of the C ABI whose solve refuses any launch of more than one workgroup,
as a refused cooperative launch does (SPK_ERR_COOPERATIVE_LAUNCH).  The size-based grid choice
must then solve with one workgroup and say so (gridFallback), any other error must not be retried,
an explicitly requested workgroup count must report the refusal, and iterates outside the size
range grid mode has run on stay on one workgroup."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STUB = r'''
#include <cstdio>
#include <cstring>
#include "spock_amd.h"
struct spk_ctx { char err[256]; };
struct spk_problem { long sizes[16]; };
struct spk_batch { spk_ctx *ctx; int wgs; };
static spk_ctx g_ctx;
static spk_problem g_prob;
static spk_batch g_batch;
long g_prim = 0;  // set by the driver
extern "C" {
int spk_ctx_create(int, spk_ctx **out) { *out = &g_ctx; return 0; }
int spk_ctx_destroy(spk_ctx *) { return 0; }
const char *spk_last_error(spk_ctx *c) { return c->err; }
int spk_problem_load(spk_ctx *, const char *, spk_problem **out) {
    std::memset(&g_prob, 0, sizeof(g_prob));
    g_prob.sizes[0] = 31; g_prob.sizes[1] = 15; g_prob.sizes[4] = 3; g_prob.sizes[5] = 2;
    g_prob.sizes[6] = g_prim; g_prob.sizes[7] = g_prim;
    *out = &g_prob;
    return 0;
}
int spk_problem_destroy(spk_problem *) { return 0; }
int spk_problem_sizes(const spk_problem *p, long *out) { std::memcpy(out, p->sizes, sizeof(p->sizes)); return 0; }
int spk_problem_scaling(const spk_problem *, double *out) { for (int k = 0; k < 5; k++) out[k] = 1.0; return 0; }
double spk_problem_step_size(const spk_problem *) { return 0.5; }
int spk_batch_create(spk_ctx *c, int, spk_problem *const *, int, spk_batch **out) {
    g_batch.ctx = c; g_batch.wgs = 0; *out = &g_batch; return 0;
}
int spk_batch_destroy(spk_batch *) { return 0; }
int spk_batch_set_workgroups(spk_batch *b, int n) { b->wgs = n; std::printf("workgroups %d\n", n); return 0; }
int spk_batch_set_x0(spk_batch *, const double *) { return 0; }
int spk_batch_set_iterate(spk_batch *, int, const double *) { return 0; }
int g_other_error = 0;  // set by the driver: fail every solve with a plain error
int spk_batch_solve(spk_batch *b, const spk_options *, int) {
    std::printf("solve %d\n", b->wgs);
    if (g_other_error) {
        std::snprintf(b->ctx->err, sizeof(b->ctx->err), "hipMalloc: out of memory");
        return SPK_ERR;
    }
    if (b->wgs > 1) {
        std::snprintf(b->ctx->err, sizeof(b->ctx->err), "cooperative launch: the operation cannot be performed in the present state");
        return SPK_ERR_COOPERATIVE_LAUNCH;
    }
    return 0;
}
int spk_batch_results(spk_batch *, spk_result *r, double *) { std::memset(r, 0, sizeof(*r)); r->iters = 7; return 0; }
int spk_batch_get_iterate(spk_batch *, int, double *) { return 0; }
int spk_batch_reset(spk_batch *) { return 0; }
int spk_batch_debug_op(spk_batch *, int, const double *, double *, double *) { return 0; }
int spk_batch_set_debug(spk_batch *, long) { return 0; }
int spk_batch_debug_records(spk_batch *, int, long, double *, long *n, long *bad) { *n = 0; *bad = -1; return 0; }
void spk_default_options(spk_options *o) { std::memset(o, 0, sizeof(*o)); }
}
'''

MAIN = r'''
#include <cstdio>
#include <cstdlib>
#include "spock_amd.hpp"
extern long g_prim;
extern int g_other_error;
using namespace spock_amd;
int main(int argc, char **argv) {
    g_prim = std::atol(argv[2]);
    const int wgs = std::atoi(argv[3]);
    g_other_error = argc > 4 ? std::atoi(argv[4]) : 0;
    ScenarioTree<double> tree(argv[1]);
    ProblemData<double> data(tree);
    CacheBuilder<double> builder(tree, data);
    Cache<double> cache = builder.workgroups(wgs).build();
    std::vector<double> x0(3, 0.1);
    try {
        for (int k = 0; k < 2; k++) {
            const int status = cache.runSpock(x0);
            std::printf("status %d iters %zu workgroups %d fallback %d\n", status, cache.solveIter(),
                        cache.workgroups(), (int) cache.gridFallback());
        }
    } catch (const std::invalid_argument &e) {
        std::printf("error %s\n", e.what());
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("fallback")
    (d / "stub.cpp").write_text(STUB)
    (d / "main.cpp").write_text(MAIN)
    (d / "data").mkdir()
    (d / "data" / "data.json").write_text('{"numStates": 3, "numInputs": 2, "numNodes": 31}')
    exe = d / "driver"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", str(d / "main.cpp"),
                    str(d / "stub.cpp"), "-o", str(exe)], check=True, capture_output=True, timeout=300)
    return str(exe), str(d / "data")


def run(driver, prim, wgs, other_error=0):
    exe, folder = driver
    out = subprocess.run([exe, folder, str(prim), str(wgs), str(other_error)], check=True, capture_output=True,
                         text=True, timeout=60)
    return out.stdout.strip().split("\n")


def test_size_based_grid_falls_back_to_one_workgroup(driver):
    # an iterate of 60,000 elements: the Cache picks 32 workgroups, the launch is refused, the
    # solve is repeated with one workgroup, and later solves use one workgroup directly
    assert run(driver, 30000, 0) == ["workgroups 32", "solve 32", "workgroups 1", "solve 1",
                                     "status 0 iters 7 workgroups 1 fallback 1",
                                     "solve 1", "status 0 iters 7 workgroups 1 fallback 1"]


def test_small_iterates_use_one_workgroup(driver):
    assert run(driver, 1000, 0) == ["workgroups 1", "solve 1", "status 0 iters 7 workgroups 1 fallback 0",
                                    "solve 1", "status 0 iters 7 workgroups 1 fallback 0"]


def test_large_iterates_stay_on_one_workgroup(driver):
    # 2.9M elements (the 5,461-node tree): outside the range grid mode has run on, so no automatic grid
    assert run(driver, 1500000, 0) == ["workgroups 1", "solve 1", "status 0 iters 7 workgroups 1 fallback 0",
                                       "solve 1", "status 0 iters 7 workgroups 1 fallback 0"]


def test_other_errors_are_not_retried(driver):
    lines = run(driver, 30000, 0, other_error=1)
    assert lines == ["workgroups 32", "solve 32", "error [runSpock] hipMalloc: out of memory"]


def test_requested_workgroups_report_the_refusal(driver):
    lines = run(driver, 30000, 64)
    assert lines[:2] == ["workgroups 64", "solve 64"]
    assert lines[2].startswith("error [runSpock] cooperative launch") and len(lines) == 3
