"""Debug traces of the HIP engine (spk_batch_set_debug / spk_batch_debug_records: the reference's
enableDebug caches, src/cache.cuh:947-959, and its isFinite check, :912) against the oracle's.

The records of a solve {idx, callsL, err0, err1, err2, err3} must equal the oracle's: the check
indices and the counts of L/L* calls exactly, the errors to 1e-8 relative in FP64 (one workgroup
and grid mode), and to 1e-3 over the first checks in FP32.  err0's running vector carries from
solve to solve and spk_batch_reset zeroes it, as in the reference.  A NaN initial state stops
the solve at iteration 0 with status 3 and the index of the first non-finite candidate entry."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from spock_amd import runtime as rt
from spock_amd.api import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return rt.Context(0)


def _close(a, b, rtol):
    """Relative to each column's largest magnitude (errors near zero are compared at that scale)."""
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return np.max(np.abs(a - b) / scale)


def _run(ctx, d, x0, dtype, cap, workgroups=1, **kw):
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p], dtype)
    if workgroups > 1:
        b.set_workgroups(workgroups)
    b.set_debug(cap)
    b.set_x0([x0])
    # 256 threads, as the fixed-iteration parity tests: the iteration windows below are the ones
    # FP64 rounding with that summation order keeps within the bounds
    b.solve(rt.options(block_size=256, **kw))
    return b


# outer iterations compared per family: as long as the fixed-iteration iterate parity runs for it
# (at tol 1e-30 the FP64 arithmetics of host and device later take different SuperMann branches)
ITERS = {"small": 75, "mid": 75, "poly": 75, "pwi": 50, "linear": 25}


@pytest.mark.parametrize("name", list(ITERS))
def test_spock_records_match_the_oracle(ctx, problem_dirs, name):
    d = problem_dirs[name]
    o = Oracle(d)
    x0 = configs.initial_states(21, o.nx, 1, "normal")[0]
    n = ITERS[name]
    ref = o.solve_debug(x0, np.zeros(o.n_prim), tol_abs=1e-30, max_iters=n)
    b = _run(ctx, d, x0, rt.F64, 16, tol_abs=1e-30, max_iters=n)
    res, _ = b.results()
    rec, bad = b.debug_records(0)
    assert bad == -1 and res[0]["status"] == 1
    assert rec.shape == ref["records"].shape == (n // 25 + 1, 6)
    np.testing.assert_array_equal(rec[:, :2], ref["records"][:, :2])
    assert _close(rec[:, 2:], ref["records"][:, 2:], 1e-8) < 1e-8
    assert rec[-1, 1] == 2 * res[0]["num_T"]


def test_cp_records_every_iteration(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = configs.initial_states(22, o.nx, 1, "uniform")[0]
    ref = o.solve_debug(x0, np.zeros(o.n_prim), method="cp", tol_abs=1e-30, max_iters=40, check_period=1)
    b = _run(ctx, d, x0, rt.F64, 64, tol_abs=1e-30, max_iters=40, check_period=1, method=rt.CP)
    rec, _ = b.debug_records(0)
    assert rec.shape == (41, 6)
    np.testing.assert_array_equal(rec[:, :2], ref["records"][:, :2])
    assert _close(rec[:, 2:], ref["records"][:, 2:], 1e-8) < 1e-8


def test_grid_mode_records(ctx, problem_dirs):
    d = problem_dirs["mid"]
    o = Oracle(d)
    x0 = configs.initial_states(23, o.nx, 1, "normal")[0]
    ref = o.solve_debug(x0, np.zeros(o.n_prim), tol_abs=1e-30, max_iters=75)
    b = _run(ctx, d, x0, rt.F64, 8, workgroups=16, tol_abs=1e-30, max_iters=75)
    rec, _ = b.debug_records(0)
    np.testing.assert_array_equal(rec[:, :2], ref["records"][:, :2])
    assert _close(rec[:, 2:], ref["records"][:, 2:], 1e-8) < 1e-8


def test_fp32_records(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = configs.initial_states(24, o.nx, 1, "uniform")[0]
    ref = o.solve_debug(x0, np.zeros(o.n_prim), tol_abs=1e-30, max_iters=25)
    b = _run(ctx, d, x0, rt.F32, 4, tol_abs=1e-30, max_iters=25)
    rec, _ = b.debug_records(0)
    np.testing.assert_array_equal(rec[:, 0], ref["records"][:, 0])
    assert _close(rec[:, 2:], ref["records"][:, 2:], 1e-3) < 1e-3


def test_err0_carry_and_reset(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = configs.initial_states(25, o.nx, 1, "uniform")[0]
    kw = dict(method="cp", tol_abs=1e-30, max_iters=6, check_period=1)
    acc = np.zeros(o.n_prim)
    o.solve_debug(x0, acc, **kw)
    second = o.solve_debug(x0, acc, **kw)["records"]
    fresh = o.solve_debug(x0, np.zeros(o.n_prim), **kw)["records"]
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p], rt.F64)
    b.set_debug(16)
    b.set_x0([x0])
    opts = rt.options(tol_abs=1e-30, max_iters=6, check_period=1, method=rt.CP, block_size=256)
    b.solve(opts)
    b.results()
    b.reset()  # zeroes err0's vector, and the iterate: the next solve is cold again
    b.solve(opts)
    b.results()
    np.testing.assert_allclose(b.debug_records(0)[0][:, 2], fresh[:, 2], rtol=1e-9, atol=1e-9 * fresh[:, 2].max())
    b.reset()
    b.solve(opts)
    b.results()
    b.set_iterates(np.zeros(p.n_iter))  # a cold start without a reset keeps err0's vector
    b.solve(opts)
    b.results()
    # (the alternating sum nearly cancels at some checks: compared at the scale of the largest)
    np.testing.assert_allclose(b.debug_records(0)[0][:, 2], second[:, 2], rtol=1e-9, atol=1e-9 * second[:, 2].max())


def test_nonfinite_initial_state_stops_the_solve(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = configs.initial_states(26, o.nx, 1, "uniform")[0]
    x0[1] = np.nan
    ref = o.solve_debug(x0, np.zeros(o.n_prim), tol_abs=1e-30, max_iters=100)
    b = _run(ctx, d, x0, rt.F64, 8, tol_abs=1e-30, max_iters=100)
    res, _ = b.results()
    rec, bad = b.debug_records(0)
    assert res[0]["status"] == 3 and res[0]["iters"] == 0
    assert bad == ref["nonfinite"] >= 0
    assert rec.shape[0] == 0


def test_shared_batches_refuse_debug(ctx, problem_dirs):
    p = rt.Problem(ctx, problem_dirs["small"])
    b = rt.Batch(ctx, [p] * 4, rt.F32, shared=True)
    with pytest.raises(RuntimeError, match="per-instance"):
        b.set_debug(8)
