"""Continuous batching of shared batches (spk_batch_refill*, spk_mlp_predict_lanes): new solves
started in the slots of a running chunked solve, on the GPU, against solves of the same initial
states in fresh batches.

A slot's arithmetic does not depend on the other slots of its workgroup, but a refilled slot joins
the workgroup's Anderson ring at whatever position it has reached, so its least-squares system is a
symmetric permutation of a fresh solve's and agrees to rounding, not bit for bit.  Fixed-iteration
solves are therefore compared to 1e-10 relative with identical SuperMann branch counts, on both
engines (tile, schedule 16; lane, schedule 0)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle.oracle import Oracle
from spock_amd import runtime as rt
from spock_amd.api import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return rt.Context(0)


def fresh(ctx, p, x0s, iters, engine, warm=None):
    b = rt.Batch(ctx, [p] * len(x0s), rt.F64, shared=True, engine=engine)
    b.set_x0(list(x0s))
    if warm is not None:
        b.set_iterates(np.concatenate(warm))
    b.solve(rt.options(tol_abs=1e-30, max_iters=iters), warm=warm is not None)
    res, _ = b.results()
    return res, b.iterates().reshape(len(x0s), -1)


def run_out(b, slots=None):
    """Resume a chunked solve until none of `slots` (default: all) is running; the final results."""
    for _ in range(1000):
        res, _ = b.results()
        if all(res[k]["status"] != -1 for k in (slots if slots is not None else range(len(res)))):
            return res
        b.resume()
    raise AssertionError("the solve did not end")


def counts(r):
    return (r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"])


@pytest.mark.parametrize("engine", [17, 16, 0])
def test_refilled_slots_match_fresh_solves(ctx, problem_dirs, engine):
    """20 slots solve 50 iterations (two chunks of 25); 12 of them are refilled with new initial
    states (cold) and, from the accepted iterates of another batch, 5 more with warm starts; after
    two more chunks every refilled slot equals a fresh 50-iteration solve of its new start, and the
    three untouched slots keep their first results."""
    d = problem_dirs["small"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    count, iters = 20, 50
    x0a = configs.initial_states(1, o.nx, count, "uniform")
    x0b = configs.initial_states(17, o.nx, count, "uniform")
    b = rt.Batch(ctx, [p] * count, rt.F64, shared=True, engine=engine)
    b.set_x0(list(x0a))
    opts = rt.options(tol_abs=1e-30, max_iters=iters, chunk_iters=25)
    b.solve(opts)
    res, _ = b.results()
    assert all(r["status"] == -1 for r in res)
    res_a = run_out(b)
    z_a = b.iterates().reshape(count, -1)
    assert all(r["status"] == 1 and r["iters"] == iters for r in res_a)
    # warm starts: accepted iterates of a second batch that ran 30 iterations from x0b
    src = rt.Batch(ctx, [p] * count, rt.F64, shared=True, engine=engine)
    src.set_x0(list(x0b))
    src.solve(rt.options(tol_abs=1e-30, max_iters=30))
    src.results()
    z_src = src.iterates().reshape(count, -1)
    cold = list(range(0, 12))
    warm = list(range(12, 17))
    src_lanes = [3, 9, 4, 0, 11]
    b.refill(cold, x0b[cold])
    b.refill_from(warm, x0b[warm], src, src_lanes)
    b.resume()
    res, _ = b.results()
    assert all(res[k]["status"] == -1 for k in cold + warm)
    res = run_out(b)
    z = b.iterates().reshape(count, -1)
    ref_cold, zc = fresh(ctx, p, x0b[cold], iters, engine)
    for i, k in enumerate(cold):
        assert counts(res[k]) == counts(ref_cold[i]), k
        assert rel_err(z[k], zc[i]) < 1e-10
    ref_warm, zw = fresh(ctx, p, x0b[warm], iters, engine, warm=[z_src[s] for s in src_lanes])
    for i, k in enumerate(warm):
        assert counts(res[k]) == counts(ref_warm[i]), k
        assert rel_err(z[k], zw[i]) < 1e-10
    for k in range(17, count):  # untouched slots
        assert counts(res[k]) == counts(res_a[k])
        assert np.array_equal(z[k], z_a[k])


def test_refill_rejects_bad_requests(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p] * 4, rt.F64, shared=True)
    x0 = configs.initial_states(1, o.nx, 4, "uniform")
    b.set_x0(list(x0))
    with pytest.raises(rt.SpockError, match="chunked solve"):
        b.refill([0], x0[:1])
    b.solve(rt.options(tol_abs=1e-30, max_iters=25, chunk_iters=25))
    b.results()
    with pytest.raises(rt.SpockError, match="out of range"):
        b.refill([4], x0[:1])
    with pytest.raises(rt.SpockError, match="twice"):
        b.refill([1, 1], x0[:2])


@pytest.mark.parametrize("engine", [17, 0])
def test_empty_pool_runs_only_filled_slots(ctx, problem_dirs, engine):
    """spk_batch_pool_open: every slot idle (SPK_NOT_RUN) until a refill; 9 slots spread over the
    workgroups get cold solves of 50 iterations and match the oracle's solves of the same initial
    states (branch counts exactly, iterates to 1e-9, as the fixed-iteration parity tests); the idle
    slots stay idle.  (A fresh GPU batch of the same states is no better a reference: its lanes
    share a workgroup with different neighbours, and a SuperMann decision within rounding of its
    threshold may go the other way there.)"""
    d = problem_dirs["small"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    count, iters = 40, 50
    b = rt.Batch(ctx, [p] * count, rt.F64, shared=True, engine=engine)
    b.pool_open(rt.options(tol_abs=1e-30, max_iters=iters, chunk_iters=20))
    slots = [0, 3, 15, 16, 17, 31, 32, 38, 39]
    x0 = configs.initial_states(23, o.nx, len(slots), "uniform")
    b.refill(slots, x0)
    b.resume()
    res = run_out(b, slots)
    ref = [o.solve(x, tol_abs=1e-30, max_iters=iters, ls="gram", adjoint_dots=True) for x in x0]
    z = b.iterates().reshape(count, -1)

    def oracle_counts(r):
        return (r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"])

    for i, k in enumerate(slots):
        assert counts(res[k]) == oracle_counts(ref[i]), k
        assert rel_err(z[k], ref[i]["z"]) < 1e-9
    for k in set(range(count)) - set(slots):
        assert res[k]["status"] == -1 and res[k]["iters"] == 0
    # a second round in other slots, with a shorter chunk
    b.set_chunk(10)
    b.refill([1, 2], x0[:2])
    b.resume()
    res = run_out(b, [1, 2])
    z = b.iterates().reshape(count, -1)
    for i, k in enumerate([1, 2]):
        assert counts(res[k]) == oracle_counts(ref[i])
        assert rel_err(z[k], ref[i]["z"]) < 1e-9
