"""spk_batch_solve_noalloc: the allocation-free variant of spk_batch_solve, called through the C ABI
on the GPU, against spk_batch_solve on identical batches.  Both run the same kernel on the same
inputs, so iterates and SuperMann counts must agree bit for bit.  Two calls in a row share the
device's scratch, the second with a smaller input (one instance of the 31-node problem after three
of the 229-node one), so the second also runs over what the first left in the scratch."""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import Oracle
from spock_amd import runtime as rt
from spock_amd.api import configs

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope="module")
def ctx():
    return rt.Context(0)


def counts(r):
    return (r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"], r["num_T"], r["num_ell"])


def pair(ctx, folder, x0s, dtype, opts):
    """The same batch solved by spk_batch_solve and by spk_batch_solve_noalloc (the latter called
    directly on the library handle)."""
    p = rt.Problem(ctx, folder)
    out = []
    for noalloc in (False, True):
        b = rt.Batch(ctx, [p] * len(x0s), dtype)
        b.set_x0(list(x0s))
        if noalloc:
            ctx.check(ctx.lib.spk_batch_solve_noalloc(b.h, C.byref(opts), 0), "solve_noalloc")
        else:
            b.solve(opts)
        res, _ = b.results()
        out.append((res, b.iterates()))
    return out


def test_noalloc_matches_solve_two_calls(ctx, problem_dirs):
    opts = rt.options(tol_abs=1e-30, max_iters=150)
    # first call: three instances of the 229-node tree, FP64
    d = problem_dirs["mid"]
    x0 = configs.initial_states(5, Oracle(d).nx, 3, "normal")
    (ra, za), (rb, zb) = pair(ctx, d, x0, rt.F64, opts)
    assert [counts(r) for r in ra] == [counts(r) for r in rb]
    assert all(r["iters"] == 150 for r in rb)
    assert np.array_equal(za, zb)
    # second call, smaller: one instance of the 31-node tree, FP64 then FP32
    d = problem_dirs["small"]
    x0 = configs.initial_states(6, Oracle(d).nx, 1, "uniform")
    for dtype in (rt.F64, rt.F32):
        (ra, za), (rb, zb) = pair(ctx, d, x0, dtype, rt.options(tol_abs=1e-30, max_iters=50))
        assert counts(ra[0]) == counts(rb[0])
        assert rb[0]["iters"] == 50
        assert np.array_equal(za, zb)


def test_noalloc_rejects_what_it_cannot_run(ctx, problem_dirs):
    d = problem_dirs["mid"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    opts = rt.options(tol_abs=1e-30, max_iters=25)
    # (3 m + 1) iterate-sized words per instance: 20 FP64 instances of the 229-node tree exceed 64 MiB
    assert 20 * (3 * 3 + 1) * p.n_iter * 8 > SCRATCH_BYTES
    big = rt.Batch(ctx, [p] * 20, rt.F64)
    big.set_x0(list(configs.initial_states(7, o.nx, 20, "normal")))
    with pytest.raises(rt.SpockError, match="bytes of scratch"):
        big.solve_noalloc(opts)
    shared = rt.Batch(ctx, [p] * 4, rt.F64, shared=True)
    with pytest.raises(rt.SpockError, match="per-instance"):
        shared.solve_noalloc(opts)
    # a rejected call leaves the library usable: a small solve afterwards still matches
    x0 = configs.initial_states(8, o.nx, 1, "normal")
    (ra, za), (rb, zb) = pair(ctx, d, x0, rt.F64, opts)
    assert counts(ra[0]) == counts(rb[0]) and np.array_equal(za, zb)
