"""The 5,461-node polyhedral i.i.d. tree at full size (4 events, horizon 6, nx=100, nu=40,
quadratic-plus-linear costs, 200 polyhedral rows, AV@R 0.95) in grid mode: T against the oracle
(1e-12) and an FP32 solve to 1e-3 whose FP64 polish is accepted against the same threshold.

An earlier probe of this workload (256 and 512 cooperating workgroups) ended in a segmentation
fault at process exit, after every result had been printed.  The cause was in the C ABI: at
interpreter exit Python may finalize a context before the batches made on it, and
spk_batch_destroy then synchronized the freed context's stream.  Handles are now reference
counted (a batch or problem keeps its context alive; tests/test_handles_host.py destroys them in
every order under AddressSanitizer), so these tests run again."""
import numpy as np
import pytest

from conftest import rel_err
from oracle.oracle import Oracle
from spock_amd import runtime as rt
from spock_amd.api import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("large"))
    configs.generate("large", d)
    return d


def test_large_T_matches_oracle(large):
    ctx = rt.Context(0)
    o = Oracle(large)
    b = rt.Batch(ctx, [rt.Problem(ctx, large)], rt.F64)
    b.set_workgroups(128)
    x0 = configs.initial_states(3, o.nx, 1, "small")[0]
    b.set_x0([x0])
    z = np.random.default_rng(7).standard_normal(o.n_iter)
    z[o.sU + o.sX + o.sY] = 0.0
    iv0 = o.n_prim + o.sI + o.sII + o.sIII
    z[iv0:iv0 + o.nx + o.info["numInputs"] + 2] = 0.0
    out, _ = b.debug_op(rt.OP_T, z)
    assert rel_err(out, o.T(o.scaled_x0(x0), z)) < 1e-12


def test_large_fp32_solve_and_fp64_polish(large):
    ctx = rt.Context(0)
    p = rt.Problem(ctx, large)
    x0 = configs.initial_states(3, p.nx, 1, "small")[0]
    b32 = rt.Batch(ctx, [p], rt.F32)
    b32.set_workgroups(128)
    b32.set_x0([x0])
    b32.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=5000))
    r32 = b32.results()[0][0]
    assert r32["status"] == 0
    b64 = rt.Batch(ctx, [p], rt.F64)
    b64.set_workgroups(128)
    b64.set_x0([x0])
    b64.copy_iterates_from(b32)
    b64.set_tolerances([r32["tol"]])
    b64.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=2000), warm=True)
    assert b64.results()[0][0]["status"] == 0
