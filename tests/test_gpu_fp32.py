"""The FP32 path against the FP64 CPU oracle, operator by operator, through the C ABI.

The FP32 bulk iterations (both engines: one workgroup per instance, and shared batches with the
instances interleaved) run every operator of the CP step in single precision with FP32-accumulated
products.  Each operator is checked on inputs that are exactly representable in FP32 (so only the
operator's own rounding counts) against the FP64 oracle, with bounds relative to the output's
largest entry:
  * L, L*, the kernel projection, <x, M y> and the error measure: 1e-5;
  * the dynamics projection (a backward sweep of 10-stage depth and a forward rollout through the
    unstable dynamics) and T, which contains it: 5e-5.
FP32 carries 6e-8 relative per operation; a row of L sums at most a few hundred products, a
stage of the sweep a few thousand, so these bounds leave one to two decades of margin for plain
rounding while a wrong index, sign or dropped term shows as an O(1) relative error.
"""
import numpy as np
import pytest

from conftest import rel_err
from oracle.oracle import Oracle
from spock_amd import runtime as rt
from test_gpu_parity import inert_zero, x0_for

pytestmark = pytest.mark.gpu

BOUND_OP = 1e-5
BOUND_SWEEP = 5e-5


@pytest.fixture(scope="module")
def ctx():
    return rt.Context(0)


def f32(v):
    return np.asarray(v, dtype=np.float32).astype(np.float64)


def check_ops(o, b, zs, ys, x0s, count, name):
    n = o.n_iter
    Z, Y = np.concatenate(zs), np.concatenate(ys)

    def rows(v):
        return [v[k * n:(k + 1) * n] for k in range(count)]

    errs = {}
    out = rows(b.debug_op(rt.OP_L, Z)[0])
    errs["L"] = max(rel_err(out[k][o.n_prim:], o.L(zs[k][:o.n_prim])) for k in range(count))
    out = rows(b.debug_op(rt.OP_LADJ, Z)[0])
    errs["Ladj"] = max(rel_err(out[k][:o.n_prim], o.Ladj(zs[k][o.n_prim:])) for k in range(count))
    out = rows(b.debug_op(rt.OP_DYN, Z)[0])
    e = 0.0
    for k in range(count):
        xo, uo = o.dyn_project(o.scaled_x0(f32(x0s[k])), zs[k][o.sU:o.sU + o.sX], zs[k][:o.sU])
        e = max(e, rel_err(out[k][o.sU:o.sU + o.sX], xo), rel_err(out[k][:o.sU], uo))
    errs["dyn"] = e
    out = rows(b.debug_op(rt.OP_KERNEL, Z)[0])
    errs["kernel"] = max(rel_err(out[k][o.sU + o.sX:o.n_prim], o.kernel_project(zs[k][:o.n_prim])[o.sU + o.sX:])
                         for k in range(count))
    out = rows(b.debug_op(rt.OP_T, Z)[0])
    errs["T"] = max(rel_err(out[k], o.T(o.scaled_x0(f32(x0s[k])), zs[k])) for k in range(count))
    _, sc = b.debug_op(rt.OP_ELL, Z, Y)
    ed, em = 0.0, 0.0
    for k in range(count):
        y = ys[k]
        ell = np.concatenate([o.Ladj(y[o.n_prim:]), o.L(y[:o.n_prim])])
        ref = o.dotM(zs[k], y)
        scale = np.sum(np.abs(zs[k] * y)) + o.gamma * np.sum(np.abs(zs[k] * ell))
        ed = max(ed, abs(sc[2 * k] - ref) / scale)
        mref = np.max(np.abs(y / o.gamma - ell))
        em = max(em, abs(sc[2 * k + 1] - mref) / np.max(np.abs(y / o.gamma)))
    errs["dotM"], errs["err"] = ed, em
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    for key in ("L", "Ladj", "kernel", "dotM", "err"):
        assert errs[key] < BOUND_OP, (key, errs[key])
    for key in ("dyn", "T"):
        assert errs[key] < BOUND_SWEEP, (key, errs[key])


@pytest.mark.parametrize("name", ["fixture", "small", "mid", "poly", "nocon", "affine", "pwi", "linear"])
def test_fp32_per_instance_operators_match_fp64_oracle(ctx, problem_dirs, name):
    d = problem_dirs[name]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    count = 2
    b = rt.Batch(ctx, [p] * count, rt.F32)
    rng = np.random.default_rng(21)
    zs = [f32(inert_zero(o, 3.0 * rng.standard_normal(o.n_iter))) for _ in range(count)]
    ys = [f32(inert_zero(o, rng.standard_normal(o.n_iter))) for _ in range(count)]
    x0s = [f32(x0_for(name, o.nx, k)) for k in range(count)]
    b.set_x0(x0s)
    check_ops(o, b, zs, ys, x0s, count, name)


@pytest.mark.parametrize("engine", [17, 16, 0])
@pytest.mark.parametrize("name", ["fixture", "small", "mid", "poly", "nocon", "affine", "pwi", "linear"])
def test_fp32_shared_operators_match_fp64_oracle(ctx, problem_dirs, name, engine):
    """engine 17, 16: the tile engine (matrix-core products; 1024- and 512-thread builds); 0: the
    lane engine."""
    d = problem_dirs[name]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    count = 3
    b = rt.Batch(ctx, [p] * count, rt.F32, shared=True, engine=engine)
    rng = np.random.default_rng(22)
    zs = [f32(inert_zero(o, 3.0 * rng.standard_normal(o.n_iter))) for _ in range(count)]
    ys = [f32(inert_zero(o, rng.standard_normal(o.n_iter))) for _ in range(count)]
    x0s = [f32(x0_for(name, o.nx, k)) for k in range(count)]
    b.set_x0(x0s)
    check_ops(o, b, zs, ys, x0s, count, name)
