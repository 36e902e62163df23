"""Shared-problem batches vs the FP64 CPU oracle, through the C ABI: the tile engine (16 instances
per workgroup, per-node products on the matrix cores, schedule 16, the default) and the lane
engine (64 instances per wave, schedule 0), both in an interleaved layout; the tile engine in its
1024-thread build (schedule 17, the default) and its 512-thread build (16).

The lane engine solves its Anderson least squares from the Gram matrix and forms SuperMann's
M-inner products from adjoints only (<x, M y> = x'y - gamma (x_p' L* y_d + (L* x_d)' y_p)), so
whole-iteration parity is checked against the oracle run the same way (``ls="gram"``,
``adjoint_dots=True``).  The Gram solve is
less well conditioned than Householder, so rounding differences in the accumulated Gram entries
grow faster along SuperMann; iterate parity is checked over the first 50 SPOCK iterations and 250
CP iterations (25 for the linear-cost family, whose iteration diverges under the reference's
linear-cost projection and amplifies rounding faster still).  Operators agree to FP64 rounding.  Batches of 70 instances exercise two
workgroups and the padding lanes of the last one.
"""
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


def x0s_for(name, nx, count):
    if name == "fixture":
        return [np.full(nx, 1.1) * (1.0 - 0.01 * k) for k in range(count)]
    if name == "small":
        return list(configs.initial_states(1, nx, count, "uniform"))
    if name == "poly":
        return list(configs.initial_states(33, nx, count, "small"))
    if name == "nocon":
        return list(configs.initial_states(7, nx, count, "uniform"))
    if name == "affine":
        return list(configs.initial_states(8, nx, count, "uniform"))
    if name == "pwi":
        return list(configs.initial_states(33, nx, count, "uniform"))
    if name == "linear":
        return list(configs.initial_states(61, nx, count, "small"))
    return list(configs.initial_states(2, nx, count))


def inert_zero(o, z):
    z = z.copy()
    z[o.sU + o.sX + o.sY] = 0.0
    iv0 = o.n_prim + o.sI + o.sII + o.sIII
    z[iv0:iv0 + o.nx + o.info["numInputs"] + 2] = 0.0
    return z


def shared(ctx, d, count, dtype=rt.F64, engine=None):
    p = rt.Problem(ctx, d)
    return p, rt.Batch(ctx, [p] * count, dtype, shared=True, engine=engine)


ENGINES = [17, 16, 0]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name", ["fixture", "small", "mid", "poly", "affine", "nocon", "pwi", "linear"])
def test_shared_operators_match_oracle(ctx, problem_dirs, name, engine):
    d = problem_dirs[name]
    o = Oracle(d)
    count = 3
    _, b = shared(ctx, d, count, engine=engine)
    rng = np.random.default_rng(11)
    zs = [inert_zero(o, 3.0 * rng.standard_normal(o.n_iter)) for _ in range(count)]
    ys = [inert_zero(o, rng.standard_normal(o.n_iter)) for _ in range(count)]
    x0s = x0s_for(name, o.nx, count)
    b.set_x0(x0s)
    Z = np.concatenate(zs)
    n = o.n_iter

    def rows(v):
        return [v[k * n:(k + 1) * n] for k in range(count)]

    out = rows(b.debug_op(rt.OP_L, Z)[0])
    for k in range(count):
        assert rel_err(out[k][o.n_prim:], o.L(zs[k][:o.n_prim])) < 1e-13
    out = rows(b.debug_op(rt.OP_LADJ, Z)[0])
    for k in range(count):
        assert rel_err(out[k][:o.n_prim], o.Ladj(zs[k][o.n_prim:])) < 1e-13
    out = rows(b.debug_op(rt.OP_DYN, Z)[0])
    for k in range(count):
        xo, uo = o.dyn_project(o.scaled_x0(x0s[k]), zs[k][o.sU:o.sU + o.sX], zs[k][:o.sU])
        assert rel_err(out[k][o.sU:o.sU + o.sX], xo) < 1e-12
        assert rel_err(out[k][:o.sU], uo) < 1e-12
    out = rows(b.debug_op(rt.OP_KERNEL, Z)[0])
    for k in range(count):
        ref = o.kernel_project(zs[k][:o.n_prim])
        assert rel_err(out[k][o.sU + o.sX:o.n_prim], ref[o.sU + o.sX:]) < 1e-13
    out = rows(b.debug_op(rt.OP_T, Z)[0])
    for k in range(count):
        assert rel_err(out[k], o.T(o.scaled_x0(x0s[k]), zs[k])) < 1e-12
    _, sc = b.debug_op(rt.OP_ELL, Z, np.concatenate(ys))
    for k in range(count):
        y = ys[k]
        ell = np.concatenate([o.Ladj(y[o.n_prim:]), o.L(y[:o.n_prim])])
        ref = o.dotM(zs[k], y)
        assert abs(sc[2 * k] - ref) <= 1e-12 * max(1.0, abs(ref))
        assert abs(sc[2 * k + 1] - np.max(np.abs(y / o.gamma - ell))) <= 1e-12 * np.max(np.abs(y / o.gamma))


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("name,iters,method", [("fixture", 50, "spock"), ("small", 50, "spock"),
                                               ("mid", 50, "spock"), ("small", 250, "cp"), ("poly", 50, "spock"),
                                               ("nocon", 50, "spock"), ("affine", 50, "spock"),
                                               ("pwi", 50, "spock"), ("linear", 25, "spock")])
def test_shared_fixed_iterations_match_oracle(ctx, problem_dirs, name, iters, method, engine):
    d = problem_dirs[name]
    o = Oracle(d)
    count = 3
    x0s = x0s_for(name, o.nx, count)
    _, b = shared(ctx, d, count, engine=engine)
    b.set_x0(x0s)
    b.solve(rt.options(tol_abs=1e-30, max_iters=iters, method=rt.SPOCK if method == "spock" else rt.CP))
    res, _ = b.results()
    for k in range(count):
        ref = o.solve(x0s[k], method=method, tol_abs=1e-30, max_iters=iters, ls="gram", adjoint_dots=True)
        r = res[k]
        assert r["status"] == ref["status"] == 1
        assert r["iters"] == ref["iters"]
        assert (r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"]) == (ref["k0"], ref["k1"], ref["k2"], ref["k2bt"],
                                                                   ref["k3"])
        assert rel_err(b.iterate(k), ref["z"]) < 1e-9


def test_shared_matches_per_instance_gram_engine(ctx, problem_dirs):
    """Both GPU engines with the Gram least squares, same 50 iterations, same data."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 5
    x0s = x0s_for("small", o.nx, count)
    p, b = shared(ctx, d, count)
    b.set_x0(x0s)
    opts = rt.options(tol_abs=1e-30, max_iters=50, ls_method=rt.LS_GRAM)
    b.solve(opts)
    res, _ = b.results()
    f = rt.Batch(ctx, [p] * count, rt.F64)
    f.set_x0(x0s)
    f.solve(opts)
    resf, _ = f.results()
    for k in range(count):
        assert res[k]["iters"] == resf[k]["iters"]
        assert (res[k]["k0"], res[k]["k1"], res[k]["k2"]) == (resf[k]["k0"], resf[k]["k1"], resf[k]["k2"])
        assert rel_err(b.iterate(k), f.iterate(k)) < 1e-9


def test_shared_converged_solves(ctx, problem_dirs):
    """70 instances (two workgroups, padded lanes): every solve converges like the oracle's."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 70
    x0s = x0s_for("small", o.nx, count)
    _, b = shared(ctx, d, count)
    b.set_x0(x0s)
    b.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=4000))
    res, _ = b.results()
    z = b.iterates().reshape(count, -1)
    for k in range(0, count, 7):
        ref = o.solve(x0s[k], tol_abs=1e-3, tol_rel=1e-3, max_iters=4000, ls="gram")
        assert res[k]["status"] == ref["status"]
        if ref["status"] == 0:
            s0, s0_ref = z[k][o.n_prim - o.sS], ref["z"][o.n_prim - o.sS]
            assert abs(s0 - s0_ref) <= 1e-2 * max(1.0, abs(s0_ref))
            zk = z[k]
            fp = np.max(np.abs(zk - o.T(o.scaled_x0(x0s[k]), zk))) / np.max(np.abs(zk))
            fp_ref = np.max(np.abs(ref["z"] - o.T(o.scaled_x0(x0s[k]), ref["z"]))) / np.max(np.abs(ref["z"]))
            assert fp <= 5 * fp_ref + 1e-6


def test_shared_fp32_bulk_and_fp64_polish(ctx, problem_dirs):
    """FP32 shared solve, then an FP64 polish accepted against the FP32 solve's own threshold."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 8
    x0s = x0s_for("small", o.nx, count)
    p, b32 = shared(ctx, d, count, rt.F32)
    b32.set_x0(x0s)
    b32.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=4000))
    r32, _ = b32.results()
    b64 = rt.Batch(ctx, [p] * count, rt.F64, shared=True)
    b64.set_x0(x0s)
    b64.copy_iterates_from(b32)
    z32 = b32.iterates()
    assert rel_err(b64.iterates(), z32) < 1e-15
    b64.set_tolerances([r["tol"] for r in r32])
    b64.solve(rt.options(max_iters=2000), warm=True)
    r64, _ = b64.results()
    z = b64.iterates().reshape(count, -1)
    assert sum(r["status"] == 0 for r in r32) >= count - 2
    for k in range(count):
        if r32[k]["status"] != 0:  # an instance the FP32 bulk did not finish is not checked here
            continue
        assert r64[k]["status"] == 0
        assert r64[k]["tol"] == pytest.approx(r32[k]["tol"], rel=1e-6)
        ref = o.solve(x0s[k], tol_abs=1e-3, tol_rel=1e-3, max_iters=4000)
        s0, s0_ref = z[k][o.n_prim - o.sS], ref["z"][o.n_prim - o.sS]
        assert abs(s0 - s0_ref) <= 2e-2 * max(1.0, abs(s0_ref))


@pytest.mark.parametrize("engine,dtype", [(17, rt.F64), (17, rt.F32), (0, rt.F64)])
def test_single_instance_iterate_round_trips(ctx, problem_dirs, engine, dtype):
    """spk_batch_get_iterate / spk_batch_set_iterate on a shared batch (one instance's column of
    the interleaved layout, 16- and 64-instance blocks) against the whole-batch copies."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count, n = 67, o.n_iter
    _, b = shared(ctx, d, count, dtype, engine)
    rng = np.random.default_rng(4)
    Z = rng.standard_normal(count * n)
    if dtype == rt.F32:
        Z = Z.astype(np.float32).astype(np.float64)  # exactly representable
    b.set_iterates(Z)
    for k in range(count):
        assert np.array_equal(b.iterate(k), Z[k * n:(k + 1) * n]), k
    lib = ctx.lib
    for k in (0, 15, 16, 63, 64, 66):
        v = np.ascontiguousarray(np.arange(n, dtype=np.float64) + 0.5 * k)
        ctx.check(lib.spk_batch_set_iterate(b.h, k, v.ctypes.data_as(rt.C.POINTER(rt.C.c_double))), "set_iterate")
        Z[k * n:(k + 1) * n] = v
        assert np.array_equal(b.iterate(k), v)
    assert np.array_equal(b.iterates(), Z)  # the other instances are untouched


def test_shared_layout_round_trips(ctx, problem_dirs):
    """Host and device copies between the interleaved layout and per-instance batches."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 67
    p, b = shared(ctx, d, count)
    rng = np.random.default_rng(3)
    Z = rng.standard_normal(count * o.n_iter)
    b.set_iterates(Z)
    assert np.array_equal(b.iterates(), Z)
    assert np.array_equal(b.iterate(5), Z[5 * o.n_iter:6 * o.n_iter])
    f = rt.Batch(ctx, [p] * count, rt.F64)
    f.copy_iterates_from(b)
    assert np.array_equal(f.iterates(), Z)
    b2 = rt.Batch(ctx, [p] * count, rt.F32, shared=True)
    b2.copy_iterates_from(f)
    assert rel_err(b2.iterates(), Z) < 1e-7
    b.reset()
    assert not np.any(b.iterates())


@pytest.mark.parametrize("dtype", [rt.F64, rt.F32])
def test_chunked_solve_equals_uninterrupted(ctx, problem_dirs, dtype):
    """A solve cut into launches of 37 outer iterations (resumed from the saved device state) gives
    bit-identical iterates, statuses and branch counts."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 70
    x0s = x0s_for("small", o.nx, count)
    p, a = shared(ctx, d, count, dtype)
    a.set_x0(x0s)
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=500)
    a.solve(rt.options(**opts))
    ra, _ = a.results()
    b = rt.Batch(ctx, [p] * count, dtype, shared=True)
    b.set_x0(x0s)
    chunks = []
    rb, _ = b.solve_chunked(rt.options(chunk_iters=37, **opts), on_chunk=lambda r, s: chunks.append(r))
    assert len(chunks) > 3
    for k in range(count):
        for f in ("status", "iters", "k0", "k1", "k2", "k2bt", "k3", "num_T"):
            assert ra[k][f] == rb[k][f], (k, f)
    assert np.array_equal(a.iterates(), b.iterates())


@pytest.mark.parametrize("engine,dtype", [(17, rt.F32), (17, rt.F64), (16, rt.F32), (0, rt.F64)])
def test_time_sliced_chunks_equal_uninterrupted(ctx, problem_dirs, engine, dtype):
    """Launches cut by a wall-clock slice (spk_batch_set_chunk_time) end at whatever iteration each
    workgroup has reached; the solve still equals an uninterrupted one bit for bit."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 70
    x0s = x0s_for("small", o.nx, count)
    p, a = shared(ctx, d, count, dtype, engine)
    a.set_x0(x0s)
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=500)
    a.solve(rt.options(**opts))
    ra, _ = a.results()
    b = rt.Batch(ctx, [p] * count, dtype, shared=True, engine=engine)
    b.set_chunk_time(2e-4)
    b.set_x0(x0s)
    chunks = []
    rb, _ = b.solve_chunked(rt.options(chunk_iters=400, **opts), on_chunk=lambda r, s: chunks.append(r))
    print(f"engine {engine}: {len(chunks)} launches of a 200 us slice")
    assert len(chunks) > 3  # 400-iteration chunks alone would end this solve in at most 2 launches
    for k in range(count):
        for f in ("status", "iters", "k0", "k1", "k2", "k2bt", "k3", "num_T"):
            assert ra[k][f] == rb[k][f], (k, f)
    assert np.array_equal(a.iterates(), b.iterates())
    with pytest.raises(rt.SpockError, match="finite"):
        b.set_chunk_time(-1.0)
    q = rt.Batch(ctx, [p], dtype)
    with pytest.raises(rt.SpockError, match="shared batch"):
        q.set_chunk_time(1.0)


@pytest.mark.parametrize("name,dtype", [("small", rt.F64), ("mid", rt.F64), ("mid", rt.F32), ("poly", rt.F64)])
def test_lds_schedule_is_bitwise_identical(ctx, problem_dirs, name, dtype):
    """The LDS-staged matrix-vector schedule accumulates in the same order as the scalar one."""
    d = problem_dirs[name]
    o = Oracle(d)
    count = 70
    x0s = x0s_for(name, o.nx, count)
    out = []
    for mv in (0, 8):
        p, b = shared(ctx, d, count, dtype, engine=mv)
        b.set_x0(x0s)
        b.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=100))
        res, _ = b.results()
        out.append((b.iterates(), [(r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k3"]) for r in res]))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_autotune_keeps_the_solution(ctx, problem_dirs):
    """After the autotuner's timed runs the batch solves exactly as a fresh batch given the chosen
    schedule (the workgroup size sets the order of the per-lane sums, so it is part of the result)."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 100
    x0s = x0s_for("small", o.nx, count)
    _, b = shared(ctx, d, count)
    b.set_x0(x0s)
    mv, block, secs = b.autotune(iters=20)
    assert (mv, block) in secs and all(s > 0 for s in secs.values())
    assert secs[(mv, block)] == min(secs.values())
    b.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=500, block_size=0))
    res, _ = b.results()
    _, ref = shared(ctx, d, count)
    ref.set_schedule(mv, block)
    ref.set_x0(x0s)
    ref.solve(rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=500, block_size=0))
    res_ref, _ = ref.results()
    assert [r["iters"] for r in res] == [r["iters"] for r in res_ref]
    assert np.array_equal(b.iterates(), ref.iterates())


@pytest.mark.parametrize("dtype", [rt.F64, rt.F32])
def test_graph_captured_chunked_solve_equals_uninterrupted(ctx, problem_dirs, dtype):
    """A chunked solve captured as one HIP graph: too few launches leave instances running (a
    resume finishes them), enough launches finish in one replay; both equal one uninterrupted
    solve bit for bit, and extra launches after the end change nothing."""
    d = problem_dirs["small"]
    o = Oracle(d)
    count = 100
    x0s = x0s_for("small", o.nx, count)
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=600)
    _, ref = shared(ctx, d, count, dtype)
    ref.set_x0(x0s)
    ref.solve(rt.options(**opts))
    res_ref, _ = ref.results()
    _, g = shared(ctx, d, count, dtype)
    g.set_x0(x0s)
    g.solve_graph(rt.options(chunk_iters=50, **opts), launches=3)
    res, _ = g.results()
    assert any(r["status"] == -1 for r in res)
    g.resume_graph(2)  # two more chunks as one graph, then launch by launch
    res, _ = g.results()
    while any(r["status"] == -1 for r in res):
        g.resume()
        res, _ = g.results()
    assert [(r["status"], r["iters"]) for r in res] == [(r["status"], r["iters"]) for r in res_ref]
    assert np.array_equal(g.iterates(), ref.iterates())
    for launches in (None, 40):
        _, g2 = shared(ctx, d, count, dtype)
        g2.set_x0(x0s)
        g2.solve_graph(rt.options(chunk_iters=50, **opts), launches=launches)
        res2, _ = g2.results()
        assert [(r["status"], r["iters"]) for r in res2] == [(r["status"], r["iters"]) for r in res_ref]
        assert np.array_equal(g2.iterates(), ref.iterates())
