"""HIP solver vs the FP64 CPU oracle, through the C ABI (libspock_amd.so).

Operators (L, L*, dynamics and kernel projections, T, the M-inner product): agreement to
FP64 rounding (1e-12 relative).  Whole iterations: the iterates agree to 1e-9 relative
with identical SuperMann branch counts.  SuperMann with Anderson directions amplifies
last-bit differences by about 10x per 25 outer iterations (the oracle perturbed by one
ulp in x0 drifts 5e-12 after 100 iterations and 6e-6 after 250), so SPOCK iterate parity
is checked over the first 75 outer iterations and plain CP over 250; converged solves
are checked on status, cost s0 and the oracle's own residual of the GPU solution.
Batched solves must reproduce single-instance GPU solves bit for bit.
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


def x0_for(name, nx, k=0):
    if name == "fixture":
        return np.full(nx, 1.1)
    if name == "small":
        return configs.initial_states(1, nx, k + 1, "uniform")[k]
    if name == "poly":
        return configs.initial_states(33, nx, k + 1, "small")[k]
    if name == "nocon":
        return configs.initial_states(7, nx, k + 1, "uniform")[k]
    if name == "affine":
        return configs.initial_states(8, nx, k + 1, "uniform")[k]
    if name == "pwi":
        return configs.initial_states(33, nx, k + 1, "uniform")[k]
    if name == "linear":
        return configs.initial_states(61, nx, k + 1, "small")[k]
    return configs.initial_states(2, nx, k + 1)[k]


def inert_zero(o, z):
    """t[0] and node 0's IV block are never written by the reference from a zero start."""
    z = z.copy()
    z[o.sU + o.sX + o.sY] = 0.0
    iv0 = o.n_prim + o.sI + o.sII + o.sIII
    z[iv0:iv0 + o.nx + o.info["numInputs"] + 2] = 0.0
    return z


@pytest.mark.parametrize("name,wgs", [("fixture", 1), ("small", 1), ("mid", 1), ("poly", 1),
                                      ("mid", 7), ("poly", 3), ("mid", 128), ("nocon", 1), ("affine", 1),
                                      ("affine", 5), ("pwi", 1), ("pwi", 3), ("linear", 1), ("linear", 3)])
def test_operators_match_oracle(ctx, problem_dirs, name, wgs):
    """wgs > 1: grid mode, the instance spread over a cooperative launch of wgs workgroups."""
    d = problem_dirs[name]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p], rt.F64)
    b.set_workgroups(wgs)
    rng = np.random.default_rng(7)
    z = inert_zero(o, 3.0 * rng.standard_normal(o.n_iter))
    x0 = x0_for(name, o.nx)
    b.set_x0([x0])
    x0s = o.scaled_x0(x0)
    out, _ = b.debug_op(rt.OP_L, z)
    assert rel_err(out[o.n_prim:], o.L(z[:o.n_prim])) < 1e-13
    out, _ = b.debug_op(rt.OP_LADJ, z)
    assert rel_err(out[:o.n_prim], o.Ladj(z[o.n_prim:])) < 1e-13
    out, _ = b.debug_op(rt.OP_DYN, z)
    xo, uo = o.dyn_project(x0s, z[o.sU:o.sU + o.sX], z[:o.sU])
    assert rel_err(out[o.sU:o.sU + o.sX], xo) < 1e-12
    assert rel_err(out[:o.sU], uo) < 1e-12
    out, _ = b.debug_op(rt.OP_KERNEL, z)
    ref = o.kernel_project(z[:o.n_prim])
    assert rel_err(out[o.sU + o.sX:o.n_prim], ref[o.sU + o.sX:]) < 1e-13
    out, _ = b.debug_op(rt.OP_T, z)
    assert rel_err(out, o.T(x0s, z)) < 1e-12
    y = inert_zero(o, rng.standard_normal(o.n_iter))
    _, sc = b.debug_op(rt.OP_ELL, z, y)
    ell = np.concatenate([o.Ladj(y[o.n_prim:]), o.L(y[:o.n_prim])])
    assert abs(sc[0] - o.dotM(z, y)) <= 1e-12 * max(1.0, abs(o.dotM(z, y)))
    assert abs(sc[1] - np.max(np.abs(y / o.gamma - ell))) <= 1e-12 * np.max(np.abs(y / o.gamma))


@pytest.mark.parametrize("name", ["fixture", "mid", "poly", "pwi", "linear"])
def test_get_dual_matches_oracle(ctx, problem_dirs, name):
    """Cache::getDual (src/cache.cuh:371-376): the dual step from the iterate with a given primal
    point; with the primal prox of T(z) as that point it is the dual part of T(z)."""
    d = problem_dirs[name]
    o = Oracle(d)
    b = rt.Batch(ctx, [rt.Problem(ctx, d)], rt.F64)
    x0 = x0_for(name, o.nx)
    b.set_x0([x0])
    z = inert_zero(o, 2.0 * np.random.default_rng(11).standard_normal(o.n_iter))
    tz = o.T(o.scaled_x0(x0), z)
    out = np.zeros(o.n_iter)
    out[:o.n_prim] = tz[:o.n_prim]
    got, _ = b.debug_op(rt.OP_DUAL, z, out)
    assert rel_err(got[o.n_prim:], tz[o.n_prim:]) < 1e-12


@pytest.mark.parametrize("name,iters,method,wgs", [
    ("fixture", 75, "spock", 1), ("small", 75, "spock", 1), ("small", 250, "cp", 1), ("fixture", 250, "cp", 1),
    ("mid", 50, "spock", 1), ("poly", 75, "spock", 1), ("poly", 250, "cp", 1),
    ("mid", 50, "spock", 64), ("poly", 75, "spock", 5), ("small", 250, "cp", 16), ("nocon", 75, "spock", 1),
    ("affine", 75, "spock", 1), ("affine", 250, "cp", 1), ("pwi", 50, "spock", 1), ("pwi", 250, "cp", 1),
    ("pwi", 50, "spock", 3), ("linear", 50, "spock", 1), ("linear", 100, "cp", 1), ("linear", 50, "spock", 3)])
def test_fixed_iterations_match_oracle(ctx, problem_dirs, name, iters, method, wgs):
    d = problem_dirs[name]
    o = Oracle(d)
    x0 = x0_for(name, o.nx)
    ref = o.solve(x0, method=method, tol_abs=1e-30, max_iters=iters)
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p], rt.F64)
    b.set_workgroups(wgs)
    b.set_x0([x0])
    b.solve(rt.options(tol_abs=1e-30, max_iters=iters, method=rt.SPOCK if method == "spock" else rt.CP,
                       block_size=256))
    res, _ = b.results()
    r = res[0]
    assert r["status"] == ref["status"] == 1
    assert r["iters"] == ref["iters"]
    assert (r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"]) == (ref["k0"], ref["k1"], ref["k2"], ref["k2bt"], ref["k3"])
    assert rel_err(b.iterate(0), ref["z"]) < 1e-9


def fixed_point_residual(o, x0, z):
    """||z - T(z)||_inf / ||z||_inf computed by the oracle for a GPU iterate."""
    return float(np.max(np.abs(z - o.T(o.scaled_x0(x0), z))) / np.max(np.abs(z)))


def gpu_solve(ctx, problems, x0s, dtype=rt.F64, **opts):
    b = rt.Batch(ctx, problems, dtype)
    b.set_x0(x0s)
    b.solve(rt.options(**opts))
    res, sec = b.results()
    return b, res


@pytest.mark.parametrize("name", ["fixture", "small", "poly", "affine", "pwi"])
def test_converged_solve_matches_oracle(ctx, problem_dirs, name):
    d = problem_dirs[name]
    o = Oracle(d)
    x0 = x0_for(name, o.nx)
    ref = o.solve(x0, tol_abs=1e-3, max_iters=3000)
    b, res = gpu_solve(ctx, [rt.Problem(ctx, d)], [x0], tol_abs=1e-3, max_iters=3000)
    r = res[0]
    z = b.iterate(0)
    assert ref["status"] == 0 and r["status"] == 0
    assert abs(r["iters"] - ref["iters"]) <= 50
    s0, s0_ref = z[o.n_prim - o.sS], ref["z"][o.n_prim - o.sS]
    assert abs(s0 - s0_ref) <= 1e-2 * max(1.0, abs(s0_ref))
    assert fixed_point_residual(o, x0, z) <= 5 * fixed_point_residual(o, x0, ref["z"]) + 1e-6


def test_batch_of_initial_states_matches_single_solves(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    p = rt.Problem(ctx, d)
    B = 6
    x0s = [x0_for("small", o.nx, k) for k in range(B)]
    opts = dict(tol_abs=1e-3, max_iters=3000)
    b, res = gpu_solve(ctx, [p] * B, x0s, **opts)
    for k in range(B):
        bk, rk = gpu_solve(ctx, [p], [x0s[k]], **opts)
        assert res[k]["status"] == rk[0]["status"]
        assert res[k]["iters"] == rk[0]["iters"]
        assert np.array_equal(b.iterate(k), bk.iterate(0))
        ref = o.solve(x0s[k], tol_abs=1e-3, max_iters=3000)
        assert res[k]["status"] == ref["status"]


def test_uneven_batch_matches_single_solves(ctx, problem_dirs):
    names = ["fixture", "small", "mid", "small", "fixture"]
    probs = {n: rt.Problem(ctx, problem_dirs[n]) for n in set(names)}
    oracles = {n: Oracle(problem_dirs[n]) for n in set(names)}
    x0s = [x0_for(n, oracles[n].nx, k) for k, n in enumerate(names)]
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=400)
    b, res = gpu_solve(ctx, [probs[n] for n in names], x0s, **opts)
    for k, n in enumerate(names):
        bk, rk = gpu_solve(ctx, [probs[n]], [x0s[k]], **opts)
        assert res[k]["status"] == rk[0]["status"]
        assert res[k]["iters"] == rk[0]["iters"]
        assert np.array_equal(b.iterate(k), bk.iterate(0))


def test_fp32_solve_tracks_fp64(ctx, problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = x0_for("small", o.nx)
    ref = o.solve(x0, tol_abs=1e-3, max_iters=3000)
    b, res = gpu_solve(ctx, [rt.Problem(ctx, d)], [x0], rt.F32, tol_abs=1e-3, max_iters=3000)
    assert res[0]["status"] == 0
    z = b.iterate(0)
    s0 = z[o.n_prim - o.sS]
    s0_ref = ref["z"][o.n_prim - o.sS]
    assert abs(s0 - s0_ref) <= 2e-2 * max(1.0, abs(s0_ref))


@pytest.fixture(scope="module")
def uneven_family(tmp_path_factory):
    """Instances 0..11 of the uneven family (horizon 5-12, 2-4 events, branching stops at 1-3,
    nx 10-40, nu 4-16; 19 to 256 nodes)."""
    base = tmp_path_factory.mktemp("uneven")
    out = []
    for k in range(12):
        d = str(base / f"i{k}")
        configs.uneven_instance(d, k)
        out.append(d)
    return out


def test_uneven_family_batch_matches_oracle(ctx, uneven_family):
    """One persistent launch over 12 differently shaped trees: every instance's iterate after 50
    SPOCK iterations matches the oracle's to 1e-9 with the same branch counts."""
    oracles = [Oracle(d) for d in uneven_family]
    probs = [rt.Problem(ctx, d) for d in uneven_family]
    x0s = [configs.initial_states(4000 + k, o.nx, 1)[0] for k, o in enumerate(oracles)]
    b = rt.Batch(ctx, probs, rt.F64)
    b.set_x0(x0s)
    b.solve(rt.options(tol_abs=1e-30, max_iters=50))
    res, _ = b.results()
    for k, o in enumerate(oracles):
        ref = o.solve(x0s[k], tol_abs=1e-30, max_iters=50)
        r = res[k]
        assert (r["iters"], r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"]) == \
            (ref["iters"], ref["k0"], ref["k1"], ref["k2"], ref["k2bt"], ref["k3"])
        assert rel_err(b.iterate(k), ref["z"]) < 1e-9


def test_uneven_family_fp32_bulk_and_fp64_polish(ctx, uneven_family):
    """FP32 solve of an uneven batch (the ten instances of at most 85 nodes), then an FP64 polish
    warm-started from it against the plain solves' thresholds: every instance the plain FP64
    solve converges on passes the FP64 residual test after the polish, at the plain solve's cost."""
    sel = [k for k in range(len(uneven_family)) if k not in (0, 5)]
    dirs = [uneven_family[k] for k in sel]
    probs = [rt.Problem(ctx, d) for d in dirs]
    x0s = [configs.initial_states(4000 + k, p.nx, 1)[0] for k, p in zip(sel, probs)]
    opts = dict(tol_abs=1e-3, tol_rel=1e-3, max_iters=4000)
    plain = rt.Batch(ctx, probs, rt.F64)
    plain.set_x0(x0s)
    plain.solve(rt.options(**opts))
    res_p, _ = plain.results()
    b32 = rt.Batch(ctx, probs, rt.F32)
    b32.set_x0(x0s)
    b32.solve(rt.options(**opts))
    b32.results()
    b64 = rt.Batch(ctx, probs, rt.F64)
    b64.set_x0(x0s)
    b64.copy_iterates_from(b32)
    b64.set_tolerances([r["tol"] for r in res_p])
    b64.solve(rt.options(**opts), warm=True)
    res64, _ = b64.results()
    ok = [i for i in range(len(probs)) if res_p[i]["status"] == 0]
    assert len(ok) >= 4  # 5 of these 10 random trees converge within 4,000 plain iterations
    for i in ok:
        assert res64[i]["status"] == 0
        o = Oracle(dirs[i])
        s = o.n_prim - o.sS
        s_p, s_64 = plain.iterate(i)[s], b64.iterate(i)[s]
        assert abs(s_64 - s_p) <= 2e-2 * max(1.0, abs(s_p))


def test_uneven_autotune_and_largest_first_order(ctx, uneven_family):
    """The uneven-batch autotuner keeps the fastest of its six schedules, and the chosen schedule
    still matches the oracle.  Largest-first dispatch only reorders the workgroups: every
    instance's result is bitwise the one of batch order at the same workgroup size."""
    oracles = [Oracle(d) for d in uneven_family]
    probs = [rt.Problem(ctx, d) for d in uneven_family]
    x0s = [configs.initial_states(4000 + k, o.nx, 1)[0] for k, o in enumerate(oracles)]
    b = rt.Batch(ctx, probs, rt.F64)
    b.set_x0(x0s)
    (blk, lf), secs = b.autotune_uneven(iters=25)
    print(f"uneven schedule {blk}, largest first {lf}: {secs}")
    assert (blk, lf) in secs and secs[(blk, lf)] == min(secs.values())
    b.solve(rt.options(tol_abs=1e-30, max_iters=50))
    res, _ = b.results()
    for k, o in enumerate(oracles):
        ref = o.solve(x0s[k], tol_abs=1e-30, max_iters=50)
        assert (res[k]["iters"], res[k]["k0"], res[k]["k2"]) == (ref["iters"], ref["k0"], ref["k2"])
        assert rel_err(b.iterate(k), ref["z"]) < 1e-9
    runs = []
    for order in (False, True):
        c = rt.Batch(ctx, probs, rt.F64)
        c.set_uneven_schedule(256, order)
        c.set_x0(x0s)
        c.solve(rt.options(tol_abs=1e-30, max_iters=50))
        r, _ = c.results()
        runs.append((c.iterates(), [(v["iters"], v["k0"], v["k1"], v["k2"], v["k3"]) for v in r]))
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]


def test_uneven_batch_graph_capture_and_replay(ctx, uneven_family):
    """The persistent solve launch of an uneven batch captured as a graph (spk_batch_solve_graph) and
    replayed (spk_batch_replay_graph) after a reset and new initial states: every replay equals a
    plain solve of the same initial states bit for bit, with eager work between the replays; a
    later plain solve drops the captured graph and a replay then fails with a message."""
    probs = [rt.Problem(ctx, d) for d in uneven_family]
    x0_sets = [[configs.initial_states(4100 + 17 * r + k, p.nx, 1)[0] for k, p in enumerate(probs)]
               for r in range(3)]
    opts = rt.options(tol_abs=1e-3, tol_rel=1e-3, max_iters=300)

    def plain(x0s):
        b = rt.Batch(ctx, probs, rt.F64)
        b.set_x0(x0s)
        b.solve(opts)
        res, _ = b.results()
        return b.iterates(), [(r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k3"]) for r in res]

    g = rt.Batch(ctx, probs, rt.F64)
    g.set_x0(x0_sets[0])
    g.solve_graph(opts)
    res, _ = g.results()
    z, counts = plain(x0_sets[0])
    assert np.array_equal(g.iterates(), z)
    assert [(r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k3"]) for r in res] == counts
    for x0s in (x0_sets[1], x0_sets[2], x0_sets[1]):
        z, counts = plain(x0s)  # eager launches between the replays
        g.reset()
        g.set_x0(x0s)
        g.replay_graph()
        res, _ = g.results()
        assert np.array_equal(g.iterates(), z)
        assert [(r["status"], r["iters"], r["k0"], r["k1"], r["k2"], r["k3"]) for r in res] == counts
    g.solve(opts)
    g.results()
    with pytest.raises(RuntimeError, match="no captured solve"):
        g.replay_graph()


def test_workgroup_autotune_single_instance(ctx, problem_dirs):
    """The workgroup autotuner of one-instance batches keeps the fastest count, leaves no state
    behind, and the tuned solve still matches the oracle over fixed iterations."""
    d = problem_dirs["mid"]
    o = Oracle(d)
    b = rt.Batch(ctx, [rt.Problem(ctx, d)], rt.F64)
    x0 = x0_for("mid", o.nx)
    b.set_x0([x0])
    w, secs = b.autotune_workgroups(iters=25)
    print(f"workgroups {w}: {secs}")
    assert w in secs and secs[w] == min(secs.values())
    b.solve(rt.options(tol_abs=1e-30, max_iters=50))
    res, _ = b.results()
    ref = o.solve(x0, tol_abs=1e-30, max_iters=50)
    assert res[0]["iters"] == ref["iters"]
    assert rel_err(b.iterate(0), ref["z"]) < 1e-9


@pytest.mark.parametrize("name,shared", [("small", False), ("mid", False), ("poly", False), ("small", True),
                                         ("mid", True)])
def test_warm_start_matches_oracle(ctx, problem_dirs, name, shared):
    """runSpock(x0, &previousSolution) (src/cache.cuh:611-632: the previous solution becomes both
    the iterate and the previous iterate, saveToPrev): 50 SPOCK iterations from a warm start
    (the oracle's iterate after 100 iterations from another initial state) against the oracle's
    warm-started solve, iterates to 1e-9 and identical branch counts.  Shared batches run the
    Gram least squares and adjoint-only dot products, and the oracle the same way."""
    d = problem_dirs[name]
    o = Oracle(d)
    x0a, x0b = x0_for(name, o.nx, 0), x0_for(name, o.nx, 1)
    kw = dict(ls="gram", adjoint_dots=True) if shared else {}
    z_prev = o.solve(x0a, tol_abs=1e-30, max_iters=100, **kw)["z"]
    ref = o.solve(x0b, warm=z_prev, tol_abs=1e-30, max_iters=50, **kw)
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p] * (2 if shared else 1), rt.F64, shared=shared)
    b.set_x0([x0b] * (2 if shared else 1))
    b.set_iterates(np.concatenate([z_prev] * (2 if shared else 1)))
    b.solve(rt.options(tol_abs=1e-30, max_iters=50, block_size=256 if not shared else 0), warm=True)
    res, _ = b.results()
    r = res[0]
    assert r["status"] == ref["status"] == 1
    assert r["iters"] == ref["iters"]
    assert (r["k0"], r["k1"], r["k2"], r["k2bt"], r["k3"]) == (ref["k0"], ref["k1"], ref["k2"], ref["k2bt"], ref["k3"])
    assert rel_err(b.iterate(0), ref["z"]) < 1e-9


def test_gpu_matches_committed_golden_vectors(ctx, tmp_path):
    """The HIP path against the committed fixture vectors (tests/golden/fixture_golden.npz, the
    oracle's outputs on the reference's unit-test problem): L, L*, T and 75 SPOCK iterations,
    without running the oracle."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixture_golden.npz"),
                allow_pickle=False)
    d = str(tmp_path / "fixture")
    configs.generate("fixture", d)
    p = rt.Problem(ctx, d)
    b = rt.Batch(ctx, [p], rt.F64)
    b.set_x0([g["x0"]])
    n_prim = g["prim"].size
    z = np.concatenate([g["prim"], g["dual"]])
    out, _ = b.debug_op(rt.OP_L, z)
    assert rel_err(out[n_prim:], g["L_prim"]) < 1e-13
    out, _ = b.debug_op(rt.OP_LADJ, z)
    assert rel_err(out[:n_prim], g["Ladj_dual"]) < 1e-13
    out, _ = b.debug_op(rt.OP_T, g["z"])
    assert rel_err(out, g["T_z"]) < 1e-12
    b.reset()  # the operators above left their input in the iterate; the solve starts cold
    b.solve(rt.options(tol_abs=1e-30, max_iters=75, block_size=256))
    r = b.results()[0][0]
    assert r["iters"] == int(g["spock75_iters"])
    assert [r[k] for k in ("k0", "k1", "k2", "k2bt", "k3")] == [int(v) for v in g["spock75_counts"]]
    assert rel_err(b.iterate(0), g["spock75"]) < 1e-9
