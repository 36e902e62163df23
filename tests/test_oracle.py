"""Pins of the FP64 CPU oracle (oracle/spock_oracle.cpp) by what the mathematics fixes and by the
known answers of the reference's tests, so that the GPU parity tests compare against a checked
checker:

* L* is the adjoint of L, and gamma ||L|| = 0.99 (the step size of api/python/spock/problem.py:402-417);
* the dynamics projection equals the KKT solution of min ||x - xbar||^2 + ||u - ubar||^2 subject to
  the dynamics (the QP of api/python/spock/problem.py:233-261, solved here with numpy);
* the kernel projection is the orthogonal projector onto ker [E', -I, -I] of every nonleaf node;
* SOC projection known answers (tests/cuda/testProjections.cu:150-170) and the textbook projection;
* the box projection (tests/cuda/testKernels.cu:14-36) and the AV@R mask (tests/cuda/testRisks.cu:46-49);
* least squares against numpy; the M inner product against its definition;
* T maps into the dynamics-feasible set and a converged solve is a fixed point of T;
* committed golden vectors (tests/golden) reproduce.
"""
import os

import numpy as np
import pytest

from conftest import rel_err
from oracle import numpy_ref as nr
from oracle.oracle import Oracle, box, lstsq, lstsq_gram, soc
from spock_amd.api import btio, build, configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def affine_constants(d, o):
    """Per-node constants e_j of affine dynamics (zeros for linear dynamics), shape (N, nx)."""
    N = o.info["numNodes"]
    if o.info["dynamics"]["type"] != "affine":
        return np.zeros((N, o.nx))
    return btio.read_bt(os.path.join(d, "dynamics_c_d.bt")).transpose(2, 0, 1).reshape(N, o.nx)


def inert_zero(o, z):
    z = z.copy()
    z[o.sU + o.sX + o.sY] = 0.0
    iv0 = o.n_prim + o.sI + o.sII + o.sIII
    z[iv0:iv0 + o.nx + o.nu + 2] = 0.0
    return z


@pytest.mark.parametrize("name", ["fixture", "small", "mid", "poly", "nocon", "affine", "linear"])
def test_adjoint(problem_dirs, name):
    o = Oracle(problem_dirs[name])
    rng = np.random.default_rng(0)
    p = rng.standard_normal(o.n_prim)
    p[o.sU + o.sX + o.sY] = 0.0  # t[0] is an inert slot (no dual block refers to it)
    d = rng.standard_normal(o.n_dual)
    lhs, rhs = p @ o.Ladj(d), o.L(p) @ d
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("name", ["small", "poly", "nocon", "affine", "pwi", "linear"])
def test_operators_match_the_generator(tmp_path, name):
    """The oracle's L and L* (a C++ transcription of src/operator.cuh) against the generator's
    (spock_amd.api.problem, written separately from the Python reference) on random vectors, for
    every constraint kind: box, polyhedron, none, and the polyhedron with identity whose dual block
    holds all identity parts before all G parts."""
    d = str(tmp_path / name)
    gen = configs.generate(name, d)
    o = Oracle(d)
    rng = np.random.default_rng(12)
    p, q = rng.standard_normal(o.n_prim), rng.standard_normal(o.n_dual)
    assert np.abs(o.L(p) - gen.op(p)).max() <= 1e-12 * max(1.0, np.abs(gen.op(p)).max())
    assert np.abs(o.Ladj(q) - gen.adj(q)).max() <= 1e-12 * max(1.0, np.abs(gen.adj(q)).max())
    if name == "pwi":
        lhs, rhs = p @ o.Ladj(q), o.L(p) @ q
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
        Ld = nr.dense_operator(o.L, o.n_prim)
        assert o.gamma * np.linalg.norm(Ld, 2) == pytest.approx(0.99, abs=2e-3)


@pytest.mark.parametrize("name", ["fixture", "small", "poly", "nocon", "affine", "linear"])
def test_step_size_is_099_over_norm(problem_dirs, name):
    o = Oracle(problem_dirs[name])
    Ld = nr.dense_operator(o.L, o.n_prim)
    assert o.gamma * np.linalg.norm(Ld, 2) == pytest.approx(0.99, abs=2e-3)


@pytest.mark.parametrize("name", ["fixture", "small", "poly", "nocon", "affine"])
def test_dynamics_projection_is_the_kkt_solution(problem_dirs, name):
    d = problem_dirs[name]
    o = Oracle(d)
    N, nl, nx, nu = o.info["numNodes"], o.info["numNonleafNodes"], o.nx, o.nu
    AB = btio.read_bt(os.path.join(d, "dynamics_AB_d.bt")).transpose(2, 0, 1)
    anc = btio.read_flat(os.path.join(d, "ancestors_u.bt")).astype(np.int64)
    rng = np.random.default_rng(1)
    xb, ub, x0 = 10 * rng.standard_normal((N, nx)), 10 * rng.standard_normal((nl, nu)), rng.standard_normal(nx)
    xk, uk = nr.kkt_dynamics_projection(AB, anc, nl, nx, nu, x0, xb, ub, affine_constants(d, o))
    xo, uo = o.dyn_project(x0, xb.reshape(-1), ub.reshape(-1))
    assert rel_err(xo, xk.reshape(-1)) < 1e-9
    assert rel_err(uo, uk.reshape(-1)) < 1e-9


@pytest.mark.parametrize("name", ["fixture", "small", "poly"])
def test_kernel_projection_is_orthogonal_onto_the_risk_kernel(problem_dirs, name):
    d = problem_dirs[name]
    o = Oracle(d)
    info = o.info
    nl, E = info["numNonleafNodes"], info["numEvents"]
    yD = 2 * E + 1
    ch_from = btio.read_flat(os.path.join(d, "childrenFrom_u.bt")).astype(int)
    num_ch = btio.read_flat(os.path.join(d, "numChildren_u.bt")).astype(int)
    cond = btio.read_flat(os.path.join(d, "conditionalProbabilities_d.bt"))
    alpha = info["risk"]["alpha"]
    rng = np.random.default_rng(2)
    v = rng.standard_normal(o.n_prim)
    kv = o.kernel_project(v)
    assert rel_err(o.kernel_project(kv), kv) < 1e-12
    w = rng.standard_normal(o.n_prim)
    assert abs(kv @ w - v @ o.kernel_project(w)) <= 1e-10 * np.linalg.norm(v) * np.linalg.norm(w)
    oY, oT, oS = o.sU + o.sX, o.sU + o.sX + o.sY, o.sU + o.sX + o.sY + o.sT
    for i in range(nl):
        n, c0 = num_ch[i], ch_from[i]
        Emat, _ = build.AVaR(alpha).make(cond[c0:c0 + n])
        y = kv[oY + i * yD:oY + i * yD + 2 * n + 1]
        t = kv[oT + c0:oT + c0 + n]
        s = kv[oS + c0:oS + c0 + n]
        assert np.max(np.abs(Emat.T @ y - t - s)) < 1e-10 * max(1.0, np.max(np.abs(v)))


def test_soc_known_answers():
    """tests/cuda/testProjections.cu:150-170."""
    assert soc([1, 2, 3, 4, 0.5]) == pytest.approx(
        [0.5456435464587639, 1.0912870929175278, 1.6369306393762917, 2.1825741858350556, 2.988612787525831],
        abs=1e-12)
    assert soc([5, 6, 7, 8, -200]) == pytest.approx([0, 0, 0, 0, 0], abs=1e-12)
    assert soc([9, -10, 11, -12, 100]) == pytest.approx([9, -10, 11, -12, 100], abs=1e-12)


def test_soc_matches_textbook_projection():
    rng = np.random.default_rng(3)
    for dim in (2, 3, 8, 72, 1034):
        for scale in (0.1, 1.0, 10.0):
            v = rng.standard_normal(dim)
            v[-1] *= scale * np.sqrt(dim)
            p = soc(v)
            assert rel_err(p, nr.soc_project(v)) < 1e-12
            assert np.linalg.norm(p[:-1]) <= p[-1] * (1 + 1e-12) + 1e-12  # inside the cone
            assert abs((v - p) @ p) <= 1e-9 * max(1.0, np.linalg.norm(v) ** 2)  # orthogonality


def test_linear_cost_projection_is_the_reference_clamp(problem_dirs):
    """The linear cost's dual block is (q'x + r'u of every node; t of every node), and its
    "projection" is k_projectRectangle with lower bound -inf and the t entry as the upper bound
    (src/costs.cuh:296-301, src/definitions.cu:39-48): a cost entry above its t becomes t, t is kept.
    The Euclidean projection onto {c <= t} would move both to their mean instead; the oracle
    follows the reference, not the textbook."""
    o = Oracle(problem_dirs["linear"])
    N = o.info["numNodes"]
    assert o.sIV == 2 * N
    oIV = o.sI + o.sII + o.sIII
    cases = [(3.0, 1.0, 1.0), (-2.0, 1.0, -2.0), (1.0, 1.0, 1.0), (0.5, -4.0, -4.0)]
    w = np.zeros(o.n_dual)
    for j, (c, t, _) in enumerate(cases):
        w[oIV + j], w[oIV + N + j] = c, t
    p = o.project_dual(w)
    for j, (c, t, want) in enumerate(cases):
        assert p[oIV + j] == want and p[oIV + N + j] == t
    assert (p[oIV + 0], p[oIV + N + 0]) != (2.0, 2.0)  # the Euclidean projection of (3, 1)
    rest = np.ones(o.n_dual, dtype=bool)
    rest[[oIV + j for j in range(len(cases))] + [oIV + N + j for j in range(len(cases))]] = False
    assert np.all(p[rest] == 0.0)


def test_box_known_answer():
    """tests/cuda/testKernels.cu:14-36 (with an infinite first bound)."""
    n, ub, oob = 1025, 4.0, 0.1
    v = np.full(n, ub + oob)
    v[::2] *= -1
    v[::3] *= 0.5
    lbs, ubs = np.full(n, -ub), np.full(n, ub)
    lbs[0], ubs[0] = -np.inf, np.inf
    out = np.array([box(v[i], lbs[i], ubs[i]) for i in range(n)])
    assert out[0] == pytest.approx(-(ub + oob) * 0.5, abs=1e-15)
    assert np.all(out >= -ub) and np.all(out <= ub)
    assert np.array_equal(out[1:], np.clip(v[1:], -ub, ub))


def test_avar_mask_known_answer(problem_dirs):
    """tests/cuda/testRisks.cu:46-49: projecting -10 leaves exactly these entries of the I block."""
    o = Oracle(problem_dirs["fixture"])
    w = np.zeros(o.n_dual)
    w[:o.sI] = -10.0
    p = o.project_dual(w)[:o.sI]
    expected = np.zeros(o.sI)
    expected[[4, 9, 14, 17, 18, 19, 22, 23, 24, 27, 28, 29, 32, 33, 34]] = -10.0
    assert np.array_equal(p, expected)


def test_least_squares_against_numpy():
    rng = np.random.default_rng(4)
    for n, m in ((10, 1), (500, 3), (5000, 5)):
        Y = rng.standard_normal((n, m))
        r = rng.standard_normal(n)
        ref = np.linalg.lstsq(Y, r, rcond=None)[0]
        assert rel_err(lstsq(Y, r), ref) < 1e-12
        assert rel_err(lstsq_gram(Y, r), ref) < 1e-9


@pytest.mark.parametrize("name", ["fixture", "small", "poly"])
def test_dotM_definition(problem_dirs, name):
    o = Oracle(problem_dirs[name])
    rng = np.random.default_rng(5)
    x = inert_zero(o, rng.standard_normal(o.n_iter))
    y = inert_zero(o, rng.standard_normal(o.n_iter))
    p = o.n_prim
    ref = x @ y - o.gamma * (x[:p] @ o.Ladj(y[p:]) + x[p:] @ o.L(y[:p]))
    assert o.dotM(x, y) == pytest.approx(ref, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("name", ["fixture", "small", "poly", "nocon", "affine", "linear"])
def test_T_lands_on_the_dynamics(problem_dirs, name):
    """The primal part of T(z) is a projection onto the dynamics with x_0 = x0."""
    d = problem_dirs[name]
    o = Oracle(d)
    N, nl, nx, nu = o.info["numNodes"], o.info["numNonleafNodes"], o.nx, o.nu
    AB = btio.read_bt(os.path.join(d, "dynamics_AB_d.bt")).transpose(2, 0, 1)
    anc = btio.read_flat(os.path.join(d, "ancestors_u.bt")).astype(np.int64)
    rng = np.random.default_rng(6)
    x0 = rng.uniform(-1, 1, nx)
    z = inert_zero(o, rng.standard_normal(o.n_iter))
    tz = o.T(o.scaled_x0(x0), z)
    x = tz[o.sU:o.sU + o.sX].reshape(N, nx)
    u = tz[:o.sU].reshape(nl, nu)
    e = affine_constants(d, o)
    assert rel_err(x[0], o.scaled_x0(x0)) < 1e-14
    for j in range(1, N):
        a = anc[j]
        res = x[j] - AB[j][:, :nx] @ x[a] - AB[j][:, nx:] @ u[a] - e[j]
        assert np.max(np.abs(res)) < 1e-10 * max(1.0, np.abs(x).max())


def test_converged_solve_is_a_fixed_point(problem_dirs):
    d = problem_dirs["small"]
    o = Oracle(d)
    x0 = configs.initial_states(1, o.nx, 1, "uniform")[0]
    r = o.solve(x0, tol_abs=1e-4, tol_rel=1e-4, max_iters=20000)
    assert r["status"] == 0
    z = r["z"]
    assert np.max(np.abs(z - o.T(o.scaled_x0(x0), z))) <= 5e-4 * np.max(np.abs(z))


def test_golden_vectors_reproduce(tmp_path):
    """tests/golden/fixture_golden.npz, written by tests/golden/make_golden.py from the oracle."""
    g = np.load(os.path.join(GOLDEN, "fixture_golden.npz"), allow_pickle=False)
    d = str(tmp_path / "fixture")
    configs.generate("fixture", d)
    o = Oracle(d)
    assert rel_err(o.L(g["prim"]), g["L_prim"]) < 1e-14
    assert rel_err(o.Ladj(g["dual"]), g["Ladj_dual"]) < 1e-14
    assert rel_err(o.T(o.scaled_x0(g["x0"]), g["z"]), g["T_z"]) < 1e-14
    r = o.solve(g["x0"], tol_abs=1e-30, max_iters=75)
    assert rel_err(r["z"], g["spock75"]) < 1e-12
    assert [r[k] for k in ("k0", "k1", "k2", "k2bt", "k3")] == list(g["spock75_counts"])


@pytest.mark.parametrize("name", ["small", "poly", "pwi"])
def test_converged_solution_solves_the_conic_program(tmp_path, name):
    """An independent optimality check of the whole iteration (CP step, dual prox with its
    translations, SuperMann and Anderson).  The RAOCP in conic form is built from the generator's
    own L (``spock_amd.api.problem``, a separate implementation from the oracle's) and the problem
    data:
      min s_0  s.t.  the dynamics with x_0 fixed, (I - NN'_i) [y_i; t_ch(i); s_ch(i)] = 0 for every
                     nonleaf i, and L z + c in C,
    with C the product of the AV@R cone (the first 2 #children entries of every y block >= 0), the
    nonnegative orthant, the boxes (the root's state bounds relaxed, src/constraints.cuh:142-148)
    and the second-order cones of the costs, c the SOC translations.  The oracle's tightly converged
    solution must be feasible for it, and SLSQP started there must not find a lower cost."""
    from scipy.optimize import minimize
    d = str(tmp_path / name)
    gen = configs.generate(name, d)
    o = Oracle(d)
    x0 = {"small": configs.initial_states(1, o.nx, 1, "uniform")[0],
          "poly": configs.initial_states(33, o.nx, 1, "small")[0],
          "pwi": configs.initial_states(33, o.nx, 1, "uniform")[0]}[name]
    # SPOCK levels off near an error of 1e-5 on these problems; 20,000 iterations reach it
    ref = o.solve(x0, tol_abs=1e-7, tol_rel=0.0, max_iters=20000)
    assert ref["err"] < 1e-4
    zp = ref["z"][:o.n_prim].copy()
    info = o.info
    N, nl, nx, nu, E = info["numNodes"], info["numNonleafNodes"], o.nx, o.nu, info["numEvents"]
    yD = 2 * E + 1
    nlf = N - nl
    Lm = np.column_stack([gen.op(np.eye(o.n_prim)[k]) for k in range(o.n_prim)])
    # translations and block offsets of the dual
    c = np.zeros(o.n_dual)
    c3 = 0 if gen.ncon.kind == "no" else gen.ncon.dim_per_node
    c5 = 0 if gen.lcon.kind == "no" else gen.lcon.dim_per_node
    oI, oII = 0, nl * yD
    oIII = oII + nl
    oIV = oIII + c3 * nl
    oV = oIV + (nx + nu + 2) * N
    oVI = oV + c5 * nlf
    d4, d6 = nx + nu + 2, nx + 2
    for j in range(1, N):
        c[oIV + j * d4:oIV + (j + 1) * d4] = gen.ncost.translation[j].ravel()
    for l in range(nlf):
        c[oVI + l * d6:oVI + (l + 1) * d6] = gen.lcost.translation[-1].ravel()
    numch = gen.tree.num_children
    # linear inequalities G z + h >= 0 from the orthant and box parts of the cone
    rows, hs = [], []
    for i in range(nl):
        for k in range(2 * numch[i]):
            rows.append(Lm[oI + i * yD + k]); hs.append(c[oI + i * yD + k])
        rows.append(Lm[oII + i]); hs.append(c[oII + i])

    def block_bounds(con, count, relax_root_state):
        """Bounds of a constraint's dual block in its layout: per node for a rectangle or a
        polyhedron; identity parts of all nodes, then G parts, for a polyhedron with identity."""
        if con.kind == "polyhedronWithIdentity":
            return (np.concatenate([np.tile(con.rect.lower_bound.ravel(), count),
                                    np.tile(con.poly.lower_bound.ravel(), count)]),
                    np.concatenate([np.tile(con.rect.upper_bound.ravel(), count),
                                    np.tile(con.poly.upper_bound.ravel(), count)]))
        lb, ub = np.tile(con.lower_bound.ravel(), count), np.tile(con.upper_bound.ravel(), count)
        if relax_root_state:
            lb[:nx], ub[:nx] = -np.inf, np.inf
        return lb, ub

    def boxes(off, lb, ub):
        for k in range(lb.size):
            r = Lm[off + k]
            if np.isfinite(lb[k]):
                rows.append(r); hs.append(-lb[k])
            if np.isfinite(ub[k]):
                rows.append(-r); hs.append(ub[k])
    if gen.ncon.kind != "no":
        boxes(oIII, *block_bounds(gen.ncon, nl, gen.ncon.kind == "rectangle"))
    if gen.lcon.kind != "no":
        boxes(oV, *block_bounds(gen.lcon, nlf, False))
    G, h = np.array(rows), np.array(hs)
    socs = [(oIV + j * d4, d4) for j in range(1, N)] + [(oVI + l * d6, d6) for l in range(nlf)]

    def soc_val(z):
        w = Lm @ z + c
        return np.array([w[a + n - 1] - np.sqrt(w[a:a + n - 1] @ w[a:a + n - 1] + 1e-18) for a, n in socs])

    def soc_jac(z):
        w = Lm @ z + c
        out = []
        for a, n in socs:
            v = w[a:a + n - 1]
            nv = np.sqrt(v @ v + 1e-18)
            out.append(Lm[a + n - 1] - (v / nv) @ Lm[a:a + n - 1])
        return np.array(out)
    # equalities: dynamics with x_0 = x0 and the risk kernel constraints
    AB = btio.read_bt(os.path.join(d, "dynamics_AB_d.bt")).transpose(2, 0, 1)
    anc = gen.tree.ancestors_arr
    oU, oX = 0, nu * nl
    oY = oX + nx * N
    oT = oY + yD * nl
    oS = oT + N
    eq, beq = [], []
    x0s = o.scaled_x0(x0)
    for k in range(nx):
        r = np.zeros(o.n_prim); r[oX + k] = 1.0; eq.append(r); beq.append(x0s[k])
    for j in range(1, N):
        a = anc[j]
        for k in range(nx):
            r = np.zeros(o.n_prim)
            r[oX + j * nx + k] = 1.0
            r[oX + a * nx:oX + (a + 1) * nx] -= AB[j][k, :nx]
            r[oU + a * nu:oU + (a + 1) * nu] -= AB[j][k, nx:]
            eq.append(r); beq.append(0.0)
    NN = btio.read_bt(os.path.join(d, "risk_NNtr_d.bt")).transpose(2, 0, 1)
    nD = 4 * E + 1
    for i in range(nl):
        sel = np.zeros((nD, o.n_prim))
        for k in range(yD):
            sel[k, oY + i * yD + k] = 1.0
        for m, ch in enumerate(gen.tree.children[i]):
            sel[yD + m, oT + ch] = 1.0
            sel[yD + E + m, oS + ch] = 1.0
        for r in (np.eye(nD) - NN[i]) @ sel:
            eq.append(r); beq.append(0.0)
    Aeq, beq = np.array(eq), np.array(beq)
    # the kernel rows are dependent (I - NN' is a projector): keep an orthonormal row basis
    U, sv, Vt = np.linalg.svd(Aeq, full_matrices=False)
    rank = int(np.sum(sv > 1e-10 * sv[0]))
    Aeq, beq = Vt[:rank], (U[:, :rank].T @ beq) / sv[:rank]
    scale = max(1.0, np.abs(zp).max())
    # the oracle's solution is feasible (to its residual) ...
    viol = (np.abs(Aeq @ zp - beq).max(), -(G @ zp + h).min(), -soc_val(zp).min())
    print(f"{name}: s0 {zp[oS]:.8f}, violations eq {viol[0]:.2e} lin {viol[1]:.2e} soc {viol[2]:.2e}")
    assert viol[0] < 1e-9 * scale and viol[1] < 1e-4 * scale and viol[2] < 1e-4 * scale
    # the nested AV@R cost of its states and inputs alone is its s_0
    assert abs(gen.trajectory_cost(zp) - zp[oS]) <= 1e-4 * max(1.0, abs(zp[oS]))
    # ... and SLSQP, started away from it, finds the same optimal cost
    cons = [{"type": "eq", "fun": lambda z: Aeq @ z - beq, "jac": lambda z: Aeq},
            {"type": "ineq", "fun": lambda z: G @ z + h, "jac": lambda z: G},
            {"type": "ineq", "fun": soc_val, "jac": soc_jac}]
    e_s0 = np.zeros(o.n_prim); e_s0[oS] = 1.0
    start = zp + 0.05 * scale * np.random.default_rng(4).standard_normal(o.n_prim)
    res = minimize(lambda z: z[oS], start, jac=lambda z: e_s0, constraints=cons, method="SLSQP",
                   options={"maxiter": 1000, "ftol": 1e-12})
    s_star = zp[oS]
    print(f"{name}: SLSQP s0 {res.fun:.8f} after {res.nit} iterations ({res.message})")
    assert res.success
    assert abs(res.fun - s_star) <= 1e-4 * max(1.0, abs(s_star))
    assert abs(gen.trajectory_cost(res.x) - res.fun) <= 1e-4 * max(1.0, abs(res.fun))
