"""The problem generator (spock.cu_amd/api) against the known answers of the reference's own tests:
tests/python/testTreeFactory.py (Markov chain and i.i.d. trees), tests/cuda/testTree.cu (the tree of
tests/cuda/test.py), and the AV@R data against the textbook definition of the risk.
"""
import os

import numpy as np
import pytest
from scipy.optimize import linprog

from spock_amd.api import btio, build, configs
from spock_amd.api.tree import IidProcess, MarkovChain


@pytest.fixture(scope="module")
def markov():
    p = np.array([[0.1, 0.8, 0.1], [0.4, 0.6, 0], [0, 0.3, 0.7]])
    v = np.array([0.5, 0.5, 0.0])
    return MarkovChain(p, v, 4, 3).build()


@pytest.fixture(scope="module")
def iid():
    return IidProcess(np.array([0.1, 0.2, 0.7]), 4, 2).build()


# tests/python/testTreeFactory.py:79-246
def test_markov_structure(markov):
    t = markov
    assert (t.num_nodes, t.num_nonleaf_nodes, t.num_events, t.num_stages) == (32, 20, 3, 5)
    for node, anc in [(1, 0), (2, 0), (3, 1), (4, 1), (5, 1), (6, 2), (7, 2), (8, 3), (9, 3), (10, 3), (13, 5)]:
        assert t.ancestor_of_node(node) == anc
    for i in range(12):
        assert t.ancestor_of_node(20 + i) == 8 + i
    for node, n in [(0, 2), (1, 3), (2, 2), (5, 2), (6, 3)]:
        assert len(t.children_of_node(node)) == n
    for idx in range(8, 20):
        assert len(t.children_of_node(idx)) == 1
    with pytest.raises(IndexError):
        t.children_of_node(20)
    assert t.stage_of_node(0) == 0
    assert [t.stage_of_node(i) for i in (1, 2)] == [1, 1]
    assert all(t.stage_of_node(i) == 2 for i in range(3, 8))
    assert all(t.stage_of_node(i) == 3 for i in range(8, 20))
    assert all(t.stage_of_node(i) == 4 for i in range(20, 32))
    with pytest.raises(ValueError):
        t.stage_of_node(-1)
    with pytest.raises(IndexError):
        t.stage_of_node(32)
    assert [len(t.nodes_of_stage(s)) for s in range(5)] == [1, 2, 5, 12, 12]
    assert list(t.nodes_of_stage(2)) == list(range(3, 8))
    assert list(t.nodes_of_stage(4)) == list(range(20, 32))


def test_markov_probabilities_and_events(markov):
    t = markov
    expect = {0: 1.0, 1: 0.5, 2: 0.5, 3: 0.05, 4: 0.4, 5: 0.05, 6: 0.2, 7: 0.3, 8: 0.005, 20: 0.005,
              29: 0.5 * 0.4 * 0.1}
    for node, pr in expect.items():
        assert t.probability_of_node(node) == pytest.approx(pr, abs=1e-10)
    assert [len(t.siblings_of_node(i)) for i in (0, 1, 2, 3, 7, 11)] == [1, 2, 2, 3, 2, 2]
    assert all(len(t.siblings_of_node(i)) == 1 for i in range(20, 32))
    assert list(t.event_of_node(range(1, 3))) == [0, 1]
    assert list(t.event_of_node(range(3, 8))) == [0, 1, 2, 0, 1]
    assert list(t.event_of_node(range(8, 20))) == [0, 1, 2, 0, 1, 1, 2, 0, 1, 2, 0, 1]
    assert list(t.event_of_node(range(8, 20))) == list(t.event_of_node(range(20, 32)))
    for stage in range(t.num_stages - 1):
        for node in t.nodes_of_stage(stage):
            assert sum(t.cond_prob_of_children_of_node(node)) == pytest.approx(1.0, abs=1e-10)


def test_markov_large_tree_and_stopping_stage():
    rng = np.random.default_rng(0)
    p = rng.random((4, 4))
    p /= p.sum(1, keepdims=True)
    v = rng.random(4)
    v /= v.sum()
    t = MarkovChain(p, v, 20, 5).build()
    for stage in range(t.num_stages - 1):
        for node in t.nodes_of_stage(stage):
            assert sum(t.cond_prob_of_children_of_node(node)) == pytest.approx(1.0, abs=1e-10)
    with pytest.raises(ValueError):
        MarkovChain(p[:3, :3] / p[:3, :3].sum(1, keepdims=True), v[:3] / v[:3].sum(), 4, 5).build()
    q = np.array([[0.1, 0.8, 0.1], [0.4, 0.6, 0], [0, 0.3, 0.7]])
    MarkovChain(q, np.array([0.5, 0.4, 0.1]), 1, 1).build()
    MarkovChain(q, np.array([0.5, 0.4, 0.1]), 3, 1).build()


# tests/python/testTreeFactory.py:248-400
def test_iid_tree(iid):
    t = iid
    assert (t.num_nodes, t.num_nonleaf_nodes, t.num_events, t.num_stages) == (31, 22, 3, 5)
    for node, anc in [(1, 0), (2, 0), (3, 0), (4, 1), (5, 1), (6, 1), (7, 2), (8, 2), (9, 2), (10, 3), (11, 3), (12, 3)]:
        assert t.ancestor_of_node(node) == anc
    for i in range(13, t.num_nodes):
        assert t.ancestor_of_node(i) == i - 9
    assert [len(t.children_of_node(i)) for i in range(4)] == [3, 3, 3, 3]
    assert all(len(t.children_of_node(i)) == 1 for i in range(4, t.num_nonleaf_nodes))
    with pytest.raises(IndexError):
        t.children_of_node(22)
    assert [len(t.nodes_of_stage(s)) for s in range(5)] == [1, 3, 9, 9, 9]
    probs = [1.0, .1, .2, .7, .01, .02, .07, .02, .04, .14, .07, .14, .49]
    for node, pr in enumerate(probs):
        assert t.probability_of_node(node) == pytest.approx(pr, abs=1e-10)
    for idx in range(13, t.num_nodes):
        assert t.probability_of_node(idx) == pytest.approx(t.probability_of_node(t.ancestor_of_node(idx)), abs=1e-10)
    assert all(len(t.siblings_of_node(i)) == 3 for i in range(1, 13))
    assert list(t.event_of_node(range(4, 13))) == [0, 1, 2] * 3
    assert list(t.event_of_node(range(4, 13))) == list(t.event_of_node(range(22, 31)))
    with pytest.raises(ValueError):
        IidProcess(np.array([0.2, 0.3, 0.5]), 4, 5).build()
    IidProcess(np.array([0.5, 0.4, 0.1]), 1, 1).build()


def test_fixture_tree_files(tmp_path):
    """tests/cuda/testTree.cu:42-176 on the files written for the tree of tests/cuda/test.py."""
    d = str(tmp_path / "fixture")
    configs.generate("fixture", d)
    rd = lambda name: btio.read_flat(os.path.join(d, name))
    assert len(rd("stages_u.bt")) == 11
    assert int(rd("numChildren_u.bt").size) == 7
    assert int(rd("stages_u.bt").max()) + 1 == 4
    assert int(rd("stages_u.bt")[5]) == 2
    assert int(rd("ancestors_u.bt")[5]) == 2
    assert rd("probabilities_d.bt")[4] == pytest.approx(0.21)
    assert rd("conditionalProbabilities_d.bt")[5] == pytest.approx(0.3)
    assert int(rd("events_u.bt")[4]) == 1
    assert int(rd("childrenFrom_u.bt")[2]) == 5
    assert int(rd("childrenTo_u.bt")[2]) == 6
    assert int(rd("numChildren_u.bt")[2]) == 2
    assert int(rd("stageFrom_u.bt")[2]) == 3
    assert int(rd("stageTo_u.bt")[2]) == 6


def test_bt_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((3, 4, 5))
    btio.write_bt(str(tmp_path / "a_d.bt"), a, "d")
    assert np.array_equal(btio.read_bt(str(tmp_path / "a_d.bt")), a)
    u = rng.integers(0, 1000, 17).astype(np.int64)
    btio.write_bt(str(tmp_path / "u_u.bt"), u, "u")
    assert np.array_equal(btio.read_flat(str(tmp_path / "u_u.bt")).astype(np.int64), u)


@pytest.mark.parametrize("alpha", [0.9, 0.95, 0.5])
def test_avar_data_gives_the_textbook_risk(alpha):
    """min b'y s.t. E'y = Z, y_1..2n >= 0, y_2n+1 free (the masked cone of the dual prox) equals
    AV@R_alpha(Z) = max{sum mu Z : 0 <= mu <= pi/alpha, sum mu = 1}, which AVaR.value computes
    greedily."""
    rng = np.random.default_rng(int(alpha * 100))
    for n in (2, 3, 5):
        pi = rng.dirichlet(np.ones(n))
        Z = rng.standard_normal(n)
        E, b = build.AVaR(alpha).make(pi)
        bounds = [(0, None)] * (2 * n) + [(None, None)]
        lp = linprog(b.ravel(), A_eq=E.T, b_eq=Z, bounds=bounds, method="highs")
        assert lp.status == 0
        assert lp.fun == pytest.approx(build.AVaR(alpha).value(pi, Z), abs=1e-9)


def test_avar_value_limits():
    """AV@R_1 is the expectation; AV@R_alpha tends to the maximum as alpha -> 0 (and is the
    maximum at 0); an outcome with zero probability never carries weight below the maximum."""
    rng = np.random.default_rng(5)
    for n in (1, 2, 4, 7):
        pi = rng.dirichlet(np.ones(n))
        Z = rng.standard_normal(n)
        assert build.AVaR(1.0).value(pi, Z) == pytest.approx(float(pi @ Z), abs=1e-12)
        assert build.AVaR(0.0).value(pi, Z) == Z.max()
        assert build.AVaR(min(pi.min(), 1e-3)).value(pi, Z) == pytest.approx(Z.max(), abs=1e-12)
        vals = [build.AVaR(a).value(pi, Z) for a in (0.9, 0.5, 0.2)]
        assert vals == sorted(vals)  # more averse, larger
    assert build.AVaR(0.5).value([0.5, 0.5, 0.0], [1.0, 2.0, 9.0]) == pytest.approx(2.0)


def test_workload_sizes(tmp_path):
    """The 229-node tree (3 events, horizon 10, branching stops at stage 3) and the 31-node tree
    (2 events, horizon 4)."""
    d = str(tmp_path / "mid")
    configs.generate("mid", d)
    st = btio.read_flat(os.path.join(d, "stages_u.bt"))
    assert len(st) == 229 and int((st < st.max()).sum()) == 202
    d = str(tmp_path / "small")
    configs.generate("small", d)
    st = btio.read_flat(os.path.join(d, "stages_u.bt"))
    assert len(st) == 31 and int((st < st.max()).sum()) == 15


def test_from_structure_matches_the_factory_tree(tmp_path):
    """FromStructure (api/python/spock/tree.py:902-951) rebuilds a Markov tree from its stage,
    ancestor and probability arrays: same tree files apart from the events (all 0 there)."""
    from spock_amd.api import btio
    from spock_amd.api.tree import FromStructure
    t = MarkovChain(np.array([[0.3, 0.7], [0.4, 0.6]]), np.array([0.5, 0.5]), 3, 2).build()
    a, b = tmp_path / "a", tmp_path / "b"
    t.write_files(str(a))
    FromStructure(t.stages_arr, t.ancestors_arr, t.prob).build(str(b))
    for name in ("stages_u", "ancestors_u", "childrenFrom_u", "childrenTo_u", "numChildren_u", "stageFrom_u",
                 "stageTo_u", "probabilities_d", "conditionalProbabilities_d"):
        assert np.array_equal(btio.read_flat(str(a / f"{name}.bt")), btio.read_flat(str(b / f"{name}.bt")))
    ev = btio.read_flat(str(b / "events_u.bt")).astype(np.int64)
    assert ev[0] == -1 and np.all(ev[1:] == 0)
    with pytest.raises(ValueError, match="do not sum"):
        FromStructure([0, 1, 1, 2, 2], [-1, 0, 0, 1, 2], [1.0, 0.5, 0.4, 0.5, 0.4]).build()
    with pytest.raises(ValueError, match="No root"):
        FromStructure([1, 1], [-1, 0], [1.0, 1.0]).build()


def test_from_data_known_answers():
    """FromData on the reference's test data (tests/python/testTreeFactory.py:30-37): 10 samples of
    a 1-D variable over 4 stages, branching [3, 2, 1, 1].  Known answers from :404-512."""
    from spock_amd.api.tree import FromData
    col = 0.1 * np.arange(10.0).reshape(10, 1)
    data = np.dstack((0.1 * col, 0.2 * col, 0.3 * col, 0.4 * col))
    t = FromData(data, [3, 2, 1, 1]).build()
    assert (t.num_nodes, t.num_nonleaf_nodes, t.num_stages) == (22, 16, 5)
    assert list(t.ancestors_arr[:10]) == [-1, 0, 0, 0, 1, 1, 2, 2, 3, 3]
    assert all(t.ancestors_arr[i] == i - 6 for i in range(10, 22))
    assert list(t.num_children) == [3, 2, 2, 2] + [1] * 12
    for s, (lo, hi) in enumerate([(0, 1), (1, 4), (4, 10), (10, 16), (16, 22)]):
        assert list(t.nodes_of_stage(s)) == list(range(lo, hi))
    np.testing.assert_allclose(t.prob[:10], [1.0, 0.4, 0.3, 0.3, 0.3, 0.1, 0.2, 0.1, 0.2, 0.1], atol=1e-10)
    for i in range(10, 22):
        assert abs(t.prob[i] - t.prob[t.ancestors_arr[i]]) < 1e-10
    assert np.all(t.events_arr[1:] == 0)


def test_from_data_conditional_probabilities_sum_to_one():
    """tests/python/testTreeFactory.py:520-531 on a seeded draw: 10 samples over 20 stages,
    branching 4 for 5 stages then 1."""
    from spock_amd.api.tree import FromData
    data = np.random.default_rng(3).random((10, 1, 20))
    branching = np.ones(20)
    branching[:5] *= 4
    t = FromData(data, branching).build()
    for i in range(t.num_nonleaf_nodes):
        assert abs(t.cond_prob[t.children[i]].sum() - 1.0) < 1e-10


def test_julia_export_is_the_unconditioned_data(tmp_path):
    """with_julia() also writes the data before preconditioning (api/python/spock/problem.py:605-684):
    the conditioned dynamics_AB of every node must be Sx [A B] diag(Sx^-1, Su^-1) of the exported
    A and B, the exported costs and bounds must be the ones given, and the linear terms appear
    only for the quadratic-plus-linear kind."""
    from spock_amd.api import problem as problem_mod
    rng = np.random.default_rng(9)
    nx, nu = 3, 2
    tr = MarkovChain(np.array([[0.6, 0.4], [0.3, 0.7]]), np.array([0.5, 0.5]), 3, 3).build()
    dyn = [build.Dynamics(np.eye(nx) + 0.1 * rng.standard_normal((nx, nx)), rng.standard_normal((nx, nu)),
                          0.1 * rng.standard_normal(nx)) for _ in range(2)]
    Q, R, q, r = 4.0 * np.eye(nx), 0.5 * np.eye(nu), rng.standard_normal(nx), rng.standard_normal(nu)
    costs = [build.CostQuadraticPlusLinear(Q, q, R, r) for _ in range(2)]
    leaf = build.CostQuadratic(9.0 * np.eye(nx), leaf=True)
    lim = np.vstack((2.0 * np.ones((nx, 1)), np.ones((nu, 1))))
    pwi = build.PolyhedronWithIdentity(build.Rectangle(-lim, lim),
                                       build.Polyhedron(rng.standard_normal((4, nx + nu)), -3 * np.ones(4), 3 * np.ones(4)))
    d = str(tmp_path / "julia")
    gen = (problem_mod.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
           .with_cost_leaf(leaf).with_constraint_nonleaf(pwi)
           .with_constraint_leaf(build.Rectangle(-2 * np.ones(nx), 2 * np.ones(nx)))
           .with_risk(build.AVaR(0.9)).with_preconditioning().with_julia().generate_problem(d))
    assert gen.preconditioned

    def rd(name):
        return btio.read_bt(os.path.join(d, name + "_d.bt"))
    A, B, c, AB = rd("uncond_dynamics_A"), rd("uncond_dynamics_B"), rd("uncond_dynamics_c"), rd("dynamics_AB")
    sx, su = gen.scaling[:nx], gen.scaling[nx:]
    events = tr.events_arr
    for j in range(1, tr.num_nodes):
        e = int(events[j])
        assert np.allclose(A[:, :, j], dyn[e].A0) and np.allclose(B[:, :, j], dyn[e].B0)
        assert np.allclose(c[:, 0, j], dyn[e].c0.ravel())
        cond = np.diag(sx) @ np.hstack((A[:, :, j], B[:, :, j])) @ np.diag(np.concatenate((1 / sx, 1 / su)))
        assert np.allclose(AB[:, :, j], cond)
    assert np.allclose(rd("uncond_cost_nonleaf_Q")[:, :, 1], Q) and np.allclose(rd("uncond_cost_nonleaf_R")[:, :, 1], R)
    assert np.allclose(rd("uncond_cost_nonleaf_q")[:, 0, 1], q) and np.allclose(rd("uncond_cost_nonleaf_r")[:, 0, 1], r)
    assert np.allclose(rd("uncond_cost_leaf_Q")[:, :, 0], 9.0 * np.eye(nx))
    assert not os.path.exists(os.path.join(d, "uncond_cost_leaf_q_d.bt"))
    assert np.allclose(rd("uncond_nonleafConstraintILB").ravel(), -lim.ravel())
    assert np.allclose(rd("uncond_nonleafConstraintGamma")[:, :, 0], pwi.poly.G0)
    assert np.allclose(rd("uncond_leafConstraintIUB").ravel(), 2 * np.ones(nx))
    # the conditioned files hold the scaled data: bounds times the scaling
    assert np.allclose(rd("nonleafConstraintILB").ravel(), -lim.ravel() * gen.scaling)
