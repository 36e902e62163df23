"""Seeded generators for the benchmark problem families and the unit-test fixture.

Every family writes the reference's file set (``data.json`` + ``*.bt``) into a
folder and returns the generated ``Problem``.  Random draws use
``numpy.random.default_rng(seed)``; initial states come from an independent
stream ``default_rng([seed, 1])`` so that the problem data do not depend on how
many initial states are requested.

Choices the family descriptions leave open, fixed here:
* Markov initial distributions and transition rows are Dirichlet(1) draws.
* The terminal weight of the 50/20 family is the discrete Riccati solution of
  event 0's (A, B) with Q = I, R = 0.1 I.
* Leaf boxes repeat the nonleaf state bounds.
* The polyhedral family (4-event i.i.d. tree, nx=100, nu=40): the i.i.d. distribution is a
  Dirichlet(1) draw, Q = I, R = 0.1 I, leaf P = I, the linear weights q, r have standard
  deviation 0.1, and the 2 nx constraint rows (over [x; u] at nonleaf nodes, over x at leaves)
  have N(0, 1) entries with bounds -5 <= row . v <= 5.  Initial states for it are drawn with
  standard deviation 0.05 (``initial_states(..., kind="small")``): at 0.5 the rows of the
  polyhedron at the root would be violated with no input able to fix them.
"""
import os
import numpy as np
import scipy.linalg as sla

from . import build, problem, tree as tree_mod


def fixture_problem(folder, dt="d"):
    """The reference's unit-test problem (``tests/cuda/test.py:15-99``).

    Its nonleaf weights are built as ``q *= i`` on one shared array, so both
    events end up with Q = 2I (``tests/cuda/test.py:57-61``); that is kept.
    """
    p = np.array([[0.3, 0.7], [0.3, 0.7]])
    v = np.array([0.3, 0.7])
    tr = tree_mod.MarkovChain(p, v, 3, 2, dt=dt).build()
    nx, nu = 3, 2
    A, B = np.eye(nx), 0.5 * np.ones((nx, nu))
    dyn = [build.Dynamics(1.5 * A, 1.5 * B), build.Dynamics(A, B)]
    Q2 = 2.0 * np.eye(nx)
    costs = [build.CostQuadratic(Q2, np.eye(nu)), build.CostQuadratic(Q2, np.eye(nu))]
    leaf = build.CostQuadratic(3 * np.eye(nx), leaf=True)
    si_lb = np.vstack((-1.0 * np.ones((nx, 1)), -5.0 * np.ones((nu, 1))))
    ncon = build.Rectangle(si_lb, -si_lb)
    lcon = build.Rectangle(-0.1 * np.ones((nx, 1)), 0.1 * np.ones((nx, 1)))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def _markov(rng, E, horizon, stop, dt):
    P = rng.dirichlet(np.ones(E), size=E)
    v = rng.dirichlet(np.ones(E))
    return tree_mod.MarkovChain(P, v, horizon, stop, dt=dt).build()


def _box_family(folder, rng, tr, nx, nu, a_std, b_std, x_lim, u_lim, alpha, leaf_dare, dt):
    dyn = []
    for _ in range(tr.num_events):
        A = np.eye(nx) + a_std * rng.standard_normal((nx, nx))
        B = b_std * rng.standard_normal((nx, nu))
        dyn.append(build.Dynamics(A, B))
    Q, R = np.eye(nx), 0.1 * np.eye(nu)
    if leaf_dare:
        Pf = sla.solve_discrete_are(dyn[0].A0, dyn[0].B0, Q, R)
        Pf = 0.5 * (Pf + Pf.T)
    else:
        Pf = np.eye(nx)
    costs = [build.CostQuadratic(Q, R) for _ in range(tr.num_events)]
    leaf = build.CostQuadratic(Pf, leaf=True)
    lim = np.vstack((x_lim * np.ones((nx, 1)), u_lim * np.ones((nu, 1))))
    ncon = build.Rectangle(-lim, lim)
    lcon = build.Rectangle(-x_lim * np.ones((nx, 1)), x_lim * np.ones((nx, 1)))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(alpha)).with_preconditioning().generate_problem(folder))


def small_markov(folder, seed=1, dt="d"):
    """2 events, horizon 4, branching at every stage; nx=3, nu=2; |x|<=5, |u|<=2; AVaR 0.95."""
    rng = np.random.default_rng(seed)
    tr = _markov(rng, 2, 4, 4, dt)
    return _box_family(folder, rng, tr, 3, 2, 0.1, 0.5, 5.0, 2.0, 0.95, False, dt)


def mid_markov(folder, seed=2, dt="d"):
    """3 events, horizon 10, branching stops at stage 3; nx=50, nu=20; |x|<=10, |u|<=1; AVaR 0.9."""
    rng = np.random.default_rng(seed)
    tr = _markov(rng, 3, 10, 3, dt)
    return _box_family(folder, rng, tr, 50, 20, 0.02, 0.1, 10.0, 1.0, 0.9, True, dt)


def uneven_instance(folder, k, dt="d"):
    """Instance k of the uneven batch: horizon U{5..12}, events U{2..4}, stop U{1..3},
    nx U{10..40}, nu U{4..16}, AVaR alpha U[0.8, 0.99]; data drawn like the 50/20 family."""
    rng = np.random.default_rng(4000 + k)
    horizon = int(rng.integers(5, 12, endpoint=True))
    E = int(rng.integers(2, 4, endpoint=True))
    stop = int(rng.integers(1, 3, endpoint=True))
    nx = int(rng.integers(10, 40, endpoint=True))
    nu = int(rng.integers(4, 16, endpoint=True))
    alpha = float(rng.uniform(0.8, 0.99))
    tr = _markov(rng, E, horizon, stop, dt)
    return _box_family(folder, rng, tr, nx, nu, 0.02, 0.1, 10.0, 1.0, alpha, True, dt)


def polyhedral_iid(folder, seed=3, dt="d", events=4, horizon=6, nx=100, nu=40):
    """i.i.d. tree with branching at every stage; quadratic-plus-linear costs; polyhedral
    constraints of 2 nx rows; AV@R 0.95.  The defaults are the 5461-node nx=100/nu=40 tree."""
    rng = np.random.default_rng(seed)
    tr = tree_mod.IidProcess(rng.dirichlet(np.ones(events)), horizon, horizon, dt=dt).build()
    dyn = []
    for _ in range(events):
        A = np.eye(nx) + 0.02 * rng.standard_normal((nx, nx))
        B = 0.1 * rng.standard_normal((nx, nu))
        dyn.append(build.Dynamics(A, B))
    Q, R = np.eye(nx), 0.1 * np.eye(nu)
    costs = [build.CostQuadraticPlusLinear(Q, 0.1 * rng.standard_normal(nx), R, 0.1 * rng.standard_normal(nu))
             for _ in range(events)]
    leaf = build.CostQuadraticPlusLinear(np.eye(nx), 0.1 * rng.standard_normal(nx), leaf=True)
    rows = 2 * nx
    ncon = build.Polyhedron(rng.standard_normal((rows, nx + nu)), -5.0 * np.ones(rows), 5.0 * np.ones(rows))
    lcon = build.Polyhedron(rng.standard_normal((rows, nx)), -5.0 * np.ones(rows), 5.0 * np.ones(rows))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def polyhedral_small(folder, seed=33, dt="d"):
    """The polyhedral family at test size: 3 events, horizon 3 (40 nodes), nx=4, nu=2."""
    return polyhedral_iid(folder, seed=seed, dt=dt, events=3, horizon=3, nx=4, nu=2)


def polyhedron_identity_small(folder, seed=44, dt="d"):
    """The reference's polyhedron-with-identity constraint on the 40-node i.i.d. tree (3 events,
    horizon 3), nx=4, nu=2, quadratic costs, AV@R 0.95: at nonleaf nodes |x| <= 2, |u| <= 0.3 and
    2 (nx + nu) random rows within +-3.3; at leaves |x| <= 2 and 2 nx rows within +-3.3.  The
    identity part keeps the root state bounds, so the bounds are loose enough for x0 in [-1, 1]^4;
    at the solution from such an x0 both parts have active rows."""
    rng = np.random.default_rng(seed)
    events, nx, nu = 3, 4, 2
    tr = tree_mod.IidProcess(rng.dirichlet(np.ones(events)), 3, 3, dt=dt).build()
    dyn = [build.Dynamics(np.eye(nx) + 0.1 * rng.standard_normal((nx, nx)), 0.5 * rng.standard_normal((nx, nu)))
           for _ in range(events)]
    costs = [build.CostQuadratic(np.eye(nx), 0.1 * np.eye(nu)) for _ in range(events)]
    leaf = build.CostQuadratic(np.eye(nx), leaf=True)
    lim = np.vstack((2.0 * np.ones((nx, 1)), 0.3 * np.ones((nu, 1))))
    rn, rl = 2 * (nx + nu), 2 * nx
    ncon = build.PolyhedronWithIdentity(build.Rectangle(-lim, lim),
                                        build.Polyhedron(rng.standard_normal((rn, nx + nu)), -3.3 * np.ones(rn),
                                                         3.3 * np.ones(rn)))
    lcon = build.PolyhedronWithIdentity(build.Rectangle(-2.0 * np.ones((nx, 1)), 2.0 * np.ones((nx, 1))),
                                        build.Polyhedron(rng.standard_normal((rl, nx)), -3.3 * np.ones(rl),
                                                         3.3 * np.ones(rl)))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def linear_cost_markov(folder, seed=61, dt="d"):
    """The reference's linear-cost problem (``tests/python/testProblemFactory.py:13-78``): the
    3-state Markov chain (P, v) of that test, horizon 4 branching until stage 3 (the same tree),
    nx=3, nu=2, per-event nonleaf costs linear with q = 10, 20, 30 and r = 1, 2, 3 (all entries),
    leaf cost 5 I, |x| <= 6 and |u| <= 0.3 at nonleaf nodes, |x| <= 0.1 at leaves, AV@R 0.95.  The
    dynamics are seeded and kept mild (A = I + 0.1 N(0, 1), B = 0.5 N(0, 1), scaled by 1, 1.05,
    0.95 per event) instead of that test's unseeded A, 2A, 3A, so the leaf bound can be met."""
    rng = np.random.default_rng(seed)
    P = np.array([[0.1, 0.8, 0.1], [0.4, 0.6, 0.], [0, 0.3, 0.7]])
    v = np.array([0.5, 0.4, 0.1])
    tr = tree_mod.MarkovChain(P, v, 4, 3, dt=dt).build()
    nx, nu = 3, 2
    A = np.eye(nx) + 0.1 * rng.standard_normal((nx, nx))
    B = 0.5 * rng.standard_normal((nx, nu))
    dyn = [build.Dynamics(k * A, k * B) for k in (1.0, 1.05, 0.95)]
    costs = [build.CostLinear(10.0 * k * np.ones(nx), 1.0 * k * np.ones(nu)) for k in (1, 2, 3)]
    leaf = build.CostQuadratic(5 * np.eye(nx), leaf=True)
    lim = np.vstack((6.0 * np.ones((nx, 1)), 0.3 * np.ones((nu, 1))))
    ncon = build.Rectangle(-lim, lim)
    lcon = build.Rectangle(-0.1 * np.ones((nx, 1)), 0.1 * np.ones((nx, 1)))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def unconstrained_markov(folder, seed=7, dt="d"):
    """The 31-node family (2 events, horizon 4; nx=3, nu=2; AV@R 0.95) with no state or input
    constraints at any node (the reference's ``no`` constraint)."""
    rng = np.random.default_rng(seed)
    tr = _markov(rng, 2, 4, 4, dt)
    nx, nu = 3, 2
    dyn = [build.Dynamics(np.eye(nx) + 0.1 * rng.standard_normal((nx, nx)), 0.5 * rng.standard_normal((nx, nu)))
           for _ in range(tr.num_events)]
    costs = [build.CostQuadratic(np.eye(nx), 0.1 * np.eye(nu)) for _ in range(tr.num_events)]
    leaf = build.CostQuadratic(np.eye(nx), leaf=True)
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(build.No()).with_constraint_leaf(build.No())
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def affine_markov(folder, seed=8, dt="d"):
    """The 31-node box-constrained family with affine dynamics x+ = A x + B u + c, c ~ N(0, 0.1^2)
    per event (the reference's ``affine`` dynamics, src/dynamics.cuh:223-238)."""
    rng = np.random.default_rng(seed)
    tr = _markov(rng, 2, 4, 4, dt)
    nx, nu = 3, 2
    dyn = [build.Dynamics(np.eye(nx) + 0.1 * rng.standard_normal((nx, nx)), 0.5 * rng.standard_normal((nx, nu)),
                          0.1 * rng.standard_normal(nx)) for _ in range(tr.num_events)]
    costs = [build.CostQuadratic(np.eye(nx), 0.1 * np.eye(nu)) for _ in range(tr.num_events)]
    leaf = build.CostQuadratic(np.eye(nx), leaf=True)
    lim = np.vstack((5.0 * np.ones((nx, 1)), 2.0 * np.ones((nu, 1))))
    return (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(build.Rectangle(-lim, lim))
            .with_constraint_leaf(build.Rectangle(-5.0 * np.ones((nx, 1)), 5.0 * np.ones((nx, 1))))
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))


def server_air_condition(folder, horizon=5, dt="d"):
    """The reference's server air-conditioning example (examples/serverAirCondition/main.py:1-86):
    2 events, Markov transitions [[.5, .5], [.5, .5]] from [.3, .7], branching at every stage;
    nx = nu = 50; A_w tridiagonal with 0.01 off the diagonal and 1 + (1 + j/50) w/2 on it, B = I;
    Q = 0.1 I, R = I, leaf 0.1 I; |x| <= 1, |u| <= 1.5 (leaf |x| <= 1); AV@R 0.95; preconditioned.
    The initial state 0.1 * ones is written as initialState, as the example does."""
    E, nu = 2, 50
    nx = nu
    tr = tree_mod.MarkovChain(np.array([[0.5, 0.5], [0.5, 0.5]]), np.array([0.3, 0.7]), horizon, horizon,
                              dt=dt).build()
    off = 0.01 * np.ones(nx - 1)
    dyn = []
    for w in range(E):
        A = np.diag(off, -1) + np.diag(off, 1)
        for j in range(nx):
            A[j, j] = 1.0 + (1.0 + j / nx) * w / E
        dyn.append(build.Dynamics(A, np.eye(nu)))
    costs = [build.CostQuadratic(0.1 * np.eye(nx), np.eye(nu)) for _ in range(E)]
    leaf = build.CostQuadratic(0.1 * np.eye(nx), leaf=True)
    x_ub, u_ub = np.ones(nx), 1.5 * np.ones(nu)
    ncon = build.Rectangle(np.hstack((-x_ub, -u_ub)), np.hstack((x_ub, u_ub)))
    lcon = build.Rectangle(-x_ub, x_ub)
    prob = (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(0.95)).with_preconditioning().generate_problem(folder))
    tr.write_to_file_fp(folder, "initialState", 0.1 * np.ones(nx))
    return prob


def random_example(folder, seed=None, lo=100, hi=1000, count="states_and_inputs", dt="d"):
    """The reference's random problem generator (examples/andersonBuffer/main.py:15-137 and
    examples/random/main.py, which differ only in how the variables are counted): sizes are drawn
    until lo < variables < hi, with horizon U{5..15}, branching stops at U{1..3}, events U{2..10},
    nu U{10..300} and nx = 2 nu; an i.i.d. tree; A_w = I + N(0, .01^2), B_w = B + N(0, .01^2) with
    B ~ N(0, 1); Q_w = (Q + N(0, .01^2))(.)' and R_w likewise from diagonal Q ~ U[0, .1], R ~ U[0, 100];
    leaf cost Q; boxes |x| <= U[1, 2], |u| <= U[0, .1]; AV@R alpha ~ U[0, 1]; preconditioned.  The
    initial state, U[-ub/2, ub/2] per state, is written as initialState.  count: "states_and_inputs"
    counts nodes * (nx + nu) (the andersonBuffer study), "random" nodes * nx + nonleaf nodes * nu.
    seed=None draws as the reference does (unseeded); a seed makes the draw reproducible."""
    rng = np.random.default_rng(seed)
    num_vars = np.inf
    while not (lo < num_vars < hi):
        horizon = int(rng.integers(5, 15, endpoint=True))
        stopping = int(rng.integers(1, 3, endpoint=True))
        E = int(rng.integers(2, 10, endpoint=True))
        nu = int(rng.integers(10, 300, endpoint=True))
        nx = 2 * nu
        nodes = (E ** (stopping + 1) - 1) / (E - 1) + E ** stopping * (horizon - stopping)
        if count == "random":
            num_vars = nodes * nx + (nodes - E ** stopping) * nu
        else:
            num_vars = nodes * (nx + nu)
    r = rng.uniform(size=E)
    tr = tree_mod.IidProcess(r / r.sum(), horizon, stopping, dt=dt).build()
    B_base = rng.normal(0., 1., size=(nx, nu))
    dyn = [build.Dynamics(np.eye(nx) + rng.normal(0., .01, size=(nx, nx)),
                          B_base + rng.normal(0., .01, size=(nx, nu))) for _ in range(E)]
    Q_base = np.diag(rng.uniform(0., .1, size=nx))
    R_base = np.diag(rng.uniform(0., 100., size=nu))
    costs = []
    for _ in range(E):
        Qw = Q_base + rng.normal(0., .01, size=(nx, nx))
        Rw = R_base + rng.normal(0., .01, size=(nu, nu))
        costs.append(build.CostQuadratic(Qw @ Qw.T, Rw @ Rw.T))
    leaf = build.CostQuadratic(Q_base, leaf=True)
    x_ub = rng.uniform(1., 2., nx)
    u_ub = rng.uniform(0., .1, nu)
    ncon = build.Rectangle(np.hstack((-x_ub, -u_ub)), np.hstack((x_ub, u_ub)))
    lcon = build.Rectangle(-x_ub, x_ub)
    alpha = rng.uniform(0., 1.)
    prob = (problem.Factory(tr, nx, nu).with_dynamics_events(dyn).with_cost_nonleaf_events(costs)
            .with_cost_leaf(leaf).with_constraint_nonleaf(ncon).with_constraint_leaf(lcon)
            .with_risk(build.AVaR(alpha)).with_preconditioning().generate_problem(folder))
    x0 = np.array([rng.uniform(-.5 * x_ub[k], .5 * x_ub[k]) for k in range(nx)])
    tr.write_to_file_fp(folder, "initialState", x0)
    return prob


def initial_states(seed, nx, count, kind="normal"):
    """Initial states: ``normal`` ~ N(0, 0.5^2 I), ``small`` ~ N(0, 0.05^2 I), ``uniform`` ~ U[-1, 1]."""
    rng = np.random.default_rng([seed, 1])
    if kind == "uniform":
        return rng.uniform(-1.0, 1.0, size=(count, nx))
    if kind == "small":
        return 0.05 * rng.standard_normal((count, nx))
    return 0.5 * rng.standard_normal((count, nx))


def generate(name, folder, **kw):
    os.makedirs(folder, exist_ok=True)
    return {"fixture": fixture_problem, "small": small_markov, "mid": mid_markov, "poly": polyhedral_small, "nocon": unconstrained_markov, "affine": affine_markov, "pwi": polyhedron_identity_small, "linear": linear_cost_markov,
            "large": polyhedral_iid, "server": server_air_condition,
            "random": random_example}[name](folder, **kw)
