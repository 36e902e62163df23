"""Host-side pieces of bench.py that run without a GPU: the cost gaps it reports for accepted
solutions (root epigraph variable s0 and the nested AV@R cost J of the states and inputs), fed
with oracle solutions of the 31-node problem."""
import importlib.util
import os
import types

import numpy as np
import pytest

from oracle.oracle import Oracle
from spock_amd.api import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_host", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_cost_gaps_of_oracle_solutions(bench, tmp_path):
    d = str(tmp_path / "small")
    gen = configs.generate("small", d)
    o = Oracle(d)
    t = gen.tree
    prob = types.SimpleNamespace(nu=o.nu, nx=o.nx, N=t.num_nodes, nl=t.num_nonleaf_nodes, E=t.num_events)
    x0s = configs.initial_states(51, o.nx, 3, "normal")
    loose = np.array([o.solve(x, tol_abs=1e-3, tol_rel=1e-3, max_iters=5000)["z"] for x in x0s])
    tight = np.array([o.solve(x, tol_abs=1e-7, tol_rel=0.0, max_iters=20000)["z"] for x in x0s])
    mask = np.array([True, True, False])
    same = bench.cost_gaps(prob, gen, tight, tight, mask)
    assert same["s0"] == same["J"] == {"mean": 0.0, "max": 0.0, "within_1e-3": 2}
    g = bench.cost_gaps(prob, gen, loose, tight, mask)
    for k in ("s0", "J"):
        assert 0.0 < g[k]["mean"] <= g[k]["max"] < 0.1 and 0 <= g[k]["within_1e-3"] <= 2
    # s0 and J of the tight solutions agree
    s_t = bench.s0_of(prob, tight)
    J_t = np.array([gen.trajectory_cost(z) for z in tight])
    assert np.allclose(s_t, J_t, rtol=1e-4, atol=1e-6)
    empty = bench.cost_gaps(prob, gen, loose, tight, np.zeros(3, dtype=bool))
    assert empty["J"] == {"mean": None, "max": None, "within_1e-3": 0}


def test_more_than_one_gpu_is_refused():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
