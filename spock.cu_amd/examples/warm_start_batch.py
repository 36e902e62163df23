"""Batched, learned-warm-start solves of many initial states of one problem, through the Python
binding of the C ABI:

  1. solve a training family of initial states cold (FP32 lane engine, chunked);
  2. fit the warm-start MLP (x0 -> primal-dual iterate) on those solutions on the GPU;
  3. for new initial states: calibrate the cold acceptance thresholds, predict starting iterates,
     solve in FP32 from them and polish in FP64 against the same thresholds.

Usage: python warm_start_batch.py [problem folder] [states]   (default: a generated 31-node problem)
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import __graft_entry__  # noqa: E402  (loads the package as spock_amd)

__graft_entry__._load_pkg()
from spock_amd import runtime as rt  # noqa: E402
from spock_amd.api import configs  # noqa: E402


def main(folder=None, states=1024, train=1024, seed=0):
    ctx = rt.Context(0)
    if folder is None:
        folder = tempfile.mkdtemp()
        configs.generate("small", folder)
    prob = rt.Problem(ctx, folder)
    tol = dict(tol_abs=1e-3, tol_rel=1e-3)  # the reference builder's tol(1e-3)

    # 1. training solutions, cold, FP32, 400 outer iterations per launch
    x_train = configs.initial_states(seed + 1, prob.nx, train, "uniform")
    tb = rt.Batch(ctx, [prob] * train, rt.F32, shared=True)
    tb.set_x0(list(x_train))
    res, _ = tb.solve_chunked(rt.options(max_iters=3000, chunk_iters=400, **tol))
    print(f"training: {sum(r['status'] == 0 for r in res)}/{train} converged")

    # 2. the warm-start network, trained on the GPU
    mlp = rt.WarmStartMLP(ctx, prob.nx, prob.n_iter, hidden=256, layers=3, seed=seed)
    mlp.dataset(tb, x_train)
    losses = mlp.train(epochs=40, minibatch=128, lr=1e-3, seed=seed)
    print(f"MLP loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    del tb

    # 3. new initial states: thresholds, prediction, FP32 solve, FP64 polish
    x0s = configs.initial_states(seed + 2, prob.nx, states, "uniform")
    b32 = rt.Batch(ctx, [prob] * states, rt.F32, shared=True)
    b64 = rt.Batch(ctx, [prob] * states, rt.F64, shared=True)
    b32.set_x0(list(x0s))
    b64.set_x0(list(x0s))
    b32.autotune(rt.options(**tol), iters=20)
    t = time.perf_counter()
    b32.solve(rt.options(max_iters=1, check_period=1, **tol))  # cold calibration of the thresholds
    tols = np.array([r["tol"] for r in b32.results()[0]])
    mlp.predict(x0s, b32)
    b32.set_tolerances(tols)
    b32.solve_graph(rt.options(max_iters=800, chunk_iters=400, **tol), warm=True)
    r32, _ = b32.results()
    b64.copy_iterates_from(b32)
    b64.set_tolerances(tols)
    b64.solve_graph(rt.options(max_iters=200, chunk_iters=200, **tol), warm=True)
    r64, _ = b64.results()
    dt = time.perf_counter() - t
    ok = sum(r["status"] == 0 for r in r64)
    print(f"{ok}/{states} accepted in {dt:.2f} s ({ok / dt:.1f} solves/s); "
          f"FP32 iterations mean {np.mean([r['iters'] for r in r32]):.0f}")
    return ok, states


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None, int(sys.argv[2]) if len(sys.argv) > 2 else 1024)
