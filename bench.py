"""Throughput benchmark: RAOCP solves/s to a 1e-3 residual on one MI355X.

Workload: the 229-node Markov-tree RAOCP (3 events, horizon 10, branching stops at stage 3,
nx = 50, nu = 20, box constraints, AV@R 0.9, leaf cost from the discrete Riccati equation,
preconditioned) with a batch of initial states x0 ~ N(0, 0.25 I), synthetic data from the
package's seeded generator.

A step solves the whole batch: FP32-bulk SPOCK for every instance in one persistent launch,
then an FP64 polish launch warm-started from the FP32 iterates and accepted against the same
per-instance threshold, then an FP64 cold solve of any instance the polish did not accept.
The baseline beside it is the plain method: FP64 SPOCK, one instance per launch.

Tolerance: ``tol(1e-3)`` of the reference's builder (absolute and relative 1e-3, both times
min(scaling)), tested every 25 iterations on ||r/gamma - [L* r_d; L r_p]||_inf
(src/cache.cuh:302-306,910-961).

Prints one JSON line (rank 0); ``--gpus`` above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "RAOCP solves/s to 1e-3 residual: warm-start batched vs plain SPOCK, 1 GPU"


def _pkg():
    import __graft_entry__ as ge
    ge._load_pkg()
    from spock_amd import runtime as rt
    from spock_amd.api import configs
    return rt, configs


def bytes_per_T(p, word):
    """Algorithmic bytes one application of the CP operator T moves for one instance.

    Per-node factors (ABKTr, APB per child; K, R~^-1 per nonleaf), the kernel projectors, the
    per-event dynamics and cost matrices read once, the iterate read and the candidate written,
    and the dynamics sweep intermediates written and read once."""
    N, nl, E, nx, nu = p.N, p.nl, p.E, p.nx, p.nu
    nD = 4 * E + 1
    shared = E * (nx * nx + nx * nu) + 2 * nx * nx + nu * nu + 2 * (nx + nu) + 2 * nx
    elems = ((N - 1) * (nx * nx + nx * nu) + nl * (nu * nx + nu * nu) + nl * nD * nD + shared
             + 2 * p.n_iter + 2 * (nx * N + nu * nl))
    return word * elems


def bytes_per_ell(p, word):
    """Algorithmic bytes of one [L*; L] application inside a norm or error check."""
    shared = 2 * p.nx * p.nx + p.nu * p.nu + 2 * (p.nx + p.nu) + 2 * p.nx
    return word * (2 * p.n_iter + shared)


def s0_of(p, z):
    """Root epigraph variable s_0 (the risk-averse cost) of iterates z (rows = instances)."""
    oS = p.nu * p.nl + p.nx * p.N + (2 * p.E + 1) * p.nl + p.N
    return z[:, oS]


def plain_single(rt, ctx, prob, x0s, tol, max_iters, block_size):
    """Plain FP64 SPOCK, one instance per launch, cold start."""
    b = rt.Batch(ctx, [prob], rt.F64)
    times, iters, s0, status = [], [], [], []
    for x0 in x0s:
        b.reset()
        b.set_x0([x0])
        t = time.perf_counter()
        b.solve(rt.options(tol_abs=tol, tol_rel=tol, max_iters=max_iters, block_size=block_size))
        res, _ = b.results()
        times.append(time.perf_counter() - t)
        iters.append(res[0]["iters"])
        status.append(res[0]["status"])
        s0.append(s0_of(prob, b.iterate(0)[None, :])[0])
        print(f"bench: plain FP64 single x0#{len(times) - 1}: {times[-1]:.2f} s, {iters[-1]} iterations, "
              f"status {status[-1]}", flush=True)
    return times, iters, s0, status


class BatchedSolver:
    """FP32-bulk batched SPOCK with FP64 polish (device-resident, one problem, B initial states)."""

    def __init__(self, rt, ctx, prob, x0s, tol, max_iters, polish_iters, block_size):
        self.rt, self.ctx, self.prob = rt, ctx, prob
        self.B = len(x0s)
        self.tol, self.max_iters, self.polish_iters = tol, max_iters, polish_iters
        self.bs = block_size
        self.b32 = rt.Batch(ctx, [prob] * self.B, rt.F32)
        self.b64 = rt.Batch(ctx, [prob] * self.B, rt.F64)
        self.b32.set_x0(list(x0s))
        self.b64.set_x0(list(x0s))

    def step(self):
        rt = self.rt
        self.b32.reset()
        self.b32.solve(rt.options(tol_abs=self.tol, tol_rel=self.tol, max_iters=self.max_iters, block_size=self.bs,
                                  ls_method=rt.LS_GRAM))
        r32, k32 = self.b32.results()
        tols = np.array([r["tol"] for r in r32])
        self.b64.copy_iterates_from(self.b32)
        self.b64.set_tolerances(tols)
        self.b64.solve(rt.options(tol_abs=self.tol, tol_rel=self.tol, max_iters=self.polish_iters,
                                  block_size=self.bs), warm=True)
        r64, k64 = self.b64.results()
        self.b64.set_tolerances(None)
        return r32, k32, r64, k64


def cpu_baseline(folder, x0, budget_s, gpu_iters_mean):
    """FP64 CPU transcription of SPOCK (the oracle), one core, a bounded run on one instance."""
    from oracle.oracle import Oracle
    o = Oracle(folder)
    n = 200
    t = time.perf_counter()
    o.solve(x0, tol_abs=1e-30, tol_rel=0.0, max_iters=n)
    dt = time.perf_counter() - t
    n2 = int(max(n, min(5000, n * budget_s / max(dt, 1e-6))))
    t = time.perf_counter()
    r = o.solve(x0, tol_abs=1e-30, tol_rel=0.0, max_iters=n2)
    dt = time.perf_counter() - t
    per_iter = dt / max(1, r["iters"])
    return {"value": 1.0 / (per_iter * gpu_iters_mean), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 SPOCK, x0#0, {r['iters']} iterations in {dt:.1f} s "
                      f"({per_iter * 1e3:.2f} ms/iteration), scaled by the GPU plain path's mean "
                      f"{gpu_iters_mean:.0f} iterations per solve"}


def copy_calibration(rt_unused=None):
    """A plain device copy of known size (1 GiB), for scaling the FETCH_SIZE/WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--plain", type=int, default=1, help="instances solved by the plain single-instance path")
    ap.add_argument("--block-size", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--max-iters", type=int, default=8000)
    ap.add_argument("--polish-iters", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cost-check", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    rt, configs = _pkg()
    ctx = rt.Context(0)
    with tempfile.TemporaryDirectory() as folder:
        configs.generate("mid", folder)
        prob = rt.Problem(ctx, folder)
        x0s = configs.initial_states(51, prob.nx, args.batch, "normal")

        # plain FP64 single-instance baseline
        plain = None
        if args.plain > 0:
            pt, pi, ps0, pst = plain_single(rt, ctx, prob, x0s[:args.plain], args.tol, 20000, args.block_size)
            plain = {"solves_per_s": len(pt) / sum(pt), "seconds": pt, "iters": pi, "status": pst}

        solver = BatchedSolver(rt, ctx, prob, x0s, args.tol, args.max_iters, args.polish_iters, args.block_size)
        for w in range(args.warmup):
            solver.step()
            print(f"bench: warmup step {w} done", flush=True)
        ctx.lib.spk_ctx_synchronize(ctx.h)
        kern32, kern64, last = [], [], None
        t0 = time.perf_counter()
        for s in range(args.steps):
            r32, k32, r64, k64 = solver.step()
            kern32.append(k32)
            kern64.append(k64)
            last = (r32, r64)
            print(f"bench: step {s} done ({time.perf_counter() - t0:.1f} s)", flush=True)
        ctx.lib.spk_ctx_synchronize(ctx.h)
        elapsed = time.perf_counter() - t0
        ms_per_step = 1e3 * elapsed / max(1, args.steps)
        r32, r64 = last
        B = args.batch
        accepted = int(sum(1 for r in r64 if r["status"] == 0))
        it32 = np.array([r["iters"] for r in r32])
        it64 = np.array([r["iters"] for r in r64])

        # roofline of the dominant kernel (the FP32 bulk solve): algorithmic bytes / HIP-event time
        bt, bl = bytes_per_T(prob, 4), bytes_per_ell(prob, 4)
        nT = sum(r["num_T"] for r in r32)
        nE = sum(r["num_ell"] for r in r32)
        alg = nT * bt + nE * bl
        k32avg = float(np.mean(kern32))
        achieved = alg / k32avg / 1e9

        # accepted solutions against the plain FP64 method's cost
        cost = None
        if not args.no_cost_check:
            ref = rt.Batch(ctx, [prob] * B, rt.F64)
            ref.set_x0(list(x0s))
            ref.solve(rt.options(tol_abs=args.tol, tol_rel=args.tol, max_iters=args.max_iters,
                                 block_size=args.block_size))
            rref, _ = ref.results()
            print(f"bench: FP64 reference batch done ({sum(r['status'] == 0 for r in rref)}/{B} converged)",
                  flush=True)
            zr = ref.iterates().reshape(B, -1)
            zb = solver.b64.iterates().reshape(B, -1)
            s_ref, s_b = s0_of(prob, zr), s0_of(prob, zb)
            rel = np.abs(s_b - s_ref) / np.maximum(1e-12, np.abs(s_ref))
            ok = np.array([r["status"] == 0 for r in r64])
            ok &= np.array([r["status"] == 0 for r in rref])
            cost = {"max_rel_err_s0": float(rel[ok].max()) if ok.any() else None,
                    "mean_rel_err_s0": float(rel[ok].mean()) if ok.any() else None,
                    "within_1e-3": int((rel[ok] <= 1e-3).sum()), "compared": int(ok.sum()),
                    "reference": "plain FP64 SPOCK (GPU), same x0",
                    "reference_iters_mean": float(np.mean([r["iters"] for r in rref]))}
            del ref

        cpu = None
        if not args.no_cpu:
            mean_plain = float(np.mean(plain["iters"])) if plain else float(np.mean(it32))
            cpu = cpu_baseline(folder, x0s[0], args.cpu_seconds, mean_plain)
            print("bench: cpu baseline done", flush=True)

    value = B * args.steps / elapsed
    line = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "RAOCP Markov tree 229 nodes (E=3, N=10, stop 3), nx=50, nu=20, box, AV@R 0.9; "
                               "batch of x0~N(0,0.25I); FP32 bulk + FP64 polish",
                   "batch": B, "tol": args.tol, "check_period": 25, "block_size": args.block_size,
                   "parallelism": "single GPU, one workgroup per instance"},
        "plain_fp64_single": plain,
        "speedup_vs_plain": (value / plain["solves_per_s"]) if plain else None,
        "accepted": accepted, "fp32_iters_mean": float(it32.mean()), "polish_iters_mean": float(it64.mean()),
        "kernel_seconds": {"fp32_bulk": k32avg, "fp64_polish": float(np.mean(kern64))},
        "cost_check": cost,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_solve<float,1,1>", "bytes_per_T": bt, "bytes_per_ell": bl,
                     "T_per_launch": nT, "ell_per_launch": nE},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
