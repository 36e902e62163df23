"""Throughput benchmark: RAOCP solves/s to a 1e-3 residual on one MI355X.

Workload: the 229-node Markov-tree RAOCP (3 events, horizon 10, branching stops at stage 3,
nx = 50, nu = 20, box constraints, AV@R 0.9, leaf cost from the discrete Riccati equation,
preconditioned), synthetic data from the package's seeded generator, and a held-out batch of
initial states x0 ~ N(0, 0.25 I).

A step solves the whole batch with the batched, learned-warm-start method:
  1. each instance's acceptance threshold is calibrated as a cold start would (two iterations);
  2. the MLP (x0 -> primal-dual iterate, trained beforehand on solver-generated solutions of the
     same family) writes every instance's starting iterate;
  3. FP32 SPOCK from the predicted iterates, all instances in one persistent launch (lane engine);
  4. FP64 polish warm-started from the FP32 iterates.
Both warm-started solves accept against kappa times the cold threshold (default kappa = 1: the
cold threshold itself).  The FP32 lane engine cannot reach much below it: at kappa = 0.1 no FP32
solve of this family converges, and FP64 needs several hundred extra iterations at kappa = 0.3.
The cost of accepted solutions is compared with plain FP64 SPOCK on a sample and reported.
With --fallback, instances the warm path does not finish within its iteration caps are solved
cold (FP32 lane engine + FP64 polish at the cold threshold); without it they are reported as not
solved and do not count in the throughput.
The baseline beside it is the plain method: FP64 SPOCK, one instance per launch, cold start.

Tolerance: ``tol(1e-3)`` of the reference's builder (absolute and relative 1e-3, both times
min(scaling)), tested every 25 iterations on ||r/gamma - [L* r_d; L r_p]||_inf
(src/cache.cuh:302-306,910-961).

Prints one JSON line; ``--gpus`` above 1 is refused: this build runs on one GPU.
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


def log(msg):
    print(f"bench: {msg}", flush=True)


def _pkg():
    import __graft_entry__ as ge
    ge._load_pkg()
    from spock_amd import runtime as rt
    from spock_amd.api import configs
    return rt, configs


def bytes_per_T(p, word, share=1):
    """Algorithmic bytes one application of the CP operator T moves for one instance: the per-node
    factors (ABKTr, APB per child; K, R~^-1 per nonleaf), the kernel projectors and the per-event
    dynamics and cost matrices, read once and shared by `share` instances (64: the instances of
    one wave of the lane engine); plus the iterate read, the candidate written, and the dynamics
    sweep intermediates written and read once."""
    N, nl, E, nx, nu = p.N, p.nl, p.E, p.nx, p.nu
    nD = 4 * E + 1
    shared = E * (nx * nx + nx * nu) + 2 * nx * nx + nu * nu + 2 * (nx + nu) + 2 * nx
    factors = (N - 1) * (nx * nx + nx * nu) + nl * (nu * nx + nu * nu) + nl * nD * nD + shared
    return word * (factors / share + 2 * p.n_iter + 2 * (nx * N + nu * nl))


def bytes_per_ell(p, word, share=1):
    """Algorithmic bytes of one [L*; L] application inside a norm or error check."""
    shared = 2 * p.nx * p.nx + p.nu * p.nu + 2 * (p.nx + p.nu) + 2 * p.nx
    return word * (2 * p.n_iter + shared / share)


def s0_of(p, z):
    """Root epigraph variable s_0 (the risk-averse cost) of iterates z (rows = instances)."""
    oS = p.nu * p.nl + p.nx * p.N + (2 * p.E + 1) * p.nl + p.N
    return z[:, oS]


class Pipeline:
    """The batched learned-warm-start method on device-resident shared batches."""

    def __init__(self, rt, ctx, prob, x0s, mlp, args):
        self.rt, self.ctx, self.prob, self.mlp, self.a = rt, ctx, prob, mlp, args
        self.x0s = np.ascontiguousarray(x0s)
        self.B = len(x0s)
        self.b32 = rt.Batch(ctx, [prob] * self.B, rt.F32, shared=True)
        self.b64 = rt.Batch(ctx, [prob] * self.B, rt.F64, shared=True)
        self.b32.set_x0(list(x0s))
        self.b64.set_x0(list(x0s))
        self.kernel_s = []
        # the lane kernels' schedule (matrix-vector staging, workgroup size), timed on these batches
        self.schedule = {}
        for name, b in (("f32", self.b32), ("f64", self.b64)):
            mv, block, secs = b.autotune(self.opts(), iters=args.tune_iters)
            self.schedule[name] = {"mv": mv, "block": block,
                                   "seconds": {f"mv{k[0]}_b{k[1]}": v for k, v in secs.items()}}
            log(f"autotuned {name}: mv {mv}, block {block}; {self.schedule[name]['seconds']}")

    def opts(self, **kw):
        o = dict(tol_abs=self.a.tol, tol_rel=self.a.tol, block_size=0)
        o.update(kw)
        return self.rt.options(**o)

    def solve(self, b, warm, label, **kw):
        """A chunked solve of a shared batch: one HIP graph of its chunk launches (--graph), or
        launch by launch with a report after every chunk."""
        t = time.perf_counter()
        launches = [0]
        if self.a.graph:
            # graphs of at most `graph` chunk launches each, with a report after every graph
            o = self.opts(chunk_iters=self.a.chunk, **kw)
            total = max(1, -(-o.max_iters // self.a.chunk) + 1)
            per = max(1, min(self.a.graph, total))
            b.solve_graph(o, warm, per)
            res, sec = b.results()
            n_launch, graphs = per, 1
            while any(r["status"] == -1 for r in res):
                log(f"  {label}: {sum(r['status'] != -1 for r in res)}/{len(res)} finished, "
                    f"{time.perf_counter() - t:.1f} s")
                more = max(1, min(per, total - n_launch))
                b.resume_graph(more)
                res, s2 = b.results()
                sec += s2
                n_launch += more
                graphs += 1
            log(f"  {label}: {sum(r['status'] != -1 for r in res)}/{len(res)} finished, "
                f"{time.perf_counter() - t:.1f} s ({graphs} graph(s), {n_launch} launches)")
            self.kernel_s.append((b.dtype, sec, sum(r["num_T"] for r in res), sum(r["num_ell"] for r in res),
                                  n_launch, label))
            return res

        def report(res, sec):
            launches[0] += 1
            done = sum(r["status"] != -1 for r in res)
            log(f"  {label}: {done}/{len(res)} finished, {time.perf_counter() - t:.1f} s")

        res, sec = b.solve_chunked(self.opts(chunk_iters=self.a.chunk, **kw), warm, report)
        self.kernel_s.append((b.dtype, sec, sum(r["num_T"] for r in res), sum(r["num_ell"] for r in res),
                              launches[0], label))
        return res

    def step(self, max32, max64, fallback=True):
        rt = self.rt
        self.kernel_s = []
        # cold-start thresholds: the error of the first iteration from zero (checks at 0 and 1)
        self.b32.reset()
        r0 = self.solve(self.b32, False, "thresholds", max_iters=1, check_period=1)
        tols = np.array([r["tol"] for r in r0])
        if self.mlp is not None:
            self.mlp.predict(self.x0s, self.b32)
        else:
            self.b32.reset()
        kappa = self.a.kappa if self.mlp is not None else 1.0
        self.b32.set_tolerances(tols * kappa)
        r32 = self.solve(self.b32, self.mlp is not None, "FP32 from the warm start", max_iters=max32)
        self.b32.set_tolerances(None)
        self.b64.copy_iterates_from(self.b32)
        self.b64.set_tolerances(tols * (self.a.polish_kappa if self.mlp is not None else 1.0))
        r64 = self.solve(self.b64, True, "FP64 polish", max_iters=max64)
        self.b64.set_tolerances(None)
        status = np.array([r["status"] for r in r64])
        failed = np.flatnonzero(status != 0)
        result = None
        if fallback and self.a.fallback and failed.size:  # cold FP32 solve + FP64 polish of what the warm path left
            f32 = rt.Batch(self.ctx, [self.prob] * failed.size, rt.F32, shared=True)
            f64 = rt.Batch(self.ctx, [self.prob] * failed.size, rt.F64, shared=True)
            f32.set_x0(list(self.x0s[failed]))
            f64.set_x0(list(self.x0s[failed]))
            f32.set_tolerances(tols[failed])
            self.solve(f32, False, "fallback FP32 cold", max_iters=self.a.max_iters)
            f64.copy_iterates_from(f32)
            f64.set_tolerances(tols[failed])
            rf = self.solve(f64, True, "fallback FP64 polish", max_iters=max64)
            result = (failed, rf, f64)
        return r32, r64, result


def plain_single(rt, ctx, prob, x0s, tol, max_iters, workgroups=1):
    """Plain FP64 SPOCK, one instance per launch, cold start (per-instance engine; workgroups > 1:
    grid mode, the instance spread over that many cooperating workgroups)."""
    b = rt.Batch(ctx, [prob], rt.F64)
    b.set_workgroups(workgroups)
    out = []
    for k, x0 in enumerate(x0s):
        b.reset()
        b.set_x0([x0])
        t = time.perf_counter()
        b.solve(rt.options(tol_abs=tol, tol_rel=tol, max_iters=max_iters, block_size=512))
        res, _ = b.results()
        dt = time.perf_counter() - t
        out.append((dt, res[0]["iters"], res[0]["status"], s0_of(prob, b.iterate(0)[None, :])[0]))
        log(f"plain FP64 single x0#{k} ({workgroups} workgroups): {dt:.2f} s, {res[0]['iters']} iterations, "
            f"status {res[0]['status']}")
    return out


def reference_costs(rt, ctx, prob, x0s, tol, max_iters):
    """Plain FP64 SPOCK (per-instance engine, one workgroup per instance) on a sample, cold.  As a
    yardstick of how far apart two solutions accepted at the same tolerance can lie, the same
    instances are also solved cold in FP32 and polished in FP64 against the plain solves' thresholds."""
    n = len(x0s)
    b = rt.Batch(ctx, [prob] * n, rt.F64)
    b.set_x0(list(x0s))
    b.solve(rt.options(tol_abs=tol, tol_rel=tol, max_iters=max_iters, block_size=512))
    res, _ = b.results()
    log("  reference: plain FP64 solves done")
    b32 = rt.Batch(ctx, [prob] * n, rt.F32)
    b32.set_x0(list(x0s))
    b32.solve(rt.options(tol_abs=tol, tol_rel=tol, max_iters=max_iters, block_size=512))
    b32.results()
    log("  reference: cold FP32 solves done")
    b64 = rt.Batch(ctx, [prob] * n, rt.F64)
    b64.set_x0(list(x0s))
    b64.copy_iterates_from(b32)
    b64.set_tolerances([r["tol"] for r in res])
    b64.solve(rt.options(tol_abs=tol, tol_rel=tol, max_iters=1000, block_size=512), warm=True)
    res_alt, _ = b64.results()
    return (np.array([r["status"] for r in res]), b.iterates().reshape(n, -1),
            np.array([r["status"] for r in res_alt]), b64.iterates().reshape(n, -1))


def cost_gaps(prob, gen, z, z_ref, mask):
    """Relative gaps of two costs of iterates z against z_ref, over the rows in mask: the root
    epigraph variable s0, and J, the nested AV@R cost of the states and inputs alone."""
    def rel(a, b):
        return (np.abs(a - b) / np.maximum(1e-12, np.abs(b)))[mask]

    def J(zz):
        return np.array([gen.trajectory_cost(row) for row in zz])
    out = {}
    for name, r in (("s0", rel(s0_of(prob, z), s0_of(prob, z_ref))), ("J", rel(J(z), J(z_ref)))):
        out[name] = {"mean": float(r.mean()) if r.size else None, "max": float(r.max()) if r.size else None,
                     "within_1e-3": int((r <= 1e-3).sum())}
    return out


def cost_check(rt, ctx, prob, gen, x0s, args, accepted, z_batch):
    """Costs of accepted batch solutions against plain FP64 SPOCK on the same sample, with the
    spread between two accepted solutions of this tolerance beside it."""
    ref_status, z_ref, alt_status, z_alt = reference_costs(rt, ctx, prob, x0s, args.tol, args.max_iters)
    both = accepted & (ref_status == 0)
    alt = (alt_status == 0) & (ref_status == 0)
    g, ga = cost_gaps(prob, gen, z_batch, z_ref, both), cost_gaps(prob, gen, z_alt, z_ref, alt)
    return {"compared": int(both.sum()), "max_rel_err_s0": g["s0"]["max"], "mean_rel_err_s0": g["s0"]["mean"],
            "within_1e-3": g["s0"]["within_1e-3"],
            "max_rel_err_J": g["J"]["max"], "mean_rel_err_J": g["J"]["mean"], "within_1e-3_J": g["J"]["within_1e-3"],
            "reference": "plain FP64 SPOCK (per-instance GPU engine), same x0, same tolerance",
            "costs": "s0: the root epigraph variable of the iterate; J: the nested AV@R cost of its states "
                     "and inputs alone (Problem.trajectory_cost); the two agree at a solution",
            "yardstick_cold_fp32_polish": {
                "compared": int(alt.sum()),
                "mean_rel_diff_s0": ga["s0"]["mean"], "max_rel_diff_s0": ga["s0"]["max"],
                "mean_rel_diff_J": ga["J"]["mean"], "max_rel_diff_J": ga["J"]["max"],
                "note": "cold FP32 solve + FP64 polish vs plain FP64, both accepted at the same thresholds: "
                        "how far apart two accepted solutions of this tolerance lie"}}


def train_warm_start(rt, ctx, prob, configs, args):
    """Solve a training family of initial states cold (FP32 lane engine, chunked) and fit the MLP
    on the solutions."""
    x_train = configs.initial_states(50, prob.nx, args.train, "normal")
    b = rt.Batch(ctx, [prob] * args.train, rt.F32, shared=True)
    b.set_x0(list(x_train))
    t = time.perf_counter()

    def report(res, sec):
        done = sum(r["status"] != -1 for r in res)
        log(f"  training solves: {done}/{len(res)} finished, {time.perf_counter() - t:.1f} s")

    res, _ = b.solve_chunked(rt.options(tol_abs=args.tol, tol_rel=args.tol, max_iters=args.train_iters,
                                        block_size=1024, chunk_iters=args.chunk), False, report)
    ok = np.array([r["status"] == 0 for r in res])
    log(f"training solves: {ok.sum()}/{args.train} converged in {time.perf_counter() - t:.1f} s")
    mlp = rt.WarmStartMLP(ctx, prob.nx, prob.n_iter, hidden=512, layers=3, seed=11)
    mlp.dataset(b, x_train)
    t = time.perf_counter()
    losses = mlp.train(epochs=args.epochs, minibatch=256, lr=1e-3, seed=3)
    log(f"MLP trained: {args.epochs} epochs on {len(x_train)} solutions in {time.perf_counter() - t:.1f} s, "
        f"loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    return mlp, int(len(x_train)), float(losses[-1])


def cpu_baseline(folder, x0, budget_s, iters_per_solve):
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
    return {"value": 1.0 / (per_iter * iters_per_solve), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 SPOCK (one core), x0#0, {r['iters']} iterations in {dt:.1f} s "
                      f"({per_iter * 1e3:.2f} ms/iteration), scaled by the {iters_per_solve:.0f} iterations the "
                      f"plain method needs per solve"}


def copy_calibration():
    """A plain device copy of known size (1 GiB each way), for scaling FETCH_SIZE/WRITE_SIZE."""
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
    ap.add_argument("--batch", type=int, default=16384, help="held-out initial states solved per step")
    ap.add_argument("--train", type=int, default=16384, help="solutions the warm-start MLP is trained on")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--train-iters", type=int, default=3000, help="iteration cap of the training solves")
    ap.add_argument("--chunk", type=int, default=400, help="outer iterations per launch of a chunked solve")
    ap.add_argument("--cold", action="store_true", help="no learned warm start (batched cold starts)")
    ap.add_argument("--plain", type=int, default=1, help="instances solved by the plain single-instance path")
    ap.add_argument("--plain-workgroups", type=int, default=64,
                    help="workgroups of the grid-mode variant of the plain single-instance solve")
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--max-iters", type=int, default=5000,
                    help="iteration cap of the cold solves of the cost sample and the fallback")
    ap.add_argument("--warm-iters", type=int, default=800, help="FP32 iteration cap from the warm start")
    ap.add_argument("--polish-iters", type=int, default=200)
    ap.add_argument("--ref-sample", type=int, default=16, help="instances checked against plain FP64 cost")
    ap.add_argument("--kappa", type=float, default=1.0, help="warm-path threshold as a fraction of the cold one")
    ap.add_argument("--polish-kappa", type=float, default=1.0,
                    help="FP64 polish threshold as a fraction of the cold one (below 1: more accurate costs, "
                         "more polish iterations)")
    ap.add_argument("--fallback", action="store_true", help="solve what the warm path leaves with cold solves")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--graph", type=int, default=3,
                    help="chunk launches per HIP graph of a solve in the step (0: launch by launch)")
    ap.add_argument("--tune-iters", type=int, default=50, help="outer iterations per schedule the autotuner times")
    ap.add_argument("--no-cpu", action="store_true")
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
        gen = configs.generate("mid", folder)
        prob = rt.Problem(ctx, folder)
        x0s = configs.initial_states(51, prob.nx, args.batch, "normal")

        plain = None
        if args.plain > 0:
            # the plain method at its fastest single-instance launch shape: one workgroup, or the
            # instance spread over 64 workgroups (grid mode)
            variants = {}
            for wgs in (1, args.plain_workgroups):
                ps = plain_single(rt, ctx, prob, x0s[:args.plain], args.tol, 20000, wgs)
                secs = [p[0] for p in ps]
                variants[wgs] = {"solves_per_s": len(secs) / sum(secs), "seconds": secs,
                                 "iters": [p[1] for p in ps], "status": [p[2] for p in ps], "workgroups": wgs}
            best = max(variants.values(), key=lambda v: v["solves_per_s"])
            plain = dict(best)
            plain["variants"] = list(variants.values())

        mlp, n_train, train_loss = None, 0, None
        if not args.cold:
            mlp, n_train, train_loss = train_warm_start(rt, ctx, prob, configs, args)

        pipe = Pipeline(rt, ctx, prob, x0s, mlp, args)
        for w in range(args.warmup):  # the same launches on a short iteration budget
            pipe.step(max32=50, max64=50, fallback=False)
            log(f"warmup step {w} done")
        ctx.lib.spk_ctx_synchronize(ctx.h)
        t0 = time.perf_counter()
        for s in range(args.steps):
            r32, r64, fallback = pipe.step(args.warm_iters, args.polish_iters)
            log(f"step {s} done ({time.perf_counter() - t0:.1f} s)")
        ctx.lib.spk_ctx_synchronize(ctx.h)
        elapsed = time.perf_counter() - t0
        B = args.batch
        kernels = pipe.kernel_s
        st64 = np.array([r["status"] for r in r64])
        it32 = np.array([r["iters"] for r in r32])
        it64 = np.array([r["iters"] for r in r64])
        ns = min(args.ref_sample, B)
        z_batch = np.array([pipe.b64.iterate(k) for k in range(ns)]) if ns else np.zeros((0, prob.n_iter))
        n_fb = 0
        accepted = st64 == 0
        if fallback is not None:
            failed, rf, f64 = fallback
            n_fb = failed.size
            for i, k in enumerate(failed):
                if rf[i]["status"] == 0:
                    accepted[k] = True
                    if k < ns:
                        z_batch[k] = f64.iterate(i)
        log(f"accepted {accepted.sum()}/{B}, fallback {n_fb}")

        # accepted solutions against plain FP64 SPOCK (cold, same tolerance) on a sample
        cost = None
        if ns:
            cost = cost_check(rt, ctx, prob, gen, x0s[:ns], args, accepted[:ns], z_batch)
            log("reference sample done")

        # roofline of the dominant kernel: k_lane_solve<float>, the FP32 solve from the warm start
        bt, bl = bytes_per_T(prob, 4, 64), bytes_per_ell(prob, 4, 64)
        main32 = [k for k in kernels if k[0] == rt.F32][1]  # [0] is the threshold calibration
        alg = main32[2] * bt + main32[3] * bl
        achieved = alg / main32[1] / 1e9
        # HBM bytes per launch: FETCH_SIZE + WRITE_SIZE per instance-T measured by rocprofv3 on this
        # kernel (scaled by a plain copy's counters), times this solve's instance-T per launch
        traffic, traffic_note = None, None
        tf = os.path.join(ROOT, "profiles", "k_lane_solve_traffic.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                meas = json.load(fh)
            traffic = meas["hbm_bytes_per_instance_T_and_ell"] * main32[2] / max(1, main32[4])
            traffic_note = ("bytes per launch: %.2f MB of HBM traffic per instance-T (rocprofv3 FETCH_SIZE + "
                            "WRITE_SIZE, copy-scaled, %s) x this solve's instance-T per launch"
                            % (meas["hbm_bytes_per_instance_T_and_ell"] / 1e6, meas["workload"]))

        cpu = None
        if not args.no_cpu:
            ipers = float(np.mean(plain["iters"])) if plain else float(it32.mean())
            cpu = cpu_baseline(folder, x0s[0], args.cpu_seconds, ipers)
            log("cpu baseline done")

    value = int(accepted.sum()) * args.steps / elapsed
    line = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / max(1, args.steps), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "RAOCP Markov tree 229 nodes (3 events, horizon 10, branching stops at stage 3), "
                               "nx=50, nu=20, box, AV@R 0.9; held-out x0~N(0,0.25I)",
                   "batch": B, "method": "learned warm start + FP32 lane engine + FP64 polish"
                   if mlp is not None else "cold start + FP32 lane engine + FP64 polish",
                   "tol": args.tol, "check_period": 25,
                   "mlp": {"hidden": 512, "layers": 3, "train_solutions": n_train, "final_loss": train_loss}
                   if mlp is not None else None,
                   "parallelism": "single GPU, 64 instances per workgroup"},
        "plain_fp64_single": plain,
        "speedup_vs_plain": (value / plain["solves_per_s"]) if plain else None,
        "accepted": int(accepted.sum()), "not_solved": int(B - accepted.sum()), "fallback": n_fb,
        "kappa": args.kappa if mlp is not None else 1.0,
        "polish_kappa": args.polish_kappa if mlp is not None else 1.0,
        "schedule": pipe.schedule,
        "fp32_iters_mean": float(it32.mean()), "fp32_iters_max": int(it32.max()),
        "polish_iters_mean": float(it64.mean()),
        "kernel_seconds": [{"solve": k[5], "dtype": "f32" if k[0] == rt.F32 else "f64", "seconds": k[1],
                            "launches": k[4]} for k in kernels],
        "cost_check": cost,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "kernel": "k_lane_solve<float>",
                     "bytes_per_T": bt, "bytes_per_ell": bl, "T": main32[2], "ell": main32[3],
                     "launches": main32[4], "seconds": main32[1], "avg_launch_ms": 1e3 * main32[1] / max(1, main32[4]),
                     "unit_note": "per instance: T bytes = factors/64 + 2|z| + 2(nx N + nu nl) words, "
                                  "ell bytes = 2|z| + shared/64 words (FP32); summed over all instances' T and "
                                  "[L*; L] applications of the FP32 solve, divided by its kernel time"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
