"""ctypes binding of libspock_amd.so (the HIP solver's C ABI, ``include/spock_amd.h``).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C spock.cu_amd/csrc``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SPK_LIB") or os.path.join(HERE, "lib", "libspock_amd.so")

F64, F32 = 0, 1
LS_HOUSEHOLDER, LS_GRAM = 0, 1
SPOCK, CP = 0, 1
OP_L, OP_LADJ, OP_DYN, OP_KERNEL, OP_T, OP_ELL, OP_DUAL = range(7)
STATUS = {-1: "notRun", 0: "converged", 1: "outOfIter", 2: "outOfTime"}

# Every entry point of include/spock_amd.h (checked against the header by tests/test_abi.py)
EXPORTS = [
    "spk_default_options", "spk_ctx_create", "spk_ctx_destroy", "spk_last_error", "spk_ctx_stream",
    "spk_ctx_synchronize",
    "spk_problem_load", "spk_problem_destroy", "spk_problem_sizes", "spk_problem_scaling",
    "spk_problem_step_size",
    "spk_batch_create", "spk_batch_create_shared", "spk_batch_is_shared", "spk_batch_destroy",
    "spk_batch_count", "spk_batch_set_workgroups", "spk_batch_set_schedule", "spk_batch_autotune",
    "spk_batch_autotune_workgroups", "spk_batch_solve_graph", "spk_batch_resume_graph", "spk_batch_reset",
    "spk_batch_set_x0", "spk_batch_set_tolerances", "spk_batch_set_iterate", "spk_batch_get_iterate",
    "spk_batch_set_iterates", "spk_batch_get_iterates", "spk_batch_iterate_ptr", "spk_batch_copy_iterates",
    "spk_batch_solve", "spk_batch_solve_noalloc", "spk_batch_resume", "spk_batch_results", "spk_batch_debug_op",
    "spk_batch_set_iterates_device", "spk_batch_get_iterates_device",
    "spk_batch_refill", "spk_batch_refill_from_batch", "spk_batch_refill_rows", "spk_batch_get_lanes_device",
    "spk_batch_pool_open", "spk_batch_set_chunk", "spk_batch_set_chunk_time", "spk_batch_set_debug", "spk_batch_debug_records",
    "spk_batch_autotune_uneven", "spk_batch_set_uneven_schedule", "spk_batch_replay_graph",
    "spk_mlp_create", "spk_mlp_destroy", "spk_mlp_num_params", "spk_mlp_get_params", "spk_mlp_set_params",
    "spk_mlp_get_grads", "spk_mlp_get_normalization", "spk_mlp_set_normalization", "spk_mlp_forward",
    "spk_mlp_loss_backward", "spk_mlp_adam", "spk_mlp_dataset", "spk_mlp_train", "spk_mlp_predict",
    "spk_mlp_predict_lanes", "spk_mlp_dataset_lanes",
]


class Options(C.Structure):
    _fields_ = [("tol_abs", C.c_double), ("tol_rel", C.c_double), ("max_time_secs", C.c_double),
                ("max_iters", C.c_long), ("max_inner", C.c_int), ("anderson_buffer", C.c_int),
                ("check_period", C.c_int), ("method", C.c_int), ("block_size", C.c_int),
                ("ls_method", C.c_int), ("dyn_inverse", C.c_int), ("chunk_iters", C.c_long)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_long), ("err", C.c_double), ("tol", C.c_double),
                ("w", C.c_double), ("k0", C.c_long), ("k1", C.c_long), ("k2", C.c_long), ("k2bt", C.c_long),
                ("k3", C.c_long), ("num_T", C.c_long), ("num_ell", C.c_long),
                ("phase_seconds", C.c_double * 10)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["phase_seconds"] = dict(zip(PHASES, list(self.phase_seconds)))
        return d


PHASES = ["adj", "proj", "dual", "res", "ell", "acol", "als", "aapply", "line", "other"]


_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_LIB = None


class SpockError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise SpockError(f"{path} is missing: build it with __graft_entry__.build() (make -C spock.cu_amd/csrc)")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.spk_default_options.argtypes = [C.POINTER(Options)]
    lib.spk_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.spk_ctx_destroy.argtypes = [vp]
    lib.spk_last_error.argtypes = [vp]
    lib.spk_last_error.restype = C.c_char_p
    lib.spk_ctx_stream.argtypes = [vp, C.POINTER(vp)]
    lib.spk_ctx_synchronize.argtypes = [vp]
    lib.spk_problem_load.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
    lib.spk_problem_destroy.argtypes = [vp]
    lib.spk_problem_sizes.argtypes = [vp, _lp]
    lib.spk_problem_scaling.argtypes = [vp, _dp]
    lib.spk_problem_step_size.argtypes = [vp]
    lib.spk_problem_step_size.restype = C.c_double
    lib.spk_batch_create.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp)]
    lib.spk_batch_create_shared.argtypes = [vp, vp, C.c_int, C.c_int, C.POINTER(vp)]
    lib.spk_batch_is_shared.argtypes = [vp]
    lib.spk_batch_destroy.argtypes = [vp]
    lib.spk_batch_count.argtypes = [vp]
    lib.spk_batch_set_workgroups.argtypes = [vp, C.c_int]
    lib.spk_batch_set_schedule.argtypes = [vp, C.c_int, C.c_int]
    lib.spk_batch_solve_graph.argtypes = [vp, vp, C.c_int, C.c_int]
    lib.spk_batch_autotune_workgroups.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int), _dp]
    lib.spk_batch_autotune_uneven.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
    lib.spk_batch_set_uneven_schedule.argtypes = [vp, C.c_int, C.c_int]
    lib.spk_batch_replay_graph.argtypes = [vp]
    lib.spk_batch_resume_graph.argtypes = [vp, C.c_int]
    lib.spk_batch_autotune.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
    lib.spk_batch_reset.argtypes = [vp]
    lib.spk_batch_set_debug.argtypes = [vp, C.c_long]
    lib.spk_batch_debug_records.argtypes = [vp, C.c_int, C.c_long, _dp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    lib.spk_batch_set_x0.argtypes = [vp, _dp]
    lib.spk_batch_set_tolerances.argtypes = [vp, _dp]
    lib.spk_batch_set_iterate.argtypes = [vp, C.c_int, _dp]
    lib.spk_batch_get_iterate.argtypes = [vp, C.c_int, _dp]
    lib.spk_batch_set_iterates.argtypes = [vp, _dp]
    lib.spk_batch_get_iterates.argtypes = [vp, _dp]
    lib.spk_batch_iterate_ptr.argtypes = [vp, C.c_int]
    lib.spk_batch_iterate_ptr.restype = vp
    lib.spk_batch_copy_iterates.argtypes = [vp, vp]
    lib.spk_batch_solve.argtypes = [vp, C.POINTER(Options), C.c_int]
    lib.spk_batch_solve_noalloc.argtypes = [vp, C.POINTER(Options), C.c_int]
    lib.spk_batch_resume.argtypes = [vp]
    lib.spk_batch_results.argtypes = [vp, C.POINTER(Result), _dp]
    lib.spk_batch_debug_op.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    lib.spk_batch_set_iterates_device.argtypes = [vp, vp, C.c_long]
    lib.spk_batch_get_iterates_device.argtypes = [vp, vp, C.c_long]
    _ip = C.POINTER(C.c_int)
    lib.spk_batch_refill.argtypes = [vp, C.c_int, _ip, _dp]
    lib.spk_batch_refill_from_batch.argtypes = [vp, C.c_int, _ip, _dp, vp, _ip]
    lib.spk_batch_refill_rows.argtypes = [vp, C.c_int, _ip, _dp, vp, C.c_long]
    lib.spk_batch_get_lanes_device.argtypes = [vp, C.c_int, _ip, vp, C.c_long]
    lib.spk_batch_pool_open.argtypes = [vp, C.POINTER(Options)]
    lib.spk_batch_set_chunk.argtypes = [vp, C.c_long]
    lib.spk_batch_set_chunk_time.argtypes = [vp, C.c_double]
    _fp = C.POINTER(C.c_float)
    lib.spk_mlp_create.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    lib.spk_mlp_destroy.argtypes = [vp]
    lib.spk_mlp_num_params.argtypes = [vp]
    lib.spk_mlp_num_params.restype = C.c_long
    for name in ("spk_mlp_get_params", "spk_mlp_set_params", "spk_mlp_get_grads"):
        getattr(lib, name).argtypes = [vp, _fp]
    lib.spk_mlp_get_normalization.argtypes = [vp, _fp, _fp, _fp, _fp]
    lib.spk_mlp_set_normalization.argtypes = [vp, _fp, _fp, _fp, _fp]
    lib.spk_mlp_forward.argtypes = [vp, _fp, C.c_int, _fp]
    lib.spk_mlp_loss_backward.argtypes = [vp, _fp, C.c_int, _dp]
    lib.spk_mlp_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
    lib.spk_mlp_dataset.argtypes = [vp, vp, _dp]
    lib.spk_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, _dp]
    lib.spk_mlp_predict.argtypes = [vp, _dp, C.c_int, vp]
    lib.spk_mlp_predict_lanes.argtypes = [vp, _dp, C.c_int, vp, _ip]
    lib.spk_mlp_dataset_lanes.argtypes = [vp, vp, C.c_int, _ip, _dp]
    _LIB = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def options(**kw):
    lib = load_library()
    o = Options()
    lib.spk_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    """A device context with its own HIP stream (``spk_ctx``)."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = C.c_void_p()
        if self.lib.spk_ctx_create(device, C.byref(h)) != 0:
            raise SpockError("spk_ctx_create failed (no HIP device?)")
        self.h = h

    def check(self, rc, what):
        if rc != 0:
            raise SpockError(f"{what}: {self.lib.spk_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.spk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Problem:
    """A problem folder loaded to the device (``spk_problem``)."""

    def __init__(self, ctx, folder):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.lib.spk_problem_load(ctx.h, folder.encode(), C.byref(h)), f"load {folder}")
        self.h = h
        s = np.zeros(16, dtype=np.int64)
        ctx.lib.spk_problem_sizes(h, s.ctypes.data_as(_lp))
        (self.N, self.nl, self.S, self.E, self.nx, self.nu, self.n_prim, self.n_dual, self.n_iter,
         self.preconditioned, self.con3, self.con5, self.distinct_factors, self.distinct_costs) = (int(v) for v in s[:14])
        self.scaling = np.zeros(self.nx + self.nu)
        ctx.lib.spk_problem_scaling(h, _d(self.scaling))
        self.step_size = ctx.lib.spk_problem_step_size(h)

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.spk_problem_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Batch:
    """A set of solve instances (``spk_batch``); one instance = one Cache of the reference.

    ``shared=True`` (all instances of one problem) selects the instance-interleaved layout solved
    in lockstep (``spk_batch_create_shared``): by the tile engine unless ``engine`` names the lane
    engine's schedule (0 or 8, see ``set_schedule``)."""

    def __init__(self, ctx, problems, dtype=F64, shared=False, engine=None):
        self.ctx = ctx
        self.problems = list(problems)
        self.dtype = dtype
        self.shared = bool(shared)
        h = C.c_void_p()
        if self.shared:
            p0 = self.problems[0]
            if any(p is not p0 for p in self.problems):
                raise SpockError("a shared batch holds instances of one problem")
            ctx.check(ctx.lib.spk_batch_create_shared(ctx.h, p0.h, len(self.problems), dtype, C.byref(h)),
                      "shared batch create")
        else:
            arr = (C.c_void_p * len(self.problems))(*[p.h for p in self.problems])
            ctx.check(ctx.lib.spk_batch_create(ctx.h, len(self.problems), arr, dtype, C.byref(h)), "batch create")
        self.h = h
        self.sizes = [p.n_iter for p in self.problems]
        self.total = int(sum(self.sizes))
        if self.shared and engine is not None:
            self.set_schedule(engine)

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.spk_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def reset(self):
        self.ctx.check(self.ctx.lib.spk_batch_reset(self.h), "reset")

    def set_debug(self, max_records):
        """Debug traces of later solves (per-instance batches; 0: off): one record per error check."""
        self.ctx.check(self.ctx.lib.spk_batch_set_debug(self.h, int(max_records)), "set debug")
        self.debug_cap = int(max_records)

    def debug_records(self, k=0):
        """Instance k's records of the last solve, rows {idx, callsL, err0, err1, err2, err3}, and the
        first non-finite candidate entry (-1: none; status SPK_NONFINITE = 3 when found)."""
        cap = getattr(self, "debug_cap", 0)
        rec = np.zeros((max(cap, 1), 6))
        n, bad = C.c_long(), C.c_long()
        self.ctx.check(self.ctx.lib.spk_batch_debug_records(self.h, k, cap, _d(rec), C.byref(n), C.byref(bad)),
                       "debug records")
        return rec[:min(n.value, cap)], bad.value

    def set_schedule(self, mv, block=0):
        """Engine and schedule of a shared batch: mv 17 (the default) or 16 the tile engine (MFMA, 16
        instances per workgroup; built for up to 1024 or 512 threads), 0 or 8 the lane engine (64 per
        workgroup; scalar or LDS-staged factor loads).  block: workgroup size (0: the schedule's
        default)."""
        self.ctx.check(self.ctx.lib.spk_batch_set_schedule(self.h, int(mv), int(block)), "set_schedule")

    AUTOTUNE_CANDIDATES = [(0, 1024), (8, 1024), (0, 512), (8, 512), (16, 256), (16, 512), (17, 1024)]

    def autotune(self, opts=None, iters=50):
        """Time the engines and schedules on this batch and keep the fastest; returns
        (mv, block, {(mv, block): seconds})."""
        mv, blk = C.c_int(), C.c_int()
        sec = np.zeros(len(self.AUTOTUNE_CANDIDATES))
        o = opts if opts is not None else options()
        self.ctx.check(self.ctx.lib.spk_batch_autotune(self.h, C.byref(o), int(iters), C.byref(mv), C.byref(blk),
                                                        _d(sec)), "autotune")
        return mv.value, blk.value, dict(zip(self.AUTOTUNE_CANDIDATES, sec.tolist()))

    def autotune_workgroups(self, opts=None, iters=50):
        """One-instance batch: time 1/8/32/64/128 cooperating workgroups and keep the fastest;
        returns (workgroups, {workgroups: seconds})."""
        w = C.c_int()
        sec = np.zeros(5)
        o = opts if opts is not None else options()
        self.ctx.check(self.ctx.lib.spk_batch_autotune_workgroups(self.h, C.byref(o), int(iters), C.byref(w), _d(sec)),
                       "autotune_workgroups")
        return w.value, dict(zip([1, 8, 32, 64, 128], sec.tolist()))

    UNEVEN_CANDIDATES = [(128, 0), (256, 0), (512, 0), (128, 1), (256, 1), (512, 1)]

    def autotune_uneven(self, opts=None, iters=25):
        """Per-instance batch: time workgroups of 128/256/512 threads, instances in batch order or
        largest first, and keep the fastest; returns ((block, largest_first), {candidate: seconds})."""
        blk, lf = C.c_int(), C.c_int()
        sec = np.zeros(len(self.UNEVEN_CANDIDATES))
        o = opts if opts is not None else options()
        self.ctx.check(self.ctx.lib.spk_batch_autotune_uneven(self.h, C.byref(o), int(iters), C.byref(blk),
                                                              C.byref(lf), _d(sec)), "autotune_uneven")
        return (blk.value, lf.value), dict(zip(self.UNEVEN_CANDIDATES, sec.tolist()))

    def set_uneven_schedule(self, block=0, largest_first=False):
        self.ctx.check(self.ctx.lib.spk_batch_set_uneven_schedule(self.h, int(block), 1 if largest_first else 0),
                       "set_uneven_schedule")

    def set_workgroups(self, n):
        """Workgroups that solve the single instance together (0 or 1: one workgroup,
        n > 1: a cooperative launch of n workgroups)."""
        self.ctx.check(self.ctx.lib.spk_batch_set_workgroups(self.h, int(n)), "set_workgroups")

    def set_x0(self, x0):
        x0 = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in x0])
                                  if isinstance(x0, (list, tuple)) else np.asarray(x0, dtype=np.float64).ravel())
        self.ctx.check(self.ctx.lib.spk_batch_set_x0(self.h, _d(x0)), "set_x0")

    def set_tolerances(self, tol):
        """Per-instance acceptance tolerances (scaled units, as ``results()[k]['tol']``); None clears."""
        if tol is None:
            self.ctx.check(self.ctx.lib.spk_batch_set_tolerances(self.h, None), "set_tolerances")
            return
        t = np.ascontiguousarray(np.asarray(tol, dtype=np.float64).ravel())
        assert t.size == len(self.problems)
        self.ctx.check(self.ctx.lib.spk_batch_set_tolerances(self.h, _d(t)), "set_tolerances")

    def set_iterates(self, z):
        z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).ravel())
        assert z.size == self.total
        self.ctx.check(self.ctx.lib.spk_batch_set_iterates(self.h, _d(z)), "set_iterates")

    def iterates(self):
        z = np.zeros(self.total)
        self.ctx.check(self.ctx.lib.spk_batch_get_iterates(self.h, _d(z)), "get_iterates")
        return z

    def iterate(self, k):
        z = np.zeros(self.sizes[k])
        self.ctx.check(self.ctx.lib.spk_batch_get_iterate(self.h, k, _d(z)), "get_iterate")
        return z

    def set_iterates_device(self, ptr, ld):
        """Iterates from device rows (FP32, instance-major, leading dimension ld)."""
        self.ctx.check(self.ctx.lib.spk_batch_set_iterates_device(self.h, C.c_void_p(ptr), ld), "set_iterates_device")

    def get_iterates_device(self, ptr, ld):
        self.ctx.check(self.ctx.lib.spk_batch_get_iterates_device(self.h, C.c_void_p(ptr), ld), "get_iterates_device")

    def copy_iterates_from(self, other):
        self.ctx.check(self.ctx.lib.spk_batch_copy_iterates(self.h, other.h), "copy_iterates")

    def solve(self, opts=None, warm=False):
        o = opts if opts is not None else options()
        self.ctx.check(self.ctx.lib.spk_batch_solve(self.h, C.byref(o), 1 if warm else 0), "solve")

    def solve_noalloc(self, opts=None, warm=False):
        """spk_batch_solve without device allocations (per-instance batches; the solve's working
        memory is the device's fixed scratch of SPK_NOALLOC_SCRATCH_BYTES, 64 MiB)."""
        o = opts if opts is not None else options()
        self.ctx.check(self.ctx.lib.spk_batch_solve_noalloc(self.h, C.byref(o), 1 if warm else 0), "solve_noalloc")

    def resume(self):
        """Continue a chunked solve (shared batches, options.chunk_iters > 0)."""
        self.ctx.check(self.ctx.lib.spk_batch_resume(self.h), "resume")

    def solve_graph(self, opts, warm=False, launches=None):
        """A chunked solve (opts.chunk_iters > 0) captured as one HIP graph of `launches` launches
        (default: enough for opts.max_iters).  Enqueues only; results() waits."""
        if launches is None:
            launches = (max(1, -(-max(1, opts.max_iters) // max(1, opts.chunk_iters)) + 1)
                        if self.shared else 1)
        self.ctx.check(self.ctx.lib.spk_batch_solve_graph(self.h, C.byref(opts), 1 if warm else 0, int(launches)),
                       "solve_graph")

    def replay_graph(self):
        """Per-instance batch: launch the solve captured by solve_graph again on the batch's current
        initial states, iterates and tolerances.  Enqueues only; results() waits."""
        self.ctx.check(self.ctx.lib.spk_batch_replay_graph(self.h), "replay_graph")

    def resume_graph(self, launches):
        """Continue a chunked solve with `launches` resume launches captured as one HIP graph."""
        self.ctx.check(self.ctx.lib.spk_batch_resume_graph(self.h, int(launches)), "resume_graph")

    def _lanes_x0(self, lanes, x0):
        lanes = np.ascontiguousarray(np.asarray(lanes, dtype=np.int32).ravel())
        x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).reshape(lanes.size, -1))
        return lanes, x0

    def refill(self, lanes, x0):
        """Continuous batching: cold-start new solves with initial states x0 (rows) in slots
        ``lanes`` of a chunked solve; they run from the next resume (``spk_batch_refill``)."""
        lanes, x0 = self._lanes_x0(lanes, x0)
        self.ctx.check(self.ctx.lib.spk_batch_refill(self.h, lanes.size, lanes.ctypes.data_as(C.POINTER(C.c_int)),
                                                     _d(x0)), "refill")

    def refill_from(self, lanes, x0, src, src_lanes):
        """Warm-start new solves in slots ``lanes`` from the accepted iterates of slots ``src_lanes``
        of the shared batch ``src`` (``spk_batch_refill_from_batch``)."""
        lanes, x0 = self._lanes_x0(lanes, x0)
        src_lanes = np.ascontiguousarray(np.asarray(src_lanes, dtype=np.int32).ravel())
        ip = C.POINTER(C.c_int)
        self.ctx.check(self.ctx.lib.spk_batch_refill_from_batch(self.h, lanes.size, lanes.ctypes.data_as(ip), _d(x0),
                                                                src.h, src_lanes.ctypes.data_as(ip)), "refill_from")

    def pool_open(self, opts):
        """An empty pool: every slot idle until a refill (``spk_batch_pool_open``)."""
        self.ctx.check(self.ctx.lib.spk_batch_pool_open(self.h, C.byref(opts)), "pool_open")

    def set_chunk(self, chunk_iters):
        """Outer iterations per launch of the following resumes (``spk_batch_set_chunk``)."""
        self.ctx.check(self.ctx.lib.spk_batch_set_chunk(self.h, int(chunk_iters)), "set_chunk")

    def set_chunk_time(self, seconds):
        """Wall-clock slice of each launch of a chunked solve, 0 for none (``spk_batch_set_chunk_time``)."""
        self.ctx.check(self.ctx.lib.spk_batch_set_chunk_time(self.h, float(seconds)), "set_chunk_time")

    def solve_chunked(self, opts, warm=False, on_chunk=None):
        """Run a chunked solve to its end; on_chunk(results, kernel_seconds) after every chunk.
        Returns the final results and the summed kernel seconds."""
        self.solve(opts, warm)
        total = 0.0
        while True:
            res, sec = self.results()
            total += sec
            if on_chunk is not None:
                on_chunk(res, sec)
            if all(r["status"] != -1 for r in res):
                return res, total
            self.resume()

    def results(self):
        res = (Result * len(self.problems))()
        sec = C.c_double()
        self.ctx.check(self.ctx.lib.spk_batch_results(self.h, res, C.byref(sec)), "results")
        return [r.as_dict() for r in res], sec.value

    def debug_op(self, op, inp, out=None):
        inp = np.ascontiguousarray(np.asarray(inp, dtype=np.float64).ravel())
        o = np.zeros(self.total) if out is None else np.ascontiguousarray(np.asarray(out, dtype=np.float64).ravel())
        sc = np.zeros(2 * len(self.problems))
        self.ctx.check(self.ctx.lib.spk_batch_debug_op(self.h, op, _d(inp), _d(o), _d(sc)), "debug_op")
        return o, sc


_fp = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


class WarmStartMLP:
    """The learned warm start (``spk_mlp``): x0 -> primal-dual iterate, trained on solved iterates."""

    def __init__(self, ctx, n_in, n_out, hidden=512, layers=3, seed=0):
        self.ctx = ctx
        self.n_in, self.n_out, self.hidden, self.layers = n_in, n_out, hidden, layers
        h = C.c_void_p()
        ctx.check(ctx.lib.spk_mlp_create(ctx.h, n_in, hidden, layers, n_out, seed, C.byref(h)), "mlp create")
        self.h = h
        self.n_params = int(ctx.lib.spk_mlp_num_params(h))

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.spk_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.spk_mlp_get_params(self.h, _f(p)), "mlp params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.ctx.check(self.ctx.lib.spk_mlp_set_params(self.h, _f(p)), "mlp set params")

    def grads(self):
        g = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx.lib.spk_mlp_get_grads(self.h, _f(g)), "mlp grads")
        return g

    def normalization(self):
        a, b = np.zeros(self.n_in, np.float32), np.zeros(self.n_in, np.float32)
        c, d = np.zeros(self.n_out, np.float32), np.zeros(self.n_out, np.float32)
        self.ctx.check(self.ctx.lib.spk_mlp_get_normalization(self.h, _f(a), _f(b), _f(c), _f(d)), "mlp norm")
        return a, b, c, d

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (in_mean, in_std, out_mean, out_std)]
        self.ctx.check(self.ctx.lib.spk_mlp_set_normalization(self.h, *[_f(v) for v in arrs]), "mlp set norm")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self.ctx.check(self.ctx.lib.spk_mlp_forward(self.h, _f(x), x.shape[0], _f(y)), "mlp forward")
        return y

    def loss_backward(self, target):
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_double()
        self.ctx.check(self.ctx.lib.spk_mlp_loss_backward(self.h, _f(t), t.shape[0], C.byref(loss)), "mlp backward")
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(self.ctx.lib.spk_mlp_adam(self.h, lr, beta1, beta2, eps), "mlp adam")

    def dataset(self, solved_batch, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self.ctx.check(self.ctx.lib.spk_mlp_dataset(self.h, solved_batch.h, _d(x0)), "mlp dataset")

    def dataset_lanes(self, solved_batch, lanes, x0):
        """Training set from slots ``lanes`` of a solved shared batch (x0: their initial states)."""
        lanes = np.ascontiguousarray(np.asarray(lanes, dtype=np.int32).ravel())
        if lanes.size == 0:
            raise ValueError("mlp dataset_lanes: no slots given (no solved instance to learn from)")
        x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).reshape(lanes.size, -1))
        self.ctx.check(self.ctx.lib.spk_mlp_dataset_lanes(self.h, solved_batch.h, lanes.size,
                                                          lanes.ctypes.data_as(C.POINTER(C.c_int)), _d(x0)),
                       "mlp dataset_lanes")

    def train(self, epochs, minibatch=256, lr=1e-3, seed=0):
        losses = np.zeros(epochs)
        self.ctx.check(self.ctx.lib.spk_mlp_train(self.h, epochs, minibatch, lr, seed, _d(losses)), "mlp train")
        return losses

    def predict(self, x0, batch):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self.ctx.check(self.ctx.lib.spk_mlp_predict(self.h, _d(x0), x0.shape[0], batch.h), "mlp predict")

    def predict_lanes(self, x0, batch, lanes):
        """Continuous batching: new solves in slots ``lanes`` of ``batch`` warm-started from the
        prediction for initial states x0 (``spk_mlp_predict_lanes``)."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        lanes = np.ascontiguousarray(np.asarray(lanes, dtype=np.int32).ravel())
        assert lanes.size == x0.shape[0]
        self.ctx.check(self.ctx.lib.spk_mlp_predict_lanes(self.h, _d(x0), x0.shape[0], batch.h,
                                                          lanes.ctypes.data_as(C.POINTER(C.c_int))), "mlp predict_lanes")

    def save(self, path):
        a, b, c, d = self.normalization()
        np.savez(path, params=self.params(), in_mean=a, in_std=b, out_mean=c, out_std=d,
                 dims=np.array([self.n_in, self.hidden, self.layers, self.n_out]))

    def load(self, path):
        z = np.load(path, allow_pickle=False)
        dims = [int(v) for v in z["dims"]]
        if dims != [self.n_in, self.hidden, self.layers, self.n_out]:
            raise SpockError(f"{path}: network dims {dims} do not match this model")
        self.set_params(z["params"])
        self.set_normalization(z["in_mean"], z["in_std"], z["out_mean"], z["out_std"])
