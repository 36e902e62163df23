"""ctypes front end of the FP64 CPU oracle (``oracle/spock_oracle.cpp``).

Test infrastructure only: used by ``tests/``, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` as the checker of the HIP solver.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libspock_oracle.so")

_COST = {"quadratic": 0, "quadraticPlusLinear": 1, "linear": 2}
_CON = {"no": 0, "rectangle": 1, "polyhedron": 2, "polyhedronWithIdentity": 3}

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _lib():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.orc_load.restype = C.c_void_p
    lib.orc_load.argtypes = [C.c_char_p, _lp, C.c_double, C.c_char_p, C.c_int]
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_sizes.argtypes = [C.c_void_p, _lp]
    for name in ("orc_L", "orc_Ladj"):
        getattr(lib, name).argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_dyn_project.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_kernel_project.argtypes = [C.c_void_p, _dp]
    lib.orc_project_dual.argtypes = [C.c_void_p, _dp]
    lib.orc_soc.argtypes = [_dp, C.c_int]
    lib.orc_box.restype = C.c_double
    lib.orc_box.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.orc_T.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_dotM.restype = C.c_double
    lib.orc_dotM.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_lstsq.argtypes = [_dp, C.c_long, C.c_int, _dp, _dp]
    lib.orc_lstsq_gram.restype = C.c_int
    lib.orc_lstsq_gram.argtypes = [_dp, C.c_long, C.c_int, _dp, _dp]
    lib.orc_solve.restype = C.c_int
    lib.orc_solve.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _lp, _dp, _lp, _dp, _ip, _dp, C.c_long]
    lib.orc_solve_debug.restype = C.c_int
    lib.orc_solve_debug.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _lp, _dp, _lp, _dp, _dp, _dp, C.c_long, _lp]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def read_info(folder):
    with open(os.path.join(folder, "data.json")) as fh:
        return json.load(fh)


class Oracle:
    """One loaded problem folder."""

    def __init__(self, folder):
        self.folder = folder
        info = read_info(folder)
        self.info = info
        ints = np.array([info["numNodes"], info["numNonleafNodes"], info["numStages"], info["numEvents"],
                         info["numStates"], info["numInputs"], _COST[info["cost"]["nonleaf"]],
                         _COST[info["cost"]["leaf"]], _CON[info["constraint"]["nonleaf"]],
                         _CON[info["constraint"]["leaf"]], int(bool(info["preconditioned"])),
                         int(info["dynamics"]["type"] == "affine")], dtype=np.int64)
        err = C.create_string_buffer(512)
        path = folder if folder.endswith("/") else folder + "/"
        self.h = lib().orc_load(path.encode(), ints.ctypes.data_as(_lp), float(info["stepSize"]), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())
        s = np.zeros(14, dtype=np.int64)
        lib().orc_sizes(self.h, s.ctypes.data_as(_lp))
        (self.n_prim, self.n_dual, self.n_iter, self.sU, self.sX, self.sY, self.sT, self.sS,
         self.sI, self.sII, self.sIII, self.sIV, self.sV, self.sVI) = (int(v) for v in s)
        self.nx, self.nu = info["numStates"], info["numInputs"]
        self.gamma = float(info["stepSize"])
        self.scaling = None
        if info["preconditioned"]:
            from spock_amd.api.btio import read_flat
            self.scaling = read_flat(os.path.join(folder, "scaling_d.bt"))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def scaled_x0(self, x0):
        x0 = np.asarray(x0, dtype=np.float64)
        return x0 * self.scaling[:self.nx] if self.scaling is not None else x0.copy()

    def L(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        out = np.zeros(self.n_dual)
        lib().orc_L(self.h, _d(prim), _d(out))
        return out

    def Ladj(self, dual):
        dual = np.ascontiguousarray(dual, dtype=np.float64)
        out = np.zeros(self.n_prim)
        lib().orc_Ladj(self.h, _d(dual), _d(out))
        return out

    def dyn_project(self, x0s, x, u):
        x = np.array(x, dtype=np.float64)
        u = np.array(u, dtype=np.float64)
        x0s = np.ascontiguousarray(x0s, dtype=np.float64)
        lib().orc_dyn_project(self.h, _d(x0s), _d(x), _d(u))
        return x, u

    def kernel_project(self, prim):
        prim = np.array(prim, dtype=np.float64)
        lib().orc_kernel_project(self.h, _d(prim))
        return prim

    def project_dual(self, w):
        w = np.array(w, dtype=np.float64)
        lib().orc_project_dual(self.h, _d(w))
        return w

    def T(self, x0s, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        x0s = np.ascontiguousarray(x0s, dtype=np.float64)
        out = np.zeros(self.n_iter)
        lib().orc_T(self.h, _d(x0s), _d(z), _d(out))
        return out

    def dotM(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return lib().orc_dotM(self.h, _d(x), _d(y))

    def solve(self, x0, method="spock", warm=None, tol_abs=1e-3, tol_rel=0.0, max_time=0.0, max_iters=0,
              max_inner=8, anderson=3, check_period=25, trace_len=0, ls="householder", tol_fixed=0.0,
              adjoint_dots=False):
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        z = np.zeros(self.n_iter)
        out = np.zeros(7, dtype=np.int64)
        outd = np.zeros(2)
        dopts = np.array([tol_abs, tol_rel, max_time, tol_fixed])
        iopts = np.array([max_iters, max_inner, anderson, check_period, 1 if ls == "gram" else 0,
                          1 if adjoint_dots else 0], dtype=np.int64)
        wp = _d(np.ascontiguousarray(warm, dtype=np.float64)) if warm is not None else None
        tb = np.zeros(trace_len, dtype=np.int32) if trace_len else None
        tw = np.zeros(trace_len) if trace_len else None
        lib().orc_solve(self.h, 0 if method == "spock" else 1, _d(x0), wp, _d(dopts), iopts.ctypes.data_as(_lp),
                        _d(z), out.ctypes.data_as(_lp), _d(outd),
                        tb.ctypes.data_as(_ip) if tb is not None else None,
                        _d(tw) if tw is not None else None, trace_len)
        res = dict(z=z, iters=int(out[0]), status=int(out[1]), k0=int(out[2]), k1=int(out[3]), k2=int(out[4]),
                   k2bt=int(out[5]), k3=int(out[6]), err=float(outd[0]), tol=float(outd[1]))
        if trace_len:
            res["branch"], res["w"] = tb, tw
        return res


    def solve_debug(self, x0, err_acc, method="spock", warm=None, tol_abs=1e-3, tol_rel=0.0, max_iters=0,
                    max_inner=8, anderson=3, check_period=25, cap=100000):
        """solve() with the reference's debug caches (src/cache.cuh:947-959).  err_acc (n_prim,
        float64, updated in place) is err0's running vector: zeros after a reset, carried from solve
        to solve otherwise.  Returns solve()'s fields plus records (rows {idx, callsL, err0, err1,
        err2, err3}), num_T and nonfinite (first non-finite candidate entry, status 3; else -1)."""
        assert err_acc.dtype == np.float64 and err_acc.size == self.n_prim and err_acc.flags.c_contiguous
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        z = np.zeros(self.n_iter)
        out = np.zeros(7, dtype=np.int64)
        outd = np.zeros(2)
        dopts = np.array([tol_abs, tol_rel, 0.0, 0.0])
        iopts = np.array([max_iters, max_inner, anderson, check_period, 0, 0], dtype=np.int64)
        wp = _d(np.ascontiguousarray(warm, dtype=np.float64)) if warm is not None else None
        rec = np.zeros((cap, 6))
        info = np.zeros(3, dtype=np.int64)
        lib().orc_solve_debug(self.h, 0 if method == "spock" else 1, _d(x0), wp, _d(dopts),
                              iopts.ctypes.data_as(_lp), _d(z), out.ctypes.data_as(_lp), _d(outd), _d(err_acc),
                              _d(rec), cap, info.ctypes.data_as(_lp))
        return dict(z=z, iters=int(out[0]), status=int(out[1]), k0=int(out[2]), k1=int(out[3]), k2=int(out[4]),
                    k2bt=int(out[5]), k3=int(out[6]), err=float(outd[0]), tol=float(outd[1]),
                    records=rec[:min(int(info[0]), cap)], num_T=int(info[2]), nonfinite=int(info[1]))


def lstsq(Y, r):
    Y = np.asfortranarray(Y, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    g = np.zeros(Y.shape[1])
    lib().orc_lstsq(_d(Y), Y.shape[0], Y.shape[1], _d(r), _d(g))
    return g


def lstsq_gram(Y, r):
    """Normal-equation least squares as the GPU engines' Gram option; None when it breaks down."""
    Y = np.asfortranarray(Y, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    g = np.zeros(Y.shape[1])
    ok = lib().orc_lstsq_gram(_d(Y), Y.shape[0], Y.shape[1], _d(r), _d(g))
    return g if ok else None


def soc(v):
    v = np.array(v, dtype=np.float64)
    lib().orc_soc(_d(v), v.size)
    return v


def box(v, lb, ub):
    return lib().orc_box(float(v), float(lb), float(ub))
