"""Offline problem generation: preconditioning, dynamics factors, kernel
projectors, step size and the problem file set.

Follows ``api/python/spock/problem.py``:

* ``Factory`` builder (``:702-879``) with event-indexed dynamics and costs;
  node 0 carries zero dynamics and a zeroed cost.
* preconditioning (``:159-190``): per-coordinate scale = max(1, sqrt(Q_ii),
  sqrt(R_ii), |q_i|, |r_i|) times sqrt(max #children).
* dynamics factors (``:192-213``): backward recursion over nonleaf nodes with
  R~ = I + sum B'PB = L L', K = -R~^-1 sum B'PA, ABKTr = (A+BK)', APB =
  (A+BK)'PB, P = I + K'K + sum (A+BK)'P(A+BK).
* kernel projectors (``:215-231``): NN' = orthogonal projector onto
  ker [E'_pad, -I_pad, -I_pad].
* b padding (``:263-267``) and the step size 0.99/||L|| (``:402-417``).
* file set and ``data.json`` keys (``:437-687``, ``api/python/spock/jinja2.py``).
"""
import json
import os
import numpy as np
import scipy.linalg as sla
from scipy.sparse.linalg import LinearOperator, eigsh

from . import build
from .btio import write_bt


def _stack(mats):
    """np.dstack of 2-D (or column) arrays -> (rows, cols, n)."""
    return np.dstack([np.asarray(m, dtype=np.float64).reshape(np.asarray(m).shape[0], -1) for m in mats])


class Problem:
    def __init__(self, tree, nx, nu, dynamics, nonleaf_cost, leaf_cost, nonleaf_con, leaf_con, risks,
                 precondition=True, julia=False):
        self.tree, self.nx, self.nu = tree, nx, nu
        self.julia = julia
        self.dyn = dynamics
        self.ncost, self.lcost = nonleaf_cost, leaf_cost
        self.ncon, self.lcon = nonleaf_con, leaf_con
        self.risk = risks[0]
        self.risks = list(risks)
        self.E_list, self.b_list = [], []
        for i in range(tree.num_nonleaf_nodes):
            E, b = risks[i].make(tree.cond_prob_of_children_of_node(i))
            self.E_list.append(E)
            self.b_list.append(b)
        self.preconditioned = False
        self.scaling = None
        if precondition:
            self._precondition()
        self._dynamics_factors()
        self._kernel_projectors()
        self._pad_b()
        self.yD = 2 * tree.num_events + 1
        self.nD = 4 * tree.num_events + 1
        self._sizes()
        self.step_size = self._step_size()

    # ------------------------------------------------------------------ offline
    def _precondition(self):
        sx = np.ones(self.nx)
        su = np.ones(self.nu)
        sx = self.ncost.scaling_states(sx)
        su = self.ncost.scaling_inputs(su)
        sx = self.lcost.scaling_states(sx)
        mul = np.sqrt(self.tree.max_num_children)
        sx, su = sx * mul, su * mul
        self.preconditioned = not (np.allclose(sx, 1.0) and np.allclose(su, 1.0))
        if not self.preconditioned:
            return
        self.scaling = np.hstack((sx, su))
        sx_inv, su_inv, sx_m = np.diag(1 / sx), np.diag(1 / su), np.diag(sx)
        self.dyn = [d.condition(sx_inv, su_inv, sx_m) for d in self.dyn]
        self.ncost.condition(sx_inv, su_inv)
        self.lcost.condition(sx_inv)
        self.ncon.condition(np.diag(self.scaling))
        self.lcon.condition(sx_m)

    def _dynamics_factors(self):
        t, nx, nu = self.tree, self.nx, self.nu
        N, nl = t.num_nodes, t.num_nonleaf_nodes
        self.P = [np.zeros((nx, nx)) for _ in range(N)]
        self.K = [np.zeros((nu, nx)) for _ in range(nl)]
        self.L = [np.zeros((nu, nu)) for _ in range(nl)]
        self.ABKTr = [np.zeros((nx, nx)) for _ in range(N)]
        self.APB = [np.zeros((nx, nu)) for _ in range(N)]
        for i in range(nl, N):
            self.P[i] = np.eye(nx)
        for i in reversed(range(nl)):
            ch = t.children_of_node(i)
            r_sum = sum(self.dyn[j].B.T @ self.P[j] @ self.dyn[j].B for j in ch)
            k_sum = sum(self.dyn[j].B.T @ self.P[j] @ self.dyn[j].A for j in ch)
            self.L[i] = sla.cholesky(np.eye(nu) + r_sum, lower=True, check_finite=True)
            self.K[i] = sla.cho_solve((self.L[i], True), -k_sum)
            p_sum = 0
            for j in ch:
                acl = self.dyn[j].A + self.dyn[j].B @ self.K[i]
                self.ABKTr[j] = acl.T
                p_sum = p_sum + acl.T @ self.P[j] @ acl
                self.APB[j] = acl.T @ self.P[j] @ self.dyn[j].B
            self.P[i] = np.eye(nx) + self.K[i].T @ self.K[i] + p_sum

    def _kernel_projectors(self):
        t = self.tree
        E = t.num_events
        self.NN = []
        self.S2 = []
        for i in range(t.num_nonleaf_nodes):
            n = int(t.num_children[i])
            fill = E - n
            et = np.pad(self.E_list[i].T, [(0, fill), (0, 2 * fill)])
            eye = np.pad(np.eye(n), [(0, fill), (0, fill)])
            s2 = np.hstack((et, -eye, -eye))
            ns = sla.null_space(s2)
            self.S2.append(s2)
            self.NN.append(ns @ ns.T)

    def _pad_b(self):
        E = self.tree.num_events
        self.b_pad = [np.pad(b, [(0, 2 * (E - b.size // 2)), (0, 0)]) for b in self.b_list]

    # ------------------------------------------------------------------ operator
    def _sizes(self):
        t = self.tree
        N, nl, nlf = t.num_nodes, t.num_nonleaf_nodes, t.num_leaf_nodes
        nx, nu = self.nx, self.nu
        self.size_prim = nu * nl + nx * N + self.yD * nl + 2 * N
        con_nl = 0 if self.ncon.kind == "no" else self.ncon.dim_per_node
        con_l = 0 if self.lcon.kind == "no" else self.lcon.dim_per_node
        self.d4, self.d6 = self.ncost.block(nx, nu), self.lcost.block(nx, 0)
        self.size_dual = self.yD * nl + nl + con_nl * nl + self.d4 * N + con_l * nlf + self.d6 * nlf

    def _stacks(self):
        t = self.tree
        anc = t.ancestors_arr.copy()
        anc[0] = 0
        nc, lc = self.ncost, self.lcost
        return dict(
            anc=anc,
            sQ=None if nc.is_linear else np.array(nc.sqrtQ), sR=None if nc.is_linear else np.array(nc.sqrtR),
            q=np.array(nc.q)[:, :, 0], r=np.array(nc.r)[:, :, 0],
            lsQ=None if lc.is_linear else np.array(lc.sqrtQ), lq=np.array(lc.q)[:, :, 0],
            b=np.array(self.b_pad)[:, :, 0],
        )

    def op(self, prim):
        """Dense-free L: prim -> dual (``src/operator.cuh:41-60`` semantics, clean node-0 rows)."""
        t, nx, nu, yD = self.tree, self.nx, self.nu, self.yD
        N, nl, nlf = t.num_nodes, t.num_nonleaf_nodes, t.num_leaf_nodes
        st = self._stacks()
        p = np.asarray(prim, dtype=np.float64).reshape(-1)
        o = 0
        u = p[o:o + nu * nl].reshape(nl, nu); o += nu * nl
        x = p[o:o + nx * N].reshape(N, nx); o += nx * N
        y = p[o:o + yD * nl].reshape(nl, yD); o += yD * nl
        tt = p[o:o + N]; o += N
        s = p[o:o + N]
        I = y
        II = s[:nl] - np.einsum("ij,ij->i", st["b"], y)
        xu = np.hstack((x[:nl], u))
        III = self._con_op(self.ncon, xu)
        xa, ua = x[st["anc"]], u[np.minimum(st["anc"], nl - 1)]
        lin = np.einsum("ni,ni->n", st["q"], xa) + np.einsum("ni,ni->n", st["r"], ua)
        if self.ncost.is_linear:  # [q'x + r'u of every node; t] (src/costs.cuh:261-267)
            lin[0] = 0.0
            IV = np.concatenate([lin, tt])
        else:
            IV = np.zeros((N, nx + nu + 2))
            IV[:, :nx] = np.einsum("nij,nj->ni", st["sQ"], xa)
            IV[:, nx:nx + nu] = np.einsum("nij,nj->ni", st["sR"], ua)
            half = 0.5 * tt - 0.5 * lin
            IV[:, nx + nu] = half
            IV[:, nx + nu + 1] = half
            IV[0] = 0.0
        xl = x[nl:]
        V = self._con_op(self.lcon, xl)
        if self.lcost.is_linear:  # [q'x of every leaf; s] (src/costs.cuh:269-272)
            VI = np.concatenate([np.einsum("ni,ni->n", st["lq"], xl), s[nl:]])
        else:
            VI = np.zeros((nlf, nx + 2))
            VI[:, :nx] = np.einsum("nij,nj->ni", st["lsQ"], xl)
            hs = 0.5 * s[nl:] - 0.5 * np.einsum("ni,ni->n", st["lq"], xl)
            VI[:, nx] = hs
            VI[:, nx + 1] = hs
        return np.concatenate([I.ravel(), II, III.ravel(), IV.ravel(), V.ravel(), VI.ravel()])

    def adj(self, dual):
        """Dense-free L*: dual -> prim (``src/operator.cuh:63-82`` semantics)."""
        t, nx, nu, yD = self.tree, self.nx, self.nu, self.yD
        N, nl, nlf = t.num_nodes, t.num_nonleaf_nodes, t.num_leaf_nodes
        st = self._stacks()
        d = np.asarray(dual, dtype=np.float64).reshape(-1)
        o = 0
        I = d[o:o + yD * nl].reshape(nl, yD); o += yD * nl
        II = d[o:o + nl]; o += nl
        c3 = 0 if self.ncon.kind == "no" else self.ncon.dim_per_node
        III = d[o:o + c3 * nl].reshape(nl, c3); o += c3 * nl
        IV = d[o:o + self.d4 * N].reshape(-1, N if self.ncost.is_linear else self.d4); o += self.d4 * N
        c5 = 0 if self.lcon.kind == "no" else self.lcon.dim_per_node
        V = d[o:o + c5 * nlf].reshape(nlf, c5); o += c5 * nlf
        VI = d[o:o + self.d6 * nlf].reshape(-1, nlf if self.lcost.is_linear else self.d6)
        s = np.zeros(N)
        s[:nl] = II
        y = I - st["b"] * II[:, None]
        xu = self._con_adj(self.ncon, III, nl, nx + nu)
        x = np.zeros((N, nx))
        u = xu[:, nx:].copy()
        x[:nl] = xu[:, :nx]
        if self.ncost.is_linear:  # src/costs.cuh:274-287: t is the second half, whole
            cx, cu = st["q"] * IV[0][:, None], st["r"] * IV[0][:, None]
            tt = IV[1].copy()
        else:
            cd = IV[:, nx + nu] + IV[:, nx + nu + 1]
            cx = np.einsum("nij,nj->ni", st["sQ"], IV[:, :nx]) - 0.5 * st["q"] * cd[:, None]
            cu = np.einsum("nij,nj->ni", st["sR"], IV[:, nx:nx + nu]) - 0.5 * st["r"] * cd[:, None]
            tt = 0.5 * cd
            tt[0] = 0.0
        cx[0] = 0.0
        cu[0] = 0.0
        np.add.at(x, st["anc"][1:], cx[1:])
        np.add.at(u, st["anc"][1:], cu[1:])
        x[nl:] = self._con_adj(self.lcon, V, nlf, nx)
        if self.lcost.is_linear:  # src/costs.cuh:289-292
            x[nl:] += st["lq"] * VI[0][:, None]
            s[nl:] = VI[1]
        else:
            cdl = VI[:, nx] + VI[:, nx + 1]
            x[nl:] += np.einsum("nij,nj->ni", st["lsQ"], VI[:, :nx]) - 0.5 * st["lq"] * cdl[:, None]
            s[nl:] = 0.5 * cdl
        return np.concatenate([u.ravel(), x.ravel(), y.ravel(), tt, s])

    def trajectory_cost(self, prim):
        """Risk-averse cost of the states and inputs of a primal iterate, in the scaled variables
        the solver iterates in.  A node's stage cost is the smallest epigraph value t that puts its
        cost block [v; h + c; h + d] = L z + translation (h = t/2 - (linear terms)/2) in the
        second-order cone, t = |v|^2 / (d - c) - (c + d) - 2 h(0): the cost the conic form
        encodes (``api/python/spock/build.py:207-216``, ``:522-535``, ``:592-604``); for a
        quadratic cost that is x'Qx + u'Ru, for the quadratic-plus-linear one it works out to
        x'Qx + u'Ru + 2 (q'x + r'u) + (q'Q^-1 q + r'R^-1 r) / 2.  A nonleaf node's cost-to-go is
        the AV@R of its children's stage cost plus cost-to-go.  Unlike the iterate's epigraph
        variable s_0 it depends on x and u alone; the two agree at a solution."""
        t, nx, nu = self.tree, self.nx, self.nu
        N, nl, nlf = t.num_nodes, t.num_nonleaf_nodes, t.num_leaf_nodes
        p = np.asarray(prim, dtype=np.float64).reshape(-1)[:self.size_prim].copy()
        p[self.size_prim - 2 * N:] = 0.0  # t = s = 0: the cost blocks hold h(0)
        w = self.op(p)
        c3 = 0 if self.ncon.kind == "no" else self.ncon.dim_per_node
        c5 = 0 if self.lcon.kind == "no" else self.lcon.dim_per_node
        o4 = self.yD * nl + nl + c3 * nl
        o6 = o4 + self.d4 * N + c5 * nlf

        def smallest_epigraph(block):
            v, a, b = block[:-2], block[-2], block[-1]
            return (v @ v) / (b - a) - (a + b)
        stage = np.zeros(N)
        for j in range(1, N):  # a linear cost's epigraph bound is the cost itself
            if self.ncost.is_linear:
                stage[j] = w[o4 + j]
                continue
            blk = w[o4 + j * (nx + nu + 2):o4 + (j + 1) * (nx + nu + 2)] + self.ncost.translation[j].ravel()
            stage[j] = smallest_epigraph(blk)
        togo = np.zeros(N)
        for k in range(nlf):
            if self.lcost.is_linear:
                togo[nl + k] = w[o6 + k]
                continue
            tr = self.lcost.translation[min(k, len(self.lcost.translation) - 1)].ravel()
            togo[nl + k] = smallest_epigraph(w[o6 + k * (nx + 2):o6 + (k + 1) * (nx + 2)] + tr)
        for i in range(nl - 1, -1, -1):  # children carry larger indices than their parent
            ch = np.asarray(t.children[i])
            togo[i] = self.risks[i].value(t.cond_prob_of_children_of_node(i), stage[ch] + togo[ch])
        return float(togo[0])

    @staticmethod
    def _con_op(con, z):
        if con.kind == "no":
            return np.zeros((z.shape[0], 0))
        if con.kind == "rectangle":
            return z
        if con.kind == "polyhedronWithIdentity":  # identity parts of all nodes, then G parts
            return np.concatenate([z.ravel(), (z @ con.poly.matrix.T).ravel()])
        return z @ con.matrix.T

    @staticmethod
    def _con_adj(con, dual, n, width):
        if con.kind == "no":
            return np.zeros((n, width))
        if con.kind == "rectangle":
            return dual.copy()
        if con.kind == "polyhedronWithIdentity":
            flat = np.asarray(dual).reshape(-1)
            r, g = con.rect.dim_per_node, con.poly.dim_per_node
            return flat[:n * r].reshape(n, r) + flat[n * r:n * (r + g)].reshape(n, g) @ con.poly.matrix
        return dual @ con.matrix

    def _step_size(self):
        npr = self.size_prim
        op = LinearOperator((npr, npr), matvec=lambda v: self.adj(self.op(v)), dtype=np.float64)
        v0 = np.ones(npr) / np.sqrt(npr)
        lam = eigsh(op, k=1, which="LM", v0=v0, return_eigenvectors=False, tol=1e-12)
        return 0.99 / np.sqrt(float(np.max(lam)))

    def _unconditioned_tensors(self):
        """The data before preconditioning under an ``uncond_`` prefix, for solvers that take the
        problem as posed (``api/python/spock/problem.py:605-684``): dynamics A, B, c of every
        node; nonleaf and leaf cost matrices, with their linear terms for the
        quadratic-plus-linear kind; constraint bounds and matrices."""
        nc, lc = self.ncost, self.lcost
        out = {
            "uncond_dynamics_A": _stack([d.A0 for d in self.dyn]),
            "uncond_dynamics_B": _stack([d.B0 for d in self.dyn]),
            "uncond_dynamics_c": _stack([d.c0 for d in self.dyn]),
        }
        if not nc.is_linear:
            out["uncond_cost_nonleaf_Q"] = _stack(nc.Q0)
            out["uncond_cost_nonleaf_R"] = _stack(nc.R0)
        if not lc.is_linear:
            out["uncond_cost_leaf_Q"] = _stack([lc.Q0[-1]])
        if not nc.is_quadratic:
            out["uncond_cost_nonleaf_q"] = _stack(nc.q0)
            out["uncond_cost_nonleaf_r"] = _stack(nc.r0)
        if not lc.is_quadratic:
            out["uncond_cost_leaf_q"] = _stack([lc.q0[-1]])
        for con, txt in ((self.ncon, "uncond_nonleafConstraint"), (self.lcon, "uncond_leafConstraint")):
            rect = con if con.kind == "rectangle" else getattr(con, "rect", None)
            poly = con if con.kind == "polyhedron" else getattr(con, "poly", None)
            if rect is not None:
                out[txt + "ILB"] = rect.lb0
                out[txt + "IUB"] = rect.ub0
            if poly is not None:
                out[txt + "Gamma"] = poly.G0
                out[txt + "GLB"] = poly.lower_bound
                out[txt + "GUB"] = poly.upper_bound
        return out

    # ------------------------------------------------------------------ files
    def write(self, folder):
        t = self.tree
        dt = t.dt
        os.makedirs(folder, exist_ok=True)
        t.write_files(folder)
        info = {
            "numEvents": t.num_events,
            "numNonleafNodes": t.num_nonleaf_nodes,
            "numNodes": t.num_nodes,
            "numStages": t.num_stages,
            "numStates": self.nx,
            "numInputs": self.nu,
            "dynamics": {"type": "affine" if self.dyn[1].is_affine else "linear"},
            "cost": {"nonleaf": self.ncost.kind, "leaf": self.lcost.kind},
            "constraint": {"nonleaf": self.ncon.kind, "leaf": self.lcon.kind},
            "risk": {"type": "avar", "alpha": self.risk.alpha},
            "rowsS2": t.num_events,
            "rowsNNtr": self.nD,
            "stepSize": float(self.step_size),
            "preconditioned": bool(self.preconditioned),
        }
        with open(os.path.join(folder, "data.json"), "w") as fh:
            json.dump(info, fh, indent=1)
        tens = {
            "dynamics_BTr": _stack([d.B.T for d in self.dyn]),
            "dynamics_AB": _stack([d.A_B for d in self.dyn]),
            "dynamics_P": _stack(self.P),
            "dynamics_K": _stack(self.K),
            "dynamics_ABKTr": _stack(self.ABKTr),
            "dynamics_APB": _stack(self.APB),
            "dynamics_lowCholesky": _stack(self.L),
            "risk_NNtr": _stack(self.NN),
            "risk_b": _stack(self.b_pad),
        }
        if self.dyn[1].is_affine:
            tens["dynamics_c"] = _stack([d.c for d in self.dyn])
        if self.preconditioned:
            tens["scaling"] = self.scaling.reshape(-1)
        nc = self.ncost
        if nc.is_linear:  # api/python/spock/problem.py:510-514
            tens["nonleafCost_gradient"] = _stack(nc.gradient)
        else:
            tens["nonleafCost_sqrtQ"] = _stack(nc.sqrtQ)
            tens["nonleafCost_sqrtR"] = _stack(nc.sqrtR)
            tens["nonleafCost_translation"] = _stack(nc.translation)
        if nc.is_quadratic_plus_linear:
            tens["nonleafCost_q"] = _stack(nc.q)
            tens["nonleafCost_r"] = _stack(nc.r)
        lc = self.lcost
        if lc.is_linear:  # api/python/spock/problem.py:541-545
            tens["leafCost_gradient"] = _stack([lc.gradient[-1]])
        else:
            tens["leafCost_sqrtQ"] = _stack([lc.sqrtQ[-1]])
            tens["leafCost_translation"] = _stack([lc.translation[-1]])
        if lc.is_quadratic_plus_linear:
            tens["leafCost_q"] = _stack([lc.q[-1]])
        for con, txt in ((self.ncon, "nonleafConstraint"), (self.lcon, "leafConstraint")):
            if con.kind == "rectangle":
                tens[txt + "ILB"] = con.lower_bound
                tens[txt + "IUB"] = con.upper_bound
            elif con.kind == "polyhedron":
                tens[txt + "Gamma"] = con.matrix
                tens[txt + "GLB"] = con.lower_bound
                tens[txt + "GUB"] = con.upper_bound
            elif con.kind == "polyhedronWithIdentity":  # api/python/spock/problem.py:578-585
                tens[txt + "ILB"] = con.rect.lower_bound
                tens[txt + "IUB"] = con.rect.upper_bound
                tens[txt + "Gamma"] = con.poly.matrix
                tens[txt + "GLB"] = con.poly.lower_bound
                tens[txt + "GUB"] = con.poly.upper_bound
        if self.julia:
            tens.update(self._unconditioned_tensors())
        for name, arr in tens.items():
            write_bt(os.path.join(folder, f"{name}_{dt}.bt"), arr, dt)
        return folder


class Factory:
    """Fluent builder mirroring ``api/python/spock/problem.py:702-879``."""

    def __init__(self, scenario_tree, num_states, num_inputs):
        self.tree, self.nx, self.nu = scenario_tree, num_states, num_inputs
        self.dyn = [None] * scenario_tree.num_nodes
        self.ncost_list = None
        self.lcost_list = None
        self.ncon, self.lcon = build.No(), build.No()
        self.risks = None
        self.precondition = True
        self.julia = False

    def with_dynamics(self, dynamics):
        self.dyn = [dynamics.zeros_like()] + [build.Dynamics(dynamics.A0, dynamics.B0, dynamics.c0 if dynamics.is_affine else None)
                                              for _ in range(1, self.tree.num_nodes)]
        return self

    def with_dynamics_events(self, dynamics):
        t = self.tree
        d = [dynamics[0].zeros_like()]
        for i in range(1, t.num_nodes):
            src = dynamics[int(t.event_of_node(i))]
            d.append(build.Dynamics(src.A0, src.B0, src.c0 if src.is_affine else None))
        self.dyn = d
        return self

    def _copy_cost(self, c):
        if c.kind == "quadratic":
            return build.CostQuadratic(c.Q, c.R, c.leaf)
        if c.kind == "linear":
            return build.CostLinear(c.q, c.r, c.leaf)
        return build.CostQuadraticPlusLinear(c.Q, c.q, c.R, c.r, c.leaf)

    def with_cost_nonleaf(self, cost):
        lst = [self._copy_cost(cost).node_zero()] + [self._copy_cost(cost) for _ in range(1, self.tree.num_nodes)]
        self.ncost_list = lst
        return self

    def with_cost_nonleaf_events(self, costs):
        t = self.tree
        lst = [self._copy_cost(costs[0]).node_zero()]
        lst += [self._copy_cost(costs[int(t.event_of_node(i))]) for i in range(1, t.num_nodes)]
        self.ncost_list = lst
        return self

    def with_cost_leaf(self, cost):
        if not cost.leaf:
            raise ValueError("[ProblemFactory] Cannot use nonleaf cost for leaf nodes!")
        self.lcost_list = [cost for _ in range(self.tree.num_leaf_nodes)]
        return self

    def with_constraint_nonleaf(self, con):
        self.ncon = con
        return self

    def with_constraint_leaf(self, con):
        self.lcon = con
        return self

    def with_risk(self, risk):
        self.risks = [risk for _ in range(self.tree.num_nonleaf_nodes)]
        return self

    def with_preconditioning(self, enable=True):
        self.precondition = enable
        return self

    def with_julia(self):
        """Also write the unconditioned data (``api/python/spock/problem.py:854-856``)."""
        self.julia = True
        return self

    def generate_problem(self, folder=None):
        p = Problem(self.tree, self.nx, self.nu, self.dyn,
                    build.CostSet(self.ncost_list, leaf=False), build.CostSet(self.lcost_list, leaf=True),
                    self.ncon, self.lcon, self.risks, self.precondition, self.julia)
        if folder is not None:
            p.write(folder)
        return p
