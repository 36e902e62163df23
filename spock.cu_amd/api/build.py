"""Problem components: dynamics, costs, constraints and the AVaR risk.

Semantics follow ``api/python/spock/build.py``:

* ``Dynamics`` (``:13-75``): x+ = A x + B u (+ c); preconditioning maps
  A -> Sx A Sx^-1, B -> Sx B Su^-1, c -> Sx c.
* ``CostQuadratic`` / ``CostQuadraticPlusLinear`` (``:184-701``): per-node
  sqrt(Q), sqrt(R), linear terms and the SOC translation; node 0 of the nonleaf
  cost is the zeroed copy (``:338-341``, ``problem.py:783-791``).
* ``Rectangle`` / ``Polyhedron`` (``:811-1014``) and ``PolyhedronWithIdentity``
  (``:1025-1103``) with their conditioning rules.
* ``AVaR`` (``:1160-1202``): E = [alpha I; -I; 1^T], b = [pi; 0; 1].
"""
import numpy as np
from scipy.linalg import sqrtm


def _spd(mat, name):
    mat = np.asarray(mat, dtype=np.float64)
    if not (np.all(np.linalg.eigvals(mat).real > 0) and np.allclose(mat, mat.T)):
        raise ValueError(f"Invalid cost. The matrix ({name}) is not symmetric positive-definite.")
    return mat


def _sqrtm(m):
    m = np.asarray(m, dtype=np.float64)
    if not np.any(m):
        return np.zeros_like(m)
    return np.real(sqrtm(m))


class Dynamics:
    def __init__(self, A, B, c=None):
        self.A0 = np.asarray(A, dtype=np.float64)
        self.B0 = np.asarray(B, dtype=np.float64)
        self.is_affine = c is not None
        self.c0 = np.asarray(c, dtype=np.float64).reshape(-1, 1) if self.is_affine else np.zeros((self.A0.shape[0], 1))
        self.A, self.B, self.c = self.A0.copy(), self.B0.copy(), self.c0.copy()

    @property
    def is_linear(self):
        return not self.is_affine

    @property
    def A_B(self):
        return np.hstack((self.A, self.B))

    def condition(self, sx_inv, su_inv, sx):
        self.A = sx @ self.A0 @ sx_inv
        self.B = sx @ self.B0 @ su_inv
        self.c = sx @ self.c0
        return self

    def zeros_like(self):
        return Dynamics(np.zeros_like(self.A0), np.zeros_like(self.B0), np.zeros(self.A0.shape[0]) if self.is_affine else None)


class CostQuadratic:
    """Quadratic stage cost x'Qx + u'Ru (leaf: x'Qx)."""
    kind = "quadratic"

    def __init__(self, Q, R=None, leaf=False):
        self.Q = _spd(Q, "Q")
        self.R = None if leaf else _spd(R, "R")
        self.q = np.zeros((self.Q.shape[0], 1))
        self.r = None if leaf else np.zeros((self.R.shape[0], 1))
        self.leaf = leaf

    def node_zero(self):
        self.Q = np.zeros_like(self.Q)
        if self.R is not None:
            self.R = np.zeros_like(self.R)
        self.q = np.zeros_like(self.q)
        if self.r is not None:
            self.r = np.zeros_like(self.r)
        return self


class CostQuadraticPlusLinear(CostQuadratic):
    """Quadratic plus linear stage cost x'Qx + u'Ru + q'x + r'u."""
    kind = "quadraticPlusLinear"

    def __init__(self, Q, q, R=None, r=None, leaf=False):
        super().__init__(Q, R, leaf)
        self.q = np.asarray(q, dtype=np.float64).reshape(-1, 1)
        self.r = None if leaf else np.asarray(r, dtype=np.float64).reshape(-1, 1)


class CostLinear:
    """Linear stage cost q'x + r'u (leaf: q'x) (``api/python/spock/build.py:351-490``).  Its cone
    block per node is the pair (q'x + r'u, t) of the node's cost and epigraph variable."""
    kind = "linear"

    def __init__(self, q, r=None, leaf=False):
        self.q = np.asarray(q, dtype=np.float64).reshape(-1, 1)
        self.r = None if leaf else np.asarray(r, dtype=np.float64).reshape(-1, 1)
        self.Q = self.R = None
        self.leaf = leaf

    def node_zero(self):
        self.q = np.zeros_like(self.q)
        if self.r is not None:
            self.r = np.zeros_like(self.r)
        return self


class CostSet:
    """A cost replicated over the nodes it applies to, with conditioning and translations
    (``api/python/spock/build.py:184-330`` and ``:470-701``)."""

    def __init__(self, costs, leaf):
        self.leaf = leaf
        self.kind = costs[-1].kind
        self.n = len(costs)
        self.start = 0 if leaf else 1
        lin = self.kind == "linear"
        self.Q0 = None if lin else [c.Q.copy() for c in costs]
        self.q0 = [c.q.copy() for c in costs]
        self.R0 = None if (leaf or lin) else [c.R.copy() for c in costs]
        self.r0 = None if leaf else [c.r.copy() for c in costs]
        self.Q = None if lin else [m.copy() for m in self.Q0]
        self.q = [m.copy() for m in self.q0]
        self.R = None if (leaf or lin) else [m.copy() for m in self.R0]
        self.r = None if leaf else [m.copy() for m in self.r0]
        self._refresh()

    @property
    def is_quadratic(self):
        return self.kind == "quadratic"

    @property
    def is_quadratic_plus_linear(self):
        return self.kind == "quadraticPlusLinear"

    @property
    def is_linear(self):
        return self.kind == "linear"

    def block(self, nx, nu):
        """Entries of the cost's dual block per node: the cone [v; a; b] of dimension nx + nu + 2
        (leaf nx + 2), or the pair (cost, epigraph variable) of a linear cost."""
        return 2 if self.is_linear else nx + (0 if self.leaf else nu) + 2

    def _refresh(self):
        if self.is_linear:  # gradient rows [q' r'] (leaf q'), src/costs.cuh:228-250
            self.sqrtQ = self.sqrtR = None
            self.dim_per_node = 2
            self.translation = [np.zeros((2, 1)) for _ in range(self.n)]
            self.gradient = [np.vstack([self.q[i]] + ([] if self.leaf else [self.r[i]])).T for i in range(self.n)]
            return
        self.sqrtQ = [_sqrtm(m) for m in self.Q]
        self.sqrtR = None if self.leaf else [_sqrtm(m) for m in self.R]
        nx = self.Q[0].shape[0]
        nu = 0 if self.leaf else self.R[0].shape[0]
        self.dim_per_node = nx + nu + 2
        self.translation = [np.zeros((self.dim_per_node, 1)) for _ in range(self.n)]
        for i in range(self.start, self.n):
            if self.is_quadratic:
                a = np.zeros((nx, 1))
                c, d = -0.5, 0.5
                b = np.zeros((nu, 1))
            else:
                a = 0.5 * np.linalg.solve(self.sqrtQ[i], self.q[i]).reshape(-1, 1)
                nq = (-0.125 * self.q[i].T @ np.linalg.solve(self.Q[i], self.q[i])).item()
                c, d = -0.5 + nq, 0.5 + nq
                if not self.leaf:
                    b = 0.5 * np.linalg.solve(self.sqrtR[i], self.r[i]).reshape(-1, 1)
                    nr = (-0.125 * self.r[i].T @ np.linalg.solve(self.R[i], self.r[i])).item()
                    c, d = c + nr, d + nr
            parts = [a] if self.leaf else [a, b]
            self.translation[i] = np.vstack(parts + [np.array([[c]]), np.array([[d]])])

    def scaling_states(self, sx):
        for e in range(sx.size):
            for i in range(self.start, self.n):
                if not self.is_linear:
                    sx[e] = max(sx[e], np.sqrt(self.Q0[i][e, e]))
                if not self.is_quadratic:  # api/python/spock/build.py:417-423
                    sx[e] = max(sx[e], abs(self.q0[i][e, 0]))
        return sx

    def scaling_inputs(self, su):
        for e in range(su.size):
            for i in range(self.start, self.n):
                if not self.is_linear:
                    su[e] = max(su[e], np.sqrt(self.R0[i][e, e]))
                if not self.is_quadratic:
                    su[e] = max(su[e], abs(self.r0[i][e, 0]))
        return su

    def condition(self, sx_inv, su_inv=None):
        for i in range(self.start, self.n):
            if not self.is_linear:
                self.Q[i] = sx_inv.T @ self.Q0[i] @ sx_inv
            self.q[i] = np.diagonal(sx_inv).reshape(-1, 1) * self.q0[i]
            if not self.leaf:
                if not self.is_linear:
                    self.R[i] = su_inv.T @ self.R0[i] @ su_inv
                self.r[i] = np.diagonal(su_inv).reshape(-1, 1) * self.r0[i]
        self._refresh()
        return self


class No:
    kind = "no"
    dim_per_node = 0

    def condition(self, s):
        return self


class Rectangle:
    kind = "rectangle"

    def __init__(self, lower_bound, upper_bound):
        lb = np.asarray(lower_bound, dtype=np.float64).reshape(-1, 1)
        ub = np.asarray(upper_bound, dtype=np.float64).reshape(-1, 1)
        if lb.size != ub.size:
            raise ValueError("Rectangle constraint - min and max bound dimensions are not equal")
        if np.any(lb > ub):
            raise ValueError("Rectangle constraint - min greater than max")
        self.lb0, self.ub0 = lb, ub
        self.lower_bound, self.upper_bound = lb.copy(), ub.copy()
        self.dim_per_node = lb.size

    def condition(self, scale):
        self.lower_bound = scale @ self.lb0
        self.upper_bound = scale @ self.ub0
        return self


class Polyhedron:
    kind = "polyhedron"

    def __init__(self, matrix, lower_bound, upper_bound):
        G = np.asarray(matrix, dtype=np.float64)
        lb = np.asarray(lower_bound, dtype=np.float64).reshape(-1, 1)
        ub = np.asarray(upper_bound, dtype=np.float64).reshape(-1, 1)
        if lb.size != ub.size or G.shape[0] != lb.size:
            raise ValueError("Polyhedron constraint - matrix and bound dimensions are not equal")
        if np.any(lb > ub):
            raise ValueError("Polyhedron constraint - min greater than max")
        self.G0, self.matrix = G, G.copy()
        self.lower_bound, self.upper_bound = lb, ub
        self.dim_per_node = lb.size

    def condition(self, scale):
        self.matrix = self.G0 @ np.diag(1.0 / np.diagonal(scale))
        return self


class PolyhedronWithIdentity:
    """lb <= [I; G] z <= ub: a rectangle and a polyhedron on the same vector
    (``api/python/spock/build.py:1025-1103``).  Its dual block holds the identity parts of all
    nodes first, then the G parts of all nodes (``src/constraints.cuh:340-373``); the two parts
    are conditioned as a rectangle and a polyhedron."""
    kind = "polyhedronWithIdentity"

    def __init__(self, rect, poly):
        if not isinstance(rect, Rectangle):
            raise ValueError("PolyhedronWithIdentity constraint - not rectangle")
        if not isinstance(poly, Polyhedron):
            raise ValueError("PolyhedronWithIdentity constraint - not polyhedron")
        if poly.matrix.shape[1] != rect.dim_per_node:
            raise ValueError("PolyhedronWithIdentity constraint - matrix and rectangle dimensions are not equal")
        self.rect, self.poly = rect, poly
        self.dim_per_node = rect.dim_per_node + poly.dim_per_node

    def condition(self, scale):
        self.rect.condition(scale)
        self.poly.condition(scale)
        return self


class AVaR:
    """Average value-at-risk; its (E, b) data (``api/python/spock/build.py:1160-1202``)."""
    kind = "avar"

    def __init__(self, alpha):
        if not 0 <= alpha <= 1:
            raise ValueError("AVaR alpha must be between 0 and 1")
        self.alpha = float(alpha)

    def make(self, cond_probs):
        n = len(cond_probs)
        E = np.vstack((self.alpha * np.eye(n), -np.eye(n), np.ones((1, n))))
        b = np.vstack((np.asarray(cond_probs).reshape(-1, 1), np.zeros((n, 1)), np.ones((1, 1))))
        return E, b

    def value(self, cond_probs, Z):
        """AV@R of the outcomes Z: max mu'Z over the set (E, b) describes, 0 <= alpha mu <= pi with
        1'mu = 1 (the last row's multiplier is free).  Greedy: the largest outcomes take their
        full weight pi/alpha until the weights sum to one."""
        pi, Z = np.asarray(cond_probs, dtype=np.float64).ravel(), np.asarray(Z, dtype=np.float64).ravel()
        if self.alpha == 0.0:
            return float(Z.max())
        left, total = 1.0, 0.0
        for k in np.argsort(-Z, kind="stable"):
            w = min(pi[k] / self.alpha, left)
            total += w * Z[k]
            left -= w
            if left <= 0.0:
                break
        return float(total)
