"""Independent numpy checks that pin the C++ oracle (test infrastructure only).

Each helper computes a quantity from mathematics rather than from the oracle's
code path:

* ``dense_operator``: the matrix of L assembled column by column by applying the
  oracle to unit vectors, for the adjoint identity and ||L||.
* ``kkt_dynamics_projection``: the projection onto the dynamics as the
  equality-constrained least-squares problem min ||x-xbar||^2 + ||u-ubar||^2,
  x_0 = x0, x_j = A_j x_anc + B_j u_anc (the QP the reference's own golden data
  solves with cvxpy, ``api/python/spock/problem.py:233-261``), solved through its
  KKT system with ``numpy.linalg.lstsq``.
* ``kernel_projector``: I - S2^+ S2, the orthogonal projector onto ker S2.
* ``soc_project``: the textbook second-order-cone projection (Parikh & Boyd,
  Proximal Algorithms, 6.3.2, cited at ``src/cones.cuh:151-152``).
"""
import numpy as np


def dense_operator(apply, n_in):
    cols = []
    for k in range(n_in):
        e = np.zeros(n_in)
        e[k] = 1.0
        cols.append(apply(e))
    return np.array(cols).T


def kkt_dynamics_projection(AB_nodes, anc, nl, nx, nu, x0, xbar, ubar, e_nodes=None):
    """AB_nodes: (N, nx, nx+nu) per-node [A B] (node 0 unused); xbar (N, nx), ubar (nl, nu);
    e_nodes: (N, nx) constants of affine dynamics x_j = A_j x_anc + B_j u_anc + e_j (or None)."""
    N = len(anc)
    nvar = N * nx + nl * nu

    def xi(n):
        return slice(n * nx, (n + 1) * nx)

    def ui(n):
        return slice(N * nx + n * nu, N * nx + (n + 1) * nu)

    rows = []
    rhs = []
    c = np.zeros((nx, nvar))
    c[:, xi(0)] = np.eye(nx)
    rows.append(c)
    rhs.append(x0)
    for j in range(1, N):
        a = anc[j]
        c = np.zeros((nx, nvar))
        c[:, xi(j)] = np.eye(nx)
        c[:, xi(a)] -= AB_nodes[j][:, :nx]
        c[:, ui(a)] -= AB_nodes[j][:, nx:]
        rows.append(c)
        rhs.append(np.zeros(nx) if e_nodes is None else np.asarray(e_nodes[j], dtype=np.float64))
    Aeq = np.vstack(rows)
    beq = np.concatenate(rhs)
    target = np.concatenate([xbar.reshape(-1), ubar.reshape(-1)])
    m = Aeq.shape[0]
    K = np.block([[np.eye(nvar), Aeq.T], [Aeq, np.zeros((m, m))]])
    sol = np.linalg.lstsq(K, np.concatenate([target, beq]), rcond=None)[0]
    z = sol[:nvar]
    return z[:N * nx].reshape(N, nx), z[N * nx:].reshape(nl, nu)


def kernel_projector(S2):
    return np.eye(S2.shape[1]) - np.linalg.pinv(S2) @ S2


def soc_project(v):
    v = np.asarray(v, dtype=np.float64)
    x, t = v[:-1], v[-1]
    nx = np.linalg.norm(x)
    if nx <= t:
        return v.copy()
    if nx <= -t:
        return np.zeros_like(v)
    a = 0.5 * (nx + t)
    return np.concatenate([a * x / nx, [a]])
