"""Scenario trees: Markov-chain and i.i.d. factories and the tree file writer.

Node numbering follows the reference factories so that every per-node tensor
lines up with the reference's file set:

* ``api/python/spock/tree.py:396-437`` (Markov) and ``:522-563`` (i.i.d.):
  stage 1 holds one node per nonzero entry of the initial distribution; up to
  the stopping stage each node of the previous stage gets one child per nonzero
  entry of its transition row ("cover"), appended in node order; after the
  stopping stage every node gets exactly one child carrying the same event.
* ``:439-462`` / ``:565-588``: probabilities multiply along the branching part
  and are copied from the parent after the stopping stage.
* ``:40-67``: children, conditional probabilities, per-stage node ranges.
* ``:240-265``: the tree file set.

Nodes are numbered stage by stage and children are contiguous, which the solver
relies on (``src/dynamics.cuh:87-92``).
"""
import os
import numpy as np
from .btio import write_bt


class Tree:
    """A scenario tree held as flat per-node arrays (``api/python/spock/tree.py:7-67``)."""

    def __init__(self, stages, ancestors, probability, events=None, dt="d"):
        self.dt = dt
        self.stages_arr = np.asarray(stages, dtype=np.int64)
        self.ancestors_arr = np.asarray(ancestors, dtype=np.int64)
        self.prob = np.asarray(probability, dtype=np.float64)
        n = len(self.ancestors_arr)
        if events is None:
            events = np.zeros(n, dtype=np.int64)
            events[0] = -1
        self.events_arr = np.asarray(events, dtype=np.int64)
        self.num_stages = int(self.stages_arr[-1]) + 1
        self.num_nodes = n
        self.num_nonleaf_nodes = int(np.sum(self.stages_arr < self.num_stages - 1))
        self.children = [np.flatnonzero(self.ancestors_arr == i) for i in range(self.num_nonleaf_nodes)]
        self.num_children = np.array([len(c) for c in self.children], dtype=np.int64)
        cond = np.empty(n)
        cond[0] = -1.0
        for i in range(1, n):
            cond[i] = self.prob[i] / self.prob[self.ancestors_arr[i]]
        self.cond_prob = cond
        self.nodes_of_stage_list = [np.flatnonzero(self.stages_arr == s) for s in range(self.num_stages)]

    # reference-compatible accessors (api/python/spock/tree.py:69-220)
    @property
    def num_events(self):
        return int(self.num_children.max())

    @property
    def max_num_children(self):
        return int(self.num_children.max())

    @property
    def num_leaf_nodes(self):
        return self.num_nodes - self.num_nonleaf_nodes

    def ancestor_of_node(self, i):
        return int(self.ancestors_arr[i])

    def children_of_node(self, i):
        return self.children[i]

    def stage_of_node(self, i):
        if i < 0:
            raise ValueError("node_idx cannot be <0")
        return int(self.stages_arr[i])

    def event_of_node(self, i):
        return self.events_arr[i]

    def nodes_of_stage(self, s):
        return self.nodes_of_stage_list[s]

    def probability_of_node(self, i):
        return self.prob[i]

    def siblings_of_node(self, i):
        if i == 0:
            return [0]
        return self.children_of_node(self.ancestor_of_node(i))

    def cond_prob_of_children_of_node(self, i):
        return self.cond_prob[self.children[i]]

    def stages(self):
        return self.stages_arr

    def ancestors(self):
        return self.ancestors_arr

    def probabilities(self):
        return self.prob

    def stage_ranges(self):
        return ([int(a[0]) for a in self.nodes_of_stage_list], [int(a[-1]) for a in self.nodes_of_stage_list])

    def write_files(self, folder):
        """Write the tree file set (``api/python/spock/tree.py:240-265``)."""
        os.makedirs(folder, exist_ok=True)
        sf, st = self.stage_ranges()
        ints = {
            "stages": self.stages_arr,
            "ancestors": self.ancestors_arr,
            "events": self.events_arr,
            "childrenFrom": [int(c[0]) for c in self.children],
            "childrenTo": [int(c[-1]) for c in self.children],
            "numChildren": self.num_children,
            "stageFrom": sf,
            "stageTo": st,
        }
        for name, v in ints.items():
            write_bt(os.path.join(folder, name + "_u.bt"), np.asarray(v, dtype=np.int64), "u")
        for name, v in {"probabilities": self.prob, "conditionalProbabilities": self.cond_prob}.items():
            write_bt(os.path.join(folder, f"{name}_{self.dt}.bt"), v, self.dt)

    def write_to_file_fp(self, folder, name, tensor):
        """A floating-point tensor in the tree's precision, e.g. the drivers' ``initialState``
        (``api/python/spock/tree.py:237-238``, ``examples/random/main.py:128``)."""
        os.makedirs(folder, exist_ok=True)
        write_bt(os.path.join(folder, f"{name}_{self.dt}.bt"), np.asarray(tensor, dtype=np.float64), self.dt)

    def __repr__(self):
        return (f"Scenario tree with {self.num_nodes} nodes, {self.num_stages} stages "
                f"and {len(self.nodes_of_stage(self.num_stages - 1))} scenarios")


def _check_prob(p, allow_zero=True):
    p = np.asarray(p, dtype=np.float64)
    if abs(p.sum() - 1.0) >= 1e-10:
        raise ValueError("probability vector does not sum up to 1")
    if np.any(p <= -1e-16):
        raise ValueError("probability vector contains negative entries")
    if not allow_zero and np.any(p <= 1e-16):
        raise ValueError("probability vector contains zero entries")
    return p


def _grow(first_values, cover_of, horizon, stopping_stage):
    """Shared node-numbering rule of both factories."""
    values = [-1] + list(first_values)
    ancestors = [-1] + [0] * len(first_values)
    stages = [0] + [1] * len(first_values)
    cursor, width = 1, len(first_values)
    for stage in range(1, stopping_stage):
        added = 0
        for i in range(width):
            node = cursor + i
            cov = cover_of(values[node])
            ancestors += [node] * len(cov)
            values += list(cov)
            added += len(cov)
        cursor += width
        width = added
        stages += [stage + 1] * added
    for stage in range(stopping_stage, horizon):
        ancestors += list(range(cursor, cursor + width))
        values += values[-width:]
        stages += [stage + 1] * width
        cursor += width
    return np.array(ancestors), np.array(values), np.array(stages)


def _probabilities(ancestors, values, stages, first_probs, branch_prob, stopping_stage):
    probs = [1.0] + list(first_probs)
    for i in range(len(first_probs) + 1, len(values)):
        a = ancestors[i]
        if stages[i] <= stopping_stage:
            probs.append(probs[a] * branch_prob(values[a], values[i]))
        else:
            probs.append(probs[a])
    return np.array(probs)


class MarkovChain:
    """Stopped Markov chain factory (``api/python/spock/tree.py:350-479``)."""

    def __init__(self, transition_prob, initial_distribution, horizon, stopping_stage=None, dt="d"):
        if stopping_stage is None:
            stopping_stage = horizon
        elif stopping_stage > horizon:
            raise ValueError("stopping time greater than number of stages")
        self.P = np.asarray(transition_prob, dtype=np.float64)
        for row in self.P:
            _check_prob(row)
        self.v = _check_prob(initial_distribution)
        self.horizon, self.stop, self.dt = horizon, stopping_stage, dt

    def build(self, folder=None):
        first = np.flatnonzero(self.v)
        anc, val, stg = _grow(first, lambda w: np.flatnonzero(self.P[w, :]), self.horizon, self.stop)
        prob = _probabilities(anc, val, stg, self.v[first], lambda a, b: self.P[a, b], self.stop)
        tree = Tree(stg, anc, prob, events=val, dt=self.dt)
        if folder is not None:
            tree.write_files(folder)
        return tree


class IidProcess:
    """i.i.d. process factory (``api/python/spock/tree.py:482-605``)."""

    def __init__(self, distribution, horizon, stopping_stage=None, dt="d"):
        if stopping_stage is None:
            stopping_stage = horizon
        elif stopping_stage > horizon:
            raise ValueError("stopping time greater than number of stages")
        self.v = _check_prob(np.asarray(distribution).reshape(-1), allow_zero=False)
        self.horizon, self.stop, self.dt = horizon, stopping_stage, dt

    def build(self, folder=None):
        cover = np.flatnonzero(self.v)
        anc, val, stg = _grow(cover, lambda w: cover, self.horizon, self.stop)
        prob = _probabilities(anc, val, stg, self.v[cover], lambda a, b: self.v[b], self.stop)
        tree = Tree(stg, anc, prob, events=val, dt=self.dt)
        if folder is not None:
            tree.write_files(folder)
        return tree


class FromStructure:
    """Tree from an explicit structure (``api/python/spock/tree.py:902-951``): the stage, ancestor
    (-1 at the root) and unconditional probability of every node; events are all 0 below the root.
    ``data`` (node x dimension) is accepted for interface parity and not used by the solver."""

    def __init__(self, stages, ancestors, probabilities, data=None, dt="d"):
        self.stages = np.asarray(stages)
        self.ancestors = np.asarray(ancestors)
        self.probs = np.asarray(probabilities, dtype=np.float64)
        n = self.stages.size
        if data is not None and not isinstance(data, np.ndarray):
            raise ValueError("[FromStructure] Data must be a numpy array!")
        self.data = np.atleast_2d(data) if data is not None else np.zeros((n, 1))
        self.dt = dt
        if n != self.ancestors.size or n != self.probs.size:
            raise ValueError("[FromStructure] Arguments have differing sizes!")
        if self.stages[0] != 0:
            raise ValueError("[FromStructure] No root node!")
        if self.ancestors[0] != -1:
            raise ValueError("[FromStructure] Ancestor of root node must be (-1)!")
        for stage in range(int(self.stages.max())):
            if not np.isclose(self.probs[self.stages == stage].sum(), 1.0):
                raise ValueError(f"[FromStructure] Probabilities of nodes at stage ({stage}) do not sum to (1.0)!")
        if self.data.shape[0] != n:
            self.data = self.data.T
        if self.data.shape[0] != n:
            raise ValueError("[FromStructure] Data does not fit tree!")

    def build(self, folder=None):
        tree = Tree(self.stages, self.ancestors, self.probs, dt=self.dt)
        if folder is not None:
            tree.write_files(folder)
        return tree


class FromData:
    """Tree from sampled data by forward scenario reduction (``api/python/spock/tree.py:608-899``).

    ``samples_dim_time`` is [samples x dimension x time]; ``branching_per_stage[k]`` caps the
    children of every node at stage k.  Each node's cluster of samples is reduced at the next
    stage to at most that many kept samples by greedy forward selection under the
    probability-weighted l2 transport cost; every sample then joins its closest kept one, which
    takes over the probability.  Samples start equally likely.  The trial set of a candidate
    follows the reference literally: it is every sample except those still eliminated, where the
    candidate's position in the eliminated list stands in as its label, so trees built from the
    same data match the reference's.  Nodes are then numbered stage by stage along the scenarios
    in lexicographic order.  Events are all 0 below the root (no per-event data)."""

    def __init__(self, samples_dim_time, branching_per_stage, dt="d", sw_scaling=False):
        data = np.asarray(samples_dim_time, dtype=np.float64)
        if data.ndim != 3:
            raise ValueError("[FromData] Data provided is not a 3D array.")
        self.data = data
        self.branching = np.asarray(branching_per_stage, dtype=np.int64)
        self.dt = dt
        self.sw_scaling = sw_scaling

    @staticmethod
    def _reduce(values, keep, probs):
        """Greedy forward selection of `keep` of the rows of `values`; returns (probabilities of
        all rows after redistribution, kept row indices, closest kept row of every row)."""
        n = values.shape[0]
        if keep > n:
            raise ValueError("[FromData] Number of desired scenarios too large.")
        dist = np.linalg.norm(values[:, None, :] - values[None, :, :], axis=2)
        everyone = set(range(n))
        remaining = list(range(n))
        kept = []
        for _ in range(keep):
            best, best_pos = np.inf, 0
            for pos in range(len(remaining)):
                trial = sorted(everyone - (set(remaining) - {pos}))
                cost = 0.0
                for jpos, j in enumerate(remaining):
                    if jpos != pos:
                        cost += probs[j] * dist[j, trial].min()
                if cost < best:
                    best, best_pos = cost, pos
            kept.append(remaining.pop(best_pos))
        kept = np.asarray(kept, dtype=np.int64)
        closest = kept[np.argmin(dist[:, kept], axis=1)]
        new_probs = np.zeros(n)
        for k in kept:
            new_probs[k] = probs[closest == k].sum()
        return new_probs, kept, closest

    def build(self, folder=None):
        data = self.data
        n_samples, n_vars, n_time = data.shape
        horizon = min(self.branching.size, n_time)
        scaled = data
        if self.sw_scaling:
            scaled = np.empty_like(data)
            for v in range(n_vars):
                x = data[:, v, :]
                scaled[:, v, :] = (x - x.mean()) / x.std()
        stage, anc, prob = [0], [-1], [1.0]
        cluster = [np.arange(n_samples)]
        leaves = [0]
        while any(stage[lf] < horizon for lf in leaves):
            for lf in list(leaves):
                k = stage[lf]
                if k >= horizon:
                    continue
                leaves.remove(lf)
                members = cluster[lf]
                p = np.full(members.size, 1.0 / members.size)
                if members.size > self.branching[k]:
                    rp, kept, closest = self._reduce(scaled[members, :, k], int(self.branching[k]), p)
                else:
                    rp, kept, closest = p, np.arange(members.size), np.arange(members.size)
                for kk in kept:
                    node = len(stage)
                    stage.append(k + 1)
                    anc.append(lf)
                    prob.append(prob[lf] * rp[kk])
                    cluster.append(members[closest == kk])
                    leaves.append(node)
        stage, anc, prob = np.asarray(stage), np.asarray(anc), np.asarray(prob)
        # number the nodes stage by stage along the scenarios sorted lexicographically by path
        paths = []
        for lf in leaves:
            path = [lf]
            while anc[path[-1]] >= 0:
                path.append(anc[path[-1]])
            paths.append(path[::-1])
        paths.sort()
        order, seen = [], set()
        for col in range(int(stage.max()) + 1):
            for path in paths:
                if col < len(path) and path[col] not in seen:
                    seen.add(path[col])
                    order.append(path[col])
        relabel = {old: new for new, old in enumerate(order)}
        order = np.asarray(order)
        new_anc = np.array([-1] + [relabel[a] for a in anc[order][1:]])
        tree = Tree(stage[order], new_anc, prob[order], dt=self.dt)
        if folder is not None:
            tree.write_files(folder)
        return tree
