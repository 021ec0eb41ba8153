"""Elimination-order optimizers (host side).

* ``greedy_peo``         -- min-degree greedy ordering: repeatedly eliminate the
  node with the fewest distinct neighbours (first such node in graph order)
  and join its neighbours into a clique (``qtensor/utils.py:12-44,77-86``).
  The treewidth reported is the largest neighbour count at elimination.
* ``OrderingOptimizer``  -- ``qtensor/optimisation/Optimizer.py:46-88``:
  order the line graph without bra/ket variables, then prepend ket+bra.
* ``SlicesOptimizer``    -- ``Optimizer.py:90-159`` with ``GreedyParvars``
  (``qtensor/optimisation/Greedy.py:78-92``): remove the highest-degree
  variable until the greedy treewidth drops below ``max_tw - tw_bias``; the
  removed variables are sliced and go last in the order.
* ``TreeTrimSplitter``   -- ``Optimizer.py:178-220``: slice the nodes that weigh most
  in the large elimination neighbourhoods, a fifth of the width excess per round.
* ``TamakiOptimizer``    -- ``Optimizer.py:161-176``: PACE-2017 heuristic via an
  external ``tw-heuristic`` executable (Java), used when it is on ``PATH``
  (greedy with a warning otherwise); ``TamakiTrimSlicing`` (``Optimizer.py:227``)
  adds tree-trimming slicing.
* ``RGreedyOptimizer``   -- ``qtensor/optimisation/RGreedy.py:15-77``: pick the next
  node with probability proportional to exp(-(degree - min degree)/temp), keep the
  narrowest of ``repeats`` samples (seeded here).
* ``RGreedyOrdering``    -- the same rule in the native planner (``qb_rgreedy_order``),
  best of the samples and the greedy order, as the order under a slicer
  (``RGreedySlicesOptimizer``, ``RGreedyTreeTrimSplitter``) or alone.
"""
import ctypes
import os
import shutil
import subprocess
import tempfile

import numpy as np

from . import _lib
from .model import Var


def _adjacency(graph):
    return {n: set(graph[n]) - {n} for n in graph.nodes()}


def _csr(graph):
    """Nodes in graph order and the CSR adjacency (int32) of ``graph``."""
    nodes = list(graph.nodes())
    index = {n: i for i, n in enumerate(nodes)}
    ptr = np.zeros(len(nodes) + 1, dtype=np.int32)
    idx = []
    for i, n in enumerate(nodes):
        nb = [index[u] for u in graph[n] if u != n]
        idx += nb
        ptr[i + 1] = ptr[i] + len(nb)
    return nodes, ptr, np.asarray(idx, dtype=np.int32)


def _i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def greedy_peo(graph):
    """Return ``(peo, degrees)`` of the min-degree greedy elimination
    (``qtensor/utils.py:12-44``): the node with the fewest distinct
    neighbours, first in graph order, is eliminated and its neighbours are
    joined into a clique.  Runs in the native planner (``qb_greedy_order``)."""
    nodes, ptr, idx = _csr(graph)
    n = len(nodes)
    order = np.zeros(max(n, 1), dtype=np.int32)
    widths = np.zeros(max(n, 1), dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.qb_greedy_order(n, _i32p(ptr), _i32p(idx if idx.size else np.zeros(1, np.int32)),
                                   _i32p(order), _i32p(widths)))
    return [nodes[k] for k in order[:n]], [int(w) for w in widths[:n]]


def rgreedy_best(graph, temp=0.002, repeats=16, seed=0):
    """``(peo, degrees)`` of the best of ``repeats`` randomized-greedy samples and
    the greedy order (``qb_rgreedy_order``): the rule of
    ``qtensor/optimisation/RGreedy.py:36-77``, kept by (largest width, sum of
    2^width).  Deterministic for a given ``seed``."""
    nodes, ptr, idx = _csr(graph)
    n = len(nodes)
    order = np.zeros(max(n, 1), dtype=np.int32)
    widths = np.zeros(max(n, 1), dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.qb_rgreedy_order(n, _i32p(ptr), _i32p(idx if idx.size else np.zeros(1, np.int32)),
                                    float(temp), int(seed), int(repeats), _i32p(order), _i32p(widths)))
    return [nodes[k] for k in order[:n]], [int(w) for w in widths[:n]]


def elimination_widths(graph, peo):
    """Neighbour count of each node when eliminated in ``peo`` order
    (``qb_elimination_widths``)."""
    nodes, ptr, idx = _csr(graph)
    index = {v: i for i, v in enumerate(nodes)}
    order = np.asarray([index[v] for v in peo] or [0], dtype=np.int32)
    widths = np.zeros(max(len(peo), 1), dtype=np.int32)
    lib = _lib.load()
    _lib.check(lib.qb_elimination_widths(len(nodes), _i32p(ptr), _i32p(idx if idx.size else np.zeros(1, np.int32)),
                                         len(peo), _i32p(order), _i32p(widths)))
    return [int(w) for w in widths[:len(peo)]]


def _vars(graph, ids):
    return [Var(v, size=graph.nodes[v]["size"], name=graph.nodes[v]["name"]) for v in ids]


def _with_clique(graph, nodes):
    g = graph.copy()
    ids = [int(v) for v in nodes]
    for a in range(len(ids)):
        for b in range(a + 1, len(ids)):
            g.add_edge(ids[a], ids[b])
    return g


class Optimizer:
    def optimize(self, tensor_net):
        raise NotImplementedError


def _line_csr(buckets, ignore_ids):
    """The CSR adjacency of the line graph ``buckets2graph`` builds (nodes in
    order of first appearance, a clique per tensor, ignored variables left
    out) without building the graph: ``(vars, ptr, idx)`` with ``vars[i]``
    the first ``Var`` seen for node i."""
    index, first, adj = {}, [], []
    for bucket in buckets:
        for t in bucket:
            ids = []
            for v in t.indices:
                k = v.identity
                if k in ignore_ids:
                    continue
                i = index.get(k)
                if i is None:
                    i = index[k] = len(first)
                    first.append(v)
                    adj.append(set())
                ids.append(i)
            if len(ids) == 2:
                adj[ids[0]].add(ids[1])
                adj[ids[1]].add(ids[0])
            elif len(ids) > 2:
                for a in ids:
                    adj[a].update(ids)
    ptr = np.zeros(len(first) + 1, dtype=np.int32)
    flat = []
    for i, nb in enumerate(adj):
        nb.discard(i)
        flat += nb
        ptr[i + 1] = ptr[i] + len(nb)
    return first, ptr, np.asarray(flat, dtype=np.int32)


class OrderingOptimizer(Optimizer):
    _graph = None
    _graph_source = None

    @property
    def graph(self):
        """The line graph the order was computed on (built on first use when
        the order came from the bucket lists directly)."""
        if self._graph is None and self._graph_source is not None:
            from .circuit import buckets2graph
            buckets, ignored = self._graph_source
            self._graph = buckets2graph(buckets, ignore_variables=ignored)
            self._graph_source = None
        return self._graph

    @graph.setter
    def graph(self, g):
        self._graph = g
        self._graph_source = None

    def _get_ordering_ints(self, graph):
        return greedy_peo(graph)

    def _get_ordering(self, graph):
        peo, path = self._get_ordering_ints(graph)
        return _vars(graph, peo), path

    def _greedy_from_buckets(self, tensor_net, ignored):
        """Greedy order straight from the bucket lists (same nodes, order and
        adjacency as the line graph, so the same order as ``greedy_peo``)."""
        first, ptr, idx = _line_csr(tensor_net.buckets, {int(v) for v in ignored})
        n = len(first)
        order = np.zeros(max(n, 1), dtype=np.int32)
        widths = np.zeros(max(n, 1), dtype=np.int32)
        lib = _lib.load()
        _lib.check(lib.qb_greedy_order(n, _i32p(ptr), _i32p(idx if idx.size else np.zeros(1, np.int32)),
                                       _i32p(order), _i32p(widths)))
        peo = [Var(first[k].identity, size=first[k].size, name=first[k].name) for k in order[:n]]
        # the bucket lists of this network are not modified later (reordering
        # and slicing build new ones), so the graph can be built from them on demand
        self._graph, self._graph_source = None, (tensor_net.buckets, list(ignored))
        return peo, [int(w) for w in widths[:n]]

    def optimize(self, tensor_net):
        free_vars = list(tensor_net.free_vars)
        ignored = list(tensor_net.ket_vars) + list(tensor_net.bra_vars)
        if not free_vars and type(self)._get_ordering_ints is OrderingOptimizer._get_ordering_ints \
                and type(self)._get_ordering is OrderingOptimizer._get_ordering:
            peo, path = self._greedy_from_buckets(tensor_net, ignored)
        else:
            graph = tensor_net.get_line_graph()
            if free_vars:
                graph = _with_clique(graph, free_vars)
            self.graph = graph
            peo, path = self._get_ordering(graph)
        self.treewidth = max(path) if path else 0
        if free_vars:
            free_ids = set(int(v) for v in free_vars)
            peo = [v for v in peo if int(v) not in free_ids] + free_vars
        self.peo_ints = [int(v) for v in peo]
        peo = ignored + peo
        self.peo = peo
        self.ignored_vars = ignored
        return peo, tensor_net


DefaultOptimizer = OrderingOptimizer


class GreedyParvars:
    """Pick slice variables greedily by highest degree; removes them from ``graph``."""

    def __init__(self, graph):
        self.graph = graph
        self.result = []

    def step(self):
        nodes = list(self.graph.nodes)
        if not nodes:
            return 1
        costs = [-self.graph.degree(n) for n in nodes]
        best = nodes[int(np.argmin(costs))]
        self.result.append(best)
        self.graph.remove_node(best)
        return 0


class SlicesOptimizer(OrderingOptimizer):
    def __init__(self, tw_bias=2, max_tw=None, **kwargs):
        self.tw_bias = tw_bias
        self.max_tw = kwargs.get("target_tw") or max_tw

    def _get_max_tw(self):
        if self.max_tw is not None:
            return self.max_tw
        import psutil
        return int(np.log2(psutil.virtual_memory().available)) - 4

    def _split_graph(self, p_graph, max_tw):
        searcher = GreedyParvars(p_graph)
        peo_ints, path = self._get_ordering_ints(p_graph)
        while True:
            if self.treewidth < max_tw:
                break
            if searcher.step():
                raise MemoryError("Estimated OOM")
            peo_ints, path = self._get_ordering_ints(p_graph)
            self.treewidth = max(path) if path else 0
        return peo_ints, searcher.result

    def optimize(self, tensor_net):
        super().optimize(tensor_net)
        graph = self.graph
        p_graph = graph.copy()
        max_tw = self._get_max_tw() - self.tw_bias
        peo_ints, par_ints = self._split_graph(p_graph, max_tw)
        self.parallel_vars = _vars(graph, par_ints)
        peo = _vars(graph, peo_ints)
        free_vars = list(tensor_net.free_vars)
        if free_vars:
            free_ids = set(int(v) for v in free_vars)
            peo = [v for v in peo if int(v) not in free_ids] + free_vars
        self.peo = self.ignored_vars + peo + self.parallel_vars
        return self.peo, self.parallel_vars, tensor_net


class RGreedyOrdering:
    """Ordering mixin: every order (the first one and each one after a slice
    variable is removed) is the best of ``repeats`` randomized-greedy samples
    and the greedy order (``rgreedy_best``, seeded), so that slicing starts
    from a narrower order.  Combined with a slicer the way the reference
    combines orderers and slicers (``TamakiTrimSlicing``,
    ``qtensor/optimisation/Optimizer.py:227``)."""

    temp = 0.002
    repeats = 16
    seed = 0

    def _get_ordering_ints(self, graph):
        return rgreedy_best(graph, self.temp, self.repeats, self.seed)


class RGreedyOrderingOptimizer(RGreedyOrdering, OrderingOptimizer):
    pass


class RGreedySlicesOptimizer(RGreedyOrdering, SlicesOptimizer):
    pass


def _trim_costs(graph, order):
    """Tree-trimming cost of each node for an elimination order: the sum, over
    the elimination steps whose neighbourhood (with the eliminated node)
    contains it, of 2^(neighbourhood size) -- the memory of those steps that
    slicing the node would halve."""
    adj = _adjacency(graph)
    cost = {n: 0 for n in adj}
    for n in order:
        nb = adj.pop(n)
        w = 1 << len(nb)
        cost[n] += w
        for u in nb:
            cost[u] += w
            s = adj[u]
            s.discard(n)
            s.update(nb)
            s.discard(u)
    return cost


class TreeTrimSplitter(SlicesOptimizer):
    """Slicing by tree trimming (``qtensor/optimisation/Optimizer.py:178-220``):
    while the greedy treewidth exceeds the target by ``delta``, slice the
    ``int(0.2 * delta) + 1`` nodes of highest tree-trimming cost under the
    current order (the reference takes that rule from qtree's
    ``split_graph_by_tree_trimming``, which is not in the checkout: the cost
    here is the one of ``_trim_costs``), remove them and re-order."""

    def _split_graph(self, p_graph, max_tw):
        # start from the order optimize() found (the solver's, for TamakiTrimSlicing)
        peo_ints = list(getattr(self, "peo_ints", None) or self._get_ordering_ints(p_graph)[0])
        tw = max(elimination_widths(p_graph, peo_ints)) if peo_ints else 0
        result = []
        self.slice_history = [(0, tw)]
        while tw > max_tw:
            delta = tw - max_tw
            target = int(delta * 0.2) + 1
            cost = _trim_costs(p_graph, peo_ints)
            pos = {n: k for k, n in enumerate(peo_ints)}
            picks = sorted(cost, key=lambda n: (-cost[n], pos[n]))[:target]
            if not picks:
                raise MemoryError("Estimated OOM")
            p_graph.remove_nodes_from(picks)
            result += picks
            peo_ints, path = self._get_ordering_ints(p_graph)
            tw = max(path) if path else 0
            self.slice_history.append((len(result), tw))
        self.treewidth = tw
        return peo_ints, result


class RGreedyTreeTrimSplitter(RGreedyOrdering, TreeTrimSplitter):
    pass


class TamakiOptimizer(OrderingOptimizer):
    """Ordering from the PACE-2017 ``tw-heuristic`` solver run for ``wait_time`` s
    (``qtensor/optimisation/Optimizer.py:161-176``).  The solver reads the line
    graph in PACE ``.gr`` format on stdin and prints tree decompositions; the
    last one printed when it is stopped is turned into an elimination order.
    Like the reference, only the first ordering of ``optimize`` comes from
    the solver (``_get_ordering``); re-orderings inside slicers use the greedy
    rule (``_get_ordering_ints``).  When ``tw-heuristic`` is not on ``PATH``,
    or returns no decomposition, the greedy order is used with a warning."""

    def __init__(self, *args, wait_time=5, **kwargs):
        super().__init__(*args, **kwargs)
        self.wait_time = wait_time

    def _get_ordering(self, graph):
        try:
            peo, path = self._tamaki_ordering_ints(graph)
        except _SolverUnavailable as exc:
            import warnings
            warnings.warn(f"TamakiOptimizer: {exc}; using the greedy order", RuntimeWarning)
            return super()._get_ordering(graph)
        return _vars(graph, peo), path

    def _tamaki_ordering_ints(self, graph):
        exe = shutil.which("tw-heuristic")
        if exe is None:
            raise _SolverUnavailable("the 'tw-heuristic' executable is not on PATH")
        nodes = list(graph.nodes())
        if not nodes:
            return [], []
        pos = {n: k + 1 for k, n in enumerate(nodes)}
        edges = [(a, b) for a, b in graph.edges() if a != b]
        lines = [f"p tw {len(nodes)} {len(edges)}"] + [f"{pos[a]} {pos[b]}" for a, b in edges]
        with tempfile.NamedTemporaryFile("w", suffix=".gr", delete=False) as f:
            f.write("\n".join(lines) + "\n")
            path = f.name
        try:
            with open(path) as fin:
                proc = subprocess.Popen([exe], stdin=fin, stdout=subprocess.PIPE, text=True)
                try:
                    out, _ = proc.communicate(timeout=self.wait_time)
                except subprocess.TimeoutExpired:
                    proc.terminate()
                    out, _ = proc.communicate()
        finally:
            os.unlink(path)
        try:
            return _peo_from_td(_last_decomposition(out), nodes, graph)
        except RuntimeError as exc:
            raise _SolverUnavailable(str(exc)) from exc


class _SolverUnavailable(RuntimeError):
    pass


def _last_decomposition(text):
    """The last complete ``s td`` block of the solver's output (it may print
    improving decompositions one after another)."""
    blocks, cur = [], None
    for line in text.splitlines():
        if line.startswith("s td"):
            cur = [line]
            blocks.append(cur)
        elif cur is not None:
            cur.append(line)
    if not blocks:
        return text
    return "\n".join(blocks[-1])


class TamakiTrimSlicing(TamakiOptimizer, TreeTrimSplitter):
    """Tamaki ordering, then tree-trimming slicing with greedy re-ordering
    (``qtensor/optimisation/Optimizer.py:227``)."""


def _peo_from_td(text, nodes, graph):
    """Elimination order from a PACE ``.td`` tree decomposition: eliminate
    leaves of the decomposition tree first, each bag's vertices not shared
    with its parent."""
    bags, tree = {}, {}
    for line in text.splitlines():
        parts = line.split()
        if not parts or parts[0] == "c":
            continue
        if parts[0] == "s":
            continue
        if parts[0] == "b":
            bags[int(parts[1])] = [nodes[int(x) - 1] for x in parts[2:]]
        else:
            a, b = int(parts[0]), int(parts[1])
            tree.setdefault(a, set()).add(b)
            tree.setdefault(b, set()).add(a)
    if not bags:
        raise RuntimeError("tw-heuristic returned no decomposition")
    root = next(iter(bags))
    order, seen, stack = [], {root}, [(root, None)]
    parent = {}
    while stack:
        node, par = stack.pop()
        parent[node] = par
        order.append(node)
        for nb in tree.get(node, ()):
            if nb not in seen:
                seen.add(nb)
                stack.append((nb, node))
    peo, done = [], set()
    for b in reversed(order):
        keep = set(bags[parent[b]]) if parent[b] is not None else set()
        for v in bags[b]:
            if v not in keep and v not in done:
                peo.append(v)
                done.add(v)
    peo += [n for n in nodes if n not in done]
    return peo, elimination_widths(graph, peo)


def rgreedy_peo(graph, temp, rng):
    """One randomized-greedy elimination: Boltzmann choice over neighbour counts."""
    adj = _adjacency(graph)
    peo, degs = [], []
    while adj:
        nodes = list(adj)
        d = np.array([len(adj[n]) for n in nodes], dtype=float)
        w = np.exp(-(d - d.min()) / temp)
        k = int(rng.choice(len(nodes), p=w / w.sum()))
        best = nodes[k]
        nb = adj.pop(best)
        peo.append(best)
        degs.append(len(nb))
        for u in nb:
            s = adj[u]
            s.discard(best)
            s.update(nb)
            s.discard(u)
    return peo, degs


class RGreedyOptimizer(OrderingOptimizer):
    def __init__(self, temp=0.002, repeats=10, seed=None):
        self.temp = temp
        self.repeats = repeats
        self.rng = np.random.default_rng(seed)

    def _get_ordering_ints(self, graph):
        best = None
        for _ in range(self.repeats):
            peo, degs = rgreedy_peo(graph, self.temp, self.rng)
            if best is None or max(degs) < max(best[1]):
                best = (peo, degs)
        return best
