"""Per-edge lightcones in a canonical labelling, so that compiled plans are shared.

The energy term of edge (u, v) at depth p only depends on the radius-p induced
subgraph around the edge (``qtensor/CircuitComposer.py:82-96,138-151``,
``qtensor/utils.py:176-189``) up to relabelling of its nodes.  This module
relabels that subgraph canonically: colour refinement seeded with (is an edge
endpoint, distance to the edge), then individualisation of the first node of
the smallest ambiguous colour class until every node has its own colour.
Isomorphic lightcones whose refinement is unambiguous (all trees, most
lightcones of sparse random graphs) get the same key and share one plan; the
key is the exact relabelled edge list, so equal keys always mean equal
networks.

The circuit built on the canonical graph differs from the one built on the
reference's relabelling only by a permutation of qubits and the order of
commuting gates, so the edge term is the same up to FP rounding.
"""
from collections import deque

import networkx as nx

from .composer import nodes_at_distance


def _refine(nodes, adj, color):
    while True:
        sig = {v: (color[v], tuple(sorted(color[w] for w in adj[v]))) for v in nodes}
        ranks = {s: k for k, s in enumerate(sorted(set(sig.values())))}
        new = {v: ranks[sig[v]] for v in nodes}
        if len(set(new.values())) == len(set(color.values())):
            return new
        color = new


def canonical_lightcone(G, edge, p):
    """Return ``(key, n, edges, i, j)`` of the radius-``p`` lightcone of ``edge``.

    ``edges`` are pairs of canonical labels ``0..n-1`` (sorted), ``(i, j)`` the
    canonical labels of the edge endpoints, ``key`` a hashable exact description.
    """
    u, v = edge
    nodes = sorted(nodes_at_distance(G, (u, v), p), key=lambda x: (str(type(x)), x))
    node_set = set(nodes)
    adj = {x: [y for y in G[x] if y in node_set and y != x] for x in nodes}
    dist = {}
    dq = deque()
    for s in (u, v):
        dist[s] = 0
        dq.append(s)
    while dq:
        x = dq.popleft()
        for y in adj[x]:
            if y not in dist:
                dist[y] = dist[x] + 1
                dq.append(y)
    color = {x: (0 if x in (u, v) else 1, dist[x]) for x in nodes}
    ranks = {c: k for k, c in enumerate(sorted(set(color.values())))}
    color = _refine(nodes, adj, {x: ranks[color[x]] for x in nodes})
    while len(set(color.values())) < len(nodes):
        classes = {}
        for x in nodes:
            classes.setdefault(color[x], []).append(x)
        amb = min((c for c, m in classes.items() if len(m) > 1), key=lambda c: (len(classes[c]), c))
        pick = classes[amb][0]
        color = {x: 2 * color[x] + (0 if x == pick else 1) for x in nodes}
        ranks = {c: k for k, c in enumerate(sorted(set(color.values())))}
        color = _refine(nodes, adj, {x: ranks[color[x]] for x in nodes})
    label = color
    n = len(nodes)
    edges = sorted(tuple(sorted((label[a], label[b]))) for a in nodes for b in adj[a] if label[a] < label[b])
    i, j = sorted((label[u], label[v]))
    key = (n, tuple(edges), i, j)
    return key, n, edges, i, j


def canonical_graph(n, edges):
    g = nx.Graph()
    g.add_nodes_from(range(n))
    g.add_edges_from(edges)
    return g
