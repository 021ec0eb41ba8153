"""Min-degree greedy elimination, restated in plain Python (test infrastructure).

Follows ``qtensor/utils.py:12-44`` (``get_neighbours_peo_vars``, as driven by
``qtensor/optimisation/Optimizer.py:46-88``): repeatedly take the node with the
fewest distinct neighbours -- the first such node in graph iteration order, as
``np.argmin`` over the node list does -- record that neighbour count as the
step's width, remove the node and join its neighbours into a clique.
``elimination_widths`` applies the same elimination to a fixed order
(``qtensor/utils.py:77-86``).  Used only by the tests, as the checker of the
native planner.
"""


def _adjacency(graph):
    return {n: set(graph[n]) - {n} for n in graph.nodes()}


def _eliminate(adj, n):
    nb = adj.pop(n)
    for u in nb:
        s = adj[u]
        s.discard(n)
        s.update(nb)
        s.discard(u)
    return len(nb)


def greedy_peo(graph):
    adj = _adjacency(graph)
    peo, degs = [], []
    while adj:
        best = min(adj, key=lambda n: len(adj[n]))
        peo.append(best)
        degs.append(_eliminate(adj, best))
    return peo, degs


def elimination_widths(graph, peo):
    adj = _adjacency(graph)
    return [_eliminate(adj, n) for n in peo]
