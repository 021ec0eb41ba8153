"""Per-edge QAOA MaxCut terms <Z_u Z_v> by bucket elimination, restated
without engine code (oracle: test infrastructure only).

The reference computes every edge term of ``QAOASimulator.energy_expectation``
(``qtensor/QAOASimulator.py:44-66``) as the amplitude <0|U^dag Z_u Z_v U|0> of a
per-edge lightcone circuit (``QtreeQAOAComposer.energy_expectation``,
``qtensor/CircuitComposer.py:82-96,138-165``), contracted by bucket elimination
in a min-degree greedy order (``qtensor/Simulate.py:66-94``,
``qtensor/optimisation/Optimizer.py:46-88``, ``qtensor/utils.py:12-44``).

This module builds that amplitude from the mathematics, with its own circuit,
network and elimination code:

* the circuit is the edge's causal cone: with S_p = {u, v} and
  S_{l-1} = S_l + neighbours(S_l), layer l (1..p) applies the cost phase on
  every edge with an endpoint in S_l and the mixer on S_l, after Hadamards on
  S_0 (gates outside the cone cancel between U and U^dag, so the term equals
  the full-graph expectation);
* gates in QTensor's units (``qtensor/CircuitComposer.py:99-112``,
  ``qtensor/OpFactory.py:14-56``): cost exp(i pi 2 gamma (z_a xor z_b)), mixer
  XPowGate(2 beta) = e^{i pi beta} [[cos pi beta, -i sin pi beta], [-i sin pi beta,
  cos pi beta]];
* the network has one variable per wire segment: a non-diagonal one-qubit gate
  opens a new variable, diagonal gates (cost, Z) attach to the current ones,
  |0> and <0| close each wire;
* elimination: min-degree greedy order over the variables' interaction graph
  (``oracle.ordering.greedy_peo``), each variable's tensors multiplied and the
  variable summed out (numpy einsum, complex128).

Pinned by the full statevector (``oracle.statevector``) and the p = 1 closed
form (``oracle.closed_form``) on small graphs in ``tests/test_oracle.py``.
"""
import string

import networkx as nx
import numpy as np

from . import ordering as _ordering

_LETTERS = string.ascii_letters
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
_Z = np.array([1, -1], dtype=np.complex128)
_KET0 = np.array([1, 0], dtype=np.complex128)


def _mixer(beta):
    a = 2 * beta
    c, s = np.cos(np.pi * a / 2), np.sin(np.pi * a / 2)
    return np.exp(1j * np.pi * a / 2) * np.array([[c, -1j * s], [-1j * s, c]])


def _cost(gamma):
    ph = np.exp(1j * np.pi * 2 * gamma)
    return np.array([[1, ph], [ph, 1]], dtype=np.complex128)


def causal_cone(G, edge, p):
    """Layers of the edge's causal cone: ``(S_0, [(edges_l, S_l) for l = 1..p])``."""
    u, v = edge
    S = [None] * (p + 1)
    S[p] = {u, v}
    for l in range(p, 0, -1):
        grown = set(S[l])
        for x in S[l]:
            grown.update(G[x])
        S[l - 1] = grown
    layers = []
    for l in range(1, p + 1):
        es = sorted({tuple(sorted((a, b), key=repr)) for a in S[l] for b in G[a] if a != b}, key=repr)
        layers.append((es, sorted(S[l], key=repr)))
    return sorted(S[0], key=repr), layers


def edge_network(G, edge, gamma, beta):
    """Tensors ``[(vars, array)]`` of <0|U^dag Z_u Z_v U|0> over integer variables."""
    p = len(gamma)
    S0, layers = causal_cone(G, edge, p)
    # gates in time order: ("m", q, matrix) one-qubit non-diagonal, ("d", qs, array) diagonal
    fwd = [("m", q, _H) for q in S0]
    for (es, nodes), g, b in zip(layers, gamma, beta):
        fwd += [("d", (a, c), _cost(g)) for a, c in es]
        X = _mixer(b)
        fwd += [("m", q, X) for q in nodes]
    mid = [("d", (edge[0],), _Z), ("d", (edge[1],), _Z)]
    back = []
    for kind, q, a in reversed(fwd):
        back.append((kind, q, a.conj().T if kind == "m" else a.conj()))
    tensors, cur = [], {}
    nxt = 0
    for q in S0:
        cur[q] = nxt
        tensors.append(((nxt,), _KET0))
        nxt += 1
    for kind, q, a in fwd + mid + back:
        if kind == "m":
            w = nxt
            nxt += 1
            tensors.append(((cur[q], w), a.T.copy()))     # T[in, out] = A[out, in]
            cur[q] = w
        else:
            tensors.append((tuple(cur[x] for x in q), a))
    for q in S0:
        tensors.append(((cur[q],), _KET0.conj()))
    return tensors


def contract(tensors):
    """Full contraction by bucket elimination in min-degree greedy order.
    Returns ``(value, width)``, width = the largest variable count of a step."""
    g = nx.Graph()
    for vs, _ in tensors:
        g.add_nodes_from(vs)
        g.add_edges_from((a, b) for i, a in enumerate(vs) for b in vs[i + 1:])
    order, _ = _ordering.greedy_peo(g)
    live = {k: t for k, t in enumerate(tensors)}
    holders = {}
    for k, (vs, _) in live.items():
        for v in vs:
            holders.setdefault(v, set()).add(k)
    scalar = 1.0 + 0j
    nid = len(tensors)
    width = 0
    for v in order:
        ks = sorted(holders.pop(v, ()))
        if not ks:
            continue
        group = [live.pop(k) for k in ks]
        union = sorted(set().union(*[set(vs) for vs, _ in group]))
        width = max(width, len(union))
        if len(union) > len(_LETTERS):
            raise ValueError(f"bucket of {len(union)} variables exceeds einsum's labels")
        lab = {x: _LETTERS[i] for i, x in enumerate(union)}
        out_vars = tuple(x for x in union if x != v)
        expr = ",".join("".join(lab[x] for x in vs) for vs, _ in group) + "->" + "".join(lab[x] for x in out_vars)
        res = np.einsum(expr, *[a for _, a in group], optimize="greedy")
        for x in out_vars:
            holders[x] -= set(ks)
        if out_vars:
            live[nid] = (out_vars, res)
            for x in out_vars:
                holders[x].add(nid)
            nid += 1
        else:
            scalar *= complex(res)
    for vs, a in live.values():          # closed networks leave nothing behind
        scalar *= complex(a)
    return scalar, width


def edge_zz(G, edge, gamma, beta):
    """<Z_u Z_v> of the QAOA state (complex; the imaginary part is rounding)."""
    return contract(edge_network(G, edge, gamma, beta))[0]


def edge_width(G, edge, p):
    """Largest elimination step (variables) of the edge's network, angles aside."""
    return contract_width(edge_network(G, edge, [0.1] * p, [0.2] * p))


def contract_width(tensors):
    g = nx.Graph()
    for vs, _ in tensors:
        g.add_nodes_from(vs)
        g.add_edges_from((a, b) for i, a in enumerate(vs) for b in vs[i + 1:])
    _, degs = _ordering.greedy_peo(g)
    return max(degs) + 1 if degs else 0
