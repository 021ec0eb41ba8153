"""QAOA MaxCut circuit composers, per-edge lightcones and circuit simplification.

Follows ``qtensor/CircuitComposer.py`` and ``qtensor/OpFactory.py``:

* ``QtreeQAOAComposer``   -- cost term as ``cX, ZPhase(2 gamma), cX``
  (``CircuitComposer.py:108-112``), the composer of ``qtensor.QAOA_energy``
  (``qtensor/__init__.py:69-77``).
* ``ZZQtreeQAOAComposer`` -- cost term as one diagonal ``ZZ(2 gamma)`` gate
  (``CircuitComposer.py:167-169``).
* ``SimpZZQtreeComposer`` / ``DefaultQAOAComposer`` -- ZZ composer whose
  circuit is passed through commutation-based gate cancellation
  (``qtensor/__init__.py:55-63``, ``qtensor/simplify_circuit/``).

Angles are in QTensor's units (multiples of pi): mixer ``XPhase(2 beta)``
(``CircuitComposer.py:99-101``), cost ``2 gamma``.  ``energy_expectation(i, j)``
builds ``U Z_i Z_j U^dagger`` with the cone ansatz: layer ``l`` of ``p`` acts on
the radius ``p-l+1`` ball around the edge (``CircuitComposer.py:138-165``).

Every parametric gate carries ``param_terms = ((j, coeff),)`` with
``alpha = coeff * params[j]``, ``params = gamma[0..p-1] + beta[0..p-1]``, so a
compiled network can be re-evaluated at new angles.
"""
import networkx as nx

from .gates import QtreeFactory, XPhase, YPhase, ZPhase, ZZ, H, cX, Z


def nodes_at_distance(G, nodes, dist):
    """All nodes within ``dist`` hops of ``nodes`` (``qtensor/utils.py:176-185``)."""
    found = set(nodes)
    frontier = set(nodes)
    for _ in range(dist):
        nxt = set()
        for n in frontier:
            nxt.update(G[n])
        nxt -= found
        found |= nxt
        frontier = nxt
    return found


def _node_key(x):
    return (str(type(x)), x)


def get_edge_subgraph(G, edge, dist):
    """Radius-``dist`` subgraph around ``edge`` (``qtensor/utils.py:187-189``).

    Returned as a new graph whose nodes iterate in ascending order and whose
    edges iterate as ascending (a, b) pairs of node positions: a networkx
    subgraph view iterates a small subgraph in the order of its node *set*,
    which would make the gate order of the composed circuit depend on hashing.
    The gates a different order moves commute (cost gates with each other,
    mixers on distinct qubits), so the circuit's value is the same."""
    keep = nodes_at_distance(G, edge, dist)
    nodes = sorted(keep, key=_node_key)
    pos = {v: k for k, v in enumerate(nodes)}
    pairs = set()
    for a in nodes:
        for b in G[a]:
            if b in keep and b != a:
                pairs.add((a, b) if pos[a] < pos[b] else (b, a))
    sub = nx.Graph()
    sub.add_nodes_from(nodes)
    sub.add_edges_from(sorted(pairs, key=lambda e: (pos[e[0]], pos[e[1]])))
    return sub


def _dagger(gate):
    g = gate.dagger_me()
    if hasattr(gate, "param_terms"):
        g.param_terms = tuple((j, -c) for j, c in gate.param_terms)
    return g


class QtreeBuilder:
    operators = QtreeFactory

    def __init__(self, n_qubits):
        self.n_qubits = n_qubits
        self.reset()
        self.qubits = list(range(n_qubits))

    def reset(self):
        self._circuit = []

    def apply_gate(self, gate, *qubits, param_terms=None, **params):
        g = gate(*qubits, **params)
        if param_terms is not None:
            g.param_terms = tuple(param_terms)
        self._circuit.append(g)

    def inverse(self):
        self._circuit = list(reversed([_dagger(g) for g in self._circuit]))

    @property
    def circuit(self):
        return self._circuit

    @circuit.setter
    def circuit(self, c):
        self._circuit = c


class QAOAComposer:
    """Cone-ansatz QAOA composer over a networkx graph."""

    Builder = QtreeBuilder

    def __init__(self, graph, *, gamma, beta):
        self.graph = graph
        self.params = {"gamma": list(gamma), "beta": list(beta)}
        self.n_qubits = graph.number_of_nodes()
        self.builder = self.Builder(self.n_qubits)

    # -- builder passthrough
    @property
    def operators(self):
        return self.builder.operators

    @property
    def circuit(self):
        return self.builder.circuit

    @circuit.setter
    def circuit(self, c):
        self.builder.circuit = c

    @property
    def qubits(self):
        return self.builder.qubits

    def apply_gate(self, gate, *qubits, **params):
        self.builder.apply_gate(gate, *qubits, **params)

    @classmethod
    def _get_of_my_type(cls, *args, **kwargs):
        return cls(*args, **kwargs)

    # -- QAOA layers
    def layer_of_Hadamards(self):
        for q in self.qubits:
            self.apply_gate(self.operators.H, q)

    def _p(self):
        return len(self.params["gamma"])

    def x_term(self, u, beta, layer=None):
        terms = None if layer is None else ((self._p() + layer, 2.0),)
        self.apply_gate(self.operators.XPhase, u, alpha=2 * beta, param_terms=terms)

    def mixer_operator(self, beta, nodes=None, layer=None):
        if nodes is None:
            nodes = self.graph.nodes()
        for n in nodes:
            self.x_term(self.qubits[n], beta, layer)

    def append_zz_term(self, q1, q2, gamma, layer=None):
        terms = None if layer is None else ((layer, 2.0),)
        self.apply_gate(self.operators.cX, q1, q2)
        self.apply_gate(self.operators.ZPhase, q2, alpha=2 * gamma, param_terms=terms)
        self.apply_gate(self.operators.cX, q1, q2)

    def cost_operator_circuit(self, gamma, edges=None, layer=None):
        if edges is None:
            edges = self.graph.edges()
        for i, j in edges:
            self.append_zz_term(self.qubits[i], self.qubits[j], gamma, layer)

    def ansatz_state(self):
        """Full-graph ansatz: H layer then p (cost, mixer) layers."""
        beta, gamma = self.params["beta"], self.params["gamma"]
        assert len(beta) == len(gamma)
        self.layer_of_Hadamards()
        for l in range(len(beta)):
            self.cost_operator_circuit(gamma[l], layer=l)
            self.mixer_operator(beta[l], layer=l)

    def cone_ansatz(self, edge):
        beta, gamma = self.params["beta"], self.params["gamma"]
        assert len(beta) == len(gamma)
        p = len(beta)
        self.layer_of_Hadamards()
        cone_base = self.graph
        for l, (radius, g, b) in enumerate(zip(range(p, 0, -1), gamma, beta)):
            self.graph = get_edge_subgraph(cone_base, edge, radius)
            self.cost_operator_circuit(g, layer=l)
            self.mixer_operator(b, layer=l)
        self.graph = cone_base

    def energy_edge(self, i, j):
        self.apply_gate(self.operators.Z, self.qubits[i])
        self.apply_gate(self.operators.Z, self.qubits[j])

    def energy_expectation(self, i, j):
        """Circuit ``U Z_i Z_j U^dagger`` whose <0|.|0> amplitude is <Z_i Z_j>."""
        self.cone_ansatz(edge=(i, j))
        self.energy_edge(i, j)
        first = self.builder.circuit
        self.builder.reset()
        self.cone_ansatz(edge=(i, j))
        self.builder.inverse()
        second = self.builder.circuit
        self.circuit = first + second

    def energy_expectation_lightcone(self, edge):
        """Build the edge's energy circuit on its radius-p lightcone subgraph."""
        G = self.graph
        gamma, beta = self.params["gamma"], self.params["beta"]
        sub = get_edge_subgraph(G, edge, len(gamma))
        mapping = {v: k for k, v in enumerate(sub.nodes())}
        sub = nx.relabel_nodes(sub, mapping, copy=True)
        i, j = mapping[edge[0]], mapping[edge[1]]
        comp = self._get_of_my_type(sub, gamma=gamma, beta=beta)
        comp.energy_expectation(i, j)
        self.n_qubits = sub.number_of_nodes()
        self.circuit = comp.circuit


class QtreeQAOAComposer(QAOAComposer):
    pass


class ZZQtreeQAOAComposer(QAOAComposer):
    def append_zz_term(self, q1, q2, gamma, layer=None):
        terms = None if layer is None else ((layer, 2.0),)
        self.apply_gate(self.operators.ZZ, q1, q2, alpha=2 * gamma, param_terms=terms)


# ---------------------------------------------------------------------------
# Commutation-based simplification (qtensor/simplify_circuit/simplify_circuit.py:78-138)
# ---------------------------------------------------------------------------

def _eigenbasis(g):
    """Per-qubit eigenbasis labels used for the commutation test
    (``qtensor/simplify_circuit/gates.py``)."""
    if isinstance(g, ZZ):
        return {g.qubits[0]: "Z", g.qubits[1]: "Z"}
    if isinstance(g, (ZPhase, Z)):
        return {g.qubits[0]: "Z"}
    if isinstance(g, XPhase):
        return {g.qubits[0]: "X"}
    if isinstance(g, YPhase):
        return {g.qubits[0]: "Y"}
    if isinstance(g, H):
        return {g.qubits[0]: "X+Y"}
    if isinstance(g, cX):
        return {g.qubits[0]: "cnot_A", g.qubits[1]: "cnot_B"}
    raise TypeError(f"gate {g!r} not supported by simplify_circuit")


def _site(g):
    return tuple(sorted(g.qubits)) if isinstance(g, ZZ) else tuple(g.qubits)


def _commute(e1, e2):
    return all(e1[q] == e2[q] for q in set(e1) & set(e2))


def simplify_circuit(circuit):
    """Cancel or merge same-type, same-site gate pairs that commute with
    every gate between them, nearest pairs first.

    Parametric pairs whose angles sum to zero are removed (symbolically when
    the gates carry angle terms, else numerically below 1e-5); other
    parametric pairs merge into one gate with the summed angle at the later
    position.  ``Z`` and ``cX`` pairs cancel;
    ``H`` pairs are left alone (H*H is the identity, not H).
    """
    gates = list(circuit)
    if not gates:
        return []
    basis = [_eigenbasis(g) for g in gates]
    groups = {}
    for idx, g in enumerate(gates):
        groups.setdefault((type(g), _site(g)), []).append(idx)
    pairs = []
    for members in groups.values():
        for a in range(len(members)):
            for b in range(a + 1, len(members)):
                pairs.append((members[b] - members[a], members[a], members[b]))
    pairs.sort()
    alive = [True] * len(gates)
    for _, i, j in pairs:
        if not (alive[i] and alive[j]):
            continue
        gi, gj = gates[i], gates[j]
        if not gi.parametric and not isinstance(gi, (Z, cX)):
            continue
        if not all(_commute(basis[i], basis[k]) for k in range(i + 1, j) if alive[k]):
            continue
        if not gi.parametric:          # Z*Z = cX*cX = identity
            alive[i] = alive[j] = False
            continue
        ti, tj = getattr(gi, "param_terms", ()), getattr(gj, "param_terms", ())
        terms = _merge_terms(ti, tj)
        alpha = gi.alpha + gj.alpha
        # symbolic angles decide exactly (a compiled plan must stay valid for
        # every angle); gates with numeric angles only cancel as the
        # reference's do, below 1e-5
        if (terms == () if (ti or tj) else abs(alpha) < 1e-5):
            alive[i] = alive[j] = False
            continue
        merged = type(gi)(*gi.qubits, alpha=alpha)
        merged.param_terms = terms
        gates[j] = merged
        alive[i] = False
    return [g for g, a in zip(gates, alive) if a]


def _merge_terms(t1, t2):
    acc = {}
    for j, c in tuple(t1) + tuple(t2):
        acc[j] = acc.get(j, 0.0) + c
    return tuple((j, c) for j, c in sorted(acc.items()) if c != 0.0)


class SimpZZQtreeComposer(ZZQtreeQAOAComposer):
    @property
    def circuit(self):
        return simplify_circuit(self.builder.circuit)

    @circuit.setter
    def circuit(self, c):
        self.builder.circuit = c


DefaultQAOAComposer = SimpZZQtreeComposer
