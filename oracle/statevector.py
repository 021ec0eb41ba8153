"""Independent full-statevector QAOA MaxCut simulator (oracle, tests only).

Standard textbook evolution, written without any tensor-network code:
|psi> = prod_l [ prod_q Rx-type mixer(beta_l) ] [ prod_edges e^{i 2 pi gamma_l (z_u xor z_v)} ] H^n |0>.

QTensor's angle units (``qtensor/CircuitComposer.py:99-112``,
``qtensor/tests/test_qaoa_energy.py:58``): the cost gate cX.ZPhase(2 gamma).cX
multiplies |z> by exp(i pi 2 gamma (z_u xor z_v)); the mixer is Cirq's
XPowGate(2 beta) = e^{i pi beta} [[cos pi beta, -i sin pi beta], [.., ..]].
"""
import numpy as np

H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)


def xphase(alpha):
    c, s, g = np.cos(np.pi * alpha / 2), np.sin(np.pi * alpha / 2), np.exp(1j * np.pi * alpha / 2)
    return g * np.array([[c, -1j * s], [-1j * s, c]])


def _apply_1q(psi, n, q, U):
    psi = psi.reshape((2,) * n)
    psi = np.moveaxis(np.tensordot(U, psi, axes=([1], [q])), 0, q)
    return psi.reshape(-1)


def _bits(n):
    idx = np.arange(2 ** n)
    return [((idx >> (n - 1 - q)) & 1) for q in range(n)]


def qaoa_state(n, edges, gamma, beta):
    bits = _bits(n)
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    for q in range(n):
        psi = _apply_1q(psi, n, q, H)
    for g, b in zip(gamma, beta):
        phase = np.zeros(2 ** n)
        for u, v in edges:
            phase += bits[u] ^ bits[v]
        psi = psi * np.exp(1j * np.pi * 2 * g * phase)
        X = xphase(2 * b)
        for q in range(n):
            psi = _apply_1q(psi, n, q, X)
    return psi


def zz_expectations(G, gamma, beta):
    """``{edge: <Z_u Z_v>}`` for every edge of networkx graph ``G`` (N <= ~20)."""
    nodes = list(G.nodes())
    pos = {v: k for k, v in enumerate(nodes)}
    n = len(nodes)
    edges = [(pos[u], pos[v]) for u, v in G.edges()]
    psi = qaoa_state(n, edges, gamma, beta)
    prob = np.abs(psi) ** 2
    bits = _bits(n)
    out = {}
    for (u, v), e in zip(G.edges(), edges):
        z = 1 - 2 * (bits[e[0]] ^ bits[e[1]])
        out[(u, v)] = float(np.dot(prob, z))
    return out


def maxcut_energy(G, gamma, beta):
    zz = zz_expectations(G, gamma, beta)
    return (G.number_of_edges() - sum(zz.values())) / 2
