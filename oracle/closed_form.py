"""p = 1 QAOA MaxCut edge expectation in closed form (oracle, tests only).

Wang, Hadfield, Jiang, Rieffel, "Quantum approximate optimization algorithm
for MaxCut: a fermionic view", PRA 97, 022304 (2018), Theorem 1: for
U = e^{-i b B} e^{-i g C}, C = sum (1 - Z_u Z_v)/2, B = sum X,

  <C_uv> = 1/2 + 1/4 sin(4b) sin(g) (cos^{d_u} g + cos^{d_v} g)
               - 1/4 sin^2(2b) cos^{d_u + d_v - 2 lam} g (1 - cos^{lam} 2g)

with d_u = deg(u) - 1, d_v = deg(v) - 1 and lam the number of triangles on
(u, v).  QTensor's units map as g = -2 pi gamma, b = pi beta
(``qtensor/tests/test_qaoa_energy.py:58``).  Returns <Z_u Z_v> = 1 - 2 <C_uv>.
"""
import numpy as np


def zz_p1(G, edge, gamma, beta):
    u, v = edge
    g = -2 * np.pi * gamma
    b = np.pi * beta
    du = G.degree(u) - 1
    dv = G.degree(v) - 1
    lam = len(set(G[u]) & set(G[v]))
    c = 0.5 + 0.25 * np.sin(4 * b) * np.sin(g) * (np.cos(g) ** du + np.cos(g) ** dv) \
        - 0.25 * np.sin(2 * b) ** 2 * np.cos(g) ** (du + dv - 2 * lam) * (1 - np.cos(2 * g) ** lam)
    return 1 - 2 * c
