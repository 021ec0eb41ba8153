"""Write the full-size parity fixtures: complex128 oracle values of the
benchmark workloads, computed on the CPU by ``oracle.qaoa_network`` (the
causal-cone restatement, which shares no code with the engine).

Workloads (graphs from networkx with fixed seeds, angles from numpy
``default_rng``):

* C2 -- ``random_regular_graph(3, 256, seed=1)``, p=3, gamma ~ U[0,pi]^3 then
  beta ~ U[0,pi/2]^3 from ``default_rng(2)``: all 384 edge terms.  One of them
  (edge (0, 83), the first edge) is also the value of the forced
  treewidth-30 network of that edge.
* C3 -- ``random_regular_graph(3, 1000, seed=3)``, p=2, angles from
  ``default_rng(3)``: all 1500 edge terms.
* C4 -- ``erdos_renyi_graph(64, 4/63, seed=4)``, p=3, angle sets 0 and 1 of the
  16 drawn from ``default_rng(4)``: every edge whose causal-cone network the
  oracle contracts with at most ``C4_MAX_WIDTH`` variables per step (the
  others are checked complex64 against complex128 on the device).
* C5 -- the serving pool (8 random 3-regular graphs, N = 2 * U{10..1000} from
  ``default_rng(5)``, graph seeds 100..107) and a trace of ``C5_REQUESTS``
  requests (graph, p in {1..4}, angles, Poisson gaps at ``C5_RATE``/s) drawn
  from ``default_rng(55)``; oracle terms for every request with p <= 3.

Outputs ``parity.json`` (inputs, metadata) and ``parity.npz`` (term values,
arrays only; load with ``allow_pickle=False``).  Run from the repository root:
``python tests/golden/make_parity_fixtures.py`` (about 10 minutes on 8 cores).
"""
import json
import multiprocessing as mp
import os
import sys
import time

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import qaoa_network as qn  # noqa: E402

C4_MAX_WIDTH = 23
C4_SETS = (0, 1)
C5_REQUESTS = 14
C5_RATE = 200.0


def c2_inputs():
    G = nx.random_regular_graph(3, 256, seed=1)
    rng = np.random.default_rng(2)
    return G, list(rng.uniform(0, np.pi, 3)), list(rng.uniform(0, np.pi / 2, 3))


def c3_inputs():
    G = nx.random_regular_graph(3, 1000, seed=3)
    rng = np.random.default_rng(3)
    return G, list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))


def c4_inputs(n_sets=16):
    G = nx.erdos_renyi_graph(64, 4 / 63, seed=4)
    rng = np.random.default_rng(4)
    sets = [(list(rng.uniform(0, np.pi, 3)), list(rng.uniform(0, np.pi / 2, 3))) for _ in range(n_sets)]
    return G, sets


def c5_pool():
    rng = np.random.default_rng(5)
    pool = []
    for k in range(8):
        n = int(rng.integers(10, 1001)) * 2
        pool.append((f"rr3_n{n}_s{k}", n, 100 + k))
    return pool


def c5_trace():
    rng = np.random.default_rng(55)
    reqs = []
    for _ in range(C5_REQUESTS):
        gap = float(rng.exponential(1.0 / C5_RATE))
        g = int(rng.integers(8))
        p = int(rng.integers(1, 5))
        reqs.append({"graph": g, "p": p, "gamma": list(rng.uniform(0, np.pi, p)),
                     "beta": list(rng.uniform(0, np.pi / 2, p)), "gap_s": gap})
    return reqs


_JOB = {}


def _init(job):
    _JOB.update(job)


def _term(args):
    name, k = args
    G, gamma, beta = _JOB[name]
    e = list(G.edges())[k]
    return name, k, qn.edge_zz(G, e, gamma, beta)


def main():
    t0 = time.time()
    meta, arrays, jobs, graphs = {}, {}, [], {}
    G2, g2, b2 = c2_inputs()
    G3, g3, b3 = c3_inputs()
    G4, sets4 = c4_inputs()
    graphs["c2"], graphs["c3"] = (G2, g2, b2), (G3, g3, b3)
    meta["c2"] = {"graph": "random_regular_graph(3, 256, seed=1)", "p": 3, "gamma": g2, "beta": b2,
                  "edges": [list(e) for e in G2.edges()]}
    meta["c3"] = {"graph": "random_regular_graph(3, 1000, seed=3)", "p": 2, "gamma": g3, "beta": b3,
                  "edges": [list(e) for e in G3.edges()]}
    jobs += [("c2", k) for k in range(G2.number_of_edges())]
    jobs += [("c3", k) for k in range(G3.number_of_edges())]
    widths = [qn.contract_width(qn.edge_network(G4, e, [0.1] * 3, [0.2] * 3)) for e in G4.edges()]
    c4_edges = [k for k, w in enumerate(widths) if w <= C4_MAX_WIDTH]
    meta["c4"] = {"graph": "erdos_renyi_graph(64, 4/63, seed=4)", "p": 3, "max_width": C4_MAX_WIDTH,
                  "sets": [{"index": s, "gamma": sets4[s][0], "beta": sets4[s][1]} for s in C4_SETS],
                  "edges": [list(e) for e in G4.edges()], "oracle_edges": c4_edges,
                  "oracle_widths": [widths[k] for k in c4_edges]}
    for s in C4_SETS:
        graphs[f"c4_{s}"] = (G4, sets4[s][0], sets4[s][1])
        jobs += [(f"c4_{s}", k) for k in c4_edges]
    pool = c5_pool()
    trace = c5_trace()
    meta["c5"] = {"pool": [{"key": key, "n": n, "seed": seed} for key, n, seed in pool], "rate_per_s": C5_RATE,
                  "requests": trace}
    for r, req in enumerate(trace):
        if req["p"] > 3:
            continue
        _, n, seed = pool[req["graph"]]
        G = nx.random_regular_graph(3, n, seed=seed)
        graphs[f"c5_{r}"] = (G, req["gamma"], req["beta"])
        jobs += [(f"c5_{r}", k) for k in range(G.number_of_edges())]
    print(f"{len(jobs)} oracle terms to compute", flush=True)
    jobs.sort(key=lambda j: 0 if j[0].startswith("c4") else 1)      # the widest first
    out = {}
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1), initializer=_init, initargs=(graphs,)) as pool_:
        for i, (name, k, v) in enumerate(pool_.imap_unordered(_term, jobs, chunksize=4)):
            out.setdefault(name, {})[k] = v
            if i % 1000 == 0:
                print(f"  {i} terms, {time.time() - t0:.0f} s", flush=True)
    for name, vals in out.items():
        if name.startswith("c4_"):
            ks = meta["c4"]["oracle_edges"]
        else:
            ks = range(len(vals))
        arrays[name] = np.array([vals[k] for k in ks], dtype=np.complex128)
    with open(os.path.join(HERE, "parity.json"), "w") as f:
        json.dump(meta, f)
    np.savez_compressed(os.path.join(HERE, "parity.npz"), **arrays)
    print(f"wrote parity.json / parity.npz in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
