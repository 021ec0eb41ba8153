"""Loader of the full-size parity fixtures written by
``tests/golden/make_parity_fixtures.py`` (complex128 oracle values of the
benchmark workloads; see that script for the workloads and seeds)."""
import json
import os

import networkx as nx
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    with open(os.path.join(HERE, "parity.json")) as f:
        meta = json.load(f)
    with np.load(os.path.join(HERE, "parity.npz"), allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    return meta, arrays


def graph(meta_entry):
    """The workload's graph, regenerated with networkx and checked against the
    edge list the fixture holds (same edges, same order)."""
    spec = meta_entry["graph"]
    if spec.startswith("random_regular_graph"):
        d, n, seed = [int(x.split("=")[-1]) for x in spec[spec.index("(") + 1:-1].split(",")]
        G = nx.random_regular_graph(d, n, seed=seed)
    elif spec == "erdos_renyi_graph(64, 4/63, seed=4)":
        G = nx.erdos_renyi_graph(64, 4 / 63, seed=4)
    else:
        raise ValueError(spec)
    assert [list(e) for e in G.edges()] == meta_entry["edges"], "networkx regenerated another graph"
    return G


def tol(ref):
    """The metric's bar: |delta| <= 1e-5 * max(1, |ref|) per term."""
    return 1e-5 * max(1.0, abs(ref))
