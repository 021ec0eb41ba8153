"""Serving host logic on the CPU: admission model and request bookkeeping."""
import networkx as nx
import pytest

from qtensor_amd.batched import LaunchShape, LightconeEngine
from qtensor_amd.serving import Admission, predicted_bytes


def test_admission_queues_then_admits():
    adm = Admission(1000)
    assert adm.try_admit(600)
    assert not adm.try_admit(500)       # would exceed: stays queued
    adm.release(600)
    assert adm.try_admit(500) and adm.in_flight == 500
    assert not adm.fits_alone(2000)      # never fits: rejected as "Estimated OOM"


def test_predicted_bytes_tracks_treewidth_and_slices():
    G = nx.random_regular_graph(3, 40, seed=1)
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048))
    wl2 = eng.workload(G, 2)
    wl3 = eng.workload(G, 3)
    b2, b3 = predicted_bytes(wl2), predicted_bytes(wl3)
    assert 0 < b2 < b3
    # one edge: exactly 2^(max rank - 1) * 8 bytes
    one = eng.workload(G, 2, edges=[list(G.edges())[0]])
    plan = next(iter(one.plans.values()))
    assert predicted_bytes(one) == (1 << (plan.max_rank - 1)) * 8
    sliced = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=6).workload(G, 3)
    assert sum(w.n_slices for w in sliced.works) > len(sliced.works)
