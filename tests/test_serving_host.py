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


def test_plan_store_grows_and_keeps_offsets():
    """The device plan store doubles when full; plans keep their offsets and
    contents, and the replaced buffer is retained for batches in flight."""
    import numpy as np
    from qtensor_amd.plan import PROGRAM_PAD
    from qtensor_amd.serving import PlanStore

    class _P:
        def __init__(self, n, v):
            self.words = np.full(n, v, dtype=np.int32)

    store = PlanStore("cpu", capacity_words=1000)
    plans = [_P(300, k + 1) for k in range(6)]
    offs = [store.offset(p) for p in plans]
    assert offs == [0, 300, 600, 900, 1200, 1500]
    assert store.offset(plans[2]) == 600                 # uploaded once
    assert len(store.retired) >= 1 and store.buf.numel() >= 1800 + PROGRAM_PAD
    buf = store.buf.numpy()
    for p, o in zip(plans, offs):
        assert (buf[o:o + 300] == p.words).all()
    assert (buf[1800:1800 + PROGRAM_PAD] == 0).all()
