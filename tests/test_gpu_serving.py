"""Continuous-batching server: results alone vs in a mix, admission, statevector."""
import networkx as nx
import numpy as np
import pytest

from oracle import statevector
from qtensor_amd.batched import LaunchShape, LightconeEngine
from qtensor_amd.serving import EnergyServer, poisson_trace

pytestmark = pytest.mark.gpu


def _pool():
    return [(f"g{k}", nx.random_regular_graph(3, n, seed=k)) for k, n in enumerate([12, 16, 20, 40, 64])]


def test_results_same_alone_and_in_mix(gpu):
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=11)
    server = EnergyServer(eng, n_streams=2).start()
    rng = np.random.default_rng(1)
    try:
        subs = poisson_trace(server, _pool(), rate=2000.0, n_requests=40, rng=rng)
    finally:
        server.stop()
    st = server.stats()
    assert st["requests"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0
    for fut, key, G, gamma, beta in subs:
        C, zz = fut.result()
        C1, zz1 = eng.energy(G, gamma, beta)            # the same request alone
        assert max(abs(a - b) for a, b in zip(zz, zz1)) <= 1e-12
        if G.number_of_nodes() <= 16:
            sv = statevector.zz_expectations(G, gamma, beta)
            assert max(abs(z - sv[e]) for z, e in zip(zz, G.edges())) <= 1e-5


def test_admission_rejects_and_queues(gpu):
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=11)
    big = nx.random_regular_graph(3, 64, seed=5)
    small = nx.random_regular_graph(3, 12, seed=6)
    wl_small = eng.workload(small, 1)
    from qtensor_amd.serving import predicted_bytes
    budget = 3 * predicted_bytes(wl_small)
    server = EnergyServer(eng, n_streams=2, hbm_budget=budget).start()
    try:
        f_big = server.submit(big, [0.1, 0.2, 0.3], [0.3, 0.2, 0.1], key="big")
        fs = [server.submit(small, [0.1 * k], [0.2], key="small") for k in range(1, 8)]
        with pytest.raises(MemoryError):
            f_big.result(timeout=60)
        for k, f in enumerate(fs, start=1):
            C, zz = f.result(timeout=60)
            sv = statevector.zz_expectations(small, [0.1 * k], [0.2])
            assert max(abs(z - sv[e]) for z, e in zip(zz, small.edges())) <= 1e-5
    finally:
        server.stop()


def test_depth4_served_with_its_own_shape_and_intractable_rejected(gpu):
    """p = 4 requests run with their own engine (256-thread items, slicing
    below 16, global scratch) beside p <= 3 requests, match the statevector,
    and a p = 4 request whose lightcones need more slice variables than a work
    item carries is rejected with "Estimated OOM"."""
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=11)
    eng4 = LightconeEngine(shape=LaunchShape(256, 4096, 2048), max_tw=16)
    server = EnergyServer(eng, n_streams=2, engines={4: eng4}).start()
    ring = nx.cycle_graph(12)                 # p = 4 lightcones of 10 qubits, all in LDS
    ladder = nx.ladder_graph(7)               # 14 qubits, tensors spill to global scratch
    G12 = nx.random_regular_graph(3, 12, seed=8)
    rng = np.random.default_rng(3)
    reqs = []
    try:
        for G, p in [(ring, 4), (G12, 2), (ladder, 4), (G12, 3)]:
            g, b = list(rng.uniform(0, np.pi, p)), list(rng.uniform(0, np.pi / 2, p))
            reqs.append((server.submit(G, g, b), G, g, b))
        hard = server.submit(nx.random_regular_graph(3, 64, seed=102), [0.1] * 4, [0.2] * 4)
        with pytest.raises(MemoryError, match="Estimated OOM"):
            hard.result()
        for fut, G, g, b in reqs:
            C, zz = fut.result()
            sv = statevector.zz_expectations(G, g, b)
            assert max(abs(z - sv[e]) for z, e in zip(zz, G.edges())) <= 1e-5
    finally:
        server.stop()
    assert server.rejected == 1


def test_plan_upgrades_and_slice_cap(gpu):
    """upgrades={p: engine}: a (graph, p) is served first with its usual
    (greedy) plans, re-planned in the background with the randomized-greedy
    order sliced further, and later requests use those plans; both give the
    statevector values.  max_slice_vars rejects a request whose lightcones
    need more slice variables ("Estimated OOM")."""
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=11)
    up = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=6, slicer="tree_trim", ordering="rgreedy",
                         rg_repeats=4)
    G = nx.random_regular_graph(3, 16, seed=9)
    g, b = [0.4, 0.9], [0.3, 0.6]
    sv = statevector.zz_expectations(G, g, b)
    server = EnergyServer(eng, n_streams=2, upgrades={2: up}).start()
    try:
        C0, zz0 = server.submit(G, g, b, key="G").result(timeout=120)
        server.wait_upgrades(timeout=120)
        assert server.upgraded == 1
        C1, zz1 = server.submit(G, g, b, key="G").result(timeout=120)
    finally:
        server.stop()
    for zz in (zz0, zz1):
        assert max(abs(z - sv[e]) for z, e in zip(zz, G.edges())) <= 1e-5
    capped = EnergyServer(up, n_streams=1, max_slice_vars=0).start()
    try:
        with pytest.raises(MemoryError, match="Estimated OOM"):
            capped.submit(G, g, b, key="G").result(timeout=120)
    finally:
        capped.stop()


def test_upgrade_engine_plans_when_quick_plans_exceed_the_cap(gpu):
    """When a request's quick plans need more slice variables than
    max_slice_vars allows, the upgrade engine plans it at once (its own launch
    shape, here 128-thread items instead of 64) and the request is served with
    those plans, matching the statevector; without an upgrade engine the same
    request is rejected."""
    G = nx.random_regular_graph(3, 16, seed=11)
    g, b = [0.25, 0.7], [0.45, 0.15]
    quick = LightconeEngine(shape=LaunchShape(64, 4096, 2048), max_tw=4)
    up = LightconeEngine(shape=LaunchShape(128, 4096, 2048), max_tw=7, slicer="tree_trim", ordering="rgreedy",
                         rg_repeats=4)
    need_quick = max(pl.n_sliced for pl in quick.workload(G, 2).plans.values())
    need_up = max(pl.n_sliced for pl in up.workload(G, 2).plans.values())
    assert need_up < need_quick
    sv = statevector.zz_expectations(G, g, b)
    server = EnergyServer(quick, n_streams=2, upgrades={2: up}, max_slice_vars=need_up).start()
    try:
        C, zz = server.submit(G, g, b, key="G").result(timeout=120)
        assert server.upgraded == 1
    finally:
        server.stop()
    assert max(abs(z - sv[e]) for z, e in zip(zz, G.edges())) <= 1e-5
    plain = EnergyServer(quick, n_streams=1, max_slice_vars=need_up).start()
    try:
        with pytest.raises(MemoryError, match="Estimated OOM"):
            plain.submit(G, g, b, key="G").result(timeout=120)
    finally:
        plain.stop()
