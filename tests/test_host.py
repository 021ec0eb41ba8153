"""Host-side logic: composers, ordering, canonical lightcones, the plan compiler
(through a numpy interpreter of the device program) and the C-ABI library
surface.  CPU only."""
import ctypes
import os
import pickle
import re

import networkx as nx
import numpy as np
import pytest

from oracle import statevector
from oracle import np_framework as npf
from plan_interp import run_plan
from qtensor_amd import _lib
from qtensor_amd.composer import (QtreeQAOAComposer, SimpZZQtreeComposer, ZZQtreeQAOAComposer,
                                  get_edge_subgraph, simplify_circuit)
from qtensor_amd.lightcone import canonical_graph, canonical_lightcone
from qtensor_amd.ordering import OrderingOptimizer, SlicesOptimizer, elimination_widths, greedy_peo
from qtensor_amd.plan import PlanCompiler, mask_runs
from qtensor_amd.simulator import QAOAQtreeSimulator, QtreeSimulator

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_greedy_peo_path_and_clique():
    path = nx.path_graph(6)
    peo, degs = greedy_peo(path)
    assert max(degs) == 1 and sorted(peo) == list(range(6))
    K = nx.complete_graph(5)
    assert max(greedy_peo(K)[1]) == 4
    g = nx.random_regular_graph(3, 20, seed=1)
    peo, degs = greedy_peo(g)
    assert elimination_widths(g, peo) == degs


def test_greedy_tie_break_first_in_order():
    g = nx.Graph()
    g.add_nodes_from([5, 3, 9])
    g.add_edges_from([(5, 3), (3, 9)])
    peo, _ = greedy_peo(g)
    assert peo[0] == 5


def test_canonical_lightcone_invariant_under_relabel():
    G = nx.random_regular_graph(3, 60, seed=5)
    perm = np.random.default_rng(0).permutation(60)
    H = nx.relabel_nodes(G, {v: int(perm[v]) for v in G.nodes()})
    for e in list(G.edges())[:20]:
        k1 = canonical_lightcone(G, e, 2)[0]
        k2 = canonical_lightcone(H, (int(perm[e[0]]), int(perm[e[1]])), 2)[0]
        k3 = canonical_lightcone(G, (e[1], e[0]), 2)[0]
        assert k1 == k2 == k3


def test_canonical_graph_is_isomorphic_to_lightcone():
    G = nx.random_regular_graph(3, 30, seed=2)
    for e in list(G.edges())[:10]:
        key, n, edges, i, j = canonical_lightcone(G, e, 2)
        sub = get_edge_subgraph(G, e, 2)
        assert nx.is_isomorphic(sub, canonical_graph(n, edges))


def test_simplify_preserves_energy():
    G = nx.random_regular_graph(3, 10, seed=3)
    g, b = [0.3, 0.1], [0.2, 0.4]
    plain = QAOAQtreeSimulator(ZZQtreeQAOAComposer, bucket_backend=npf.NumpyBackend())
    simp = QAOAQtreeSimulator(SimpZZQtreeComposer, bucket_backend=npf.NumpyBackend())
    for e in list(G.edges())[:4]:
        c1 = ZZQtreeQAOAComposer(G, gamma=g, beta=b)
        c1.energy_expectation_lightcone(e)
        c2 = SimpZZQtreeComposer(G, gamma=g, beta=b)
        c2.energy_expectation_lightcone(e)
        assert len(c2.circuit) < len(c1.circuit)
        assert abs(plain._get_edge_energy(G, g, b, e)[0] - simp._get_edge_energy(G, g, b, e)[0]) < 1e-10


def test_max_tw_guard():
    G = nx.random_regular_graph(3, 10, seed=1)
    comp = QtreeQAOAComposer(G, gamma=[0.1, 0.2], beta=[0.3, 0.1])
    comp.ansatz_state()
    with pytest.raises(ValueError):
        QtreeSimulator(bucket_backend=npf.NumpyBackend(), max_tw=2).simulate(comp.circuit)


def test_mask_runs():
    assert mask_runs(0b1101100) == [(2, 2, 0), (5, 2, 2)]
    for m in (0, 1, 0xFFFF, 0xA5A5):
        x = 0x1234 & 0xFFFF
        off = 0
        for s, w, d in mask_runs(m):
            off |= ((x >> s) & ((1 << w) - 1)) << d
        ref, bit, mm = 0, 1, m
        while mm:
            low = mm & -mm
            if x & low:
                ref |= bit
            bit <<= 1
            mm &= mm - 1
        assert off == ref


def _edge_plan(G, e, p, Comp=QtreeQAOAComposer, **kw):
    key, n, ce, i, j = canonical_lightcone(G, e, p)
    comp = Comp(canonical_graph(n, ce), gamma=[0.0] * p, beta=[0.0] * p)
    comp.energy_expectation(i, j)
    opt = kw.pop("optimizer", None) or OrderingOptimizer()
    return PlanCompiler(**kw).compile_circuit(comp.circuit, opt)


@pytest.mark.parametrize("Comp", [QtreeQAOAComposer, ZZQtreeQAOAComposer, SimpZZQtreeComposer])
@pytest.mark.parametrize("p,lim,cap", [(1, 4, None), (2, 1 << 20, 4096), (2, 16, 2048)])
def test_plan_program_vs_oracle(Comp, p, lim, cap):
    G = nx.random_regular_graph(3, 16, seed=p)
    rng = np.random.default_rng(p)
    g, b = list(rng.uniform(0, np.pi, p)), list(rng.uniform(0, np.pi / 2, p))
    sv = statevector.zz_expectations(G, g, b)
    for e in list(G.edges())[:4]:
        plan = _edge_plan(G, e, p, Comp, lds_tensor_limit=lim, lds_capacity=cap)
        assert abs(run_plan(plan.words, g + b) - sv[e]) < 1e-10
        if cap is not None:
            assert plan.lds_need <= cap


@pytest.mark.parametrize("schedule", ["list", "first_fit", "levels", "alap"])
@pytest.mark.parametrize("lanes", [64, 256])
def test_plan_schedules_vs_statevector(schedule, lanes):
    """Every op scheduler and gather width gives the same terms; ops never
    exceed the lane count and the allocation respects the LDS capacity."""
    G = nx.random_regular_graph(3, 14, seed=5)
    rng = np.random.default_rng(5)
    g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
    sv = statevector.zz_expectations(G, g, b)
    for e in list(G.edges())[:3]:
        plan = _edge_plan(G, e, 2, lds_tensor_limit=512, lds_capacity=4096, lanes=lanes, schedule=schedule)
        assert plan.lds_need <= 4096
        assert abs(run_plan(plan.words, g + b, lanes=lanes) - sv[e]) < 1e-10


def test_plan_p3_and_slices_vs_oracle():
    G = nx.random_regular_graph(3, 40, seed=1)
    g, b = [0.3, 0.7, 0.2], [0.5, 0.1, 0.9]
    sim = QAOAQtreeSimulator(QtreeQAOAComposer, bucket_backend=npf.NumpyBackend())
    e = list(G.edges())[0]
    ref = sim._get_edge_energy(G, g, b, e)[0]
    plan = _edge_plan(G, e, 3, lds_tensor_limit=256, lds_capacity=4096)
    assert abs(run_plan(plan.words, g + b) - ref) < 1e-10
    sl = _edge_plan(G, e, 3, lds_tensor_limit=256, lds_capacity=4096,
                    optimizer=SlicesOptimizer(max_tw=plan.treewidth - 2, tw_bias=0))
    assert sl.n_sliced >= 1 and sl.max_rank < plan.max_rank
    total = sum(run_plan(sl.words, g + b, slice_bits=s) for s in range(1 << sl.n_sliced))
    assert abs(total - ref) < 1e-10


def test_tree_trim_slicing_vs_oracle():
    """TreeTrimSplitter (Optimizer.py:178-220) reaches the target width and the
    sum over its slices equals the unsliced term (test_qaoa_energy.py:106-121)."""
    from qtensor_amd.ordering import TreeTrimSplitter
    G = nx.random_regular_graph(3, 40, seed=1)
    g, b = [0.3, 0.7, 0.2], [0.5, 0.1, 0.9]
    sim = QAOAQtreeSimulator(QtreeQAOAComposer, bucket_backend=npf.NumpyBackend())
    e = list(G.edges())[2]
    ref = sim._get_edge_energy(G, g, b, e)[0]
    full = _edge_plan(G, e, 3, lds_tensor_limit=256, lds_capacity=4096)
    opt = TreeTrimSplitter(max_tw=full.treewidth - 3, tw_bias=0)
    sl = _edge_plan(G, e, 3, lds_tensor_limit=256, lds_capacity=4096, optimizer=opt)
    assert sl.n_sliced >= 1 and sl.treewidth <= full.treewidth - 3
    assert opt.slice_history[0][1] == full.treewidth and opt.slice_history[-1][1] == sl.treewidth
    total = sum(run_plan(sl.words, g + b, slice_bits=s) for s in range(1 << sl.n_sliced))
    assert abs(total - ref) < 1e-10


def test_plan_reparameterise_without_recompile():
    G = nx.random_regular_graph(3, 12, seed=4)
    e = list(G.edges())[2]
    plan = _edge_plan(G, e, 2, lds_tensor_limit=64)
    for seed in range(3):
        rng = np.random.default_rng(seed)
        g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
        assert abs(run_plan(plan.words, g + b) - statevector.zz_expectations(G, g, b)[e]) < 1e-10


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "qtensor_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(qb_\w+)\s*\(", text, re.M)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    syms = _header_symbols()
    assert syms and set(syms) == set(_lib.EXPORTED)
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.qb_version() >= 100
    # calls that need no device report "not initialised" cleanly
    n = ctypes.c_int32(0)
    assert lib.qb_exec_slots(64, 16, 0, 1, ctypes.byref(n)) == _lib.QB_ERR_NOT_INIT
    assert b"qb_init" in lib.qb_last_error()


def test_backend_pickles_without_device_state():
    from qtensor_amd.backend import HipBackend
    be = HipBackend()
    be2 = pickle.loads(pickle.dumps(be))
    assert be2.dtype == "complex64" and not be2._ready


def test_rgreedy_orders_are_valid_and_seeded():
    from qtensor_amd.ordering import RGreedyOptimizer, rgreedy_peo
    g = nx.random_regular_graph(3, 30, seed=2)
    peo, degs = rgreedy_peo(g, 1.0, np.random.default_rng(3))
    assert sorted(peo) == sorted(g.nodes()) and elimination_widths(g, peo) == degs
    p1, d1 = RGreedyOptimizer(temp=2.0, repeats=3, seed=5)._get_ordering_ints(g)
    p2, d2 = RGreedyOptimizer(temp=2.0, repeats=3, seed=5)._get_ordering_ints(g)
    assert p1 == p2 and max(d1) >= max(greedy_peo(g)[1]) - 3


def test_batch_vars_amplitudes_match_single_amplitude():
    """simulate_batch(batch_vars=2)[0] == simulate() (reference test_simulators.py:44-65)."""
    G = nx.random_regular_graph(4, 10, seed=3)
    comp = QtreeQAOAComposer(G, gamma=[np.pi / 3] * 2, beta=[np.pi / 2] * 2)
    comp.ansatz_state()
    sim = QtreeSimulator(bucket_backend=npf.NumpyBackend())
    amp = sim.simulate(comp.circuit)
    amps = sim.simulate_batch(comp.circuit, batch_vars=2)
    assert len(amps) == 4 and abs(amp[0] - amps[0]) < 1e-10
    nq = G.number_of_nodes()
    sv = statevector.qaoa_state(nq, [tuple(e) for e in G.edges()], [np.pi / 3] * 2, [np.pi / 2] * 2)
    # free qubits 0,1 are the two most significant bits of the statevector index
    assert np.allclose(amps, [sv[k << (nq - 2)] for k in range(4)], atol=1e-10)


def test_chain_plan_program_vs_oracle():
    """Wide plans with chain ops (a contraction and the contraction reading its
    output merged into one op, the intermediate never stored) interpret to the
    statevector value, sliced or not, and to the value of the same order
    compiled without them; the interpreter checks each chain op against the
    device kernel's structural limits."""
    from qtensor_amd import plan as P
    from qtensor_amd.wide import compile_wide, forced_width_order, slice_vars_for_order
    from qtensor_amd.tensornet import QtreeTensorNet
    G = nx.random_regular_graph(3, 14, seed=2)
    g, b = [0.3, 0.7], [0.5, 0.1]
    e = list(G.edges())[1]
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    ref = statevector.zz_expectations(G, g, b)[e]
    graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    tw0 = compile_wide(comp.circuit).treewidth
    n_chain = 0
    for extra, n_par in ((3, 0), (5, 0), (7, 0), (7, 2)):
        order = forced_width_order(graph, target_tw=tw0 + extra, temp=3.0, seed=1)
        par = slice_vars_for_order(graph, order, n_par) if n_par else []
        for chain in (True, False):
            wp = compile_wide(comp.circuit, order=order, par_ids=par, chain=chain)
            w = np.asarray(wp.plan.words)
            pos, found = int(w[3]), 0
            for _ in range(int(w[1])):
                w0 = int(w[pos]) & 0xFFFFFFFF
                found += bool((w0 & P.OP_FUSED) and (w0 & P.OP_CHAIN))
                pos += w0 >> 16
            assert chain or found == 0
            n_chain += found
            total = sum(run_plan(wp.plan.words, g + b, slice_bits=s, lanes=256) for s in range(wp.n_slices))
            assert abs(total - ref) < 1e-10
    assert n_chain >= 3


@pytest.mark.parametrize("fuse", [False, True], ids=["steps", "fused"])
def test_wide_plan_program_vs_oracle(fuse):
    """Wide plans (all tensors global, step ops only, or with fused chains)
    interpret to the oracle value, sliced or not, with greedy and with a forced
    wide order; the interpreter also re-derives every fused offset table."""
    import functools
    from qtensor_amd import wide as W
    forced_width_order, slice_vars_for_order = W.forced_width_order, W.slice_vars_for_order
    compile_wide = functools.partial(W.compile_wide, fuse=fuse)
    G = nx.random_regular_graph(3, 14, seed=2)
    g, b = [0.3, 0.7], [0.5, 0.1]
    e = list(G.edges())[1]
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    ref = statevector.zz_expectations(G, g, b)[e]
    wp = compile_wide(comp.circuit)
    assert wp.plan.lds_need == 0 and wp.plan.n_gather == 0
    assert abs(run_plan(wp.plan.words, g + b, lanes=256) - ref) < 1e-10
    from qtensor_amd.tensornet import QtreeTensorNet
    graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    order = forced_width_order(graph, target_tw=wp.treewidth + 3, temp=3.0, seed=1)
    par = slice_vars_for_order(graph, order, 2)
    ws = compile_wide(comp.circuit, order=order, par_ids=par)
    assert ws.n_slices == 4
    total = sum(run_plan(ws.plan.words, g + b, slice_bits=s, lanes=256) for s in range(4))
    assert abs(total - ref) < 1e-10
    if fuse:
        assert wp.plan.n_fused > 0 and ws.plan.n_fused > 0
        assert wp.plan.n_ops < compile_wide(comp.circuit, fuse=False).plan.n_ops


def test_fuse_steps_limits():
    """Fused groups respect the kernel's limits and cover every bucket step
    exactly once."""
    from qtensor_amd import plan as P
    from qtensor_amd.wide import compile_wide
    G = nx.random_regular_graph(3, 30, seed=4)
    comp = QtreeQAOAComposer(G, gamma=[0.1, 0.2, 0.3], beta=[0.4, 0.5, 0.6])
    comp.energy_expectation_lightcone(list(G.edges())[0])
    captured = []
    orig = P.fuse_steps

    def spy(steps):
        out = orig(steps)
        captured.append((steps, out))
        return out
    P.fuse_steps = spy
    try:
        compile_wide(comp.circuit, fuse=True)
    finally:
        P.fuse_steps = orig
    steps, out = captured[0]
    covered = []
    for st in out:
        if getattr(st, "fused", False):
            k = len(st.union) - len(st.out.dev)
            assert len(st.inputs) <= P.FUSED_MAX_IN and k <= P.FUSED_MAX_SUM and len(st.union) <= 30
            assert len(st.out.dev) >= 14 or k <= 4
            covered += st.members
        else:
            covered.append(steps.index(st))
    assert sorted(covered) == list(range(len(steps)))


def test_autotuner_two_stage_search(monkeypatch):
    """Stage 1 times every slicing threshold with each slicer under the base
    shape, stage 2 the other shapes at the best (threshold, slicer) pairs; the
    fastest candidate wins."""
    import networkx as nx
    from qtensor_amd import autotune
    from qtensor_amd.batched import LaunchShape
    G = nx.random_regular_graph(3, 16, seed=7)
    calls = []

    def fake_time(self, G, p, gamma, beta, shape, mtw, slicer="greedy"):
        calls.append((shape, mtw, slicer))
        if shape.key() == (64, 4096, 2048):
            raise ValueError("LDS capacity exhausted")
        cost = {None: 5.0, 13: 4.0, 12: 3.0, 11: 2.0, 10: 2.5}[mtw] - 0.6 * (slicer == "tree_trim" and mtw == 10)
        return cost - 0.1 * (shape.item_threads == 256) - 0.01 * (shape.lds_capacity == 2048), 0.0, None, None

    monkeypatch.setattr(autotune.Autotuner, "time_config", fake_time)
    monkeypatch.setattr(autotune, "slicing_candidates", lambda max_tw: [None, 13, 12, 11, 10])
    shape, mtw, slicer, table = autotune.Autotuner().tune(G, 2, [0.1, 0.2], [0.3, 0.4])
    base = [c for c in calls if c[0] == autotune.BASE_SHAPE]
    # every threshold with each slicer, unsliced once
    assert [(m, sl) for _, m, sl in base] == [(None, "greedy")] + [(m, sl) for m in (13, 12, 11, 10)
                                                                  for sl in autotune.SLICERS]
    rest = [c for c in calls if c[0] != autotune.BASE_SHAPE]
    # the best (threshold, slicer) pairs of stage 1 under every other shape
    # (10, tree_trim) costs 1.9, then (11, greedy) and (11, tree_trim) at 2.0
    assert sorted({(m, sl) for _, m, sl in rest}) == [(10, "tree_trim"), (11, "greedy"), (11, "tree_trim")]
    assert len(rest) == autotune.STAGE2_PAIRS * len(autotune.OTHER_SHAPES)
    assert shape == LaunchShape(256, 2048, 1024) and mtw == 10 and slicer == "tree_trim"
    assert sum("error" in row for row in table) == autotune.STAGE2_PAIRS    # the (64, ...) shape


def test_native_greedy_matches_oracle():
    """qb_greedy_order / qb_elimination_widths against the plain-Python
    restatement: random graphs (self loops, isolated nodes, shuffled labels)
    and QAOA lightcone line graphs."""
    from oracle import ordering as oo
    from qtensor_amd.composer import QtreeQAOAComposer
    from qtensor_amd.tensornet import QtreeTensorNet
    rng = np.random.default_rng(11)
    graphs = [nx.Graph(), nx.empty_graph(3), nx.complete_graph(6)]
    for k in range(30):
        g = nx.gnp_random_graph(int(rng.integers(2, 60)), float(rng.uniform(0.02, 0.4)), seed=int(rng.integers(1 << 30)))
        labels = list(rng.permutation(g.number_of_nodes()) * 7 + 3)
        g = nx.relabel_nodes(g, dict(zip(g.nodes(), labels)))
        if g.number_of_nodes() and k % 3 == 0:
            v = next(iter(g.nodes()))
            g.add_edge(v, v)
        graphs.append(g)
    G = nx.random_regular_graph(3, 24, seed=2)
    for e in list(G.edges())[:4]:
        comp = QtreeQAOAComposer(G, gamma=[0.1, 0.2], beta=[0.3, 0.4])
        comp.energy_expectation_lightcone(e)
        graphs.append(QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph())
    for g in graphs:
        assert greedy_peo(g) == oo.greedy_peo(g)
        order = list(rng.permutation(list(g.nodes()))) if g.number_of_nodes() else []
        assert elimination_widths(g, order) == oo.elimination_widths(g, order)
    with pytest.raises(ValueError):
        elimination_widths(nx.path_graph(3), [0, 0])


def test_native_rgreedy_rules():
    """qb_rgreedy_order: with no samples it is the greedy order of the oracle;
    with samples, the kept order is a permutation whose widths are the oracle's
    widths of that order, no worse than greedy by (largest width, sum of
    2^width), and at a temperature whose non-minimal weights underflow to 0
    every pick has the fewest neighbours (the rule of
    qtensor/optimisation/RGreedy.py:47-62); the same seed gives the same order."""
    from oracle import ordering as oo
    from qtensor_amd.composer import QtreeQAOAComposer
    from qtensor_amd.ordering import rgreedy_best
    from qtensor_amd.tensornet import QtreeTensorNet
    rng = np.random.default_rng(17)
    graphs = [nx.Graph(), nx.complete_graph(5)]
    for k in range(12):
        graphs.append(nx.gnp_random_graph(int(rng.integers(2, 50)), float(rng.uniform(0.05, 0.3)),
                                          seed=int(rng.integers(1 << 30))))
    G = nx.random_regular_graph(3, 16, seed=3)
    for e in list(G.edges())[:3]:
        comp = QtreeQAOAComposer(G, gamma=[0.1, 0.2], beta=[0.3, 0.4])
        comp.energy_expectation_lightcone(e)
        graphs.append(QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph())

    def cost(w):
        return (max(w) if w else 0, sum(2.0 ** x for x in w))

    for g in graphs:
        assert rgreedy_best(g, repeats=0) == oo.greedy_peo(g)
        peo, w = rgreedy_best(g, temp=1e-3, repeats=8, seed=5)
        assert sorted(peo, key=str) == sorted(g.nodes(), key=str)
        assert w == oo.elimination_widths(g, peo)
        assert cost(w) <= cost(oo.greedy_peo(g)[1])
        assert rgreedy_best(g, temp=1e-3, repeats=8, seed=5) == (peo, w)
        adj = {n: set(g[n]) - {n} for n in g.nodes()}
        for v in peo:
            assert len(adj[v]) == min(len(s) for s in adj.values())
            nb = adj.pop(v)
            for u in nb:
                adj[u] |= nb
                adj[u] -= {u, v}
    with pytest.raises(ValueError):
        rgreedy_best(nx.path_graph(3), temp=0.0)
    with pytest.raises(ValueError):
        rgreedy_best(nx.path_graph(3), repeats=-1)


def test_native_place_intervals_matches_restatement():
    """qb_place_intervals against the plain-Python allocator (tests/place_ref.py)
    on random buffer sets, and directly: buffers whose op intervals overlap get
    disjoint element ranges, all at or above the base."""
    import numpy as np
    from qtensor_amd.plan import _place_order
    from place_ref import place_order
    rng = np.random.default_rng(21)
    for trial in range(40):
        m = int(rng.integers(1, 300))
        sizes = rng.integers(1, 64, m) * 2
        starts = rng.integers(0, 100, m)
        ends = starts + rng.integers(0, 30, m)
        order = rng.permutation(m)
        base = int(rng.integers(0, 16))
        offs, peak = _place_order(sizes, starts, ends, order, base)
        ref_offs, ref_peak = place_order(sizes, starts, ends, order, base)
        assert list(offs) == ref_offs and peak == ref_peak
        assert offs.min() >= base and peak == max(base, int((offs + sizes).max()))
        for i in range(m):
            for j in range(i + 1, m):
                if starts[i] <= ends[j] and starts[j] <= ends[i]:
                    assert offs[i] + sizes[i] <= offs[j] or offs[j] + sizes[j] <= offs[i]


def test_order_items_serpentine_bands():
    """Longest first; with a band, every other run of ``band`` items reversed;
    word 3 always carries the item's index in the given order."""
    import numpy as np
    from qtensor_amd.batched import order_items
    items = [(k, 0, 0, 0) for k in range(10)]
    costs = [5, 9, 1, 7, 3, 8, 2, 6, 4, 0]
    plain = order_items(items, costs)
    assert list(plain[:, 0]) == [1, 5, 3, 7, 0, 8, 4, 6, 2, 9]
    assert list(plain[:, 3]) == list(plain[:, 0])
    snake = order_items(items, costs, band=4)          # runs [1 5 3 7] [0 8 4 6] [2 9], the second reversed
    assert list(snake[:, 0]) == [1, 5, 3, 7, 6, 4, 8, 0, 2, 9]
    assert list(snake[:, 3]) == list(snake[:, 0])


def test_group_items_pairs_items_of_one_plan():
    """Groups of two: both items of a group share a plan, every real item
    appears once with its own slot, padding items copy their group's last item
    and take the slots after the real ones; groups are longest first."""
    import numpy as np
    from qtensor_amd.batched import group_items
    # plan offsets 100 (3 items, cost 5), 200 (2 items, cost 9), 300 (1 item, cost 1)
    items = [(100, 0, 0, 0), (200, 4, 0, 0), (100, 0, 1, 0), (300, 8, 0, 0), (100, 0, 2, 0), (200, 4, 1, 0)]
    costs = [5, 9, 5, 1, 5, 9]
    out, n_pad = group_items(items, costs, 2)
    assert n_pad == 2 and out.shape == (8, 4)
    for g in range(0, 8, 2):
        assert out[g, 0] == out[g + 1, 0]
    assert list(out[::2, 0]) == [200, 100, 100, 300]
    slots = sorted(out[:, 3])
    assert slots == list(range(8))
    real = out[out[:, 3] < 6]
    assert sorted(map(tuple, real[:, :3])) == sorted(items_[:3] for items_ in items)
    for k in np.nonzero(out[:, 3] >= 6)[0]:
        assert tuple(out[k, :3]) == tuple(out[k - 1, :3])    # a copy of its group's last real item
    one, n0 = group_items(items, costs, 1)
    assert n0 == 0 and one.shape == (6, 4)


def test_check_program_groups():
    """With groups, every group's items share a plan and the item count is a
    multiple of the group size."""
    plans, prog, items, lds, glb = _checked_program()
    pair = np.array([items[0], items[0], items[1], items[1]], dtype=np.int32)
    pair[:, 3] = [0, 1, 2, 3]
    _lib.check_program(prog, pair, 4, 64, lds, glb, group=2)
    mixed = pair.copy()
    mixed[1, 0] = items[1, 0]
    with pytest.raises(ValueError):
        _lib.check_program(prog, mixed, 4, 64, lds, glb, group=2)
    with pytest.raises(ValueError):
        _lib.check_program(prog, pair[:3], 4, 64, lds, glb, group=2)
    with pytest.raises(ValueError):
        _lib.check_program(prog, pair, 4, 64, lds, glb, group=3)


def _checked_program(lanes=64, lim=64, cap=4096, p=2, seed=3):
    """A small valid batched program (two plans, their items) for qb_check_program."""
    G = nx.random_regular_graph(3, 12, seed=seed)
    plans = [_edge_plan(G, e, p, lds_tensor_limit=lim, lds_capacity=cap, lanes=lanes) for e in list(G.edges())[:2]]
    words, items, pos = [], [], 0
    for k, pl in enumerate(plans):
        words.append(pl.words)
        items.append([pos, 0, 0, k])
        pos += pl.words.size
    words.append(np.zeros(320, dtype=np.int32))
    prog = np.concatenate(words).astype(np.int32)
    lds = max(pl.lds_need for pl in plans)
    glb = max(pl.glb_need for pl in plans)
    return plans, prog, np.asarray(items, dtype=np.int32), lds, glb


def test_check_program_accepts_compiled_plans():
    for lim in (64, 4):       # all-LDS plans, and plans with intermediates in scratch
        plans, prog, items, lds, glb = _checked_program(lim=lim)
        _lib.check_program(prog, items, 2 * 2, 64, lds, glb)


def test_check_program_rejects_corruptions():
    """Every corruption that would make qb_execute read or write out of bounds
    is rejected on the host (no device work)."""
    plans, prog, items, lds, glb = _checked_program()
    ok = dict(n_params=4, item_threads=64, lds_elems=lds, scratch_elems=glb)

    def rejects(prog_, items_, **kw):
        args = dict(ok)
        args.update(kw)
        with pytest.raises(ValueError):
            _lib.check_program(prog_, items_, args["n_params"], args["item_threads"], args["lds_elems"],
                               args["scratch_elems"])

    _lib.check_program(prog, items, **ok)
    rejects(prog, items, lds_elems=lds - 1)                      # plan needs more LDS than reserved
    bad = items.copy(); bad[1, 0] = prog.size                    # plan offset past the program
    rejects(prog, bad)
    bad = items.copy(); bad[1, 3] = 0                            # output slots not a permutation
    rejects(prog, bad)
    bad = items.copy(); bad[0, 1] = 5                            # angle offset past the angles
    rejects(prog, bad)
    rejects(prog, items, n_params=1)                             # gate angle index past the angles
    p0 = plans[0].words
    g0 = int(p0[2])
    bad = prog.copy(); bad[g0] = 0                               # kConst gate: no constant pool
    rejects(bad, items)
    bad = prog.copy(); bad[g0 + 1] = lds                         # gate tensor outside the LDS data
    rejects(bad, items)
    # the first op record of plan 0: break its fields one at a time
    op = int(p0[3])
    w0 = int(p0[op])
    if w0 & (1 << 15):                                           # gather op: lane table entry
        tab = op + int(p0[op + 2])
        bad = prog.copy(); bad[tab] = lds                        # lane writes past the LDS data
        rejects(bad, items)
        bad = prog.copy(); bad[tab + 1] = (lds << 16) | 4        # lane reads past the LDS data
        rejects(bad, items)
        bad = prog.copy(); bad[op + 1] = 65                      # more active lanes than item threads
        rejects(bad, items)
    bad = prog.copy(); bad[op] = (w0 & 0xFFFF) | (1 << 28)       # record longer than the program
    rejects(bad, items)
    bad = prog.copy(); bad[4] = lds                              # result outside the LDS data
    rejects(bad, items)


def test_check_program_rejects_step_mask_runs_and_top():
    """Step ops: a runs word that disagrees with its mask, a top that is not
    2^popcount(mask), or scratch locations without scratch are rejected."""
    G = nx.random_regular_graph(3, 12, seed=3)
    pl = _edge_plan(G, list(G.edges())[0], 2, lds_tensor_limit=4, lds_capacity=4096, gather=False)
    prog = np.concatenate([pl.words, np.zeros(320, np.int32)]).astype(np.int32)
    items = np.array([[0, 0, 0, 0]], dtype=np.int32)
    _lib.check_program(prog, items, 4, 64, pl.lds_need, pl.glb_need)
    pos, steps = int(pl.words[3]), []
    for _ in range(int(pl.words[1])):
        w0 = int(prog[pos])
        if not w0 & ((1 << 15) | (1 << 13)):
            steps.append(pos)
        pos += w0 >> 16
    s = next(q for q in steps if int(prog[q + 5]) != 0)          # a step whose input 0 has a mask
    bad = prog.copy(); bad[s + 6] ^= 1 << 5                       # runs word no longer matches
    with pytest.raises(ValueError):
        _lib.check_program(bad, items, 4, 64, pl.lds_need, pl.glb_need)
    bad = prog.copy(); bad[s + 4] = 3                             # top not 0 or 2^popcount
    with pytest.raises(ValueError):
        _lib.check_program(bad, items, 4, 64, pl.lds_need, pl.glb_need)
    if pl.glb_need:
        with pytest.raises(ValueError):
            _lib.check_program(prog, items, 4, 64, pl.lds_need, 0)


def test_oracle_transpose_and_slice_vs_loops():
    """The oracle's own transpose + slice (no engine code) against explicit
    index loops: ints drop an axis, slices keep it with the new size."""
    from qtensor_amd.model import Var
    rng = np.random.default_rng(0)
    for trial in range(20):
        r = int(rng.integers(1, 5))
        ids = [int(x) for x in rng.choice(50, size=r, replace=False)]
        vs = [Var(i) for i in ids]
        data = rng.normal(size=(2,) * r) + 1j * rng.normal(size=(2,) * r)
        sd = {}
        for v in vs:
            c = int(rng.integers(0, 3))
            if c == 1:
                sd[v] = int(rng.integers(0, 2))
            elif c == 2:
                b = int(rng.integers(0, 2))
                sd[v] = slice(b, b + 1)
        out_vars, out = npf.transpose_and_slice(tuple(vs), data, sd)
        order = sorted(range(r), key=lambda a: ids[a])
        kept = [a for a in order if not isinstance(sd.get(vs[a]), int)]
        assert [int(v) for v in out_vars] == [ids[a] for a in kept]
        ranges = [range(2) if vs[a] not in sd else
                  ([sd[vs[a]]] if isinstance(sd[vs[a]], int) else range(2)[sd[vs[a]]]) for a in range(r)]
        for v, a in zip(out_vars, kept):
            assert v.size == len(ranges[a])
        for full in np.ndindex(*(len(x) for x in ranges)):
            idx = tuple(ranges[a][full[a]] for a in range(r))
            out_idx = tuple(full[a] for a in kept)
            assert out[out_idx] == data[idx]


def _td(bags, tree_edges, n, width):
    lines = [f"s td {len(bags)} {width + 1} {n}"]
    lines += [f"b {k + 1} " + " ".join(map(str, b)) for k, b in enumerate(bags)]
    lines += [f"{a} {b}" for a, b in tree_edges]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("name", ["path", "cycle", "grid"])
def test_peo_from_td_hand_written_decompositions(name):
    """PACE .td -> elimination order (leaves of the decomposition first): for
    these hand-written decompositions the order's width equals the largest
    bag size minus one, and every node is eliminated once."""
    from qtensor_amd.ordering import _peo_from_td
    if name == "path":
        g = nx.path_graph(6)
        bags = [[k, k + 1] for k in range(1, 6)]
    elif name == "cycle":
        g = nx.cycle_graph(6)
        bags = [[1, 2, 6], [2, 3, 6], [3, 4, 6], [4, 5, 6]]
    else:
        g = nx.grid_2d_graph(3, 3)
        g = nx.convert_node_labels_to_integers(g, ordering="sorted")     # row-major 0..8
        bags = [[k, k + 1, k + 2, k + 3] for k in range(1, 7)]
    tree = [(k, k + 1) for k in range(1, len(bags))]
    nodes = list(g.nodes())
    width = max(len(b) for b in bags) - 1
    peo, widths = _peo_from_td(_td(bags, tree, len(nodes), width), nodes, g)
    assert sorted(peo) == sorted(nodes)
    assert max(widths) == width
    assert widths == elimination_widths(g, peo)


def _stub_solver(tmp_path):
    """A stand-in ``tw-heuristic`` on PATH: reads a .gr on stdin and prints a
    tree decomposition built from a min-degree elimination (a valid .td)."""
    exe = tmp_path / "tw-heuristic"
    exe.write_text(
        "#!/usr/bin/env python3\n"
        "import sys\n"
        "adj = {}\n"
        "for line in sys.stdin:\n"
        "    p = line.split()\n"
        "    if not p or p[0] in ('c',):\n"
        "        continue\n"
        "    if p[0] == 'p':\n"
        "        adj = {v: set() for v in range(1, int(p[2]) + 1)}\n"
        "        continue\n"
        "    a, b = int(p[0]), int(p[1])\n"
        "    adj[a].add(b); adj[b].add(a)\n"
        "bags, owner = [], {}\n"
        "live = {v: set(s) for v, s in adj.items()}\n"
        "while live:\n"
        "    v = min(live, key=lambda x: (len(live[x]), x))\n"
        "    nb = live.pop(v)\n"
        "    for u in nb:\n"
        "        live[u] |= nb - {u}; live[u].discard(v)\n"
        "    bags.append(sorted(nb | {v})); owner[v] = len(bags)\n"
        "edges = []\n"
        "for k, b in enumerate(bags, start=1):\n"
        "    later = [owner[u] for u in b if owner[u] > k]\n"
        "    if later: edges.append((k, min(later)))\n"
        "print('s td', len(bags), max(len(b) for b in bags), len(adj))\n"
        "for k, b in enumerate(bags, start=1): print('b', k, *b)\n"
        "for a, b in edges: print(a, b)\n"
        "sys.stdout.flush()\n")
    exe.chmod(0o755)
    return str(tmp_path)


def test_tamaki_falls_back_to_greedy_without_solver(monkeypatch, tmp_path):
    from qtensor_amd.ordering import TamakiOptimizer
    monkeypatch.setenv("PATH", str(tmp_path))           # no tw-heuristic anywhere
    G = nx.random_regular_graph(3, 12, seed=2)
    comp = QtreeQAOAComposer(G, gamma=[0.2], beta=[0.3])
    comp.energy_expectation_lightcone(list(G.edges())[0])
    from qtensor_amd.tensornet import QtreeTensorNet
    tn = QtreeTensorNet.from_qtree_gates(comp.circuit)
    with pytest.warns(RuntimeWarning, match="tw-heuristic"):
        peo, _ = TamakiOptimizer(wait_time=1).optimize(tn)
    ref, _ = OrderingOptimizer().optimize(QtreeTensorNet.from_qtree_gates(comp.circuit))
    assert [int(v) for v in peo] == [int(v) for v in ref]


def test_tamaki_and_tamaki_trim_slicing_with_stub_solver(monkeypatch, tmp_path):
    """The solver's decomposition drives the order (TamakiOptimizer) and the
    first order of TamakiTrimSlicing, whose slices sum to the statevector term
    (the sliced == unsliced invariant, test_qaoa_energy.py:106-121)."""
    from qtensor_amd.ordering import TamakiOptimizer, TamakiTrimSlicing
    from qtensor_amd.simulator import FeynmanSimulator
    monkeypatch.setenv("PATH", _stub_solver(tmp_path) + os.pathsep + os.environ.get("PATH", ""))
    G = nx.random_regular_graph(3, 14, seed=3)
    g, b = [0.21, 0.43], [0.37, 0.12]
    sv = statevector.zz_expectations(G, g, b)
    opt = TamakiOptimizer(wait_time=20)
    sim = QAOAQtreeSimulator(QtreeQAOAComposer, bucket_backend=npf.NumpyBackend(), optimizer=opt)
    e = list(G.edges())[1]
    assert abs(sim._get_edge_energy(G, g, b, e)[0] - sv[e]) < 1e-10
    assert opt.treewidth == max(elimination_widths(opt.graph, opt.peo_ints))
    tts = TamakiTrimSlicing(max_tw=opt.treewidth - 2, tw_bias=0, wait_time=20)
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    fs = FeynmanSimulator(bucket_backend=npf.NumpyBackend(), optimizer=tts)
    val = fs.simulate(comp.circuit, tw_bias=0)
    assert len(fs.parallel_vars) >= 1 and tts.treewidth <= opt.treewidth - 2
    assert abs(complex(val[0]) - sv[e]) < 1e-10


@pytest.mark.parametrize("p,lanes,lim", [(1, 64, 2048), (2, 128, 64), (3, 256, 2048)])
def test_native_planner_equals_python_compiler(p, lanes, lim):
    """The native lightcone planner gives the Python planner's canonical
    lightcone keys and, for every composer, order and slicer, the same program
    words and plan figures (the Python compiler is its checker)."""
    from qtensor_amd import native_planner as npl
    from qtensor_amd.ordering import (RGreedyOrderingOptimizer, RGreedySlicesOptimizer,
                                      RGreedyTreeTrimSplitter, TreeTrimSplitter)
    G = nx.random_regular_graph(3, 50, seed=p)
    edges = list(G.edges())[:16]
    edge_class, classes = npl.lightcone_classes(G, edges, p)
    assert [classes[c][0] for c in edge_class] == [canonical_lightcone(G, e, p)[0] for e in edges]
    configs = [("greedy", None, None), ("rgreedy", None, None), ("greedy", "greedy", 6),
               ("rgreedy", "tree_trim", 6), ("greedy", "tree_trim", 5)]
    fields = ("n_steps", "n_ops", "n_gates", "treewidth", "max_rank", "lds_need", "glb_need", "n_sliced",
              "bytes_step_model", "flops", "n_barriers", "n_gather", "n_reduce", "steps_words")
    n_cmp = 0
    for Comp in (QtreeQAOAComposer, ZZQtreeQAOAComposer, SimpZZQtreeComposer):
        for ordering, slicer, mtw in configs:
            for key, n, ce, i, j, rec in classes[:3]:
                comp = Comp(canonical_graph(n, ce), gamma=[0.0] * p, beta=[0.0] * p)
                comp.energy_expectation(i, j)
                rg = ordering == "rgreedy"
                if mtw is None:
                    opt = RGreedyOrderingOptimizer() if rg else OrderingOptimizer()
                elif slicer == "tree_trim":
                    opt = (RGreedyTreeTrimSplitter if rg else TreeTrimSplitter)(max_tw=mtw, tw_bias=0)
                else:
                    opt = (RGreedySlicesOptimizer if rg else SlicesOptimizer)(max_tw=mtw, tw_bias=0)
                res = []
                for build in (lambda: PlanCompiler(lds_tensor_limit=lim, lds_capacity=4096, lanes=lanes)
                              .compile_circuit(comp.circuit, opt),
                              lambda: npl.compile_lightcone(rec, p, Comp, ordering=ordering, slicer=slicer,
                                                            max_tw=mtw, lanes=lanes, lds_tensor_limit=lim,
                                                            lds_capacity=4096)):
                    try:
                        res.append(build())
                    except (ValueError, MemoryError) as exc:
                        res.append(type(exc))
                py, nat = res
                if isinstance(py, type) or isinstance(nat, type):
                    assert py == nat
                    continue
                assert np.array_equal(py.words, nat.words)
                for f in fields:
                    assert getattr(py, f) == getattr(nat, f), f
                n_cmp += 1
    assert n_cmp >= 30


def test_native_planner_rejects_bad_input():
    from qtensor_amd import native_planner as npl
    with pytest.raises(ValueError):
        npl.compile_lightcone(np.array([3, 1, 0, 5, 0, 1], np.int32), 1, QtreeQAOAComposer)   # j outside
    with pytest.raises(ValueError):
        npl.compile_lightcone(npl.lightcone_record(2, [(0, 1)], 0, 1), 1, QtreeQAOAComposer, lanes=96)
    # a width no slicing can reach below: "Estimated OOM", as the reference's slicer
    with pytest.raises(MemoryError):
        npl.compile_lightcone(npl.lightcone_record(4, [(0, 1), (1, 2), (2, 3), (0, 3)], 0, 1), 2,
                              QtreeQAOAComposer, slicer="greedy", max_tw=0, lds_capacity=None)


def test_greedy_from_buckets_equals_line_graph_greedy():
    """The optimizer's bucket-list path (no networkx graph) gives the order,
    treewidth and (lazily built) line graph of the graph path, on lightcones of
    three composers."""
    import networkx as nx
    from qtensor_amd.composer import DefaultQAOAComposer, QtreeQAOAComposer, ZZQtreeQAOAComposer
    from qtensor_amd.ordering import OrderingOptimizer, greedy_peo
    from qtensor_amd.tensornet import QtreeTensorNet
    for seed in range(6):
        G = nx.random_regular_graph(3, 24 + 2 * seed, seed=seed)
        for Comp in (QtreeQAOAComposer, ZZQtreeQAOAComposer, DefaultQAOAComposer):
            for e in list(G.edges())[:2]:
                comp = Comp(G, gamma=[0.1, 0.2], beta=[0.3, 0.4])
                comp.energy_expectation_lightcone(e)
                tn = QtreeTensorNet.from_qtree_gates(comp.circuit)
                opt = OrderingOptimizer()
                peo, _ = opt.optimize(tn)
                g = tn.get_line_graph()
                ref, path = greedy_peo(g)
                assert [int(v) for v in peo][len(tn.ket_vars) + len(tn.bra_vars):] == ref
                assert opt.treewidth == max(path)
                assert list(opt.graph.nodes) == list(g.nodes)
                assert {frozenset(x) for x in opt.graph.edges} == {frozenset(x) for x in g.edges}


@pytest.mark.parametrize("fuse", [False, True], ids=["steps", "fused"])
def test_wide_plan_open_variables_vs_statevector(fuse):
    """Wide plans with b open output variables (simulate_batch(batch_vars=b),
    qtensor/Simulate.py:66-94): the interpreted program gives the 2^b
    amplitudes of the statevector (qubits 0..b-1 most significant, the rest
    0), with angles from the parameter vector or baked in, and sliced plans
    sum to the same amplitudes."""
    from plan_interp import run_plan
    from qtensor_amd.tensornet import QtreeTensorNet
    from qtensor_amd.wide import compile_wide, slice_vars_for_order
    from qtensor_amd.ordering import greedy_peo
    for seed, b in ((3, 1), (4, 2), (5, 3)):
        G = nx.random_regular_graph(3, 10, seed=seed)
        gamma, beta = [0.3, 0.5], [0.2, 0.7]
        comp = QtreeQAOAComposer(G, gamma=gamma, beta=beta)
        comp.ansatz_state()
        nq = G.number_of_nodes()
        sv = statevector.qaoa_state(nq, [tuple(e) for e in G.edges()], gamma, beta)
        ref = np.array([sv[k << (nq - b)] for k in range(1 << b)])
        prm = np.array(gamma + beta + [0.0] * 8)
        for bake in (False, True):
            wp = compile_wide(comp.circuit, batch_vars=b, fuse=fuse, bake_angles=bake)
            assert wp.plan.words[11] == b
            got = run_plan(wp.plan.words, np.zeros(8) if bake else prm)
            assert np.allclose(got, ref, atol=1e-12)
        graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
        order, _ = greedy_peo(graph)
        par = slice_vars_for_order(graph, order, 2)
        ws = compile_wide(comp.circuit, order=order, par_ids=par, batch_vars=b, fuse=fuse)
        got = sum(run_plan(ws.plan.words, prm, slice_bits=s) for s in range(ws.n_slices))
        assert ws.n_slices == 4 and np.allclose(got, ref, atol=1e-12)


def test_small_inputs_multiplied_first_keeps_the_value(monkeypatch):
    """plan.split_small_inputs: in a wide p=2 lightcone plan under a wide order,
    steps of three or more inputs get their small inputs multiplied first (the
    pass fires), and the interpreted program still gives the statevector value;
    the same plan without the pass agrees."""
    from qtensor_amd import plan as P
    from qtensor_amd import wide as W
    from qtensor_amd.tensornet import QtreeTensorNet
    counts = []
    real = P.split_small_inputs

    def counting(steps):
        out = real(steps)
        counts.append(len(out) - len(steps))
        return out

    G = nx.random_regular_graph(3, 16, seed=4)
    g, b = [0.3, 0.7], [0.5, 0.1]
    e = list(G.edges())[0]
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    ref = statevector.zz_expectations(G, g, b)[e]
    graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    order = W.forced_width_order(graph, target_tw=16, temp=3.0, seed=0)      # greedy: width 10
    monkeypatch.setattr(P, "split_small_inputs", counting)
    with_pass = W.compile_wide(comp.circuit, order=order)
    assert counts and counts[-1] > 0
    monkeypatch.setattr(P, "split_small_inputs", lambda steps: steps)
    without = W.compile_wide(comp.circuit, order=order)
    assert with_pass.plan.n_ops > without.plan.n_ops
    for wp in (with_pass, without):
        assert abs(run_plan(wp.plan.words, g + b, lanes=256) - ref) < 1e-10
