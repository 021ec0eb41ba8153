"""Wide plans (global-memory tensors, one grid-wide step kernel per step)."""
import ctypes

import networkx as nx
import numpy as np
import pytest

from oracle import statevector
from qtensor_amd import _lib
from qtensor_amd.composer import QtreeQAOAComposer
from qtensor_amd.tensornet import QtreeTensorNet
from qtensor_amd.wide import WideEngine, compile_wide, forced_width_order, slice_vars_for_order

pytestmark = pytest.mark.gpu


def _edge_circuit(n=14, seed=2, p=2):
    G = nx.random_regular_graph(3, n, seed=seed)
    rng = np.random.default_rng(seed)
    g, b = list(rng.uniform(0, np.pi, p)), list(rng.uniform(0, np.pi / 2, p))
    e = list(G.edges())[1]
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    return G, e, g, b, comp.circuit


@pytest.mark.parametrize("fuse", [False, True], ids=["steps", "fused"])
@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-5), ("complex128", 1e-11)])
def test_wide_greedy_and_forced_sliced(gpu, dtype, tol, fuse):
    G, e, g, b, circ = _edge_circuit()
    ref = statevector.zz_expectations(G, g, b)[e]
    eng = WideEngine(dtype=dtype)
    assert abs(eng.run(compile_wide(circ, fuse=fuse), g + b) - ref) <= tol
    graph = QtreeTensorNet.from_qtree_gates(circ).get_line_graph()
    order = forced_width_order(graph, target_tw=14, temp=3.0, seed=1)
    wp = compile_wide(circ, order=order, par_ids=slice_vars_for_order(graph, order, 3), fuse=fuse)
    assert wp.n_slices == 8
    assert abs(eng.run(wp, g + b) - ref) <= tol


def test_fused_large_rank_matches_steps(gpu):
    """A p=3 lightcone under a wide order: fused plans (outputs of rank >= 14,
    several blocks per output, high output bits computed on the device) give
    the step-by-step value in complex128."""
    G = nx.random_regular_graph(3, 64, seed=3)
    rng = np.random.default_rng(3)
    g, b = list(rng.uniform(0, np.pi, 3)), list(rng.uniform(0, np.pi / 2, 3))
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(list(G.edges())[0])
    graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    order = forced_width_order(graph, target_tw=20, temp=2.0, seed=0)
    fused = compile_wide(comp.circuit, order=order, fuse=True)
    plain = compile_wide(comp.circuit, order=order, fuse=False)
    assert fused.plan.n_fused > 0 and fused.plan.n_ops < plain.plan.n_ops
    eng = WideEngine(dtype="complex128")
    assert abs(eng.run(fused, g + b) - eng.run(plain, g + b)) < 1e-12


def test_wide_rejects_inconsistent_fused_table(gpu):
    """The host recomputes every fused offset table from the masks before any
    launch: a plan whose table disagrees is refused."""
    import torch
    from qtensor_amd import plan as P
    G, e, g, b, circ = _edge_circuit()
    wp = compile_wide(circ, fuse=True)
    words = np.ascontiguousarray(wp.plan.words, dtype=np.int32).copy()
    pos, tampered = int(words[3]), False
    for _ in range(int(words[1])):
        w0 = int(words[pos]) & 0xFFFFFFFF
        if w0 & P.OP_FUSED:
            n_in = w0 & 15
            words[pos + P.FUSED_HDR + 4 * n_in] += 1          # first output-table entry
            tampered = True
            break
        pos += w0 >> 16
    assert tampered
    lib = _lib.init(0)
    dev = torch.from_numpy(words).cuda()
    arena = torch.empty(max(wp.plan.glb_need, 1), dtype=torch.complex64, device="cuda")
    prm = torch.tensor(g + b, dtype=torch.float64, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    d = _lib.WideDesc()
    d.plan_host, d.plan_words, d.plan_dev = words.ctypes.data, words.size, dev.data_ptr()
    d.params, d.arena, d.arena_elems = prm.data_ptr(), arena.data_ptr(), arena.numel()
    d.slice_bits, d.dtype, d.out = 0, _lib.QB_C64, out.data_ptr()
    rc = lib.qb_wide_execute(ctypes.byref(d), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.QB_ERR_INVALID and b"table" in lib.qb_last_error()


def test_wide_rejects_arena_overflow(gpu):
    import torch
    G, e, g, b, circ = _edge_circuit()
    wp = compile_wide(circ)
    lib = _lib.init(0)
    words = np.ascontiguousarray(wp.plan.words, dtype=np.int32)
    dev = torch.from_numpy(words).cuda()
    arena = torch.empty(8, dtype=torch.complex64, device="cuda")      # far too small
    prm = torch.tensor(g + b, dtype=torch.float64, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    d = _lib.WideDesc()
    d.plan_host, d.plan_words, d.plan_dev = words.ctypes.data, words.size, dev.data_ptr()
    d.params, d.arena, d.arena_elems = prm.data_ptr(), arena.data_ptr(), arena.numel()
    d.slice_bits, d.dtype, d.out = 0, _lib.QB_C64, out.data_ptr()
    rc = lib.qb_wide_execute(ctypes.byref(d), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.QB_ERR_INVALID and b"arena" in lib.qb_last_error()


@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-5), ("complex128", 1e-11)])
def test_wide_simulator_amplitude_batches(gpu, dtype, tol):
    """WideQtreeSimulator.simulate_batch(batch_vars=b) (qtensor/Simulate.py:66-94)
    on the device: the 2^b amplitudes of the statevector, and a sliced plan with
    open variables (4 slices on 4 streams, element-wise compensated sums)."""
    from oracle import statevector
    from qtensor_amd.simulator import WideQtreeSimulator
    from qtensor_amd.ordering import greedy_peo
    from qtensor_amd.tensornet import QtreeTensorNet
    from qtensor_amd.wide import WideEngine, compile_wide, slice_vars_for_order
    for seed, b in ((3, 0), (4, 2), (5, 4)):
        G = nx.random_regular_graph(3, 12, seed=seed)
        gamma, beta = [0.3, 0.5], [0.2, 0.7]
        comp = QtreeQAOAComposer(G, gamma=gamma, beta=beta)
        comp.ansatz_state()
        nq = G.number_of_nodes()
        sv = statevector.qaoa_state(nq, [tuple(e) for e in G.edges()], gamma, beta)
        ref = np.array([sv[k << (nq - b)] for k in range(1 << b)])
        got = WideQtreeSimulator(dtype=dtype).simulate_batch(comp.circuit, batch_vars=b)
        assert got.shape == ref.shape and np.max(np.abs(got - ref)) <= tol
        graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
        order, _ = greedy_peo(graph)
        ws = compile_wide(comp.circuit, order=order, par_ids=slice_vars_for_order(graph, order, 2), batch_vars=b)
        val = np.atleast_1d(WideEngine(dtype=dtype).run(ws, gamma + beta))
        assert np.max(np.abs(val - ref)) <= tol


def test_wide_graph_replays_match_eager(gpu):
    """WideEngine.run(graph=True): the slices' launches captured once per
    stream and replayed with new angles.  Every replay, with eager device work
    and synchronisations between replays, equals the eager evaluation of the
    same angles, and the first matches the statevector."""
    import torch
    G, e, g, b, circ = _edge_circuit()
    graph = QtreeTensorNet.from_qtree_gates(circ).get_line_graph()
    order = forced_width_order(graph, target_tw=14, temp=3.0, seed=1)
    wp = compile_wide(circ, order=order, par_ids=slice_vars_for_order(graph, order, 3))
    eng = WideEngine(dtype="complex128")
    first = eng.run(wp, g + b, graph=True)
    assert abs(first - statevector.zz_expectations(G, g, b)[e]) <= 1e-11
    rng = np.random.default_rng(11)
    for _ in range(4):
        prm = list(rng.uniform(0, np.pi, 2)) + list(rng.uniform(0, np.pi / 2, 2))
        eager = eng.run(wp, prm)
        x = torch.randn(1 << 16, device="cuda")
        (x * x).sum().item()                   # unrelated device work between replays
        torch.cuda.synchronize()
        replay = eng.run(wp, prm, graph=True)
        assert abs(replay - eager) <= 1e-13, (replay, eager)


@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-5), ("complex128", 1e-11)])
def test_wide_run_many_matches_run(gpu, dtype, tol):
    """Several wide plans at once (their (plan, slice) networks spread over the
    streams) give each plan's value: equal to running the plans one by one
    (the same slices, reduced in the same order) and to the statevector."""
    G = nx.random_regular_graph(3, 14, seed=2)
    rng = np.random.default_rng(7)
    g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
    sv = statevector.zz_expectations(G, g, b)
    plans, refs = [], []
    for k, e in enumerate(list(G.edges())[:6]):
        comp = QtreeQAOAComposer(G, gamma=g, beta=b)
        comp.energy_expectation_lightcone(e)
        if k % 2:
            graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
            order = forced_width_order(graph, target_tw=12, temp=3.0, seed=k)
            plans.append(compile_wide(comp.circuit, order=order, par_ids=slice_vars_for_order(graph, order, 2)))
        else:
            plans.append(compile_wide(comp.circuit))
        refs.append(sv[e])
    assert any(wp.n_slices > 1 for wp in plans)
    eng = WideEngine(dtype=dtype)
    vals, dt = eng.run_many(plans, g + b, timing=True)
    assert dt > 0 and len(vals) == len(plans)
    for wp, v, r in zip(plans, vals, refs):
        assert v == eng.run(wp, g + b)
        assert abs(v - r) <= tol


def _chain_count(wp):
    from qtensor_amd import plan as P
    w = np.asarray(wp.plan.words)
    pos, n = int(w[3]), 0
    for _ in range(int(w[1])):
        w0 = int(w[pos]) & 0xFFFFFFFF
        if (w0 & P.OP_FUSED) and (w0 & P.OP_CHAIN):
            n += 1
        pos += w0 >> 16
    return n


def _forced_p3_plans(tw=20, n_slice=0):
    """A p=3 lightcone of N=64 under a deliberately wide order, compiled with
    and without chain ops."""
    G = nx.random_regular_graph(3, 64, seed=3)
    rng = np.random.default_rng(3)
    g, b = list(rng.uniform(0, np.pi, 3)), list(rng.uniform(0, np.pi / 2, 3))
    e = list(G.edges())[0]
    comp = QtreeQAOAComposer(G, gamma=g, beta=b)
    comp.energy_expectation_lightcone(e)
    graph = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    order = forced_width_order(graph, target_tw=tw, temp=2.0, seed=0)
    par = slice_vars_for_order(graph, order, n_slice) if n_slice else []
    chained = compile_wide(comp.circuit, order=order, par_ids=par)
    plain = compile_wide(comp.circuit, order=order, par_ids=par, chain=False)
    return G, e, g, b, chained, plain


@pytest.mark.parametrize("n_slice", [0, 2])
def test_chain_ops_match_separate_contractions(gpu, n_slice):
    """Chain ops (a contraction and its consumer in one kernel, the
    intermediate never stored) give the value of the same plan compiled
    without them: complex128 to 1e-12, and both precisions match the oracle's
    contraction of the edge's network."""
    from oracle import qaoa_network
    G, e, g, b, chained, plain = _forced_p3_plans(n_slice=n_slice)
    assert _chain_count(chained) > 0 and _chain_count(plain) == 0
    e128 = WideEngine(dtype="complex128")
    v_chain, v_plain = e128.run(chained, g + b), e128.run(plain, g + b)
    assert abs(v_chain - v_plain) < 1e-12
    ref = qaoa_network.edge_zz(G, e, g, b)
    assert abs(v_chain - ref) < 1e-11
    assert abs(WideEngine(dtype="complex64").run(chained, g + b) - ref) < 1e-5


@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-6), ("complex128", 1e-13)])
def test_slices_per_launch_match_one_slice_per_launch(gpu, dtype, tol):
    """Slices run in lockstep (one launch per op covering 2, 4 or all 8 slices,
    their arenas side by side), eagerly and replayed from captured graphs,
    equal one launch per slice; also for a plan with open variables (each
    slice's amplitudes at their own stride) and for several plans at once."""
    G, e, g, b, circ = _edge_circuit()
    graph = QtreeTensorNet.from_qtree_gates(circ).get_line_graph()
    order = forced_width_order(graph, target_tw=14, temp=3.0, seed=1)
    wp = compile_wide(circ, order=order, par_ids=slice_vars_for_order(graph, order, 3))
    assert wp.n_slices == 8
    eng = WideEngine(dtype=dtype)
    base = eng.run(wp, g + b, n_streams=4, slices_per_launch=1)
    assert abs(base - statevector.zz_expectations(G, g, b)[e]) <= max(tol, 1e-11) * 10
    for ns, per in ((1, 8), (2, 4), (4, 2), (1, 3)):
        assert abs(eng.run(wp, g + b, n_streams=ns, slices_per_launch=per) - base) <= tol
        assert abs(eng.run(wp, g + b, n_streams=ns, slices_per_launch=per, graph=True) - base) <= tol
    from qtensor_amd.ordering import greedy_peo
    G2 = nx.random_regular_graph(3, 12, seed=4)
    comp = QtreeQAOAComposer(G2, gamma=[0.3, 0.5], beta=[0.2, 0.7])
    comp.ansatz_state()
    graph2 = QtreeTensorNet.from_qtree_gates(comp.circuit).get_line_graph()
    order2, _ = greedy_peo(graph2)
    ws = compile_wide(comp.circuit, order=order2, par_ids=slice_vars_for_order(graph2, order2, 2), batch_vars=2)
    amp1 = np.atleast_1d(eng.run(ws, [0.3, 0.5, 0.2, 0.7], slices_per_launch=1))
    amp4 = np.atleast_1d(eng.run(ws, [0.3, 0.5, 0.2, 0.7], n_streams=1, slices_per_launch=4))
    assert amp1.shape == (4,) and np.max(np.abs(amp4 - amp1)) <= tol
    many1 = eng.run_many([wp, wp], g + b, slices_per_launch=1)
    many4 = eng.run_many([wp, wp], g + b, slices_per_launch=4)
    assert max(abs(x - base) for x in many1 + many4) <= tol


def test_wide_rejects_malformed_chain_op(gpu):
    """The chain flag on a fused op without a third input is refused on the host."""
    import torch
    from qtensor_amd import plan as P
    G, e, g, b, chained, plain = _forced_p3_plans()
    words = np.ascontiguousarray(plain.plan.words, dtype=np.int32).copy()
    pos, tampered = int(words[3]), False
    for _ in range(int(words[1])):
        w0 = int(words[pos]) & 0xFFFFFFFF
        if (w0 & P.OP_FUSED) and (w0 & 15) == 2:
            words[pos] = np.int32(np.uint32(w0 | P.OP_CHAIN).view(np.int32))
            tampered = True
            break
        pos += w0 >> 16
    assert tampered
    lib = _lib.init(0)
    dev = torch.from_numpy(words).cuda()
    arena = torch.empty(max(plain.plan.glb_need, 1), dtype=torch.complex64, device="cuda")
    prm = torch.tensor(g + b, dtype=torch.float64, device="cuda")
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    d = _lib.WideDesc()
    d.plan_host, d.plan_words, d.plan_dev = words.ctypes.data, words.size, dev.data_ptr()
    d.params, d.arena, d.arena_elems = prm.data_ptr(), arena.data_ptr(), arena.numel()
    d.slice_bits, d.dtype, d.out = 0, _lib.QB_C64, out.data_ptr()
    rc = lib.qb_wide_execute(ctypes.byref(d), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rc == _lib.QB_ERR_INVALID and b"chain op" in lib.qb_last_error()
