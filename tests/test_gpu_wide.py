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


@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-5), ("complex128", 1e-11)])
def test_wide_greedy_and_forced_sliced(gpu, dtype, tol):
    G, e, g, b, circ = _edge_circuit()
    ref = statevector.zz_expectations(G, g, b)[e]
    eng = WideEngine(dtype=dtype)
    assert abs(eng.run(compile_wide(circ), g + b) - ref) <= tol
    graph = QtreeTensorNet.from_qtree_gates(circ).get_line_graph()
    order = forced_width_order(graph, target_tw=14, temp=3.0, seed=1)
    wp = compile_wide(circ, order=order, par_ids=slice_vars_for_order(graph, order, 3))
    assert wp.n_slices == 8
    assert abs(eng.run(wp, g + b) - ref) <= tol


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
