"""Batched persistent executor against the oracle and the statevector."""
import networkx as nx
import numpy as np
import pytest

from oracle import closed_form, statevector
from oracle import np_framework as npf
from qtensor_amd.batched import LaunchShape, LightconeEngine
from qtensor_amd.composer import QtreeQAOAComposer, SimpZZQtreeComposer, ZZQtreeQAOAComposer
from qtensor_amd.simulator import QAOAQtreeSimulator

pytestmark = pytest.mark.gpu


def _tol(ref):
    return 1e-5 * max(1.0, abs(ref))


def test_config1_batched(gpu):
    G = nx.random_regular_graph(3, 10, seed=42)
    g, b = [np.pi / 3], [np.pi / 2]
    C, zz = LightconeEngine().energy(G, g, b)
    for e, z in zip(G.edges(), zz):
        ref = closed_form.zz_p1(G, e, g[0], b[0])
        assert abs(z - ref) <= _tol(ref)
    assert abs(C - statevector.maxcut_energy(G, g, b)) < 1e-4


SHAPES = [LaunchShape(64, 5120, 2048), LaunchShape(64, 5120, 8), LaunchShape(256, 10240, 4096),
          LaunchShape(128, 4096, 64), LaunchShape(512, 16384, 16384)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"t{s.item_threads}_c{s.lds_capacity}_l{s.lds_tensor_limit}")
@pytest.mark.parametrize("Comp", [QtreeQAOAComposer, SimpZZQtreeComposer])
def test_p2_vs_statevector(gpu, shape, Comp):
    G = nx.random_regular_graph(3, 16, seed=3)
    rng = np.random.default_rng(5)
    g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
    C, zz = LightconeEngine(composer=Comp, shape=shape).energy(G, g, b)
    sv = statevector.zz_expectations(G, g, b)
    for e, z in zip(G.edges(), zz):
        assert abs(z - sv[e]) <= _tol(sv[e])


@pytest.mark.parametrize("dtype,tol", [("complex64", 1e-5), ("complex128", 1e-11)])
def test_p3_vs_oracle(gpu, dtype, tol):
    G = nx.random_regular_graph(3, 64, seed=1)
    rng = np.random.default_rng(2)
    g, b = list(rng.uniform(0, np.pi, 3)), list(rng.uniform(0, np.pi / 2, 3))
    edges = list(G.edges())[:6]
    eng = LightconeEngine(dtype=dtype, shape=LaunchShape(256, 8192, 1024))
    _, zz = eng.energy(G, g, b, edges=edges)
    sim = QAOAQtreeSimulator(QtreeQAOAComposer, bucket_backend=npf.NumpyBackend())
    for e, z in zip(edges, zz):
        ref = sim._get_edge_energy(G, g, b, e)[0]
        assert abs(z - ref) <= tol * max(1.0, abs(ref))


@pytest.mark.parametrize("slicer,ordering", [("greedy", "greedy"), ("tree_trim", "greedy"),
                                             ("greedy", "rgreedy"), ("tree_trim", "rgreedy")])
def test_sliced_items_sum_to_unsliced(gpu, slicer, ordering):
    """Sliced work items (max-degree or tree-trimming slice variables, under
    the greedy or the randomized-greedy order) sum to the unsliced greedy-order
    terms (the sliced == unsliced invariant of tests/test_qaoa_energy.py:106-121;
    orders change the work, not the values)."""
    G = nx.random_regular_graph(3, 40, seed=7)
    g, b = [0.3, 0.7, 0.2], [0.5, 0.1, 0.9]
    edges = list(G.edges())[:4]
    shape = LaunchShape(256, 8192, 512)
    _, ref = LightconeEngine(shape=shape, dtype="complex128").energy(G, g, b, edges=edges)
    eng = LightconeEngine(shape=shape, max_tw=9, dtype="complex128", slicer=slicer, ordering=ordering)
    wl = eng.workload(G, 3, edges)
    assert any(w.n_slices > 1 for w in wl.works)
    batch = eng.batch([(wl, g, b)])
    batch.run()
    for z, r in zip(batch.terms()[0], ref):
        assert abs(z - r) < 1e-11


def test_reparameterise_and_graph_replay(gpu):
    import torch
    G = nx.random_regular_graph(3, 24, seed=9)
    eng = LightconeEngine()
    wl = eng.workload(G, 2)
    g0, b0 = [0.1, 0.2], [0.3, 0.4]
    batch = eng.batch([(wl, g0, b0)])
    batch.capture()
    for k in range(3):
        rng = np.random.default_rng(k)
        g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
        batch.set_params([(g, b)])
        batch.run()
        sv = statevector.zz_expectations(G, g, b) if G.number_of_nodes() <= 20 else None
        _, eager = LightconeEngine().energy(G, g, b)
        torch.cuda.synchronize()
        for z, r in zip(batch.terms()[0], eager):
            assert abs(z - r) < 1e-6
        if sv:
            for e, z in zip(G.edges(), batch.terms()[0]):
                assert abs(z - sv[e]) < 1e-5


def test_mixed_requests_match_alone(gpu):
    """Each request's terms are the same alone and inside a mixed batch."""
    eng = LightconeEngine()
    reqs = []
    for k, (n, seed) in enumerate([(12, 1), (20, 2), (16, 3)]):
        G = nx.random_regular_graph(3, n, seed=seed)
        rng = np.random.default_rng(10 + k)
        reqs.append((eng.workload(G, 2), list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))))
    mixed = eng.batch(reqs)
    mixed.run()
    for r, terms in zip(reqs, mixed.terms()):
        alone = eng.batch([r])
        alone.run()
        for a, m in zip(alone.terms()[0], terms):
            assert abs(a - m) <= 1e-12


@pytest.mark.parametrize("shape,max_tw", [(LaunchShape(128, 4096, 2048), 7), (LaunchShape(64, 2048, 64), None)],
                         ids=["sliced_lds", "global_scratch"])
def test_plan_handle_matches_batch(gpu, shape, max_tw):
    """qb_plan_create / qb_plan_execute (library-owned buffers) give the batch's
    terms bit for bit, for the batch's angles and for new ones, and agree with
    the statevector; the first case has more items than one row of blocks
    needs, the second spills tensors to global scratch."""
    import torch
    from qtensor_amd.batched import PlanHandle
    G = nx.random_regular_graph(3, 14, seed=9)
    rng = np.random.default_rng(9)
    g, b = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
    eng = LightconeEngine(shape=shape, max_tw=max_tw)
    wl = eng.workload(G, 2)
    batch = eng.batch([(wl, g, b)])
    batch.run()
    ref = batch.terms()[0]
    h = PlanHandle(batch)
    for rep in range(2):
        terms, total = h.execute(batch.params)
        assert terms == ref
        assert abs(total - sum(ref)) < 1e-12
    g2, b2 = list(rng.uniform(0, np.pi, 2)), list(rng.uniform(0, np.pi / 2, 2))
    prm = torch.tensor(g2 + b2 + [0.0] * 8, dtype=torch.float64, device="cuda")
    terms2, _ = h.execute(prm)
    sv = statevector.zz_expectations(G, g2, b2)
    for e, z in zip(G.edges(), terms2):
        assert abs(z - sv[e]) <= _tol(sv[e])
    h.close()
    h.close()


@pytest.mark.parametrize("n_seg,max_len", [(1, 1), (384, 12), (1500, 3), (3000, 40), (4096, 5), (4097, 1),
                                           (9000, 2)])
def test_reduce_terms_vs_fsum(gpu, n_seg, max_len):
    """qb_reduce_terms (compensated FP64 sums of item results per term, then of
    the terms) against math.fsum, for segment counts from 1 to 9000 and
    segments of up to 40 items, with values spanning ten decades."""
    import ctypes
    import math

    import torch

    from qtensor_amd import _lib
    lib = _lib.init(0)
    rng = np.random.default_rng(n_seg)
    lens = rng.integers(1, max_len + 1, n_seg)
    seg = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    vals = rng.standard_normal((int(seg[-1]), 2)) * 10.0 ** rng.integers(-8, 3, (int(seg[-1]), 1))
    d_items = torch.tensor(vals, dtype=torch.float64, device=gpu)
    d_seg = torch.tensor(seg, dtype=torch.int32, device=gpu)
    d_terms = torch.zeros((n_seg, 2), dtype=torch.float64, device=gpu)
    d_total = torch.zeros(2, dtype=torch.float64, device=gpu)
    _lib.check(lib.qb_reduce_terms(ctypes.c_void_p(d_items.data_ptr()), ctypes.c_void_p(d_seg.data_ptr()),
                                   n_seg, ctypes.c_void_p(d_terms.data_ptr()),
                                   ctypes.c_void_p(d_total.data_ptr()), None))
    torch.cuda.synchronize()
    terms = d_terms.cpu().numpy()
    for c in range(2):
        ref = np.array([math.fsum(vals[seg[s]:seg[s + 1], c]) for s in range(n_seg)])
        assert np.max(np.abs(terms[:, c] - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
        tot = math.fsum(vals[:, c])
        assert abs(d_total[c].item() - tot) <= 1e-12 * max(1.0, np.sum(np.abs(vals[:, c])))


@pytest.mark.parametrize("n", [1, 1500, 5000])
def test_reduce_terms_one_item_per_term(gpu, n):
    """qb_reduce_terms without a segment table (item s is term s): the terms
    are the item values unchanged, the total matches math.fsum."""
    import ctypes
    import math

    import torch

    from qtensor_amd import _lib
    lib = _lib.init(0)
    rng = np.random.default_rng(n)
    vals = rng.standard_normal((n, 2)) * 10.0 ** rng.integers(-8, 3, (n, 1))
    d_items = torch.tensor(vals, dtype=torch.float64, device=gpu)
    d_terms = torch.zeros((n, 2), dtype=torch.float64, device=gpu)
    d_total = torch.zeros(2, dtype=torch.float64, device=gpu)
    _lib.check(lib.qb_reduce_terms(ctypes.c_void_p(d_items.data_ptr()), None, n,
                                   ctypes.c_void_p(d_terms.data_ptr()), ctypes.c_void_p(d_total.data_ptr()), None))
    torch.cuda.synchronize()
    assert np.array_equal(d_terms.cpu().numpy(), vals)
    for c in range(2):
        assert abs(d_total[c].item() - math.fsum(vals[:, c])) <= 1e-12 * max(1.0, np.sum(np.abs(vals[:, c])))


def test_autotuner_pick_matches_statevector(gpu):
    """The autotuner's staged search on a small p=2 workload returns a
    configuration whose energy terms match the statevector."""
    from qtensor_amd.autotune import Autotuner
    G = nx.random_regular_graph(3, 14, seed=11)
    g, b = [0.4, 0.9], [0.3, 0.6]
    shape, max_tw, slicer, table = Autotuner(reps=3, ordering="rgreedy").tune(G, 2, g, b)
    assert table and any("eval_ms" in row for row in table)
    C, zz = LightconeEngine(shape=shape, max_tw=max_tw, slicer=slicer, ordering="rgreedy").energy(G, g, b)
    sv = statevector.zz_expectations(G, g, b)
    for e, z in zip(G.edges(), zz):
        assert abs(z - sv[e]) <= _tol(sv[e])


def test_library_loaded_before_torch_touches_the_device(gpu):
    """A fresh process that plans (loading the library) before torch has used
    the GPU still finds the device: the library binds to torch's HIP runtime."""
    import os
    import subprocess
    import sys
    code = ("import networkx as nx\n"
            "from qtensor_amd.batched import LightconeEngine\n"
            "eng = LightconeEngine()\n"
            "G = nx.random_regular_graph(3, 10, seed=42)\n"
            "wl = eng.workload(G, 1)\n"
            "C, zz = eng.energy(G, [0.3], [0.2])\n"
            "print(len(zz))\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr[-2000:]
    assert res.stdout.strip().splitlines()[-1] == "15"
