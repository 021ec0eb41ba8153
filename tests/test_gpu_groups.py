"""Executor blocks that run two work items of one plan in lockstep (``group``
= 2, LDS data interleaved) against single-item blocks: every item performs the
same operations in the same order either way, so item values and terms must be
bit-identical.  Covers 64-, 128- and 256-thread items, plans with intermediates
in global scratch, slicing, mixed requests (items of one plan with different
angles share a block), padding (plans with an odd item count), the C-ABI plan
handle and the statevector."""
import networkx as nx
import numpy as np
import pytest

from oracle import statevector
from qtensor_amd.batched import LaunchShape, LightconeEngine, PlanHandle

pytestmark = pytest.mark.gpu

SHAPES = [LaunchShape(64, 4096, 2048), LaunchShape(128, 4096, 64), LaunchShape(256, 8192, 4096),
          LaunchShape(256, 4096, 16)]


def _pair(shape):
    return LaunchShape(shape.item_threads, shape.lds_capacity, shape.lds_tensor_limit, group=2)


def _requests(eng, p=3, seed=1, n=40, edges=8, sets=2):
    G = nx.random_regular_graph(3, n, seed=seed)
    wl = eng.workload(G, p, list(G.edges())[:edges])
    rng = np.random.default_rng(seed)
    return [(wl, list(rng.uniform(0, np.pi, p)), list(rng.uniform(0, np.pi / 2, p))) for _ in range(sets)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"t{s.item_threads}_c{s.lds_capacity}_l{s.lds_tensor_limit}")
def test_pairs_equal_single_items(gpu, shape):
    kw = dict(max_tw=9, slicer="tree_trim", ordering="rgreedy")
    one = LightconeEngine(shape=shape, **kw)
    two = LightconeEngine(shape=_pair(shape), **kw)
    reqs = _requests(one)
    a = one.batch(reqs)
    a.run()
    b = two.batch(reqs)
    b.run()
    assert b.group == 2 and b.n_items % 2 == 0
    assert b.n_items == a.n_items + b.n_padding
    va, vb = a.item_values(), b.item_values()
    # slots of real items are the same in both launches
    assert np.array_equal(va, vb[:a.n_items])
    for ta, tb in zip(a.terms(), b.terms()):
        assert ta == tb
    if shape.lds_tensor_limit <= 64:
        assert a.scratch_elems > 0       # the global-scratch variant ran


def test_pairs_padding_and_statevector(gpu):
    """An odd number of items per plan: the padding item's value is discarded
    and every term still matches the statevector."""
    G = nx.random_regular_graph(3, 16, seed=3)
    g, b = [0.31, 0.62], [0.44, 0.17]
    eng = LightconeEngine(shape=LaunchShape(64, 4096, 2048, group=2))
    wl = eng.workload(G, 2)
    batch = eng.batch([(wl, g, b)])
    batch.run()
    assert batch.n_padding > 0
    sv = statevector.zz_expectations(G, g, b)
    for e, z in zip(G.edges(), batch.terms()[0]):
        assert abs(z - sv[e]) <= 1e-5 * max(1.0, abs(sv[e]))


def test_pairs_through_plan_handle(gpu):
    """qb_plan_create / qb_plan_execute with groups of two equal the batch."""
    eng = LightconeEngine(shape=LaunchShape(256, 4096, 2048, group=2), max_tw=9, slicer="tree_trim")
    reqs = _requests(eng, sets=1)
    batch = eng.batch(reqs)
    batch.run()
    h = PlanHandle(batch)
    try:
        terms, total = h.execute(batch.params)
    finally:
        h.close()
    assert terms == batch.terms()[0]
