"""qb_execute_noalloc: the executor entry point that allocates nothing.

Its work-queue ticket and global scratch slots live in the library's static
device workspace.  Every test calls it through the C ABI and compares its item
results with qb_execute's (same kernels, caller workspace) bit for bit, on
consecutive calls whose inputs shrink, with and without global scratch, and
with more items than resident blocks (queue tickets).
"""
import ctypes

import networkx as nx
import numpy as np
import pytest

from qtensor_amd import _lib
from qtensor_amd.batched import LaunchShape, LightconeEngine

pytestmark = pytest.mark.gpu


def _batch(G, p, shape, seed, max_tw=None):
    rng = np.random.default_rng(seed)
    g, b = list(rng.uniform(0, np.pi, p)), list(rng.uniform(0, np.pi / 2, p))
    eng = LightconeEngine(shape=shape, max_tw=max_tw, slicer="tree_trim")
    return eng.batch([(eng.workload(G, p), g, b)])


def _noalloc_items(batch):
    """Item results of qb_execute_noalloc for ``batch`` (fresh output buffer)."""
    import torch
    out = torch.full_like(batch.item_out, float("nan"))
    d = batch._make_desc()
    d.scratch = None
    d.n_slots = 0
    d.queue = None
    d.item_out = out.data_ptr()
    st = torch.cuda.current_stream()
    _lib.check(batch.lib.qb_execute_noalloc(ctypes.byref(d), ctypes.c_void_p(st.cuda_stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _reference_items(batch):
    import torch
    batch.item_out.fill_(float("nan"))
    batch.launch_exec()
    torch.cuda.synchronize()
    return batch.item_out.cpu().numpy()


def test_noalloc_matches_execute_on_shrinking_calls(gpu):
    """Two calls in a row, the second with fewer items; then a batch whose
    plans keep intermediates in global scratch; each equals qb_execute."""
    big = _batch(nx.random_regular_graph(3, 60, seed=1), 2, LaunchShape(64, 4096, 2048), seed=1)
    small = _batch(nx.random_regular_graph(3, 12, seed=2), 1, LaunchShape(64, 4096, 2048), seed=2)
    assert small.n_items < big.n_items
    for b in (big, small):
        got = _noalloc_items(b)
        ref = _reference_items(b)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    glb = _batch(nx.random_regular_graph(3, 30, seed=3), 2, LaunchShape(64, 4096, 16), seed=3)
    assert glb.scratch_elems > 0, "this batch should keep intermediates in global scratch"
    assert glb.scratch_elems * 8 * glb.n_items < _lib.QB_NOALLOC_SCRATCH_BYTES
    for _ in range(2):
        got = _noalloc_items(glb)
        assert np.array_equal(got, _reference_items(glb))
    # the batch's terms are the reduced item results: check one energy as well
    glb.run()
    terms = np.array(glb.terms()[0])
    assert np.all(np.abs(terms.imag) < 1e-5)


def test_noalloc_pairs_with_scratch(gpu):
    """Pairs of items per block (interleaved LDS data and scratch slots): a
    plan set with intermediates in global scratch, then a smaller one, each
    equal to qb_execute's pairs."""
    for n, seed in ((30, 5), (12, 6)):
        b = _batch(nx.random_regular_graph(3, n, seed=seed), 2, LaunchShape(128, 4096, 16, group=2), seed=seed)
        assert b.group == 2 and b.scratch_elems > 0
        got = _noalloc_items(b)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, _reference_items(b))


def test_noalloc_with_queue_tickets(gpu):
    """More items than resident blocks: later items come from the static
    queue ticket, which the call resets with a fill kernel."""
    shape = LaunchShape(512, 4096, 2048)
    b = _batch(nx.random_regular_graph(3, 1000, seed=4), 1, shape, seed=4)
    assert b.n_items > b.n_slots, (b.n_items, b.n_slots)
    for _ in range(2):
        assert np.array_equal(_noalloc_items(b), _reference_items(b))


def test_noalloc_rejects_oversize_and_workspace_arguments(gpu):
    import torch
    b = _batch(nx.random_regular_graph(3, 12, seed=5), 1, LaunchShape(64, 4096, 2048), seed=5)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    d = b._make_desc()        # carries the caller workspace of qb_execute
    with pytest.raises(ValueError):
        _lib.check(b.lib.qb_execute_noalloc(ctypes.byref(d), st))
    d.scratch = None
    d.n_slots = 0
    d.queue = None
    d.scratch_elems = (_lib.QB_NOALLOC_SCRATCH_BYTES // 8)       # one slot alone fills the workspace
    rc = b.lib.qb_execute_noalloc(ctypes.byref(d), st)
    assert rc == _lib.QB_ERR_UNSUPPORTED
    assert b"static scratch" in b.lib.qb_last_error()
