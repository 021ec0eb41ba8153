"""Plain GPU backend (one qb_step launch per bucket) against the oracle."""
import json
import os

import networkx as nx
import numpy as np
import pytest

from oracle import closed_form, statevector
from oracle import np_framework as npf
from qtensor_amd.backend import HipBackend, PerfBackend
from qtensor_amd.composer import QtreeQAOAComposer, SimpZZQtreeComposer, ZZQtreeQAOAComposer
from qtensor_amd.model import Tensor, Var
from qtensor_amd.simulator import QAOAQtreeSimulator, QtreeSimulator

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


def _to_dev(be, bucket):
    import torch
    return [Tensor(t.name, t.indices, data=torch.as_tensor(np.asarray(t.data), dtype=be.torch_dtype,
                                                              device=be.device)) for t in bucket]


@pytest.mark.parametrize("dtype,tol", [("complex64", 2e-5), ("complex128", 1e-12)])
@pytest.mark.parametrize("seed", range(8))
def test_step_random_buckets(gpu, dtype, tol, seed):
    rng = np.random.default_rng(seed)
    n_vars = int(rng.integers(1, 12))
    sizes = [1 if (rng.random() < 0.2 and i > 0) else 2 for i in range(n_vars)]
    vs = [Var(i, size=sizes[i]) for i in range(n_vars)]
    k = int(rng.integers(1, 7))
    bucket = []
    for j in range(k):
        ids = sorted({0} | set(int(x) for x in rng.choice(n_vars, size=rng.integers(0, n_vars + 1))))
        idx = [vs[i] for i in ids]
        shape = [v.size for v in idx]
        bucket.append(Tensor(f"T{j}", idx, data=rng.normal(size=shape) + 1j * rng.normal(size=shape)))
    be = HipBackend(dtype=dtype)
    be._ensure()
    for no_sum in (False, True):
        got = be.process_bucket(_to_dev(be, bucket), no_sum=no_sum)
        ref = npf.process_bucket(bucket, no_sum=no_sum)
        assert [int(v) for v in got.indices] == [int(v) for v in ref.indices]
        assert got.name == ref.name
        g = be.get_result_data(got)
        r = np.asarray(ref.data)
        assert g.shape == r.shape
        assert np.max(np.abs(g - r)) <= tol * max(1.0, np.max(np.abs(r)))


def test_step_wide_bucket_and_rank(gpu):
    """More tensors than one launch takes, and a rank-22 union."""
    rng = np.random.default_rng(1)
    vs = [Var(i) for i in range(22)]
    bucket = [Tensor("A", vs[:12], data=rng.normal(size=(2,) * 12) + 0j),
              Tensor("B", [vs[0]] + vs[12:], data=rng.normal(size=(2,) * 11) + 0j)]
    for j in range(18):
        a = [vs[0], vs[1 + j]]
        bucket.append(Tensor(f"g{j}", a, data=rng.normal(size=(2, 2)) + 1j))
    be = HipBackend(dtype="complex128")
    be._ensure()
    got = be.get_result_data(be.process_bucket(_to_dev(be, bucket)))
    ref = np.asarray(npf.process_bucket(bucket).data)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11)


def test_f1_f2_through_plain_backend(gpu):
    f1, f2 = FIX["F1"], FIX["F2"]
    G = nx.Graph()
    G.add_edges_from([tuple(e) for e in f1["edges"]])
    comp = QtreeQAOAComposer(G, gamma=f1["gamma"], beta=f1["beta"])
    comp.ansatz_state()
    amp = QtreeSimulator(bucket_backend=HipBackend()).simulate(comp.circuit)[0]
    assert abs(amp - complex(*f1["amplitude"])) < 1e-6
    for Comp in (QtreeQAOAComposer, ZZQtreeQAOAComposer, SimpZZQtreeComposer):
        E = QAOAQtreeSimulator(Comp, bucket_backend=HipBackend()).energy_expectation(
            G, f2["gamma"], f2["beta"])
        assert abs(E[0] - f2["energy"]) < 1e-5


def test_config1_plain_backend(gpu):
    """N=10 3-regular (seed 42), p=1, gamma=pi/3, beta=pi/2: all 15 terms."""
    G = nx.random_regular_graph(3, 10, seed=42)
    g, b = [np.pi / 3], [np.pi / 2]
    sim = QAOAQtreeSimulator(QtreeQAOAComposer, bucket_backend=HipBackend())
    E = sim.energy_expectation(G, g, b)
    for e, z in zip(G.edges(), sim.edge_terms):
        ref = closed_form.zz_p1(G, e, g[0], b[0])
        assert abs(z - ref) <= 1e-5 * max(1.0, abs(ref))
    assert abs(E[0] - statevector.maxcut_energy(G, g, b)) < 1e-5 * 15


def test_perf_backend_reports(gpu):
    G = nx.random_regular_graph(3, 10, seed=1)
    comp = QtreeQAOAComposer(G, gamma=[np.pi / 3], beta=[np.pi / 4])
    comp.ansatz_state()
    be = PerfBackend()
    amp = QtreeSimulator(bucket_backend=be).simulate(comp.circuit)
    ref = QtreeSimulator(bucket_backend=npf.NumpyBackend()).simulate(comp.circuit)
    assert abs(amp[0] - ref[0]) < 1e-5
    assert "Total bucket contractions" in be.gen_report()


def test_batch_vars_through_plain_backend(gpu):
    """simulate_batch(batch_vars=2): the no_sum tail bucket runs on the GPU."""
    G = nx.random_regular_graph(4, 10, seed=3)
    comp = QtreeQAOAComposer(G, gamma=[np.pi / 3] * 2, beta=[np.pi / 2] * 2)
    comp.ansatz_state()
    amps = QtreeSimulator(bucket_backend=HipBackend()).simulate_batch(comp.circuit, batch_vars=2)
    ref = QtreeSimulator(bucket_backend=npf.NumpyBackend()).simulate_batch(comp.circuit, batch_vars=2)
    assert np.allclose(amps, ref, atol=1e-5)


@pytest.mark.parametrize("dtype,tol", [("complex64", 2e-5), ("complex128", 1e-12)])
@pytest.mark.parametrize("seed", range(6))
def test_step_large_rank_register_tiling(gpu, dtype, tol, seed):
    """Rank ~20 steps with 1-4 inputs where the largest input lacks high output
    bits (register-reuse path) or carries all of them."""
    rng = np.random.default_rng(100 + seed)
    n_vars = 21
    vs = [Var(i) for i in range(n_vars)]
    k = 1 + seed % 4
    big_ids = sorted({0} | set(int(x) for x in rng.choice(np.arange(1, n_vars), size=14, replace=False)))
    bucket = [Tensor("A", [vs[i] for i in big_ids],
                     data=rng.normal(size=(2,) * len(big_ids)) + 1j * rng.normal(size=(2,) * len(big_ids)))]
    rest = [i for i in range(1, n_vars) if i not in big_ids]
    for j in range(1, k):
        ids = sorted({0} | set(rest[(j - 1)::max(1, k - 1)]) | {int(rng.integers(1, n_vars))})
        bucket.append(Tensor(f"T{j}", [vs[i] for i in ids],
                             data=rng.normal(size=(2,) * len(ids)) + 1j * rng.normal(size=(2,) * len(ids))))
    be = HipBackend(dtype=dtype)
    be._ensure()
    for no_sum in (False, True):
        got = be.get_result_data(be.process_bucket(_to_dev(be, bucket), no_sum=no_sum))
        ref = np.asarray(npf.process_bucket(bucket, no_sum=no_sum).data)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref)))


def test_feynman_thread_pool_shares_one_backend(gpu):
    """FeynmanSimulator(pool_type='thread') (``qtensor/FeynmanSimulator.py:36-41``):
    four threads run the 2^s slices through one shared HipBackend at once; the
    sum equals the serial numpy oracle's, slice by slice and in total."""
    from qtensor_amd.ordering import SlicesOptimizer
    from qtensor_amd.simulator import FeynmanSimulator
    G = nx.random_regular_graph(3, 16, seed=7)
    comp = QtreeQAOAComposer(G, gamma=[0.3, 0.7], beta=[0.2, 0.4])
    comp.energy_expectation_lightcone(list(G.edges())[0])
    be = HipBackend()
    sim = FeynmanSimulator(bucket_backend=be, optimizer=SlicesOptimizer(tw_bias=0, max_tw=6),
                           pool_type="thread", n_processes=4)
    got = sim.simulate(comp.circuit)
    assert len(sim.parallel_vars) >= 3
    ref_sim = FeynmanSimulator(bucket_backend=npf.NumpyBackend(), optimizer=SlicesOptimizer(tw_bias=0, max_tw=6))
    ref = ref_sim.simulate(comp.circuit)
    assert [int(v) for v in ref_sim.parallel_vars] == [int(v) for v in sim.parallel_vars]
    whole = QtreeSimulator(bucket_backend=npf.NumpyBackend()).simulate(comp.circuit)
    assert abs(got[0] - ref[0]) <= 1e-5 and abs(got[0] - whole[0]) <= 1e-5
    parts = [sim._parallel_unit(k)[0] for k in range(2 ** len(sim.parallel_vars))]
    ref_parts = [ref_sim._parallel_unit(k)[0] for k in range(2 ** len(sim.parallel_vars))]
    assert np.allclose(parts, ref_parts, atol=1e-5)


def test_sliced_buckets_are_device_spans(gpu):
    """get_sliced_buckets puts every tensor in one device buffer; a span reads
    back as the transposed and sliced gate, and repeated gates reuse the host
    slicing."""
    from qtensor_amd.backend import DeviceSpan
    from qtensor_amd.circuit import slice_tensor_data
    from qtensor_amd.tensornet import QtreeTensorNet
    G = nx.random_regular_graph(3, 10, seed=2)
    comp = QtreeQAOAComposer(G, gamma=[0.5], beta=[0.25])
    comp.energy_expectation_lightcone(list(G.edges())[0])
    tn = QtreeTensorNet.from_qtree_gates(comp.circuit)
    be = HipBackend(dtype="complex128")
    sd = {v: slice(0, 1) for v in tn.ket_vars + tn.bra_vars}
    out = be.get_sliced_buckets(tn.buckets, tn.data_dict, sd)
    n = 0
    for bucket, ob in zip(tn.buckets, out):
        for t, s in zip(bucket, ob):
            ix, ref = slice_tensor_data(t, tn.data_dict[t.data_key], sd)
            assert [int(v) for v in s.indices] == [int(v) for v in ix]
            assert [v.size for v in s.indices] == [v.size for v in ix]
            assert type(s.data) is DeviceSpan
            assert np.array_equal(np.asarray(s.data), ref)
            n += 1
    assert n > 0 and len(be._slice_cache) < n
