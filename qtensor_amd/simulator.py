"""Simulators that drive a bucket backend (host side).

* ``QtreeSimulator``   -- ``qtensor/Simulate.py:22-107``: circuit -> buckets ->
  order (+ ``max_tw`` guard) -> reorder -> slice dict -> ``get_sliced_buckets``
  -> ``bucket_elimination(process_bucket)`` -> ``get_result_data``.
* ``QAOAQtreeSimulator`` -- ``qtensor/QAOASimulator.py:9-66,114-115``: sum of
  per-edge lightcone amplitudes ``<Z_u Z_v>``, post-processed to the MaxCut
  expectation ``C = (|E| - Re sum)/2``.
* ``FeynmanSimulator`` -- ``qtensor/FeynmanSimulator.py:18-95``: slice the
  optimizer's parallel variables to integers and sum the 2^s eliminations.
"""
import copy
import warnings

import numpy as np

from .circuit import bucket_elimination, reorder_buckets, slice_from_bits, unravel_index
from .ordering import DefaultOptimizer, SlicesOptimizer
from .tensornet import QtreeTensorNet


def _default_backend():
    from .backend import HipBackend
    return HipBackend()


class Simulator:
    def simulate(self, qc):
        raise NotImplementedError


class QtreeSimulator(Simulator):
    FallbackOptimizer = DefaultOptimizer

    def __init__(self, bucket_backend=None, optimizer=None, max_tw=None):
        self.bucket_backend = bucket_backend if bucket_backend is not None else _default_backend()
        self.optimizer = optimizer if optimizer is not None else self.FallbackOptimizer()
        self.max_tw = max_tw

    # -- internal helpers (Simulate.py:32-59)
    def _new_circuit(self, qc):
        self.all_gates = qc

    def _create_buckets(self):
        self.tn = QtreeTensorNet.from_qtree_gates(self.all_gates)
        self.tn.bucket_backend = self.bucket_backend

    def _set_free_qubits(self, free_final_qubits):
        self.tn.free_vars = [self.tn.bra_vars[i] for i in free_final_qubits]
        self.tn.bra_vars = [v for v in self.tn.bra_vars if v not in self.tn.free_vars]

    def _optimize_buckets(self):
        self.peo = self.optimize_buckets()

    def _reorder_buckets(self):
        perm_buckets, perm_dict = reorder_buckets(self.tn.buckets, self.peo)
        self.tn.ket_vars = sorted([perm_dict[v] for v in self.tn.ket_vars], key=str)
        self.tn.bra_vars = sorted([perm_dict[v] for v in self.tn.bra_vars], key=str)
        self.tn.free_vars = [perm_dict[v] for v in self.tn.free_vars]
        self.tn.buckets = perm_buckets
        return perm_dict

    def _get_slice_dict(self, initial_state=0, target_state=0):
        sd = slice_from_bits(initial_state, self.tn.ket_vars)
        sd.update(slice_from_bits(target_state, self.tn.bra_vars))
        sd.update({v: slice(None) for v in self.tn.free_vars})
        return sd

    def optimize_buckets(self):
        peo, self.tn = self.optimizer.optimize(self.tn)
        return peo

    def simulate_batch(self, qc, batch_vars=0, peo=None):
        self._new_circuit(qc)
        self._create_buckets()
        self._set_free_qubits(list(range(batch_vars)))
        if peo is None:
            self._optimize_buckets()
            if self.max_tw and self.optimizer.treewidth > self.max_tw:
                raise ValueError(
                    f"Treewidth {self.optimizer.treewidth} is larger than max_tw={self.max_tw}.")
        else:
            self.peo = peo
        self._reorder_buckets()
        slice_dict = self._get_slice_dict()
        sliced = self.tn.slice(slice_dict)
        result = bucket_elimination(sliced, self.bucket_backend.process_bucket,
                                    n_var_nosum=len(self.tn.free_vars))
        return np.asarray(self.bucket_backend.get_result_data(result)).flatten()

    def simulate(self, qc):
        return self.simulate_state(qc)

    def simulate_state(self, qc, peo=None):
        return self.simulate_batch(qc, peo=peo, batch_vars=0)


class QAOASimulator(Simulator):
    def __init__(self, composer, profile=False, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.composer = composer
        self.profile = profile

    def _edge_energy_circuit(self, G, gamma, beta, edge):
        comp = self.composer(G, gamma=gamma, beta=beta)
        comp.energy_expectation_lightcone(edge)
        return comp.circuit

    def _get_edge_energy(self, G, gamma, beta, edge):
        return self.simulate(self._edge_energy_circuit(G, gamma, beta, edge))

    def _post_process_energy(self, G, E):
        if np.imag(E) > 1e-6:
            warnings.warn(f"Energy result imaginary part was: {np.imag(E)}")
        return (G.number_of_edges() - np.real(E)) / 2

    def energy_expectation(self, G, gamma, beta):
        """MaxCut expectation summed edge by edge (``QAOASimulator.py:44-66``)."""
        total = 0
        self.edge_terms = []
        for edge in G.edges():
            E = self._get_edge_energy(G, gamma, beta, edge)
            self.edge_terms.append(complex(np.asarray(E).ravel()[0]))
            total += E
        if self.profile and hasattr(self.bucket_backend, "gen_report"):
            print(self.bucket_backend.gen_report())
        return self._post_process_energy(G, total)


class QAOAQtreeSimulator(QAOASimulator, QtreeSimulator):
    pass


class WideQtreeSimulator(QtreeSimulator):
    """``QtreeSimulator`` whose elimination runs on the device as one compiled
    wide plan (gates built in FP64 on the device, one grid kernel per op)
    instead of one ``process_bucket`` launch per bucket: same order rule
    (greedy, open variables last), same ``simulate_batch`` result (the 2^b
    amplitudes of the first b output qubits, ``qtensor/Simulate.py:66-94``).
    ``max_tw`` guards the order's width as the reference does."""

    def __init__(self, dtype="complex64", max_tw=None, device=0):
        self.dtype = dtype
        self.max_tw = max_tw
        self.device = device

    def simulate_batch(self, qc, batch_vars=0, peo=None):
        from .wide import WideEngine, compile_wide
        wp = compile_wide(qc, order=peo, batch_vars=batch_vars, bake_angles=True)
        if self.max_tw and wp.treewidth > self.max_tw:
            raise ValueError(f"Treewidth {wp.treewidth} is larger than max_tw={self.max_tw}.")
        val = WideEngine(device=self.device, dtype=self.dtype).run(wp, [])
        return np.atleast_1d(np.asarray(val, dtype=np.complex128)).flatten()


def int_slice(value, vars_to_slice):
    """``{var: int}`` for the multi-index ``value`` over ``vars_to_slice``
    (``qtensor/FeynmanSimulator.py:18-25``)."""
    dims = [v.size for v in vars_to_slice]
    return dict(zip(vars_to_slice, unravel_index(value, dims)))


class FeynmanSimulator(QtreeSimulator):
    """Sliced simulation: 2^s independent eliminations summed.

    ``pool_type`` is ``'serial'`` (default) or ``'thread'``; a thread pool
    shares one backend, which must then be thread-safe.
    """
    FallbackOptimizer = SlicesOptimizer

    def __init__(self, *args, pool_type="serial", n_processes=None, **kwargs):
        super().__init__(*args, **kwargs)
        self.n_processes = 2 if n_processes is None else n_processes
        self.pool_type = pool_type

    def optimize_buckets(self):
        peo, par_vars, self.tn = self.optimizer.optimize(self.tn)
        self.parallel_vars = par_vars
        return peo

    def _reorder_buckets(self):
        perm_dict = super()._reorder_buckets()
        self.parallel_vars = sorted([perm_dict[v] for v in self.parallel_vars], key=str)
        return perm_dict

    def _get_slice_dict(self, initial_state=0, target_state=0, par_state=0):
        sd = super()._get_slice_dict(initial_state, target_state)
        sd.update(int_slice(par_state, self.parallel_vars))
        return sd

    def _parallel_unit(self, par_idx):
        sd = self._get_slice_dict(par_state=par_idx)
        tn = copy.copy(self.tn)
        sliced = tn.slice(sd)
        result = bucket_elimination(sliced, self.bucket_backend.process_bucket,
                                    n_var_nosum=len(tn.free_vars + self.parallel_vars))
        return np.asarray(self.bucket_backend.get_result_data(result)).flatten()

    def simulate(self, qc, batch_vars=0, tw_bias=2):
        return self.simulate_batch_adaptive(qc, batch_vars, tw_bias=tw_bias)

    def simulate_batch_adaptive(self, qc, batch_vars=0, tw_bias=2):
        self.tw_bias = tw_bias
        self._new_circuit(qc)
        self._create_buckets()
        self._set_free_qubits(list(range(batch_vars)))
        self._optimize_buckets()
        self._reorder_buckets()
        total = 2 ** len(self.parallel_vars)
        if self.pool_type == "thread":
            from multiprocessing.dummy import Pool as ThreadPool
            with ThreadPool(self.n_processes) as p:
                parts = list(p.imap(self._parallel_unit, range(total)))
        else:
            parts = [self._parallel_unit(k) for k in range(total)]
        return _neumaier_sum(parts)


def _neumaier_sum(parts):
    """Compensated (Neumaier) sum of equally shaped complex arrays, in complex128."""
    s = np.zeros_like(np.asarray(parts[0], dtype=np.complex128))
    c = np.zeros_like(s)
    for x in parts:
        x = np.asarray(x, dtype=np.complex128)
        for comp in ("real", "imag"):
            sv, cv, xv = getattr(s, comp), getattr(c, comp), getattr(x, comp)
            t = sv + xv
            big = np.abs(sv) >= np.abs(xv)
            cv += np.where(big, (sv - t) + xv, (xv - t) + sv)
            sv[...] = t
    return s + c
