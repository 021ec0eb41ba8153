"""Bucket tensor network built from a circuit (``qtensor/optimisation/TensorNet.py:21-83``)."""
from .circuit import buckets2graph, circ2buckets
from .ordering import elimination_widths


class QtreeTensorNet:
    def __init__(self, buckets, data_dict, bra_vars, ket_vars, free_vars=None,
                 bucket_backend=None):
        self.buckets = buckets
        self.data_dict = data_dict
        self.bra_vars = bra_vars
        self.ket_vars = ket_vars
        self.free_vars = list(free_vars or [])
        self.bucket_backend = bucket_backend

    def set_free_qubits(self, free):
        self.free_vars = [self.bra_vars[i] for i in free]
        self.bra_vars = [v for v in self.bra_vars if v not in self.free_vars]

    @property
    def tensors(self):
        return sum(self.buckets, [])

    def __len__(self):
        return len(self.tensors)

    def slice(self, slice_dict):
        self.buckets = self.bucket_backend.get_sliced_buckets(
            self.buckets, self.data_dict, slice_dict)
        return self.buckets

    def get_line_graph(self):
        return buckets2graph(self.buckets, ignore_variables=self.bra_vars + self.ket_vars)

    def simulation_cost(self, peo):
        """Per-step memory (elements) and flop estimates of eliminating in ``peo``."""
        ignore = set(int(v) for v in self.bra_vars + self.ket_vars + self.free_vars)
        order = [int(v) for v in peo if int(v) not in ignore]
        g = self.get_line_graph()
        widths = elimination_widths(g, order)
        mems = [2 ** w for w in widths]
        flops = [2 ** (w + 1) for w in widths]
        return mems, flops

    @classmethod
    def from_qtree_gates(cls, qc, bucket_backend=None):
        n_qubits = len(set(q for g in qc for q in g.qubits))
        buckets, data_dict, bra, ket = circ2buckets(n_qubits, [[g] for g in qc])
        return cls(buckets, data_dict, bra, ket, bucket_backend=bucket_backend)
