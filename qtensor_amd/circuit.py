"""Circuit -> bucket tensor network, relabelling, slicing and the elimination loop.

Host-side counterparts of the ``qtree`` functions QTensor drives its bucket
backends with (``qtree`` is an un-vendored submodule of the reference; the
behaviour below is the one its call sites and notebook outputs fix):

* ``circ2buckets``      -- call site ``qtensor/optimisation/TensorNet.py:60-68``;
  layout shown at ``examples/Qtree Basic usage.ipynb:186``.
* ``reorder_buckets``   -- call site ``qtensor/Simulate.py:47-52``.
* ``slice_from_bits``   -- call site ``qtensor/Simulate.py:54-59``.
* ``bucket_elimination``-- call sites ``qtensor/Simulate.py:90-93``,
  ``qtensor/FeynmanSimulator.py:54-56``; behaviour at
  ``examples/Qtree Basic usage.ipynb:296,476,755``.
* ``buckets2graph``     -- call site ``qtensor/optimisation/TensorNet.py:53-58``.
* ``slice_tensor_data`` -- the per-tensor transpose+slice step of the numpy
  backend's ``get_sliced_buckets`` (``qtensor/ProcessingFrameworks.py:33-34``).
"""
import itertools

import networkx as nx
import numpy as np

from .gates import M
from .model import Tensor, Var


class GateDataDict(dict):
    """``data_key -> ndarray`` plus the symbolic description of each gate tensor.

    ``gate_info[data_key] = (kind, alpha, terms)`` where ``terms`` is a tuple of
    ``(param_index, coeff)`` with ``alpha == sum(coeff * params[param_index])``
    (empty for constant gates).  The plan compiler uses it to rebuild gate
    tensors from new (gamma, beta) on the device.
    """

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.gate_info = {}


def circ2buckets(qubit_count, circuit):
    """Build buckets for the amplitude <bra|circuit|ket>.

    The network is built from the output side: variables ``o_q`` (ids
    ``0..n-1``) are the bra wires, an ``M`` tensor closes each output wire,
    gates are visited in reverse time order, and a final ``M`` tensor per
    qubit closes the input wire with a ket variable ``i_q``.  A gate that
    changes a qubit's basis opens a new variable on that wire and is indexed
    ``(out, in)`` on it; a diagonal qubit contributes its current wire
    variable only.

    Returns ``(buckets, data_dict, bra_vars, ket_vars)``.
    """
    data_dict = GateDataDict()
    layer_vars = [Var(q, name=f"o_{q}") for q in range(qubit_count)]
    bra_vars = list(layer_vars)
    buckets = [[] for _ in range(qubit_count)]
    next_id = qubit_count

    gate_info = data_dict.gate_info

    def place(gate, closing_names=None):
        nonlocal next_id
        changed = gate._changes_qubits
        indices = []
        low = None
        for pos, q in enumerate(gate.qubits):
            cur = layer_vars[q]
            if low is None or cur.identity < low:
                low = cur.identity
            if pos in changed:
                name = closing_names[q] if closing_names else None
                new = Var(next_id, name=name)
                next_id += 1
                buckets.append([])
                indices += [cur, new]
                layer_vars[q] = new
            else:
                indices.append(cur)
        label = gate.label
        key = (label, len(data_dict))
        data_dict[key] = gate.tensor()
        gate_info[key] = (gate.kind, gate.parameters.get("alpha", 0.0), tuple(getattr(gate, "param_terms", ())))
        # a new variable's id exceeds every current one, so the smallest index
        # is the smallest of the wires' current variables
        buckets[low].append(Tensor(label, indices, data_key=key))

    for q in range(qubit_count):
        place(M(q))
    for layer in reversed(list(circuit)):
        for gate in layer:
            place(gate)
    ket_names = {q: f"i_{q}" for q in range(qubit_count)}
    ket_vars = []
    for q in range(qubit_count):
        place(M(q), closing_names=ket_names)
        ket_vars.append(layer_vars[q])
    return buckets, data_dict, bra_vars, ket_vars


def reorder_buckets(old_buckets, permutation):
    """Relabel variables by their position in ``permutation`` and re-bucket.

    Index order inside each tensor is kept (it still matches the tensor's
    data layout); each tensor moves to the bucket of its smallest new id.
    Returns ``(perm_buckets, perm_dict)``.
    """
    perm_dict = {}
    by_id = {}
    for n, var in enumerate(permutation):
        nv = var.copy(identity=n)
        perm_dict[var] = nv
        by_id[var.identity] = nv
    perm_buckets = [[] for _ in range(len(permutation))]
    for bucket in old_buckets:
        for tensor in bucket:
            new_indices = tuple([by_id[v.identity] for v in tensor.indices])
            b = min([v.identity for v in new_indices]) if new_indices else 0
            perm_buckets[b].append(Tensor(tensor.name, new_indices, tensor.data_key, tensor.data))
    return perm_buckets, perm_dict


def slice_from_bits(value, variables):
    """``{var: slice(b, b+1)}`` with ``b`` the bits of ``value``, most significant first."""
    n = len(variables)
    bits = [(int(value) >> (n - 1 - k)) & 1 for k in range(n)]
    return {var: slice(b, b + 1) for var, b in zip(variables, bits)}


def unravel_index(value, dims):
    out = []
    for d in reversed(dims):
        out.append(int(value) % d)
        value = int(value) // d
    return list(reversed(out))


def slice_tensor_data(tensor, data, slice_dict):
    """Transpose ``data`` to ascending variable order and apply ``slice_dict``.

    ``slice(a, b)`` keeps the axis with a new size; an ``int`` drops it.
    Returns ``(indices, data)`` with the new variable sizes.
    """
    order = sorted(range(len(tensor.indices)), key=lambda a: int(tensor.indices[a]))
    data = np.transpose(data, order)
    idx_sorted = [tensor.indices[a] for a in order]
    bounds = tuple(slice_dict.get(v, slice(None)) for v in idx_sorted)
    data = data[bounds]
    kept = [v for v, b in zip(idx_sorted, bounds) if not isinstance(b, (int, np.integer))]
    indices = tuple(v.copy(size=s) for v, s in zip(kept, data.shape))
    assert len(indices) == data.ndim
    return indices, data


def bucket_elimination(buckets, process_bucket_fn, n_var_nosum=0):
    """Eliminate buckets in order with ``process_bucket_fn``.

    Each processed bucket's result goes to the bucket of its first index, or
    is multiplied into the scalar result.  The last ``n_var_nosum`` buckets
    are concatenated and processed once with ``no_sum=True``.
    """
    n_contract = len(buckets) - n_var_nosum
    result = None
    for bucket in buckets[:n_contract]:
        if not bucket:
            continue
        t = process_bucket_fn(bucket)
        if len(t.indices) > 0:
            buckets[int(t.indices[0])].append(t)
        else:
            result = t if result is None else result * t
    rest = list(itertools.chain.from_iterable(buckets[n_contract:]))
    if rest:
        t = process_bucket_fn(rest, no_sum=True)
        result = t if result is None else result * t
    return result


def buckets2graph(buckets, ignore_variables=()):
    """Interaction (line) graph: one node per variable, a clique per tensor."""
    ignore = set(int(v) for v in ignore_variables)
    graph = nx.Graph()
    for bucket in buckets:
        for tensor in bucket:
            nodes = [v for v in tensor.indices if int(v) not in ignore]
            for v in nodes:
                if int(v) not in graph:
                    graph.add_node(int(v), name=v.name, size=v.size)
            graph.add_edges_from(itertools.combinations([int(v) for v in nodes], 2))
    return graph
