"""Numpy restatement of the reference's numpy bucket backend (oracle, tests only).

``process_bucket``: multiply all tensors of a bucket over the union of their
indices (ascending variable id, C order, size-1 axes broadcast) and sum the
first (bucket) axis unless ``no_sum`` -- the semantics that
``qtensor/ProcessingFrameworks.py:29-31`` delegates to ``qtree.np_framework``
and that ``CMKLExtendedBackend.process_bucket`` restates pairwise at
``ProcessingFrameworks.py:45-147`` (result tag ``E<first id>`` / ``Ef``, the
ordering assertion at ``:137``).

``get_sliced_buckets``: transpose each tensor's data to ascending variable id
and apply the slice dict (``ProcessingFrameworks.py:33-34``; size-1 retention at
``examples/Qtree Basic usage.ipynb:330,468``).

The functions take and return objects of the engine's data model
(``qtensor_amd.model``) but compute with numpy only.
"""
import string

import numpy as np

from qtensor_amd.circuit import slice_tensor_data
from qtensor_amd.model import Tensor

_LETTERS = string.ascii_letters


def _einsum_multiply(ix_a, a, ix_b, b):
    union = sorted(set(ix_a) | set(ix_b))
    lab = {v: _LETTERS[k] for k, v in enumerate(union)}
    expr = "".join(lab[v] for v in ix_a) + "," + "".join(lab[v] for v in ix_b) + "->" + \
        "".join(lab[v] for v in union)
    return union, np.einsum(expr, a, b)


def process_bucket(bucket, no_sum=False, dtype=np.complex128):
    ix = [int(v) for v in bucket[0].indices]
    vars_by_id = {int(v): v for t in bucket for v in t.indices}
    data = np.asarray(bucket[0].data, dtype=dtype)
    for t in bucket[1:]:
        ix, data = _einsum_multiply(ix, data, [int(v) for v in t.indices],
                                    np.asarray(t.data, dtype=dtype))
    indices = [vars_by_id[i] for i in ix]
    if indices:
        tag = indices[0].identity
        if not no_sum:
            data = data.sum(axis=0)
            indices = indices[1:]
    else:
        tag = "f"
    return Tensor(f"E{tag}", indices, data=data)


def get_sliced_buckets(buckets, data_dict, slice_dict):
    out = []
    for bucket in buckets:
        nb = []
        for t in bucket:
            indices, data = slice_tensor_data(t, data_dict[t.data_key], slice_dict)
            nb.append(t.copy(indices=indices, data=data))
        out.append(nb)
    return out


class NumpyBackend:
    """Oracle bucket backend with the reference interface."""

    def __init__(self, dtype=np.complex128):
        self.dtype = dtype

    def process_bucket(self, bucket, no_sum=False):
        return process_bucket(bucket, no_sum=no_sum, dtype=self.dtype)

    def get_sliced_buckets(self, buckets, data_dict, slice_dict):
        return get_sliced_buckets(buckets, data_dict, slice_dict)

    def get_result_data(self, result):
        return result.data


def brute_force_bucket(bucket, no_sum=False):
    """Explicit loop over every assignment of the union's variables (tiny inputs)."""
    union = sorted({int(v): v for t in bucket for v in t.indices}.items())
    ids = [i for i, _ in union]
    sizes = [v.size for _, v in union]
    full = np.zeros(sizes, dtype=np.complex128)
    for assign in np.ndindex(*sizes):
        val = 1.0 + 0j
        where = dict(zip(ids, assign))
        for t in bucket:
            val *= np.asarray(t.data)[tuple(where[int(v)] for v in t.indices)]
        full[assign] = val
    if not no_sum and ids:
        return full.sum(axis=0)
    return full
