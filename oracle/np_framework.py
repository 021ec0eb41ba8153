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

The functions take and return tensors of any class with the reference data
model's attributes (``name``, ``indices``, ``data_key``, ``data`` and a
``copy(indices=..., data=...)`` method; variables with ``int(v)``, ``size`` and
``copy(size=...)``, ``qtensor/ProcessingFrameworks.py:53,69-72,96``).  Results
are built with the input tensors' own class.  This module imports nothing from
the engine: the transpose and slice below are its own.
"""
import string

import numpy as np

_LETTERS = string.ascii_letters


def _einsum_multiply(ix_a, a, ix_b, b):
    union = sorted(set(ix_a) | set(ix_b))
    lab = {v: _LETTERS[k] for k, v in enumerate(union)}
    expr = "".join(lab[v] for v in ix_a) + "," + "".join(lab[v] for v in ix_b) + "->" + \
        "".join(lab[v] for v in union)
    return union, np.einsum(expr, a, b)


def process_bucket(bucket, no_sum=False, dtype=np.complex128):
    tensor_cls = type(bucket[0])
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
    return tensor_cls(f"E{tag}", indices, data=data)


def transpose_and_slice(indices, data, slice_dict):
    """Axes of ``data`` (one per variable of ``indices``, in that order) sorted
    by ascending variable id, then ``slice_dict`` applied axis by axis: a
    ``slice`` keeps the axis with the sliced size, an ``int`` selects one
    entry and drops the axis (``qtensor/ProcessingFrameworks.py:33-34``; the
    kept size-1 axes of ``examples/Qtree Basic usage.ipynb:330,468``).
    Returns ``(new indices, data)``."""
    data = np.asarray(data)
    perm = sorted(range(len(indices)), key=lambda a: int(indices[a]))
    data = np.transpose(data, perm)
    out_vars = []
    axis = 0
    for a in perm:
        v = indices[a]
        sel = slice_dict.get(v, None)
        if sel is None:
            out_vars.append(v)
            axis += 1
        elif isinstance(sel, slice):
            data = np.take(data, np.arange(data.shape[axis])[sel], axis=axis)
            out_vars.append(v.copy(size=data.shape[axis]))
            axis += 1
        else:
            data = np.take(data, int(sel), axis=axis)
    if len(out_vars) != data.ndim:
        raise AssertionError("sliced data and indices disagree")
    return tuple(out_vars), data


def get_sliced_buckets(buckets, data_dict, slice_dict):
    out = []
    for bucket in buckets:
        nb = []
        for t in bucket:
            indices, data = transpose_and_slice(t.indices, data_dict[t.data_key], slice_dict)
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
