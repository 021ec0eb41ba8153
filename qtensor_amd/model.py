"""Bucket-elimination data model: variables, tensors and buckets.

Mirrors the data model that QTensor's bucket backends consume (the ``qtree``
``Var``/``Tensor`` types used at ``qtensor/ProcessingFrameworks.py:53,69-72,96``
and ``qtensor/optimisation/Optimizer.py:33,60-62``):

* ``Var(identity, size=2, name=None)``: ``int(v)`` is the identity and the sort key
  of every index list; ``name`` is ``o_<q>`` for output (bra) wires, ``i_<q>`` for
  input (ket) wires and ``v_<id>`` otherwise.
* ``Tensor(name, indices, data_key=None, data=None)``: ``indices`` is a tuple of
  ``Var``; ``data`` is either ``None`` (symbolic, data lives in a ``data_dict``
  under ``data_key``), a numpy array, or a device buffer owned by a backend.
* A bucket list is ``list[list[Tensor]]`` with one bucket per variable; a tensor
  lives in the bucket of its smallest variable.
"""


class Var:
    __slots__ = ("identity", "size", "name")

    def __init__(self, identity, size=2, name=None):
        self.identity = int(identity)
        self.size = int(size)
        self.name = name if name is not None else f"v_{self.identity}"

    def copy(self, identity=None, size=None, name=None):
        ident = self.identity if identity is None else identity
        if name is None:
            # relabelled internal variables take the new identity in their name
            name = self.name if (identity is None or not self.name.startswith("v")) else None
        return Var(ident, size=self.size if size is None else size, name=name)

    def __int__(self):
        return self.identity

    def __index__(self):
        return self.identity

    def __hash__(self):
        return hash(self.identity)

    def __eq__(self, other):
        if isinstance(other, Var):
            return self.identity == other.identity
        return NotImplemented

    def __lt__(self, other):
        return self.identity < int(other)

    def __repr__(self):
        return self.name

    __str__ = __repr__


class Tensor:
    __slots__ = ("name", "indices", "data_key", "data")

    def __init__(self, name, indices, data_key=None, data=None):
        self.name = name
        self.indices = tuple(indices)
        self.data_key = data_key
        self.data = data

    def copy(self, name=None, indices=None, data_key=None, data=None):
        return Tensor(self.name if name is None else name,
                      self.indices if indices is None else indices,
                      self.data_key if data_key is None else data_key,
                      self.data if data is None else data)

    @property
    def shape(self):
        return tuple(v.size for v in self.indices)

    def __mul__(self, other):
        """Scalar product of two rank-0 results (``bucket_elimination`` uses it)."""
        return Tensor(self.name, self.indices, data=self.data * other.data)

    def __repr__(self):
        return f"{self.name}({','.join(map(str, self.indices))})"
