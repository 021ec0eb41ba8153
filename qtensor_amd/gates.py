"""Quantum gates as tensors for bucket elimination.

Gate conventions follow the ``qtree.operators`` gates that QTensor's composers
use (``qtensor/OpFactory.py:30-60``; Cirq equivalence at ``OpFactory.py:14-28``):

* ``H``       = [[1, 1], [1, -1]] / sqrt(2)
* ``cX``      changes qubit 1 only; tensor indexed (control, target_out, target_in)
* ``Z``       = diag(1, -1)                           (diagonal, tensor shape (2,))
* ``ZPhase``  = diag(1, exp(i pi alpha))              (diagonal, tensor shape (2,))
* ``XPhase``  = exp(i pi alpha/2) [[c, -i s], [-i s, c]], c = cos(pi alpha/2),
  s = sin(pi alpha/2)  (Cirq ``XPowGate(exponent=alpha)``)
* ``YPhase``  = exp(i pi alpha/2) [[c, -s], [s, c]]    (Cirq ``YPowGate``)
* ``ZZ``      = [[m, p], [p, m]] over (z_a, z_b), p = exp(i pi alpha/2),
  m = exp(-i pi alpha/2)  (diagonal on both qubits, ``OpFactory.py:44-56``)
* ``M``       = 2x2 identity, complex64; closes every wire in ``circ2buckets``.

A gate's tensor index order is, per qubit in ``qubits`` order, ``(out, in)`` for
a qubit whose basis it changes and ``(wire,)`` for a qubit it is diagonal on.
``dagger_me`` returns the adjoint gate; parametric gates negate ``alpha``.

Every gate also reports ``kind`` (a small integer shared with the device-side
gate builder) so that plans can rebuild gate tensors from (gamma, beta) on the
GPU in FP64.
"""
import numpy as np

# device gate-kind codes (must match qtensor_amd/csrc/qtensor_amd.hip)
KIND_CONST = 0      # data copied verbatim from the host
KIND_ZPHASE = 1     # [1, e^{i pi a}]
KIND_XPHASE = 2     # e^{i pi a/2} [[c, -is], [-is, c]]
KIND_ZZ = 3         # [[e^{-i pi a/2}, e^{i pi a/2}], [e^{i pi a/2}, e^{-i pi a/2}]]
KIND_YPHASE = 4     # e^{i pi a/2} [[c, -s], [s, c]]


class Gate:
    name = "Gate"
    _changes_qubits = ()
    kind = KIND_CONST
    parametric = False

    def __init__(self, *qubits, **parameters):
        self._qubits = tuple(map(int, qubits))
        self._parameters = parameters          # a fresh dict per call
        self._dagger = False

    @property
    def qubits(self):
        return self._qubits

    @property
    def changed_qubits(self):
        return tuple(self._qubits[i] for i in self._changes_qubits)

    @property
    def parameters(self):
        return self._parameters

    def is_parametric(self):
        return self.parametric

    def gen_tensor(self):
        raise NotImplementedError

    def tensor(self):
        """The gate's tensor.  Equal gates (same class, adjoint flag and
        parameters) share one read-only array."""
        key = (self.__class__, self._dagger, tuple(sorted(self._parameters.items())))
        t = _TENSORS.get(key)
        if t is None:
            t = self.gen_tensor()
            if self._dagger and not self.parametric:
                t = _adjoint(t, self)
            t = np.array(t)
            t.setflags(write=False)
            if len(_TENSORS) >= 1 << 14:
                _TENSORS.clear()
            _TENSORS[key] = t
        return t

    def dagger_me(self):
        g = self.__class__(*self._qubits, **self._parameters)
        g._dagger = not self._dagger
        if self.parametric:
            g._parameters = dict(self._parameters)
            g._parameters["alpha"] = -self._parameters["alpha"]
            g._dagger = False
            g._daggered_name = not getattr(self, "_daggered_name", False)
        return g

    @property
    def label(self):
        daggered = self._dagger or getattr(self, "_daggered_name", False)
        return self.name + ("+" if daggered else "")

    def __repr__(self):
        args = ",".join(map(str, self._qubits))
        if self.parametric:
            args += f",a={self._parameters['alpha']:.4g}"
        return f"{self.label}({args})"


_TENSORS = {}


def _adjoint(t, gate):
    """Adjoint of a constant gate tensor in the (out,in)/(wire) index layout."""
    # every constant gate here is its own adjoint up to conjugation of a real
    # tensor: H, cX, Z, M are real symmetric unitaries.
    return np.conj(t)


class H(Gate):
    name = "H"
    _changes_qubits = (0,)

    def gen_tensor(self):
        return np.array([[1.0, 1.0], [1.0, -1.0]]) / np.sqrt(2.0)


class Z(Gate):
    name = "Z"
    _changes_qubits = ()

    def gen_tensor(self):
        return np.array([1.0, -1.0])


class cX(Gate):
    name = "cX"
    _changes_qubits = (1,)

    def gen_tensor(self):
        t = np.zeros((2, 2, 2))
        for c in range(2):
            for tin in range(2):
                t[c, tin ^ c, tin] = 1.0
        return t


class M(Gate):
    name = "M"
    _changes_qubits = (0,)

    def gen_tensor(self):
        return np.eye(2, dtype=np.complex64)


class ParametricGate(Gate):
    parametric = True

    def __init__(self, *qubits, alpha=0.0, **kw):
        super().__init__(*qubits, alpha=float(alpha), **kw)

    @property
    def alpha(self):
        return self._parameters["alpha"]


class ZPhase(ParametricGate):
    name = "ZPhase"
    _changes_qubits = ()
    kind = KIND_ZPHASE

    def gen_tensor(self):
        return np.array([1.0, np.exp(1j * np.pi * self.alpha)])


class XPhase(ParametricGate):
    name = "XPhase"
    _changes_qubits = (0,)
    kind = KIND_XPHASE

    def gen_tensor(self):
        a = self.alpha
        c, s, g = np.cos(np.pi * a / 2), np.sin(np.pi * a / 2), np.exp(1j * np.pi * a / 2)
        return g * np.array([[c, -1j * s], [-1j * s, c]])


class YPhase(ParametricGate):
    name = "YPhase"
    _changes_qubits = (0,)
    kind = KIND_YPHASE

    def gen_tensor(self):
        a = self.alpha
        c, s, g = np.cos(np.pi * a / 2), np.sin(np.pi * a / 2), np.exp(1j * np.pi * a / 2)
        return g * np.array([[c, -s], [s, c]])


class ZZ(ParametricGate):
    name = "ZZ"
    _changes_qubits = ()
    kind = KIND_ZZ

    def gen_tensor(self):
        p = np.exp(1j * np.pi * self.alpha / 2)
        m = np.exp(-1j * np.pi * self.alpha / 2)
        return np.array([[m, p], [p, m]])


class QtreeFactory:
    """Operator namespace the composers apply gates from (``QtreeBuilder.operators``)."""
    H = H
    cX = cX
    Z = Z
    ZPhase = ZPhase
    XPhase = XPhase
    YPhase = YPhase
    ZZ = ZZ
    M = M
