"""Python binding of the native lightcone planner (``qb_lightcone_classes``,
``qb_compile_lightcone`` in ``csrc/lightcone_planner.cpp``).

``lightcone_classes(G, edges, p)`` gives every edge's canonical lightcone class
(the keys ``lightcone.canonical_lightcone`` computes); ``compile_lightcone``
compiles one class into the executor plan that ``plan.PlanCompiler`` produces
for the same lightcone, composer, order, slicer and launch shape -- the same
program words, computed natively.
"""
import ctypes

import numpy as np

from . import _lib
from .plan import Plan

COMPOSERS = {"QtreeQAOAComposer": 0, "ZZQtreeQAOAComposer": 1, "SimpZZQtreeComposer": 2}
SLICERS = {None: 0, "greedy": 1, "tree_trim": 2}
ORDERINGS = {"greedy": 0, "rgreedy": 1}


class PlanOpts(ctypes.Structure):
    _fields_ = [("composer", ctypes.c_int32), ("ordering", ctypes.c_int32), ("slicer", ctypes.c_int32),
                ("max_tw", ctypes.c_int32), ("tw_bias", ctypes.c_int32), ("lanes", ctypes.c_int32),
                ("lds_tensor_limit", ctypes.c_int64), ("lds_capacity", ctypes.c_int64),
                ("rg_temp", ctypes.c_double), ("rg_repeats", ctypes.c_int32), ("rg_seed", ctypes.c_uint64)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("n_steps", ctypes.c_int32), ("n_ops", ctypes.c_int32), ("n_gates", ctypes.c_int32),
                ("treewidth", ctypes.c_int32), ("max_rank", ctypes.c_int32), ("n_sliced", ctypes.c_int32),
                ("lds_need", ctypes.c_int64), ("glb_need", ctypes.c_int64),
                ("bytes_step_model", ctypes.c_int64), ("flops", ctypes.c_int64),
                ("n_barriers", ctypes.c_int32), ("n_gather", ctypes.c_int32), ("n_reduce", ctypes.c_int32),
                ("steps_words", ctypes.c_int32)]


_bound = False


def _lib_bound():
    global _bound
    lib = _lib.load()
    if not _bound:
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.qb_lightcone_classes.argtypes = [ctypes.c_int32, i32p, i32p, ctypes.c_int32, i32p, ctypes.c_int32,
                                             i32p, i32p, i32p, ctypes.c_int64, i64p]
        lib.qb_compile_lightcone.argtypes = [i32p, ctypes.c_int32, ctypes.POINTER(PlanOpts), i32p, ctypes.c_int64,
                                             i64p, ctypes.POINTER(PlanInfo)]
        _bound = True
    return lib


def _i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def supports(composer):
    return composer.__name__ in COMPOSERS


def lightcone_classes(G, edges, p):
    """``(edge_class, classes)``: class index of each edge and, per class, the
    key ``(n, edges, i, j)`` of ``lightcone.canonical_lightcone`` and the class
    record the compiler takes."""
    from .composer import _node_key
    nodes = sorted(G.nodes(), key=_node_key)
    index = {v: k for k, v in enumerate(nodes)}
    n = len(nodes)
    ptr = np.zeros(n + 1, dtype=np.int32)
    idx = []
    for k, v in enumerate(nodes):
        nb = [index[u] for u in G[v] if u != v]
        idx += nb
        ptr[k + 1] = ptr[k] + len(nb)
    idx = np.asarray(idx or [0], dtype=np.int32)
    e = np.asarray([(index[a], index[b]) for a, b in edges], dtype=np.int32).reshape(-1)
    if e.size == 0:
        return [], []
    lib = _lib_bound()
    edge_class = np.zeros(len(edges), dtype=np.int32)
    n_classes = ctypes.c_int32(0)
    used = ctypes.c_int64(0)
    cap = 64 * len(edges) + 64
    while True:
        words = np.zeros(cap, dtype=np.int32)
        rc = lib.qb_lightcone_classes(n, _i32p(ptr), _i32p(idx), len(edges), _i32p(e), int(p), _i32p(edge_class),
                                      ctypes.byref(n_classes), _i32p(words), cap, ctypes.byref(used))
        if rc == _lib.QB_ERR_UNSUPPORTED and used.value > cap:
            cap = used.value
            continue
        _lib.check(rc)
        break
    classes, pos = [], 0
    for _ in range(n_classes.value):
        nn, ne, i, j = (int(x) for x in words[pos:pos + 4])
        rec = words[pos:pos + 4 + 2 * ne].copy()
        pairs = words[pos + 4:pos + 4 + 2 * ne].reshape(-1, 2)
        ce = [(int(a), int(b)) for a, b in pairs]
        classes.append(((nn, tuple(ce), i, j), nn, ce, i, j, rec))
        pos += 4 + 2 * ne
    return edge_class.tolist(), classes


def lightcone_record(n, edges, i, j):
    """Class record of one canonical lightcone (the compiler's input)."""
    rec = [n, len(edges), i, j]
    for a, b in edges:
        rec += [a, b]
    return np.asarray(rec, dtype=np.int32)


def compile_lightcone(record, p, composer, ordering="greedy", slicer=None, max_tw=None, tw_bias=0, lanes=64,
                      lds_tensor_limit=2048, lds_capacity=None, rg_temp=0.002, rg_repeats=16, rg_seed=0):
    """Executor plan of one lightcone class (a ``plan.Plan``).  Raises
    ``ValueError`` (e.g. "LDS capacity exhausted") or ``MemoryError``
    ("Estimated OOM", a width the slicer cannot reach)."""
    lib = _lib_bound()
    o = PlanOpts()
    o.composer = COMPOSERS[composer.__name__]
    o.ordering = ORDERINGS[ordering]
    o.slicer = SLICERS[slicer if max_tw is not None else None]
    o.max_tw = int(max_tw) if max_tw is not None else 0
    o.tw_bias = int(tw_bias)
    o.lanes = int(lanes)
    o.lds_tensor_limit = int(lds_tensor_limit)
    o.lds_capacity = -1 if lds_capacity is None else int(lds_capacity)
    o.rg_temp = float(rg_temp)
    o.rg_repeats = int(rg_repeats)
    o.rg_seed = int(rg_seed)
    record = np.ascontiguousarray(record, dtype=np.int32)
    info = PlanInfo()
    used = ctypes.c_int64(0)
    cap = 1 << 16
    while True:
        words = np.zeros(cap, dtype=np.int32)
        rc = lib.qb_compile_lightcone(_i32p(record), int(p), ctypes.byref(o), _i32p(words), cap, ctypes.byref(used),
                                      ctypes.byref(info))
        if rc == _lib.QB_ERR_UNSUPPORTED and used.value > cap:
            cap = used.value
            continue
        if rc == _lib.QB_ERR_UNSUPPORTED and b"Estimated OOM" in lib.qb_last_error():
            raise MemoryError("Estimated OOM")
        _lib.check(rc)
        break
    return Plan(words=words[:used.value].copy(), n_steps=info.n_steps, n_ops=info.n_ops, n_gates=info.n_gates,
                treewidth=info.treewidth, max_rank=info.max_rank, lds_need=int(info.lds_need),
                glb_need=int(info.glb_need), n_sliced=info.n_sliced, bytes_step_model=int(info.bytes_step_model),
                flops=int(info.flops), n_barriers=info.n_barriers, n_gather=info.n_gather, n_reduce=info.n_reduce,
                steps_words=info.steps_words)
