"""Batched QAOA energy engine: every edge lightcone of a graph in one persistent launch.

This is the batched counterpart of the per-edge loop of
``QAOASimulator.energy_expectation`` (``qtensor/QAOASimulator.py:44-66``): the
same per-edge ``<Z_u Z_v>`` networks (composer + greedy ordering +
bucket elimination), compiled once per canonical lightcone into static plans
(``plan.py``) and executed by ``qb_execute`` -- a persistent kernel whose
work items are (edge, slice) pairs -- followed by ``qb_reduce_terms`` (FP64,
compensated) for the per-edge terms and the total.

Gate tensors are rebuilt on the device from the request's (gamma, beta) in
FP64, so a compiled batch can be re-run with new angles without any host
work (``Batch.set_params`` + ``Batch.run``).  The launch sequence can also be
captured in a HIP graph (``Batch.capture``), though stream launches are
faster on ROCm here.
"""
import ctypes
import math
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib
from .composer import QtreeQAOAComposer
from .lightcone import canonical_graph, canonical_lightcone
from .ordering import (OrderingOptimizer, RGreedyOrderingOptimizer, RGreedySlicesOptimizer,
                       RGreedyTreeTrimSplitter, SlicesOptimizer, TreeTrimSplitter)
from .plan import PARAM_PAD, PROGRAM_PAD, PlanCompiler


def _torch():
    import torch
    return torch


@dataclass
class LaunchShape:
    item_threads: int = 64
    lds_capacity: int = 4096          # elements per item (complex64: 32 KiB)
    lds_tensor_limit: int = 2048      # largest intermediate kept in LDS
    group: int = 1                    # items of one plan per executor block, in lockstep

    def key(self):
        """What the compiled plans depend on (the group size only changes the
        launch)."""
        return (self.item_threads, self.lds_capacity, self.lds_tensor_limit)


DEFAULT_SHAPES = {
    "small": LaunchShape(64, 4096, 2048),
    "large": LaunchShape(64, 4096, 2048),
}


@dataclass
class EdgeWork:
    edge: tuple
    key: tuple
    n_slices: int


class GraphWorkload:
    """Plans for every edge of one graph at depth p (host side, reusable)."""

    def __init__(self, G, p, edges, works, plans):
        self.G = G
        self.p = p
        self.edges = edges
        self.works = works
        self.plans = plans      # key -> Plan
        self.shape_key = None   # launch shape the plans were compiled for (set by the server)

    @property
    def n_terms(self):
        return len(self.works)

    def step_model_bytes(self):
        return sum(self.plans[w.key].bytes_step_model * w.n_slices for w in self.works)

    def flops(self):
        return sum(self.plans[w.key].flops * w.n_slices for w in self.works)

    def max_treewidth(self):
        return max(self.plans[w.key].treewidth for w in self.works)


# compiled plans, shared by every engine in the process (plans do not depend on
# the angles or the precision, only on network structure and launch shape)
_PLAN_CACHE = {}


def clear_plan_cache():
    """Forget every compiled plan (the next workload plans from scratch)."""
    _PLAN_CACHE.clear()


_PLAN_LOCK = threading.Lock()
_POOL = None
_POOL_SIZE = 0


def planning_pool(workers):
    """Process pool for host planning (started once, spawn: the workers carry no
    GPU state).  Canonical lightcones and plan compilation are independent per
    edge / per distinct lightcone, as the reference's edge-parallel energy is
    (``qtensor/QAOASimulator.py:72-88``)."""
    global _POOL, _POOL_SIZE
    import multiprocessing as mp
    if _POOL is None or _POOL_SIZE != workers:
        if _POOL is not None:
            _POOL.terminate()
        _POOL = mp.get_context("spawn").Pool(workers)
        _POOL_SIZE = workers
    return _POOL


def _keys_task(args):
    G, edges, p = args
    return [(e,) + canonical_lightcone(G, e, p) for e in edges]


def _plan_task(args):
    composer_name, dtype, shape, max_tw, tw_bias, slicer, ordering, rg_repeats, tasks = args
    from . import composer as _composer
    eng = LightconeEngine(composer=getattr(_composer, composer_name), dtype=dtype, shape=shape,
                          max_tw=max_tw, tw_bias=tw_bias, slicer=slicer, ordering=ordering,
                          planner="python", rg_repeats=rg_repeats)
    out = []
    for key, n, ce, i, j, p in tasks:
        try:
            out.append((key, eng.plan(key, n, ce, i, j, p, shape), None))
        except ValueError as exc:
            out.append((key, None, str(exc)))
    return out


class LightconeEngine:
    """Compile lightcones and run them in batches on one GPU.

    ``composer``: QAOA composer class (default ``QtreeQAOAComposer``, the one
    ``qtensor.QAOA_energy`` uses).  ``max_tw``: lightcones whose greedy
    treewidth reaches it are sliced into 2^s work items, the slice variables
    chosen by ``slicer``: "greedy" (``SlicesOptimizer``, max-degree variables,
    qtensor/optimisation/Optimizer.py:90-159) or "tree_trim"
    (``TreeTrimSplitter``, Optimizer.py:178-228).  ``ordering``: "greedy"
    (min-degree, the reference's default ``OrderingOptimizer``) or "rgreedy"
    (best of ``rg_repeats`` seeded randomized-greedy samples, 16 by default,
    and the greedy order, ``qtensor/optimisation/RGreedy.py:15-77``), also the
    order the slicer works from.  Orders change the work, not the values (up
    to rounding).
    """

    def __init__(self, composer=QtreeQAOAComposer, device=0, dtype="complex64", shape=None,
                 max_tw=None, tw_bias=0, slicer="greedy", ordering="greedy", planner="native", rg_repeats=16):
        if slicer not in ("greedy", "tree_trim"):
            raise ValueError(f"unknown slicer {slicer!r}")
        if ordering not in ("greedy", "rgreedy"):
            raise ValueError(f"unknown ordering {ordering!r}")
        if planner not in ("native", "python"):
            raise ValueError(f"unknown planner {planner!r}")
        # "native": lightcone classes and plans from csrc/lightcone_planner.cpp
        # (same program words as the Python compiler, which "python" runs and
        # which other composers always use)
        self.planner = planner
        self.ordering = ordering
        if int(rg_repeats) < 1:
            raise ValueError("rg_repeats must be at least 1")
        self.rg_repeats = int(rg_repeats)
        self.composer = composer
        self.device_index = device
        self.dtype = dtype
        self.shape = shape
        self.max_tw = max_tw
        self.tw_bias = tw_bias
        self.slicer = slicer
        self._plans = _PLAN_CACHE
        self._lock = _PLAN_LOCK

    # -- host planning -----------------------------------------------------
    def _optimizer(self):
        rg = self.ordering == "rgreedy"
        if self.max_tw is None:
            opt = RGreedyOrderingOptimizer() if rg else OrderingOptimizer()
        else:
            if self.slicer == "tree_trim":
                cls = RGreedyTreeTrimSplitter if rg else TreeTrimSplitter
            else:
                cls = RGreedySlicesOptimizer if rg else SlicesOptimizer
            opt = cls(max_tw=self.max_tw, tw_bias=self.tw_bias)
        if rg:
            opt.repeats = self.rg_repeats
        return opt

    def _cache_key(self, key, p, shape):
        return (key, p, shape.key(), self.composer.__name__, self.max_tw, self.tw_bias, self.slicer,
                self.ordering, self.rg_repeats if self.ordering == "rgreedy" else 0)

    def _check_fits(self, plan, shape):
        """The plan's LDS data fit the launch shape in this engine's precision."""
        if plan.lds_need > shape.lds_capacity:
            raise ValueError(f"plan needs {plan.lds_need} LDS elements (> {shape.lds_capacity}); "
                             "its gate tensors alone do not fit the launch shape")
        elem = 16 if self.dtype == "complex128" else 8
        if plan.lds_elems(elem) * elem > 160 * 1024:
            raise ValueError("plan data exceed 160 KiB of LDS")

    def shape_for(self, p):
        if self.shape is not None:
            return self.shape
        return DEFAULT_SHAPES["small"] if p <= 2 else DEFAULT_SHAPES["large"]

    def plan(self, key, n, edges, i, j, p, shape):
        ck = self._cache_key(key, p, shape)
        with self._lock:
            hit = self._plans.get(ck)
        if hit is not None:
            # plans are shared across precisions: the LDS check depends on it
            self._check_fits(hit, shape)
            return hit
        g = canonical_graph(n, edges)
        comp = self.composer(g, gamma=[0.0] * p, beta=[0.0] * p)
        comp.energy_expectation(i, j)
        lim = shape.lds_tensor_limit
        while True:
            pc = PlanCompiler(lds_tensor_limit=lim, lds_capacity=shape.lds_capacity,
                              lanes=shape.item_threads)
            try:
                plan = pc.compile_circuit(comp.circuit, self._optimizer())
                break
            except ValueError as exc:
                # LDS arena too small: keep fewer tensors in LDS, spill the rest
                if "LDS capacity" not in str(exc) or lim <= shape.item_threads:
                    raise
                lim //= 2
        self._check_fits(plan, shape)
        with self._lock:
            self._plans[ck] = plan
        return plan

    def _native(self):
        from . import native_planner
        return self.planner == "native" and native_planner.supports(self.composer)

    def plan_native(self, key, record, p, shape):
        """``plan`` through the native planner (same plan, same cache)."""
        from . import native_planner
        ck = self._cache_key(key, p, shape)
        with self._lock:
            hit = self._plans.get(ck)
        if hit is not None:
            self._check_fits(hit, shape)
            return hit
        lim = shape.lds_tensor_limit
        while True:
            try:
                plan = native_planner.compile_lightcone(
                    record, p, self.composer, ordering=self.ordering,
                    slicer=self.slicer if self.max_tw is not None else None, max_tw=self.max_tw,
                    tw_bias=self.tw_bias, lanes=shape.item_threads, lds_tensor_limit=lim,
                    lds_capacity=shape.lds_capacity, rg_repeats=self.rg_repeats)
                break
            except ValueError as exc:
                if "LDS capacity" not in str(exc) or lim <= shape.item_threads:
                    raise
                lim //= 2
        self._check_fits(plan, shape)
        with self._lock:
            self._plans[ck] = plan
        return plan

    def workload(self, G, p, edges=None, workers=None):
        """Plans of every edge's lightcone.  The native planner computes the
        canonical lightcone classes of all edges in one call and compiles each
        class not yet cached (on ``workers`` threads, default the host's cores
        up to 16: the planner runs without the interpreter lock).  The Python planner, with
        ``workers`` > 1, computes canonical lightcones and plans in a process
        pool of that size."""
        shape = self.shape_for(p)
        edges = list(G.edges()) if edges is None else list(edges)
        if self._native():
            return self._workload_native(G, p, edges, shape, workers)
        if workers and workers > 1 and len(edges) > 1:
            self._plan_parallel(G, p, edges, shape, workers)
        keyed = self._keyed.pop((id(G), p), None) if hasattr(self, "_keyed") else None
        works, plans = [], {}
        for e in edges:
            if keyed is not None and e in keyed:
                key, n, ce, i, j = keyed[e]
            else:
                key, n, ce, i, j = canonical_lightcone(G, e, p)
            plan = self.plan(key, n, ce, i, j, p, shape)
            pk = (key, p)
            plans[pk] = plan
            works.append(EdgeWork(e, pk, 1 << plan.n_sliced))
        return GraphWorkload(G, p, edges, works, plans)

    def _workload_native(self, G, p, edges, shape, workers):
        from . import native_planner
        edge_class, classes = native_planner.lightcone_classes(G, edges, p)

        def one(c):
            key, n, ce, i, j, rec = c
            return self.plan_native(key, rec, p, shape)

        if workers is None:       # default: the host's cores, up to 16
            import os
            workers = min(16, len(os.sched_getaffinity(0)))
        if workers > 1 and len(classes) >= 8:
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(min(workers, len(classes))) as ex:
                plans_c = list(ex.map(one, classes))
        else:
            plans_c = [one(c) for c in classes]
        works, plans = [], {}
        for e, c in zip(edges, edge_class):
            pk = (classes[c][0], p)
            plan = plans_c[c]
            plans[pk] = plan
            works.append(EdgeWork(e, pk, 1 << plan.n_sliced))
        return GraphWorkload(G, p, edges, works, plans)

    def _plan_parallel(self, G, p, edges, shape, workers):
        pool = planning_pool(workers)
        chunk = max(1, -(-len(edges) // (4 * workers)))
        parts = pool.map(_keys_task, [(G, edges[k:k + chunk], p) for k in range(0, len(edges), chunk)])
        keyed = {row[0]: row[1:] for part in parts for row in part}
        todo, seen = [], set()
        for e in edges:
            key, n, ce, i, j = keyed[e]
            ck = self._cache_key(key, p, shape)
            if key in seen:
                continue
            seen.add(key)
            with self._lock:
                if ck in self._plans:
                    continue
            todo.append((key, n, ce, i, j, p))
        if todo:
            chunk = max(1, -(-len(todo) // (2 * workers)))
            jobs = [(self.composer.__name__, self.dtype, shape, self.max_tw, self.tw_bias, self.slicer,
                     self.ordering, self.rg_repeats, todo[k:k + chunk])
                    for k in range(0, len(todo), chunk)]
            for part in pool.map(_plan_task, jobs):
                for key, plan, err in part:
                    if plan is not None:
                        with self._lock:
                            self._plans[self._cache_key(key, p, shape)] = plan
        if not hasattr(self, "_keyed"):
            self._keyed = {}
        self._keyed[(id(G), p)] = keyed

    # -- device batches -----------------------------------------------------
    def batch(self, requests):
        """Device batch for ``requests = [(workload, gamma, beta), ...]``."""
        return Batch(self, requests)

    def energy(self, G, gamma, beta, edges=None):
        """MaxCut expectation and per-edge <Z_u Z_v> terms (complex) of one graph."""
        p = len(gamma)
        wl = self.workload(G, p, edges)
        b = self.batch([(wl, gamma, beta)])
        b.run()
        terms = b.terms()
        zz = terms[0]
        C = (G.number_of_edges() - math.fsum(z.real for z in zz)) / 2
        return C, zz


def _cu_count(device):
    return _torch().cuda.get_device_properties(device).multi_processor_count


def order_items(items, costs, band=None):
    """Work items in launch order: longest first (by ``costs``, stable), each
    carrying its output slot (its index in the given order) in word 3.  With
    ``band`` (the device's CU count), every other run of ``band`` items is
    reversed: blocks b and b + band tend to share a CU, so each CU then gets a
    long item of one run with a short one of the next."""
    items = np.array(items, dtype=np.int32).reshape(-1, 4)
    items[:, 3] = np.arange(items.shape[0], dtype=np.int32)
    order = np.argsort(-np.asarray(costs, dtype=np.int64), kind="stable")
    if band:
        for k in range(band, len(order), 2 * band):
            order[k:k + band] = order[k:k + band][::-1].copy()
    return np.ascontiguousarray(items[order])


def group_items(items, costs, group, band=None):
    """Work items in launch order for blocks of ``group`` items each: items of
    one plan (word 0) form groups of ``group`` consecutive items, the last group
    of a plan padded with copies of its last item whose output slots follow the
    real items' (they belong to no term).  Groups are ordered longest first and
    serpentine over ``band`` groups, as ``order_items`` orders single items.
    Returns the items (slot in word 3) and the number of padding items."""
    if group <= 1:
        return order_items(items, costs, band), 0
    items = np.array(items, dtype=np.int32).reshape(-1, 4)
    items[:, 3] = np.arange(items.shape[0], dtype=np.int32)
    costs = np.asarray(costs, dtype=np.int64)
    by_plan = {}
    for k in range(items.shape[0]):
        by_plan.setdefault(int(items[k, 0]), []).append(k)
    groups, gcost, pad = [], [], []
    next_slot = items.shape[0]
    for members in by_plan.values():
        for a in range(0, len(members), group):
            chunk = [items[k] for k in members[a:a + group]]
            while len(chunk) < group:
                extra = chunk[-1].copy()
                extra[3] = next_slot
                next_slot += 1
                chunk.append(extra)
                pad.append(1)
            groups.append(np.stack(chunk))
            gcost.append(int(costs[members[a]]))
    order = np.argsort(-np.asarray(gcost, dtype=np.int64), kind="stable")
    if band:
        for k in range(band, len(order), 2 * band):
            order[k:k + band] = order[k:k + band][::-1].copy()
    out = np.concatenate([groups[g] for g in order]).astype(np.int32)
    return np.ascontiguousarray(out), len(pad)


class Batch:
    """Device buffers and launch of one set of (workload, angles) requests."""

    def __init__(self, engine, requests):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the batched engine needs a ROCm GPU")
        self.engine = engine
        self.lib = _lib.init(engine.device_index)
        dev = torch.device("cuda", engine.device_index)
        self.device = dev
        self.qb_dtype = _lib.QB_C64 if engine.dtype == "complex64" else _lib.QB_C128
        elem = 8 if self.qb_dtype == _lib.QB_C64 else 16
        shapes = {engine.shape_for(wl.p).key() for wl, _, _ in requests}
        if len(shapes) != 1:
            raise ValueError("all requests of a batch must use one launch shape")
        self.shape = engine.shape_for(requests[0][0].p)
        # program: every distinct plan once
        plan_off, words = {}, []
        pos = 0
        for wl, _, _ in requests:
            for pk, plan in wl.plans.items():
                ident = id(plan)
                if ident not in plan_off:
                    plan_off[ident] = pos
                    words.append(plan.words)
                    pos += plan.words.size
        words.append(np.zeros(PROGRAM_PAD, dtype=np.int32))   # record fetches read past the last op
        program = np.concatenate(words).astype(np.int32)
        # params, items, segments (one segment per edge term)
        params, items, seg = [], [], [0]
        self.term_slices = []
        order = []
        lds_need, glb_need = 0, 0
        self.bytes_step_model = 0
        self.flops = 0
        for r, (wl, gamma, beta) in enumerate(requests):
            if len(gamma) != wl.p or len(beta) != wl.p:
                raise ValueError("gamma and beta must have p entries")
            poff = len(params)
            params += [float(x) for x in gamma] + [float(x) for x in beta]
            start = len(seg) - 1
            for w in wl.works:
                plan = wl.plans[w.key]
                lds_need = max(lds_need, plan.lds_elems(elem))
                glb_need = max(glb_need, plan.glb_need)
                self.bytes_step_model += plan.bytes_step_model * w.n_slices
                self.flops += plan.flops * w.n_slices
                for s in range(w.n_slices):
                    items.append((plan_off[id(plan)], poff, s, 0))
                    order.append(plan.n_ops)
                seg.append(len(items))
            self.term_slices.append((start, len(seg) - 1))
        self.n_terms = len(seg) - 1
        # no slicing anywhere: item slot s holds term s and the reduction needs
        # no segment table (one launch)
        self.one_item_per_term = len(items) == self.n_terms
        self.lds_elems = max(lds_need, 2)
        self.scratch_elems = int(glb_need)
        self.group = max(1, int(self.shape.group))
        slots = ctypes.c_int32(0)
        _lib.check(self.lib.qb_exec_slots(self.shape.item_threads, self.lds_elems, self.qb_dtype, self.group,
                                          ctypes.byref(slots)))
        self.n_slots = slots.value
        band = _cu_count(self.device)
        items_np, self.n_padding = group_items(items, order, self.group, band=band)
        # every item the launch runs, padding included
        self.n_items = items_np.shape[0]
        # qb_execute trusts the program and items: check them on the host first
        _lib.check_program(program, items_np, len(params), self.shape.item_threads, self.lds_elems,
                           self.scratch_elems, self.group)
        self._host = (program, items_np, np.asarray(seg, dtype=np.int32), len(params))
        self.program = torch.from_numpy(program).to(dev)
        self.items = torch.from_numpy(np.ascontiguousarray(items_np)).to(dev)
        # the executor reads 8 angles from every item's offset in one load
        self.params = torch.tensor(params + [0.0] * PARAM_PAD, dtype=torch.float64, device=dev)
        self.n_params = len(params)
        self.seg = torch.tensor(seg, dtype=torch.int32, device=dev)
        self.item_out = torch.zeros(2 * self.n_items, dtype=torch.float64, device=dev)
        self.term_out = torch.zeros(2 * self.n_terms, dtype=torch.float64, device=dev)
        self.total = torch.zeros(2, dtype=torch.float64, device=dev)
        self.queue = torch.zeros(1, dtype=torch.int32, device=dev)
        self.pool = torch.zeros(2, dtype=torch.float64, device=dev)
        launched = max(self.n_items, 1)
        self.scratch_slots = min(self.n_slots, launched) if self.scratch_elems else 0
        n_scr = self.scratch_elems * self.scratch_slots
        self.scratch = torch.empty(max(n_scr, 1) * (elem // 8), dtype=torch.complex64, device=dev)
        self.graph = None
        self._desc = self._make_desc()

    def _make_desc(self):
        d = _lib.ExecDesc()
        d.program = self.program.data_ptr()
        d.items = self.items.data_ptr()
        d.n_items = self.n_items
        d.item_threads = self.shape.item_threads
        d.lds_elems = self.lds_elems
        d.dtype = self.qb_dtype
        d.params = self.params.data_ptr()
        d.const_pool = self.pool.data_ptr()
        d.scratch = self.scratch.data_ptr()
        d.scratch_elems = self.scratch_elems
        d.n_slots = self.scratch_slots if self.scratch_elems else self.n_slots
        d.program_words = self.program.numel()
        d.item_out = self.item_out.data_ptr()
        d.queue = self.queue.data_ptr()
        d.group = self.group
        return d

    def set_params(self, requests_params):
        """New (gamma, beta) per request, same structure; no host planning."""
        flat = []
        for gamma, beta in requests_params:
            flat += [float(x) for x in gamma] + [float(x) for x in beta]
        if len(flat) != self.n_params:
            raise ValueError("parameter count does not match the batch")
        self.params[:self.n_params].copy_(_torch().tensor(flat, dtype=_torch().float64))

    def launch_exec(self, stream=None):
        """Launch only the executor (one persistent kernel)."""
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        _lib.check(self.lib.qb_execute(ctypes.byref(self._desc), ctypes.c_void_p(st.cuda_stream)))

    def launch(self, stream=None):
        torch = _torch()
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        h = ctypes.c_void_p(st.cuda_stream)
        _lib.check(self.lib.qb_execute(ctypes.byref(self._desc), h))
        _lib.check(self.lib.qb_reduce_terms(ctypes.c_void_p(self.item_out.data_ptr()),
                                            None if self.one_item_per_term else ctypes.c_void_p(self.seg.data_ptr()),
                                            self.n_terms,
                                            ctypes.c_void_p(self.term_out.data_ptr()),
                                            ctypes.c_void_p(self.total.data_ptr()), h))

    def capture(self):
        """Capture the launch sequence in a HIP graph (replayed by ``run``)."""
        torch = _torch()
        self.launch()     # warm up outside capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        return g

    def run(self, sync=True):
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()
        if sync:
            _torch().cuda.synchronize(self.device)

    def terms(self):
        """Per request: list of complex edge terms (FP64)."""
        t = self.term_out.cpu().numpy().reshape(-1, 2)
        out = []
        for s, e in self.term_slices:
            out.append([complex(a, b) for a, b in t[s:e]])
        return out

    def item_values(self):
        return self.item_out.cpu().numpy().reshape(-1, 2)


class PlanHandle:
    """A batch's program and items behind the C-ABI plan handle
    (``qb_plan_create`` / ``qb_plan_execute`` / ``qb_plan_destroy``): the
    library owns every device buffer; only the angle buffer is passed in."""

    def __init__(self, batch):
        program, items, seg, n_params = batch._host
        self.lib = batch.lib
        self.device = batch.device
        self.n_terms = batch.n_terms
        self._arrays = (np.ascontiguousarray(program, dtype=np.int32), np.ascontiguousarray(items, dtype=np.int32),
                        np.ascontiguousarray(seg, dtype=np.int32))
        d = _lib.PlanDesc()
        d.program = self._arrays[0].ctypes.data
        d.program_words = self._arrays[0].size
        d.items = self._arrays[1].ctypes.data
        d.n_items = batch.n_items
        d.n_terms = batch.n_terms
        d.seg = self._arrays[2].ctypes.data
        d.n_params = n_params
        d.item_threads = batch.shape.item_threads
        d.lds_elems = batch.lds_elems
        d.dtype = batch.qb_dtype
        d.scratch_elems = batch.scratch_elems
        d.group = batch.group
        h = ctypes.c_void_p()
        _lib.check(self.lib.qb_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h

    def execute(self, params, stream=None):
        """Complex terms and total for ``params`` (a device float64 tensor with
        at least n_params + 8 entries)."""
        torch = _torch()
        term_out = torch.empty(2 * self.n_terms, dtype=torch.float64, device=self.device)
        total = torch.empty(2, dtype=torch.float64, device=self.device)
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        _lib.check(self.lib.qb_plan_execute(self.handle, ctypes.c_void_p(params.data_ptr()),
                                            ctypes.c_void_p(term_out.data_ptr()),
                                            ctypes.c_void_p(total.data_ptr()), ctypes.c_void_p(st.cuda_stream)))
        st.synchronize()
        t = term_out.cpu().numpy().reshape(-1, 2)
        tot = total.cpu().numpy()
        return [complex(a, b) for a, b in t], complex(tot[0], tot[1])

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.check(self.lib.qb_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

