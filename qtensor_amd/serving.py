"""Continuous-batching server for QAOA energy requests on one GPU.

Requests ``(G, gamma, beta)`` arrive at any time.  A scheduler thread
compiles each new graph's lightcone plans once (plans are cached by canonical
lightcone and uploaded once into a device-resident plan store), admits
pending requests into the next batch while their predicted memory fits the
budget, and launches the batch -- one persistent ``qb_execute`` plus
``qb_reduce_terms`` -- on one of several streams.  Batches on different
streams overlap on the device; a request that arrives while others run joins
the next batch as soon as a stream slot and memory are free.  A completion
thread turns finished batches into results and latencies.

Each work item is contracted independently and reduced in a fixed order, so a
request's terms are the same whether it runs alone or in a mix.

Admission: a request's predicted footprint is ``sum over edges of 2^tw * 8 B``
(times the slice count), the reference's memory model (``Optimizer.py:99-107``);
requests are admitted while in-flight footprints stay within ``hbm_budget``
(default: a fraction of free HBM), and a request that can never fit is
rejected with ``MemoryError`` (the reference's "Estimated OOM",
``Optimizer.py:122-124``).
"""
import ctypes
import math
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .batched import group_items
from .plan import PARAM_PAD, PROGRAM_PAD


def predicted_bytes(workload, elem_bytes=8):
    """Sum over edges of 2^tw * elem_bytes * slices (the admission model)."""
    total = 0
    for w in workload.works:
        plan = workload.plans[w.key]
        total += (1 << max(plan.max_rank - 1, 0)) * elem_bytes * w.n_slices
    return total


class Admission:
    """Byte budget shared by in-flight requests (host-only logic)."""

    def __init__(self, budget_bytes):
        self.budget = int(budget_bytes)
        self.in_flight = 0
        self._lock = threading.Lock()

    def fits_alone(self, nbytes):
        return nbytes <= self.budget

    def try_admit(self, nbytes):
        with self._lock:
            if self.in_flight + nbytes > self.budget:
                return False
            self.in_flight += nbytes
            return True

    def release(self, nbytes):
        with self._lock:
            self.in_flight -= nbytes
            assert self.in_flight >= 0


@dataclass
class Request:
    G: object
    gamma: list
    beta: list
    key: object
    future: Future
    t_arrival: float
    workload: object = None
    nbytes: int = 0
    t_launch: float = 0.0
    t_done: float = 0.0


@dataclass
class _Slot:
    """Per-stream device buffers, reused batch after batch."""
    stream: object
    cap_items: int = 0
    cap_terms: int = 0
    cap_params: int = 0
    scratch_elems: int = 0
    scratch_slots: int = 0
    busy: bool = False
    event: object = None
    tensors: dict = field(default_factory=dict)


class PlanStore:
    """Device-resident program buffer; every distinct plan is uploaded once.
    When full, the buffer grows (doubling) and the old one is kept until the
    store is dropped, since batches in flight may still read it."""

    def __init__(self, device, capacity_words=1 << 22):
        import torch
        self.device = device
        self.buf = torch.zeros(capacity_words + PROGRAM_PAD, dtype=torch.int32, device=device)
        self.used = 0
        self.offsets = {}
        self.retired = []

    def offset(self, plan):
        import torch
        off = self.offsets.get(id(plan))
        if off is not None:
            return off
        n = plan.words.size
        if self.used + n + PROGRAM_PAD > self.buf.numel():
            cap = max(2 * (self.buf.numel() - PROGRAM_PAD), self.used + n)
            new = torch.zeros(cap + PROGRAM_PAD, dtype=torch.int32, device=self.device)
            new[:self.used].copy_(self.buf[:self.used])
            self.retired.append(self.buf)
            self.buf = new
        self.buf[self.used:self.used + n].copy_(torch.from_numpy(plan.words))
        self.offsets[id(plan)] = self.used
        self.used += n
        return self.offsets[id(plan)]


class EnergyServer:
    """Serve ``submit(G, gamma, beta)`` requests with continuous batching.

    ``engine``: a ``LightconeEngine`` (its launch shape and slicing are used for
    every batch).  ``n_streams``: batches that may be in flight at once.
    ``max_batch_items``: work items per batch.  ``hbm_budget``: bytes for the
    admission model (None: 80% of free HBM at start).  ``engines``: optional
    ``{p: LightconeEngine}`` overriding ``engine`` for some depths.
    ``planning_workers``: process-pool size for planning new graphs.
    ``max_slice_vars``: a request whose lightcones need more slice variables
    than this is rejected (``MemoryError("Estimated OOM ...")``, as the
    reference's slicer stops when it cannot reach its width), which bounds the
    device time one request can take; None: up to what a work item carries.
    ``upgrades``: optional ``{p: LightconeEngine}``: a (graph, p) is served
    first with the plans of its usual engine, which are quick to make, and
    planned again in the background with the upgrade engine (for example a
    randomized-greedy order sliced further, slower to plan but faster to run);
    requests that arrive after that plan is done use it.  Each request runs
    whole with one of the two plans.
    """

    def __init__(self, engine, n_streams=2, max_batch_items=1 << 16, hbm_budget=None, engines=None,
                 planning_workers=None, max_slice_vars=None, upgrades=None):
        self.engine = engine
        # per-p engines (launch shape and slicing), e.g. wider items and less
        # slicing for p = 4; requests of one batch share one launch shape
        self.engines = dict(engines or {})
        self.planning_workers = planning_workers
        self.max_slice_vars = max_slice_vars
        self.upgrades = dict(upgrades or {})
        self.upgraded = 0
        self.rejected = 0
        self.n_streams = n_streams
        self.max_batch_items = max_batch_items
        self.hbm_budget = hbm_budget
        self._pending = queue.Queue()
        self._done_q = queue.Queue()
        self._workloads = {}
        self._planning = {}           # (key, p) -> Future of a workload being planned
        self._stop = threading.Event()
        self._progress = threading.Event()
        self._lock = threading.Lock()
        self.finished = []
        self._threads = []

    # -- public API ------------------------------------------------------------
    def start(self):
        import torch
        self.device = torch.device("cuda", self.engine.device_index)
        torch.cuda.set_device(self.device)
        self.lib = _lib.init(self.engine.device_index)
        self.qb_dtype = _lib.QB_C64 if self.engine.dtype == "complex64" else _lib.QB_C128
        self.elem = 8 if self.qb_dtype == _lib.QB_C64 else 16
        if self.hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            self.hbm_budget = int(0.8 * free)
        self.admission = Admission(self.hbm_budget)
        self.store = PlanStore(self.device)
        self.slots = [_Slot(stream=torch.cuda.Stream(device=self.device)) for _ in range(self.n_streams)]
        self._slot_free = threading.Semaphore(self.n_streams)
        from concurrent.futures import ThreadPoolExecutor
        # one planner thread per engine (per launch configuration), so that
        # slow plans of one engine (or plans that end up rejected) do not hold
        # up the quick plans of another; each workload's classes compile on
        # the native planner's own threads
        self._planners = {}
        for eng in [self.engine] + list(self.engines.values()) + list(self.upgrades.values()):
            if id(eng) not in self._planners:
                self._planners[id(eng)] = ThreadPoolExecutor(1)
        self._threads = [threading.Thread(target=self._schedule_loop, daemon=True),
                         threading.Thread(target=self._complete_loop, daemon=True)]
        for t in self._threads:
            t.start()
        return self

    def submit(self, G, gamma, beta, key=None):
        """Queue one energy request; returns a Future of ``(C, terms)``."""
        fut = Future()
        self._pending.put(Request(G, list(gamma), list(beta), key if key is not None else id(G), fut,
                                  time.perf_counter()))
        return fut

    def stop(self):
        self._stop.set()
        self._pending.put(None)
        for t in self._threads:
            t.join(timeout=60)
        for pool in getattr(self, "_planners", {}).values():
            pool.shutdown(wait=True)

    def wait_upgrades(self, timeout=None):
        """Block until the upgrade plans queued so far are done."""
        for eng in self.upgrades.values():
            self._planners[id(eng)].submit(lambda: None).result(timeout=timeout)

    def stats(self):
        with self._lock:
            done = list(self.finished)
        if not done:
            return {}
        lat = np.array([r.t_done - r.t_arrival for r in done]) * 1e3
        terms = sum(r.workload.n_terms for r in done)
        span = max(r.t_done for r in done) - min(r.t_arrival for r in done)
        return {"requests": len(done), "rejected": self.rejected, "terms": terms,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "terms_per_s": terms / span if span > 0 else float("nan")}

    # -- scheduler ---------------------------------------------------------------
    def engine_for(self, p):
        return self.engines.get(p, self.engine)

    def _plan_with(self, eng, G, p, workers=None):
        try:
            wl = eng.workload(G, p, workers=self.planning_workers if workers is None else workers)
        except ValueError as exc:
            # e.g. more slice variables than a work item can carry: the
            # reference's planner stops the same way ("Estimated OOM")
            raise MemoryError(f"Estimated OOM: {exc}") from exc
        wl.shape = eng.shape_for(p)
        wl.shape_key = wl.shape.key()
        if self.max_slice_vars is not None:
            need = max((pl.n_sliced for pl in wl.plans.values()), default=0)
            if need > self.max_slice_vars:
                raise MemoryError(f"Estimated OOM: a lightcone needs {need} slice variables "
                                  f"(limit {self.max_slice_vars})")
        return wl

    def _plan(self, G, p, k):
        """Plan one (graph, depth) on its engine's planner thread; with an
        upgrade engine for this depth, then plan it again with that one on its
        own thread, and serve later requests with the upgraded plans.  When the
        quick plans exceed the limits (slice variables), the upgrade engine's
        plans are made at once instead and the request waits for them."""
        up = self.upgrades.get(p)
        try:
            wl = self._plan_with(self.engine_for(p), G, p)
        except MemoryError:
            if up is None:
                raise
            wl = self._plan_with(up, G, p)
            self._workloads[k] = wl
            self.upgraded += 1
            self._progress.set()
            return wl
        self._workloads[k] = wl
        self._progress.set()
        if up is not None:
            self._planners[id(up)].submit(self._upgrade, up, G, p, k)
        return wl

    def _upgrade(self, eng, G, p, k):
        # background work: a quarter of the planning threads, so that the
        # quick first plans of new requests keep the rest of the host
        workers = max(1, (self.planning_workers or 4) // 4)
        try:
            wl = self._plan_with(eng, G, p, workers)
        except MemoryError:
            return                      # keep the first plans
        self._workloads[k] = wl
        self.upgraded += 1

    def _workload(self, req):
        """The request's workload if planned; otherwise start planning it
        (once per graph and depth) on the planner thread and return None."""
        p = len(req.gamma)
        k = (req.key, p)
        wl = self._workloads.get(k)
        if wl is not None:
            return wl
        fut = self._planning.get(k)
        if fut is None:
            fut = self._planners[id(self.engine_for(p))].submit(self._plan, req.G, p, k)
            self._planning[k] = fut
        if not fut.done():
            return None
        return fut.result()          # raises the planning error, if any

    def _schedule_loop(self):
        import torch
        torch.cuda.set_device(self.device)
        waiting = []
        while True:
            # gather everything that has arrived
            block = not waiting
            try:
                item = self._pending.get(timeout=0.05 if block else 0.0)
                if item is None:
                    if not waiting:
                        break
                else:
                    waiting.append(item)
                while True:
                    item = self._pending.get_nowait()
                    if item is not None:
                        waiting.append(item)
            except queue.Empty:
                pass
            if self._stop.is_set() and not waiting:
                break
            if not waiting:
                continue
            # new graphs are planned on the planner thread while planned
            # requests go on to the device; a request waits only for its own plan
            ready, planning = [], []
            for req in waiting:
                if req.workload is None:
                    try:
                        wl = self._workload(req)
                    except Exception as exc:      # planning error goes to the caller
                        self.rejected += 1
                        req.future.set_exception(exc)
                        continue
                    if wl is None:
                        planning.append(req)
                        continue
                    req.workload = wl
                    req.nbytes = predicted_bytes(wl, self.elem)
                    if not self.admission.fits_alone(req.nbytes):
                        self.rejected += 1
                        req.future.set_exception(MemoryError(
                            f"Estimated OOM: request needs {req.nbytes} B, budget {self.admission.budget} B"))
                        continue
                ready.append(req)
            waiting = ready + planning
            if not ready:
                if planning:          # nothing to launch until a plan is done
                    self._progress.clear()
                    self._progress.wait(timeout=0.005)
                continue
            if not self._slot_free.acquire(timeout=0.05):
                continue
            self._progress.clear()
            batch, n_items = [], 0
            shape_key = ready[0].workload.shape_key
            for req in list(ready):
                if req.workload.shape_key != shape_key:
                    continue              # another launch shape: a later batch
                items = sum(w.n_slices for w in req.workload.works)
                if batch and n_items + items > self.max_batch_items:
                    break
                if not self.admission.try_admit(req.nbytes):
                    break
                batch.append(req)
                n_items += items
                waiting.remove(req)
            if not batch:
                # memory is held by batches in flight: wait for one to finish
                self._slot_free.release()
                self._progress.wait(timeout=0.05)
                continue
            slot = next(s for s in self.slots if not s.busy)
            slot.busy = True
            try:
                self._launch(slot, batch)
            except Exception as exc:
                for r in batch:
                    self.admission.release(r.nbytes)
                    r.future.set_exception(exc)
                slot.busy = False
                self._slot_free.release()
        self._done_q.put(None)

    def _ensure(self, slot, name, n, dtype):
        import torch
        t = slot.tensors.get(name)
        if t is None or t.numel() < n:
            t = torch.empty(max(n, 1) * 2, dtype=dtype, device=self.device)
            slot.tensors[name] = t
            hn = name + "_host"
            slot.tensors[hn] = torch.empty(t.numel(), dtype=dtype, pin_memory=True)
        return slot.tensors[name], slot.tensors[name + "_host"]

    def _items_of(self, req):
        """Cached per-workload item rows (plan offset, 0, slice, 0) and segment sizes."""
        wl = req.workload
        cached = getattr(wl, "_serving_items", None)
        if cached is not None and cached[0] is self.store:
            return cached[1:]
        rows, sizes, costs = [], [], []
        lds_need, glb_need = 2, 0
        for w in wl.works:
            plan = wl.plans[w.key]
            off = self.store.offset(plan)
            lds_need = max(lds_need, plan.lds_elems(self.elem))
            glb_need = max(glb_need, plan.glb_need)
            rows += [(off, 0, s, 0) for s in range(w.n_slices)]
            costs += [plan.n_ops] * w.n_slices
            sizes.append(w.n_slices)
        arr = np.asarray(rows, dtype=np.int32).reshape(-1, 4)
        wl._serving_items = (self.store, arr, np.asarray(sizes, dtype=np.int64), lds_need, glb_need,
                             np.asarray(costs, dtype=np.int64))
        return wl._serving_items[1:]

    def _launch(self, slot, batch):
        import torch
        parts, sizes, params, costs = [], [], [], []
        lds_need, glb_need = 2, 0
        for req in batch:
            rows, seg_sizes, ln, gn, cost = self._items_of(req)
            rows = rows.copy()
            rows[:, 1] = len(params)
            params += req.gamma + req.beta
            parts.append(rows)
            costs.append(cost)
            sizes.append(seg_sizes)
            lds_need, glb_need = max(lds_need, ln), max(glb_need, gn)
        # the launch shape the batch's plans were compiled for (every request of
        # a batch has the same shape key; an upgrade engine may differ from the
        # depth's usual engine)
        p0 = len(batch[0].gamma)
        shape = getattr(batch[0].workload, "shape", None) or self.engine_for(p0).shape_for(p0)
        group = max(1, int(shape.group))
        # items of one plan (requests on one graph) share a block in groups
        items, _ = group_items(np.concatenate(parts), np.concatenate(costs), group,
                               band=torch.cuda.get_device_properties(self.device).multi_processor_count)
        params += [0.0] * PARAM_PAD      # the executor reads 8 angles per item at once
        seg = np.concatenate([[0], np.cumsum(np.concatenate(sizes))]).astype(np.int32)
        n_items, n_terms = items.shape[0], seg.size - 1
        n_real = int(seg[-1])            # items of terms (the rest pad the last groups)
        # newly stored plans were copied on this thread's stream
        torch.cuda.current_stream(self.device).synchronize()
        slots_c = ctypes.c_int32(0)
        _lib.check(self.lib.qb_exec_slots(shape.item_threads, lds_need, self.qb_dtype, group,
                                          ctypes.byref(slots_c)))
        used_slots = min(slots_c.value, n_items)
        d_items, h_items = self._ensure(slot, "items", 4 * n_items, torch.int32)
        d_params, h_params = self._ensure(slot, "params", len(params), torch.float64)
        d_seg, h_seg = self._ensure(slot, "seg", len(seg), torch.int32)
        d_out, _ = self._ensure(slot, "item_out", 2 * n_items, torch.float64)
        d_terms, h_terms = self._ensure(slot, "terms", 2 * n_terms, torch.float64)
        d_total, _ = self._ensure(slot, "total", 2, torch.float64)
        d_queue = slot.tensors.get("queue")
        if d_queue is None:          # reset by qb_execute whenever a launch takes tickets
            d_queue = slot.tensors["queue"] = torch.zeros(1, dtype=torch.int32, device=self.device)
        scr_elems = glb_need * used_slots if glb_need else 1
        d_scr, _ = self._ensure(slot, "scratch", scr_elems * (self.elem // 8), torch.complex64)
        h_items[:4 * n_items].copy_(torch.from_numpy(items.reshape(-1)))
        h_params[:len(params)].copy_(torch.tensor(params, dtype=torch.float64))
        h_seg[:len(seg)].copy_(torch.from_numpy(seg))
        st = slot.stream
        now = time.perf_counter()
        for r in batch:
            r.t_launch = now
        with torch.cuda.stream(st):
            d_items[:4 * n_items].copy_(h_items[:4 * n_items], non_blocking=True)
            d_params[:len(params)].copy_(h_params[:len(params)], non_blocking=True)
            d_seg[:len(seg)].copy_(h_seg[:len(seg)], non_blocking=True)
            d = _lib.ExecDesc()
            d.program = self.store.buf.data_ptr()
            d.items = d_items.data_ptr()
            d.n_items = n_items
            d.item_threads = shape.item_threads
            d.lds_elems = lds_need
            d.dtype = self.qb_dtype
            d.params = d_params.data_ptr()
            d.const_pool = d_params.data_ptr()
            d.scratch = d_scr.data_ptr()
            d.scratch_elems = glb_need
            d.n_slots = used_slots if glb_need else slots_c.value
            d.program_words = self.store.buf.numel()
            d.item_out = d_out.data_ptr()
            d.queue = d_queue.data_ptr()
            d.group = group
            h = ctypes.c_void_p(st.cuda_stream)
            _lib.check(self.lib.qb_execute(ctypes.byref(d), h))
            _lib.check(self.lib.qb_reduce_terms(ctypes.c_void_p(d_out.data_ptr()),
                                                None if n_real == n_terms else ctypes.c_void_p(d_seg.data_ptr()),
                                                n_terms,
                                                ctypes.c_void_p(d_terms.data_ptr()),
                                                ctypes.c_void_p(d_total.data_ptr()), h))
            h_terms[:2 * n_terms].copy_(d_terms[:2 * n_terms], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        slot.event = ev
        self._done_q.put((slot, batch, n_terms))

    # -- completion ----------------------------------------------------------------
    def _complete_loop(self):
        while True:
            item = self._done_q.get()
            if item is None:
                break
            slot, batch, n_terms = item
            slot.event.synchronize()
            t = slot.tensors["terms_host"][:2 * n_terms].numpy().reshape(-1, 2).copy()
            now = time.perf_counter()
            pos = 0
            for req in batch:
                n = req.workload.n_terms
                zz = [complex(a, b) for a, b in t[pos:pos + n]]
                pos += n
                C = (req.G.number_of_edges() - math.fsum(z.real for z in zz)) / 2
                req.t_done = now
                self.admission.release(req.nbytes)
                with self._lock:
                    self.finished.append(req)
                req.future.set_result((C, zz))
            slot.busy = False
            self._slot_free.release()
            self._progress.set()


def poisson_trace(server, pool, rate, n_requests, rng, p_choices=(1, 2, 3)):
    """Submit ``n_requests`` at Poisson times (``rate`` per second) drawing a graph
    from ``pool`` (list of (key, G)) and random angles; wait for all results.
    Returns the list of futures with their requests' parameters."""
    out = []
    t0 = time.perf_counter()
    t_next = t0
    for _ in range(n_requests):
        t_next += rng.exponential(1.0 / rate)
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        key, G = pool[int(rng.integers(len(pool)))]
        p = int(rng.choice(p_choices))
        gamma = list(rng.uniform(0, np.pi, p))
        beta = list(rng.uniform(0, np.pi / 2, p))
        out.append((server.submit(G, gamma, beta, key=key), key, G, gamma, beta))
    for f, *_ in out:
        try:
            f.result()
        except MemoryError:      # rejected at admission ("Estimated OOM"); counted by the server
            pass
    return out


def trace_failures(out):
    """Messages of the requests of a trace that ended in an exception."""
    msgs = []
    for f, *_ in out:
        exc = f.exception()
        if exc is not None:
            msgs.append(f"{type(exc).__name__}: {exc}")
    return msgs
