"""Autotuner for the batched executor: launch shape and slice count per workload.

For one (graph, p) workload it compiles the lightcone plans under each
candidate (item threads, LDS capacity, LDS tensor limit, slicing treewidth,
slice-variable rule) and times one evaluation (the executor launch and the
term reductions, as a caller runs it) on the device with HIP events, keeping
the fastest.  The search runs in two stages: every slicing
threshold below the workload's largest greedy treewidth, with each slicer,
under a base shape; then the other shapes (item width, LDS budget per item)
at the three best (threshold, slicer) pairs.  Every candidate is compiled under
the tuner's elimination order (``ordering``: "greedy" or "rgreedy", as in
LightconeEngine).

Results are memoised per (graph, p, composer, dtype, order) in-process and
can be saved to a JSON file.
"""
import json
import sys
import time
from dataclasses import asdict

import numpy as np

from .batched import LaunchShape, LightconeEngine


def _signature(G, p, composer, dtype, ordering):
    """Memo key of a tuning result: the exact graph (node order and edge set,
    so two different graphs of one size and degree sequence are tuned
    separately), depth, composer, precision and order."""
    nodes = list(G.nodes())
    pos = {v: k for k, v in enumerate(nodes)}
    edges = tuple(sorted(tuple(sorted((pos[a], pos[b]))) for a, b in G.edges()))
    return (len(nodes), edges, p, composer.__name__, dtype, ordering)


BASE_SHAPE = LaunchShape(128, 4096, 2048)
# the last two run two items of one plan per block (more items resident, so
# finer slicing and shorter op chains fit one round)
OTHER_SHAPES = (LaunchShape(128, 2048, 1024), LaunchShape(256, 4096, 2048), LaunchShape(256, 2048, 1024),
                LaunchShape(64, 4096, 2048), LaunchShape(256, 8192, 4096), LaunchShape(256, 4096, 2048, group=2),
                LaunchShape(128, 4096, 2048, group=2))
# slice-variable rules (LightconeEngine's ``slicer``)
SLICERS = ("greedy", "tree_trim")
# (threshold, slicer) pairs of stage 1 timed again under every other shape:
# the ranking changes with the shape (a slicer that needs more LDS per item
# can lose under the base shape and win with wider items)
STAGE2_PAIRS = 3


def slicing_candidates(max_tw):
    """Slicing thresholds worth timing for a workload whose greedy lightcone
    treewidth reaches ``max_tw`` (None = no slicing)."""
    slicing = [None]
    for t in (13, 12, 11, 10, 9, 8):
        if t < max_tw:
            slicing.append(t)
    return slicing


def _with_slicers(thresholds):
    """(threshold, slicer) pairs: no slicer choice without slicing."""
    return [(m, sl) for m in thresholds for sl in ((SLICERS[0],) if m is None else SLICERS)]


def candidates(max_tw):
    """Every launch shape x slicing threshold x slicer (the exhaustive search space)."""
    return [(s, m, sl) for s in (BASE_SHAPE,) + OTHER_SHAPES for m, sl in _with_slicers(slicing_candidates(max_tw))]


class Autotuner:
    def __init__(self, composer=None, dtype="complex64", reps=10, device=0, verbose=False, workers=None,
                 ordering="greedy"):
        from .composer import QtreeQAOAComposer
        self.composer = composer or QtreeQAOAComposer
        self.ordering = ordering        # LightconeEngine's ``ordering`` for every candidate
        self.workers = workers          # planning processes per candidate (None: in-process)
        self.dtype = dtype
        self.reps = reps
        self.device = device
        self.verbose = verbose
        self.memo = {}

    def time_config(self, G, p, gamma, beta, shape, max_tw, slicer="greedy"):
        import torch
        eng = LightconeEngine(composer=self.composer, device=self.device, dtype=self.dtype,
                              shape=shape, max_tw=max_tw, slicer=slicer, ordering=self.ordering)
        t0 = time.perf_counter()
        wl = eng.workload(G, p, workers=self.workers)
        plan_s = time.perf_counter() - t0
        batch = eng.batch([(wl, gamma, beta)])
        batch.launch()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        evs = []
        for _ in range(self.reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            batch.launch(st)
            b.record(st)
            evs.append((a, b))
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
        return ms, plan_s, eng, wl

    def tune(self, G, p, gamma, beta, cands=None):
        """Return ``(shape, max_tw, slicer, table)`` of the fastest candidate;
        ``cands`` (optional) lists (shape, max_tw, slicer) to time instead of
        the staged search."""
        key = _signature(G, p, self.composer, self.dtype, self.ordering)
        if key in self.memo:
            return self.memo[key]
        table = []
        staged = cands is None
        if staged:
            # stage 1: slicing thresholds and slicers under the base shape;
            # stage 2: the other shapes at the best (threshold, slicer) pairs
            probe = LightconeEngine(composer=self.composer, dtype=self.dtype, ordering=self.ordering)
            max_tw = probe.workload(G, p).max_treewidth()
            cands = [(BASE_SHAPE, m, sl) for m, sl in _with_slicers(slicing_candidates(max_tw))]
        best = None
        timed = []

        def run(shape, mtw, sl):
            nonlocal best
            try:
                ms, plan_s, _, _ = self.time_config(G, p, gamma, beta, shape, mtw, sl)
            except ValueError as exc:      # shape cannot hold these plans
                table.append({"shape": asdict(shape), "max_tw": mtw, "slicer": sl, "error": str(exc)})
                return
            table.append({"shape": asdict(shape), "max_tw": mtw, "slicer": sl, "eval_ms": ms, "plan_s": plan_s})
            timed.append((ms, (mtw, sl)))
            if self.verbose:
                print(f"[autotune] {shape.key()} group={shape.group} max_tw={mtw} {sl}: {ms:.4f} ms",
                      file=sys.stderr, flush=True)
            if best is None or ms < best[0]:
                best = (ms, shape, mtw, sl)

        for shape, mtw, sl in cands:
            run(shape, mtw, sl)
        if staged:
            top = []
            for _, m in sorted(timed, key=lambda x: x[0]):
                if m not in top:
                    top.append(m)
            for shape in OTHER_SHAPES:
                for mtw, sl in top[:STAGE2_PAIRS]:
                    run(shape, mtw, sl)
        if best is None:
            raise RuntimeError("no executor configuration fits this workload")
        res = (best[1], best[2], best[3], table)
        self.memo[key] = res
        return res

    def save(self, path):
        out = [{"key": list(map(str, k)), "shape": asdict(v[0]), "max_tw": v[1], "slicer": v[2]}
               for k, v in self.memo.items()]
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
