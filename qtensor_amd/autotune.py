"""Autotuner for the batched executor: launch shape and slice count per workload.

For one (graph, p) workload it compiles the lightcone plans under each
candidate (item threads, LDS capacity, LDS tensor limit, slicing treewidth)
and times the executor kernel on the device with HIP events, keeping the
fastest.  The search runs in two stages: every slicing threshold below the
workload's largest greedy treewidth under a base shape, then the other
shapes (item width, LDS budget per item) at the two best thresholds.

Results are memoised per (graph signature, p, composer, dtype) in-process and
can be saved to / loaded from a JSON file.
"""
import json
import sys
import time
from dataclasses import asdict

import numpy as np

from .batched import LaunchShape, LightconeEngine


def _signature(G, p, composer, dtype):
    degs = tuple(sorted(d for _, d in G.degree()))
    return (G.number_of_nodes(), G.number_of_edges(), hash(degs), p, composer.__name__, dtype)


BASE_SHAPE = LaunchShape(128, 4096, 2048)
OTHER_SHAPES = (LaunchShape(128, 2048, 1024), LaunchShape(256, 4096, 2048), LaunchShape(256, 2048, 1024),
                LaunchShape(64, 4096, 2048))


def slicing_candidates(max_tw):
    """Slicing thresholds worth timing for a workload whose greedy lightcone
    treewidth reaches ``max_tw`` (None = no slicing)."""
    slicing = [None]
    for t in (13, 12, 11, 10):
        if t < max_tw:
            slicing.append(t)
    return slicing


def candidates(max_tw):
    """Every launch shape x slicing threshold (the exhaustive search space)."""
    return [(s, m) for s in (BASE_SHAPE,) + OTHER_SHAPES for m in slicing_candidates(max_tw)]


class Autotuner:
    def __init__(self, composer=None, dtype="complex64", reps=10, device=0, verbose=False):
        from .composer import QtreeQAOAComposer
        self.composer = composer or QtreeQAOAComposer
        self.dtype = dtype
        self.reps = reps
        self.device = device
        self.verbose = verbose
        self.memo = {}

    def time_config(self, G, p, gamma, beta, shape, max_tw):
        import torch
        eng = LightconeEngine(composer=self.composer, device=self.device, dtype=self.dtype,
                              shape=shape, max_tw=max_tw)
        t0 = time.perf_counter()
        wl = eng.workload(G, p)
        plan_s = time.perf_counter() - t0
        batch = eng.batch([(wl, gamma, beta)])
        batch.launch_exec()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        evs = []
        for _ in range(self.reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            batch.launch_exec(st)
            b.record(st)
            evs.append((a, b))
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
        return ms, plan_s, eng, wl

    def tune(self, G, p, gamma, beta, cands=None):
        """Return ``(shape, max_tw, table)`` of the fastest candidate."""
        key = _signature(G, p, self.composer, self.dtype)
        if key in self.memo:
            return self.memo[key]
        table = []
        staged = cands is None
        if staged:
            # stage 1: slicing thresholds under the base shape; stage 2: the
            # other shapes at the two best thresholds
            probe = LightconeEngine(composer=self.composer, dtype=self.dtype)
            max_tw = probe.workload(G, p).max_treewidth()
            cands = [(BASE_SHAPE, m) for m in slicing_candidates(max_tw)]
        best = None
        timed = []

        def run(shape, mtw):
            nonlocal best
            try:
                ms, plan_s, _, _ = self.time_config(G, p, gamma, beta, shape, mtw)
            except ValueError as exc:      # shape cannot hold these plans
                table.append({"shape": asdict(shape), "max_tw": mtw, "error": str(exc)})
                return
            table.append({"shape": asdict(shape), "max_tw": mtw, "exec_ms": ms, "plan_s": plan_s})
            timed.append((ms, mtw))
            if self.verbose:
                print(f"[autotune] {shape.key()} max_tw={mtw}: {ms:.4f} ms", file=sys.stderr, flush=True)
            if best is None or ms < best[0]:
                best = (ms, shape, mtw)

        for shape, mtw in cands:
            run(shape, mtw)
        if staged:
            top = []
            for _, m in sorted(timed, key=lambda x: x[0]):
                if m not in top:
                    top.append(m)
            for shape in OTHER_SHAPES:
                for m in top[:2]:
                    run(shape, m)
        if best is None:
            raise RuntimeError("no executor configuration fits this workload")
        res = (best[1], best[2], table)
        self.memo[key] = res
        return res

    def save(self, path):
        out = [{"key": list(map(str, k)), "shape": asdict(v[0]), "max_tw": v[1]} for k, v in self.memo.items()]
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
