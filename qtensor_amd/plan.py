"""Plan compiler: one lightcone tensor network -> a static step program.

A plan is the symbolic run of ``bucket_elimination`` with the backend's
``process_bucket`` (``qtensor/Simulate.py:84-94``,
``qtensor/ProcessingFrameworks.py:13-37``) over a network whose data is not
known yet.  Each bucket becomes one step

    out[o] = sum_{x_v} prod_j in_j[pext(u(x_v, o), mask_j)]

over the union ``U`` of the bucket's size-2 variables (ascending id, C order;
``v = U[0]`` is summed unless its size is 1).  Size-1 variables (the bra and
ket wires sliced to state 0) carry no data; they are dropped from device
layouts but kept in the symbolic indices so that every result is routed to
the same bucket as in the reference loop.  Variables sliced to an integer
(Feynman slicing, ``qtensor/FeynmanSimulator.py:18-25,49-58``) disappear from
the network; each work item selects their values through ``slice_bits``.

Leaf tensors are gates.  Each leaf records its gate kind and angle as a
linear form of the request's (gamma, beta) so the device rebuilds it in FP64,
plus a map from its device layout to the gate's natural layout (transposed
to ascending variable id and sliced).

Memory: leaves live in an LDS gate area; an intermediate goes to the LDS
arena when it has at most ``lds_tensor_limit`` elements and fits, otherwise to
the item's global scratch slot.  Both arenas are first-fit with reuse after
the consuming step.  A barrier is emitted before a step only if it reads
data written, or writes memory read or written, since the previous barrier.

The program format is shared with ``csrc/qtensor_amd.hip`` (constants below).
"""
import struct
from dataclasses import dataclass, field

import numpy as np

from .circuit import reorder_buckets
from .gates import (KIND_CONST, KIND_XPHASE, KIND_YPHASE, KIND_ZPHASE, KIND_ZZ)
from .tensornet import QtreeTensorNet

PLAN_HDR = 8
GATE_WORDS = 16
IN_WORDS = 8
MAX_IN = 8
MAX_RUNS = 4
GLOBAL = 0x80000000

K_H, K_CX, K_Z, K_M = 5, 6, 7, 8
_KIND_BY_NAME = {"H": K_H, "cX": K_CX, "Z": K_Z, "M": K_M, "ZPhase": KIND_ZPHASE,
                 "XPhase": KIND_XPHASE, "ZZ": KIND_ZZ, "YPhase": KIND_YPHASE}


def _i32(x):
    x = int(x) & 0xFFFFFFFF
    return x - (1 << 32) if x >= (1 << 31) else x


def _f32_bits(x):
    return struct.unpack("<i", struct.pack("<f", float(x)))[0]


def _f64_words(x):
    lo, hi = struct.unpack("<ii", struct.pack("<d", float(x)))
    return lo, hi


def mask_runs(mask):
    """Runs of contiguous set bits, least significant first: (src, width, dst)."""
    runs, d, b = [], 0, 0
    while b < 32:
        if not (mask >> b) & 1:
            b += 1
            continue
        s = b
        while b < 32 and (mask >> b) & 1:
            b += 1
        runs.append((s, b - s, d))
        d += b - s
    return runs


@dataclass
class Node:
    """A tensor of the symbolic elimination."""
    indices: tuple            # full symbolic indices (Var), ascending id
    dev: tuple                # ids of size-2 variables = device layout axes
    loc: int = -1             # element offset; GLOBAL bit for the global arena
    leaf: int = -1            # gate index for leaves
    step: int = -1            # producing step for intermediates

    @property
    def n(self):
        return 1 << len(self.dev)


@dataclass
class Plan:
    words: np.ndarray
    n_steps: int
    n_gates: int
    treewidth: int
    max_rank: int
    lds_need: int
    glb_need: int
    n_sliced: int
    bytes_step_model: int     # algorithmic bytes (complex64) summed over steps
    flops: int
    n_barriers: int
    step_sizes: list = field(default_factory=list)


class _Arena:
    """First-fit allocator over element offsets."""

    def __init__(self, base=0, align=1):
        self.base = base
        self.free = []          # sorted list of (start, end)
        self.top = base
        self.align = align
        self.peak = base

    def alloc(self, n, cap=None):
        """First-fit offset for ``n`` elements, or None if it would end past ``cap``."""
        n = (n + self.align - 1) // self.align * self.align
        for k, (s, e) in enumerate(self.free):
            if e - s >= n:
                if e - s == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, e)
                return s
        if cap is not None and self.top + n > cap:
            return None
        s = self.top
        self.top += n
        self.peak = max(self.peak, self.top)
        return s

    def release(self, s, n):
        n = (n + self.align - 1) // self.align * self.align
        self.free.append((s, s + n))
        self.free.sort()
        merged = []
        for a, b in self.free:
            if merged and merged[-1][1] == a:
                merged[-1] = (merged[-1][0], b)
            else:
                merged.append((a, b))
        if merged and merged[-1][1] == self.top:
            self.top = merged[-1][0]
            merged.pop()
        self.free = merged


def _overlap(a, b):
    return a[0] < b[1] and b[0] < a[1]


class PlanCompiler:
    """Compile a (reordered) bucket network into a device plan."""

    def __init__(self, lds_tensor_limit=1 << 12, lds_capacity=None, elem_bytes=8):
        self.lds_tensor_limit = lds_tensor_limit
        self.lds_capacity = lds_capacity      # elements available per item, or None
        self.elem_bytes = elem_bytes

    # -- front end: circuit -> reordered symbolic buckets
    @staticmethod
    def prepare(circuit, optimizer):
        tn = QtreeTensorNet.from_qtree_gates(circuit)
        res = optimizer.optimize(tn)
        peo = res[0]
        par_vars = list(res[1]) if len(res) == 3 else []
        buckets, perm = reorder_buckets(tn.buckets, peo)
        ket = [perm[v] for v in tn.ket_vars]
        bra = [perm[v] for v in tn.bra_vars]
        par = [perm[v] for v in par_vars]
        return buckets, tn.data_dict, ket, bra, par, getattr(optimizer, "treewidth", -1)

    def compile_circuit(self, circuit, optimizer):
        buckets, data_dict, ket, bra, par, tw = self.prepare(circuit, optimizer)
        return self.compile(buckets, data_dict, ket, bra, par, tw)

    # -- back end
    def compile(self, buckets, data_dict, ket_vars, bra_vars, par_vars=(), treewidth=-1,
                ket_state=0, bra_state=0):
        fixed = {}
        for vs, state in ((ket_vars, ket_state), (bra_vars, bra_state)):
            vs_sorted = sorted(vs, key=str)
            n = len(vs_sorted)
            for k, v in enumerate(vs_sorted):
                fixed[int(v)] = (state >> (n - 1 - k)) & 1
        par_index = {int(v): k for k, v in enumerate(par_vars)}
        if len(par_index) > 24:
            raise ValueError("more than 24 sliced variables")

        gates = []          # gate entries (lists of words)
        gate_area = 0
        sym = []            # symbolic buckets of Nodes
        for bucket in buckets:
            nb = []
            for t in bucket:
                node, entry, size = self._leaf(t, data_dict, fixed, par_index, gate_area)
                node.leaf = len(gates)
                gates.append(entry)
                gate_area += size
                nb.append(node)
            sym.append(nb)

        lds = _Arena(base=gate_area, align=2)
        glb = _Arena(base=0, align=64)
        steps = []
        barriers = 0
        since_w = []    # intervals written since last barrier: (space, start, end)
        since_r = []
        n_nosum = len(par_index)
        n_contract = len(sym) - n_nosum
        scalars = []
        total_bytes = 0
        total_flops = 0
        step_sizes = []
        max_rank = 0

        def interval(node):
            sp = 1 if node.loc & GLOBAL else 0
            s = node.loc & ~GLOBAL
            return (sp, s, s + node.n)

        def emit(inputs, no_sum):
            nonlocal barriers, total_bytes, total_flops, max_rank
            union = {}
            for nd in inputs:
                for v in nd.indices:
                    union.setdefault(int(v), v)
            U = [union[k] for k in sorted(union)]
            U2 = [int(v) for v in U if v.size == 2]
            do_sum = (not no_sum) and len(U) > 0
            sum_flag = 1 if (do_sum and U[0].size == 2) else 0
            out_indices = tuple(U[1:]) if do_sum else tuple(U)
            out_dev = tuple(U2[1:]) if sum_flag else tuple(U2)
            rank_u = len(U2)
            rank_out = rank_u - sum_flag
            max_rank = max(max_rank, rank_u)
            if rank_u > 31:
                raise ValueError(f"step rank {rank_u} exceeds 31")
            pos = {vid: rank_u - 1 - a for a, vid in enumerate(U2)}
            top_bit = (1 << (rank_u - 1)) if sum_flag else 0
            out = Node(out_indices, out_dev, step=len(steps))
            # allocate output before releasing inputs (no aliasing)
            n_out = out.n
            s = lds.alloc(n_out, self.lds_capacity) if n_out <= self.lds_tensor_limit else None
            out.loc = s if s is not None else (GLOBAL | glb.alloc(n_out))
            words = [len(inputs), rank_out, sum_flag, _i32(out.loc)]
            for nd in inputs:
                m = 0
                for vid in nd.dev:
                    m |= 1 << pos[vid]
                low = m & ~top_bit
                top = (1 << bin(low).count("1")) if (m & top_bit) else 0
                runs = mask_runs(low)
                w = [_i32(nd.loc), top, len(runs) if len(runs) <= MAX_RUNS else 255, low]
                for r in range(MAX_RUNS):
                    if r < len(runs) and len(runs) <= MAX_RUNS:
                        src, width, dst = runs[r]
                        w.append(src | (width << 5) | (dst << 11))
                    else:
                        w.append(0)
                words += w
            # hazards since the last barrier
            reads = [interval(nd) for nd in inputs]
            wr = interval(out)
            need = any(_overlap(r[1:], x[1:]) and r[0] == x[0] for r in reads for x in since_w) or \
                any(wr[0] == x[0] and _overlap(wr[1:], x[1:]) for x in since_w + since_r)
            if need:
                words[2] |= 2
                barriers += 1
                since_w.clear()
                since_r.clear()
            since_w.append(wr)
            since_r.extend(reads)
            steps.append(words)
            n_in_elems = sum(nd.n for nd in inputs)
            total_bytes += 8 * (n_in_elems + n_out)
            n_union = 1 << rank_u
            total_flops += 6 * (len(inputs) - 1) * n_union + 2 * (n_union - n_out)
            step_sizes.append(rank_u)
            for nd in inputs:
                if nd.step >= 0:
                    if nd.loc & GLOBAL:
                        glb.release(nd.loc & ~GLOBAL, nd.n)
                    else:
                        lds.release(nd.loc, nd.n)
            return out

        def process(nodes, no_sum=False):
            nodes = list(nodes)
            while len(nodes) > MAX_IN:
                head = emit(nodes[:MAX_IN], no_sum=True)
                nodes = [head] + nodes[MAX_IN:]
            # pass-through: one tensor whose bucket variable has size 1 and no sum
            if len(nodes) == 1 and nodes[0].indices and nodes[0].indices[0].size == 1 \
                    and not no_sum:
                nd = nodes[0]
                return Node(nd.indices[1:], nd.dev, loc=nd.loc, leaf=nd.leaf, step=nd.step)
            return emit(nodes, no_sum)

        for b in range(n_contract):
            if not sym[b]:
                continue
            res = process(sym[b])
            if res.indices:
                sym[int(res.indices[0])].append(res)
            else:
                scalars.append(res)
        rest = [nd for b in range(n_contract, len(sym)) for nd in sym[b]]
        if rest:
            res = process(rest, no_sum=True)
            if res.indices and any(v.size == 2 for v in res.indices):
                raise ValueError("plans with free (batch) variables are not supported")
            scalars.append(res)
        if not scalars:
            raise ValueError("network has no scalar result")
        if len(scalars) == 1:
            result = scalars[0]
        else:
            result = emit(scalars, no_sum=True)
        if result.dev:
            raise ValueError("network result is not a scalar")

        n_steps = len(steps)
        gates_off = PLAN_HDR
        steps_off = gates_off + GATE_WORDS * len(gates)
        words = [len(gates), n_steps, gates_off, steps_off, _i32(result.loc),
                 lds.peak, glb.peak, len(par_index)]
        for g in gates:
            words += g
        for s in steps:
            words += s
        return Plan(words=np.asarray(words, dtype=np.int32), n_steps=n_steps, n_gates=len(gates),
                    treewidth=treewidth, max_rank=max_rank, lds_need=lds.peak, glb_need=glb.peak,
                    n_sliced=len(par_index), bytes_step_model=total_bytes, flops=total_flops,
                    n_barriers=barriers, step_sizes=step_sizes)

    def _leaf(self, t, data_dict, fixed, par_index, dst):
        """Leaf node and its gate entry for tensor ``t`` (indices in data order)."""
        info = getattr(data_dict, "gate_info", {}).get(t.data_key)
        data = np.asarray(data_dict[t.data_key])
        r = len(t.indices)
        if data.ndim != r or any(s != 2 for s in data.shape):
            raise ValueError(f"leaf {t} has shape {data.shape}; plans need size-2 axes")
        if r > 3:
            raise ValueError(f"leaf {t} has rank {r} > 3")
        order = sorted(range(r), key=lambda a: int(t.indices[a]))
        base_off = 0
        dev_axes = []     # natural axis of each device axis (ascending var id)
        slice_axes = []   # (par index, natural stride)
        kept = []
        for a in order:
            v = t.indices[a]
            stride = 1 << (r - 1 - a)
            vid = int(v)
            if vid in par_index:
                slice_axes.append((par_index[vid], stride))
                continue
            if vid in fixed:
                base_off += fixed[vid] * stride
                kept.append(v.copy(size=1))
                continue
            dev_axes.append(stride)
            kept.append(v)
        n = 1 << len(dev_axes)
        src_map = 0
        for e in range(n):
            off = base_off
            for k, stride in enumerate(dev_axes):
                if (e >> (len(dev_axes) - 1 - k)) & 1:
                    off += stride
            src_map |= (off & 15) << (4 * e)
        kind = KIND_CONST
        alpha, terms = 0.0, ()
        if info is not None:
            kind, alpha, terms = info
        name = t.name.rstrip("+")
        if kind == KIND_CONST:
            kind = _KIND_BY_NAME.get(name, KIND_CONST)
        if kind == KIND_CONST:
            raise ValueError(f"gate {t.name} has no device builder")
        if len(terms) > 2:
            raise ValueError(f"gate {t.name} angle depends on {len(terms)} parameters (max 2)")
        j0, c0, j1, c1 = -1, 0.0, -1, 0.0
        a0 = alpha
        if terms:
            a0 = 0.0
            j0, c0 = terms[0]
            if len(terms) > 1:
                j1, c1 = terms[1]
        lo, hi = _f64_words(a0)
        sl = len(slice_axes)
        slw = sl | ((slice_axes[0][0] if sl > 0 else 0) << 8) | ((slice_axes[1][0] if sl > 1 else 0) << 16)
        if sl > 2:
            raise ValueError("a gate touches more than 2 sliced variables")
        entry = [kind, dst, n, _i32(src_map), int(j0), _f32_bits(c0), int(j1), _f32_bits(c1), lo, hi,
                 0, slw, slice_axes[0][1] if sl > 0 else 0, slice_axes[1][1] if sl > 1 else 0, 0, 0]
        node = Node(tuple(kept), tuple(int(v) for v in kept if v.size == 2), loc=dst)
        return node, entry, n
