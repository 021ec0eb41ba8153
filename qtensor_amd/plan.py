"""Plan compiler: one lightcone tensor network -> a static op program.

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

Leaf tensors are gates.  Each distinct leaf records its gate kind and angle
as a linear form of the request's (gamma, beta) so the device rebuilds it in
FP64, plus a map from its device layout to the gate's natural layout
(transposed to ascending variable id and sliced).

Scheduling: steps form a DAG (each result feeds one later bucket).  Steps
with at most one wave of outputs, at most four inputs, at most two bit-runs
per input offset and all data in LDS are packed, dependency-respecting, into
"gather" ops whose lanes each compute one output element of one step; every
other step is a "step" op executed by all lanes of the item.

Memory: leaves live in an LDS gate area (plus constants ONE, ZERO and a sink
element); an intermediate goes to the LDS arena when it has at most
``lds_tensor_limit`` elements, otherwise to the item's global scratch slot.
Both arenas are first-fit; an op's outputs are placed before its inputs are
released, so no op writes memory it also reads.

The program format is shared with ``csrc/qtensor_amd.hip`` (constants below).
"""
import ctypes
import functools
import struct
from dataclasses import dataclass, field

import numpy as np

from .circuit import reorder_buckets
from .gates import (KIND_CONST, KIND_XPHASE, KIND_YPHASE, KIND_ZPHASE, KIND_ZZ)
from .tensornet import QtreeTensorNet

PLAN_HDR = 12
GATE_WORDS = 16
IN_WORDS = 4
STEP_HDR = 3
MAX_IN = 4
# zero words after the last plan of a program: the executor fetches words 0-63
# of each op's record (one per lane) whatever the record's length
PROGRAM_PAD = 320
# zero angles after the last item's: the executor loads 8 angles per item at once
PARAM_PAD = 8
MAX_RUNS = 2
GLOBAL = 0x80000000
OP_GATHER = 1 << 15
# fused step (wide plans): several bucket steps contracted in one pass
OP_FUSED = 1 << 14
# fused step flag: a chain op (see chain_pairs)
OP_CHAIN = 1 << 12
CHAIN_MAX_X, CHAIN_MAX_Y, CHAIN_MAX_D = 4, 5, 10
# reduction op (item plans): a chain of single-input sums in one op, one wave
OP_REDUCE = 1 << 13
REDUCE_HDR = 4
REDUCE_MAX_OUT = 64
FUSED_MAX_IN = 8
FUSED_MAX_SUM = 8
FUSED_HDR = 8
FUSED_IN_WORDS = 4
GATHER_HDR = 4
GATHER_MAX_STEPS = 16
# LDS element offsets of the constants every plan reserves ahead of its gates
LOC_ONE, LOC_ZERO, LOC_SINK = 0, 1, 2
GATE_BASE = 4

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


def _runs_word(runs):
    if len(runs) > MAX_RUNS:
        return 3 << 30
    packed = [src | (width << 5) | (dst << 10) for src, width, dst in runs] + [0, 0]
    return (len(runs) << 30) | packed[0] | (packed[1] << 15)


class Buf:
    """Storage of a tensor; pass-through aliases share one Buf."""
    __slots__ = ("n", "loc", "leaf", "last_op")

    def __init__(self, n, loc=-1, leaf=False):
        self.n = n
        self.loc = loc
        self.leaf = leaf
        self.last_op = -1

    @property
    def is_global(self):
        return bool(self.loc & GLOBAL) if self.loc >= 0 else False


@dataclass
class Node:
    """A tensor of the symbolic elimination."""
    indices: tuple            # full symbolic indices (Var), ascending id
    dev: tuple                # ids of size-2 variables = device layout axes
    buf: Buf = None
    step: int = -1            # producing step (-1 for leaves)

    @property
    def n(self):
        return 1 << len(self.dev)


@dataclass
class SymStep:
    inputs: list
    out: Node
    rank_out: int
    sum_flag: int
    lows: list                # per input: mask of output bits
    tops: list                # per input: offset of the summed variable = 1
    global_res: bool = False  # output too large for LDS

    def gatherable(self, lanes, lds_limit):
        if self.out.n > lanes or len(self.inputs) > MAX_IN:
            return False
        return not any(nd.n > lds_limit and not nd.buf.leaf for nd in self.inputs)


class FusedStep:
    """Several consecutive bucket steps of a wide plan contracted in one pass:
    out[o] = sum over the summed variables x of prod_j in_j[pext(u(o, x), mask_j)],
    u = the union offset (variable a of the union at bit rank_u - 1 - a).
    ``chain``: three inputs A, B, D where D carries none of the summed
    variables of the first two's contraction (see ``chain_pairs``)."""
    fused = True

    def __init__(self, inputs, out, union_dev, members, chain=False):
        self.inputs = inputs
        self.out = out
        self.union = sorted(union_dev)
        self.members = members          # original step indices, last = the output step
        self.chain = chain

    @property
    def rank_out(self):
        return len(self.out.dev)

    def gatherable(self, lanes, lds_limit):
        return False


class ReduceStep:
    """A chain of single-input steps, each summing the top (most significant)
    variable of the previous one's output, contracted in one op:
    out[o] = sum over x < 2**k of in[o + x * n_out], for the chain's k summed
    variables.  Run by one wave (n_out <= REDUCE_MAX_OUT)."""

    def __init__(self, inp, out, k, rank_out, members):
        self.inputs = [inp]
        self.out = out
        self.k = k
        self.rank_out = rank_out
        self.members = members

    def gatherable(self, lanes, lds_limit):
        return False


def _simple_sum(st):
    """True for a single-input step summing its input's top variable with the
    other variables kept in order (out[o] = in[o] + in[o + n_out])."""
    if len(st.inputs) != 1 or not st.sum_flag:
        return False
    n_out = 1 << st.rank_out
    return st.lows[0] == n_out - 1 and st.tops[0] == n_out


def reduce_chains(steps, result_step, lds_limit):
    """Replace chains of at least two simple sums whose intermediates have no
    other reader by ReduceSteps.  Returns the new step list, with every
    node's producer index renumbered into it."""
    n = len(steps)
    readers = [0] * n
    for st in steps:
        for nd in st.inputs:
            if nd.step >= 0:
                readers[nd.step] += 1
    nxt = [-1] * n           # the simple-sum step reading step i's output alone
    for i, st in enumerate(steps):
        if _simple_sum(st):
            j = st.inputs[0].step
            if j >= 0 and readers[j] == 1 and j != result_step and _simple_sum(steps[j]):
                nxt[j] = i
    has_prev = [False] * n
    for j in range(n):
        if nxt[j] >= 0:
            has_prev[nxt[j]] = True
    replaced = {}
    for i in range(n):
        if has_prev[i] or nxt[i] < 0 or not _simple_sum(steps[i]):
            continue
        chain = [i]
        while nxt[chain[-1]] >= 0:
            chain.append(nxt[chain[-1]])
        first, last = steps[chain[0]], steps[chain[-1]]
        if (1 << last.rank_out) > REDUCE_MAX_OUT or first.inputs[0].n > lds_limit:
            continue
        replaced[chain[-1]] = ReduceStep(first.inputs[0], last.out, len(chain), last.rank_out, chain)
        for c in chain[:-1]:
            replaced[c] = None
    if not replaced:
        return steps
    new, index = [], {}
    for i, st in enumerate(steps):
        r = replaced.get(i, st)
        if r is None:
            continue
        index[i] = len(new)
        new.append(r)
    seen = set()
    for st in new:
        for nd in list(st.inputs) + [st.out]:
            if id(nd) not in seen and nd.step >= 0:
                seen.add(id(nd))
                nd.step = index[nd.step]
    return new


def _pdep(x, mask):
    r, bit = 0, 0
    while mask:
        low = mask & -mask
        if (x >> bit) & 1:
            r |= low
        bit += 1
        mask &= mask - 1
    return r


def _pext(x, mask):
    r, bit = 0, 0
    while mask:
        low = mask & -mask
        if x & low:
            r |= 1 << bit
        bit += 1
        mask &= mask - 1
    return r


@functools.lru_cache(maxsize=4096)
def _pext_table(mask, n):
    """(_pext(o, mask) for o in range(n)), increasing in o."""
    return tuple(_pext(o, mask) for o in range(n))


def _step_work(n_in, union_rank):
    """Arithmetic of a step including its loads: per union element, n_in loads,
    n_in - 1 complex products and one complex sum."""
    return (6 * max(n_in - 1, 0) + 2 + 4 * n_in) * (1 << union_rank)


# the wide path's time model: seconds per launch of a grid kernel
MODEL_LAUNCH_S = 4e-6


def _step_time(in_ranks, out_rank, union_rank, bw=4e12, rate=40e12, launch=None):
    """Time model of one step or fused step on the device (memory, arithmetic
    including its loads, launch)."""
    if launch is None:
        launch = MODEL_LAUNCH_S
    nbytes = 8 * (sum(1 << r for r in in_ranks) + (1 << out_rank))
    return max(nbytes / bw, _step_work(len(in_ranks), union_rank) / rate) + launch


def fuse_steps(steps):
    """Greedy fusion of single-consumer step chains (every intermediate of
    bucket elimination has one reader): a step absorbs the group producing one
    of its inputs when the time model says one pass beats two.  Limits: at
    most FUSED_MAX_IN external inputs and FUSED_MAX_SUM summed variables, union
    rank <= 30, and either an output of rank >= 14 or at most 4 summed
    variables (each output element loops over the summed ones)."""
    def ranks(nodes):
        return [len(x.dev) for x in nodes]

    groups = {}        # output step index -> (external inputs, union set, members, time)
    order = []
    for i, st in enumerate(steps):
        out_rank = len(st.out.dev)
        ext = list(st.inputs)
        union = set()
        for nd in st.inputs:
            union |= set(nd.dev)
        members = [i]
        absorbed = []
        for nd in list(st.inputs):
            g = groups.get(nd.step) if nd.step >= 0 else None
            if g is None:
                continue
            g_inputs, g_union, g_members, g_time = g
            cand_ext = [x for x in ext if x is not nd] + g_inputs
            cand_union = union | g_union
            k = len(cand_union) - out_rank
            if not (len(cand_ext) <= FUSED_MAX_IN and k <= FUSED_MAX_SUM and len(cand_union) <= 30
                    and (out_rank >= 14 or k <= 4)):
                continue
            t_sep = g_time + _step_time(ranks(ext), out_rank, len(union))
            t_fused = _step_time(ranks(cand_ext), out_rank, len(cand_union))
            if t_fused < 0.9 * t_sep:
                ext, union = cand_ext, cand_union
                members = g_members + members
                absorbed.append(nd.step)
        for a_ in absorbed:
            groups.pop(a_)
            order.remove(a_)
        groups[i] = (ext, union, members, _step_time(ranks(ext), out_rank, len(union)))
        order.append(i)
    out = []
    for i in order:
        ext, union, members, _ = groups[i]
        if len(members) == 1:
            out.append(steps[i])
        else:
            out.append(FusedStep(ext, steps[i].out, union, members))
    return out


# split_small_inputs: the small inputs' product is taken when its union is at
# least this many variables below the contraction's
SPLIT_MARGIN = 6


def split_small_inputs(steps):
    """Within a step or fused step of three or more inputs, multiply all inputs
    but the largest into one tensor first (summing the summed variables only
    they carry), so that the wide contraction reads two inputs per summed
    assignment instead of all of them, when the time model says the two passes
    beat one.  The product of a bucket's tensors may be formed in any order
    (``qtensor/ProcessingFrameworks.py:106-146`` multiplies them pairwise)."""
    def ranks(nodes):
        return [len(x.dev) for x in nodes]

    out = []
    for st in steps:
        if isinstance(st, ReduceStep) or len(st.inputs) < 3:
            out.append(st)
            continue
        if getattr(st, "fused", False):
            union = set(st.union)
            members = st.members
        else:
            union = set()
            for nd in st.inputs:
                union |= set(nd.dev)
            members = None
        out_set = set(st.out.dev)
        big = max(st.inputs, key=lambda nd: len(nd.dev))
        rest = [nd for nd in st.inputs if nd is not big]
        u_rest = set()
        for nd in rest:
            u_rest |= set(nd.dev)
        private = {v for v in u_rest if v not in out_set and v not in big.dev}
        p_dev = tuple(sorted(u_rest - private))
        main_union = union - private
        if len(u_rest) > 30 or len(private) > FUSED_MAX_SUM or len(p_dev) > len(st.out.dev):
            out.append(st)
            continue
        # the model's memory term hides the per-element work of a many-input
        # contraction, so a product at least 64x smaller than the contraction
        # is taken whenever it cuts the arithmetic
        t_one = _step_time(ranks(st.inputs), len(out_set), len(union))
        t_two = _step_time(ranks(rest), len(p_dev), len(u_rest)) + \
            _step_time([len(big.dev), len(p_dev)], len(out_set), len(main_union))
        small = len(u_rest) + SPLIT_MARGIN <= len(main_union) and \
            _step_work(2, len(main_union)) < 0.9 * _step_work(len(st.inputs), len(union))
        if not (t_two < 0.9 * t_one or small):
            out.append(st)
            continue
        prod = Node((), p_dev, buf=Buf(1 << len(p_dev)))
        out.append(FusedStep(rest, prod, u_rest, []))
        out.append(FusedStep([big, prod], st.out, main_union, members if members is not None else []))
    return out


def chain_pairs(steps, bw=4e12, rate=40e12, launch=None):
    """Merge a two-input contraction T = sum_x A B with the two-input
    contraction out = sum_y T D that reads T into one chain op over A, B and D
    (``FusedStep(chain=True)``), so that T, often the largest tensor of a wide
    plan, is formed in registers and never written: the kernel sums
    sum_y D (sum_x A B).  D carries no variable of x (bucket elimination
    removes a variable from every tensor when it sums it).  Taken when the time
    model (bytes at ``bw``, arithmetic at ``rate``; T's values recomputed for
    each output bit beyond the first that only D carries) says one pass beats
    the two; limits: at least one summed variable that D lacks (up to
    CHAIN_MAX_X of them are summed innermost, over A B alone), at most
    CHAIN_MAX_Y others, D of rank <= CHAIN_MAX_D, union rank <= 31, output
    rank >= 8.
    The value is the same sum as the two steps', up to rounding (the
    reference forms a bucket's product in any order,
    ``qtensor/ProcessingFrameworks.py:106-146``)."""
    def union_of(st):
        if getattr(st, "fused", False):
            return set(st.union)
        u = set()
        for nd in st.inputs:
            u |= set(nd.dev)
        return u

    producer = {id(st.out): i for i, st in enumerate(steps)}
    merged, gone = {}, set()
    for ci, c in enumerate(steps):
        if isinstance(c, ReduceStep) or getattr(c, "chain", False) or len(c.inputs) != 2:
            continue
        big = max(c.inputs, key=lambda nd: len(nd.dev))
        d = c.inputs[1] if big is c.inputs[0] else c.inputs[0]
        pi = producer.get(id(big))
        if pi is None or pi in gone or pi in merged:
            continue
        pst = steps[pi]
        if isinstance(pst, ReduceStep) or getattr(pst, "chain", False) or len(pst.inputs) != 2:
            continue
        u_p, u_c = union_of(pst), union_of(c)
        out_c = set(c.out.dev)
        union = u_p | u_c
        if (u_p - set(big.dev)) & set(d.dev):
            continue
        # the device sums up to CHAIN_MAX_X of the summed variables D lacks
        # innermost (over A B only) and the rest in the outer loop
        summed = union - out_c
        kx = min(CHAIN_MAX_X, len(summed - set(d.dev)))
        ky = len(summed) - kx
        if not (kx >= 1 and ky <= CHAIN_MAX_Y and len(d.dev) <= CHAIN_MAX_D and len(union) <= 31
                and len(out_c) >= 8 and len(summed) <= FUSED_MAX_SUM):
            continue
        # the kernel tiles up to 2 x 2 high output bits that only A (B) carries in
        # registers; a large op without such bits would load A and B per product
        a_set, b_set = set(pst.inputs[0].dev), set(pst.inputs[1].dev)
        ordered = sorted(union)
        high_out = [v for v in ordered if v in out_c][:max(len(out_c) - 8, 0)]
        tile = min(sum(1 for v in high_out if v in a_set and v not in b_set),
                   sum(1 for v in high_out if v in b_set and v not in a_set), 2)
        if tile < 2 and len(union) > 20:
            continue
        d_only = len(out_c - u_p)
        t_sep = _step_time([len(nd.dev) for nd in pst.inputs], len(big.dev), len(u_p), bw, rate, launch) + \
            _step_time([len(nd.dev) for nd in c.inputs], len(out_c), len(u_c), bw, rate, launch)
        nbytes = 8 * (sum(nd.n for nd in pst.inputs) + d.n + c.out.n)
        # products A B: every union assignment, except that one output bit only D
        # carries shares them; products with D: every (output, outer) assignment
        work = _step_work(2, len(union) - min(d_only, 1)) + _step_work(2, len(out_c) + ky)
        t_chain = max(nbytes / bw, work / rate) + (MODEL_LAUNCH_S if launch is None else launch)
        if t_chain < 0.9 * t_sep:
            a_, b_ = pst.inputs
            members = list(getattr(pst, "members", [])) + list(getattr(c, "members", []))
            merged[ci] = FusedStep([a_, b_, d], c.out, union, members, chain=True)
            gone.add(pi)
    return [merged.get(i, st) for i, st in enumerate(steps) if i not in gone]


def hoist_gate_only(steps):
    """Wide plans: steps whose inputs are all gate tensors go first (in their
    order, as fused records), so that they form one run of mutually
    independent ops at the start of the plan, which ``qb_wide_execute``
    launches together (one block per op).  Moving them earlier keeps every
    dependency: they read gates only."""
    def gate_only(st):
        return not isinstance(st, ReduceStep) and all(nd.buf is not None and nd.buf.leaf for nd in st.inputs)

    def as_fused(st):
        # a plain step is the fused op over its inputs' variables
        if getattr(st, "fused", False):
            return st
        union = set()
        for nd in st.inputs:
            union |= set(nd.dev)
        return FusedStep(list(st.inputs), st.out, union, [])
    first = [as_fused(st) for st in steps if gate_only(st)]
    return first + [st for st in steps if not gate_only(st)]


@dataclass
class Plan:
    words: np.ndarray
    n_steps: int
    n_ops: int
    n_gates: int
    treewidth: int
    max_rank: int
    lds_need: int
    glb_need: int
    n_sliced: int
    bytes_step_model: int     # algorithmic bytes (complex64) summed over steps
    flops: int
    n_barriers: int
    n_gather: int = 0
    n_fused: int = 0
    n_reduce: int = 0
    step_sizes: list = field(default_factory=list)
    steps_words: int = 0
    dag: list = field(default_factory=list)   # per step: (producer steps of inputs, n_out, n_in, sum)
    bytes_ops: int = 0        # bytes (complex64) the ops read and write: inputs and output of each op

    def lds_elems(self, elem_bytes=8):
        """LDS elements one item needs (its data; op records stay in global memory)."""
        return self.lds_need

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


def _place_order(sizes, starts, ends, order, base):
    """Offsets for buffers placed in ``order``, each at the lowest offset free
    over its whole op interval (native ``qb_place_intervals``); returns
    (offsets, peak)."""
    from . import _lib
    m = len(sizes)
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (sizes, starts, ends, order)]
    offsets = np.zeros(max(m, 1), dtype=np.int64)
    peak = ctypes.c_int64(0)
    p64 = ctypes.POINTER(ctypes.c_int64)
    _lib.check(_lib.load().qb_place_intervals(m, *(a.ctypes.data_as(p64) for a in arrs), int(base),
                                              offsets.ctypes.data_as(p64), ctypes.byref(peak)))
    return offsets[:m], int(peak.value)


def _place_intervals(items, base, align):
    """Offline interval allocation: ``items`` are (buf, first_op, last_op);
    buffers whose op intervals overlap get disjoint element ranges.  Two
    placement orders are tried (largest first; largest size x lifetime first),
    the one with the lowest peak is kept.  Sets ``buf.loc``;
    returns the peak offset."""
    if not items:
        return base
    sizes = np.array([(b.n + align - 1) // align * align for b, _, _ in items], dtype=np.int64)
    starts = np.array([s for _, s, _ in items], dtype=np.int64)
    ends = np.array([e for _, _, e in items], dtype=np.int64)
    orders = (np.lexsort((starts, -sizes)),
              np.lexsort((starts, -sizes * (ends - starts + 1))))
    best = None
    for order in orders:
        offsets, peak = _place_order(sizes, starts, ends, order, base)
        if best is None or peak < best[1]:
            best = (offsets, peak)
    for (b, _, _), off in zip(items, best[0].tolist()):
        b.loc = int(off)
    return best[1]


class PlanCompiler:
    """Compile a (reordered) bucket network into a device plan.

    ``lanes``: threads per work item (a gather op holds at most this many
    output elements); ``gather=False`` emits one step op per bucket.
    """

    def __init__(self, lds_tensor_limit=1 << 12, lds_capacity=None, elem_bytes=8, lanes=64,
                 gather=True, window=None, gate_space="lds", schedule="list", fuse=False, reduce=True,
                 bake_angles=False, chain=True):
        self.lds_tensor_limit = lds_tensor_limit
        # gates keep the angle they were composed with instead of rebuilding it
        # from the parameter vector (circuits that carry no parameter terms)
        self.bake_angles = bake_angles
        self.lds_capacity = lds_capacity      # elements available per item, or None
        self.elem_bytes = elem_bytes
        self.lanes = lanes
        self.gather = gather
        # first-fit scheduling: a step may join one of the last `window` ops
        # (None: any op after its inputs)
        self.window = window
        if schedule not in ("levels", "alap", "first_fit", "list"):
            raise ValueError(f"unknown schedule {schedule!r}")
        self.schedule = schedule
        # wide plans: contract chains of steps in one pass (see fuse_steps), and
        # producer/consumer pairs without storing the intermediate (chain_pairs)
        self.fuse = fuse
        self.chain = chain
        # item plans: chains of single-input sums become one reduction op each
        self.reduce = reduce
        # "global": gates and every intermediate live in global memory (wide plans)
        if gate_space not in ("lds", "global"):
            raise ValueError("gate_space is 'lds' or 'global'")
        self.gate_space = gate_space

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
                ket_state=0, bra_state=0, free_vars=()):
        """``free_vars`` (wide plans only): output variables kept open, whose
        buckets come last (after the sliced ones); the result is then the
        tensor over them (variables in ascending id, the first one the most
        significant bit of the element index), and header word 11 holds their
        count."""
        free_ids = [int(v) for v in free_vars]
        if free_ids and self.gate_space != "global":
            raise ValueError("open (batch) variables need a wide plan")
        fixed = {}
        for vs, state in ((ket_vars, ket_state), (bra_vars, bra_state)):
            vs_sorted = sorted(vs, key=str)
            n = len(vs_sorted)
            for k, v in enumerate(vs_sorted):
                fixed[int(v)] = (state >> (n - 1 - k)) & 1
        par_index = {int(v): k for k, v in enumerate(par_vars)}
        if len(par_index) > 24:
            raise ValueError("more than 24 sliced variables")

        # leaves: distinct gate tensors in the LDS gate area
        gates, gate_buf = [], {}
        wide = self.gate_space == "global"
        gate_area = 0 if wide else GATE_BASE
        sym = []
        for bucket in buckets:
            nb = []
            for t in bucket:
                node, entry, size = self._leaf(t, data_dict, fixed, par_index, gate_area)
                sig = tuple(entry[:1] + entry[2:])
                if sig not in gate_buf:
                    gate_buf[sig] = Buf(size, loc=(GLOBAL | gate_area) if wide else gate_area, leaf=True)
                    gates.append(entry)
                    gate_area += size
                node.buf = gate_buf[sig]
                nb.append(node)
            sym.append(nb)

        # phase 1: symbolic elimination -> steps
        steps = []

        def emit(inputs, no_sum):
            if not 1 <= len(inputs) <= MAX_IN:
                raise AssertionError(f"step with {len(inputs)} inputs (kernel takes 1..{MAX_IN})")
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
            if rank_u > 31:
                raise ValueError(f"step rank {rank_u} exceeds 31")
            pos = {vid: rank_u - 1 - a for a, vid in enumerate(U2)}
            top_bit = (1 << (rank_u - 1)) if sum_flag else 0
            lows, tops = [], []
            for nd in inputs:
                m = 0
                for vid in nd.dev:
                    m |= 1 << pos[vid]
                low = m & ~top_bit
                lows.append(low)
                tops.append((1 << bin(low).count("1")) if (m & top_bit) else 0)
            out = Node(out_indices, out_dev, buf=Buf(1 << len(out_dev)), step=len(steps))
            steps.append(SymStep(list(inputs), out, rank_u - sum_flag, sum_flag, lows, tops))
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
                return Node(nd.indices[1:], nd.dev, buf=nd.buf, step=nd.step)
            return emit(nodes, no_sum)

        n_nosum = len(par_index) + len(free_ids)
        n_contract = len(sym) - n_nosum
        scalars = []
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
            if any(v.size == 2 and int(v) not in free_ids for v in res.indices):
                raise ValueError("plans with free (batch) variables need them in free_vars")
            scalars.append(res)
        if not scalars:
            raise ValueError("network has no scalar result")
        result = scalars[0] if len(scalars) == 1 else process(scalars, no_sum=True)
        if sorted(result.dev) != sorted(free_ids):
            raise ValueError("network result is not a tensor over exactly the free variables")
        if free_ids and result.buf.leaf:
            # an open result that is a gate tensor as is: copy it into an intermediate
            result = emit([result], no_sum=True)

        # step-model figures of the bucket steps (before any fusion renumbers them)
        total_bytes, total_flops, max_rank = 0, 0, 0
        step_sizes, dag = [], []
        for st in steps:
            n_out = st.out.n
            rank_u = st.rank_out + st.sum_flag
            max_rank = max(max_rank, rank_u)
            total_bytes += 8 * (sum(nd.n for nd in st.inputs) + n_out)
            n_union = 1 << rank_u
            total_flops += 6 * (len(st.inputs) - 1) * n_union + 2 * (n_union - n_out)
            step_sizes.append(rank_u)
            dag.append(([nd.step for nd in st.inputs], n_out, len(st.inputs), st.sum_flag))

        # phase 2: schedule steps into ops.  Wide plans run one grid-wide launch
        # per op in bucket order, optionally with fused step chains; item plans
        # contract chains of single-input sums in one reduction op each.
        lanes = self.lanes
        lim = self.lds_tensor_limit
        exec_steps = steps
        if wide:
            if self.fuse:
                exec_steps = split_small_inputs(fuse_steps(steps))
                if self.chain:
                    exec_steps = chain_pairs(exec_steps)
                exec_steps = hoist_gate_only(exec_steps)
            ops, op_kind = [[i] for i in range(len(exec_steps))], [False] * len(exec_steps)
        else:
            if self.reduce:
                exec_steps = reduce_chains(steps, result.step, lim)
            if self.schedule == "first_fit":
                ops, op_kind = self._schedule_first_fit(exec_steps, lanes, lim)
            elif self.schedule == "list":
                ops, op_kind = self._schedule_list(exec_steps, lanes, lim)
            else:
                ops, op_kind = self._schedule_levels(exec_steps, lanes, lim, alap=self.schedule == "alap")
        op_of = [0] * len(exec_steps)
        for k, members in enumerate(ops):
            for i in members:
                op_of[i] = k

        # phase 3: allocation.  An intermediate lives from its op to its last
        # reading op, inclusive (the result to the end), so an op's outputs never
        # share storage with its inputs.  Offsets come from an offline interval
        # allocator; the first-fit arena only provides the gate area's base.
        for i, st in enumerate(exec_steps):
            for nd in st.inputs:
                nd.buf.last_op = max(nd.buf.last_op, op_of[i])
        if wide:
            lim = 0
        lds_items, glb_items = [], []
        end_all = len(ops)
        seen = set()
        for k, members in enumerate(ops):
            for i in members:
                b = exec_steps[i].out.buf
                assert id(b) not in seen, "a buffer is written by two steps"
                seen.add(id(b))
                end = b.last_op if b.last_op >= 0 else end_all
                (lds_items if b.n <= lim else glb_items).append((b, k, end))
        lds_base = 0 if wide else gate_area
        glb_base = (gate_area + 63) // 64 * 64 if wide else 0
        lds_peak = _place_intervals(lds_items, lds_base, 2)
        if self.lds_capacity is not None and lds_peak > self.lds_capacity:
            raise ValueError("LDS capacity exhausted; use a larger capacity, a smaller "
                             "LDS tensor limit or more slicing")
        glb_peak = _place_intervals(glb_items, glb_base, 64)
        for b, _, _ in glb_items:
            b.loc = GLOBAL | b.loc
        lds = _Arena(base=lds_base)
        lds.peak = lds_peak
        glb = _Arena(base=glb_base)
        glb.peak = glb_peak

        # phase 4: encode ops; barriers only matter for global data
        op_words = []
        glb_written_since = set()
        barriers = 0
        n_gather = 0
        n_fused = 0
        n_reduce = 0
        for k, members in enumerate(ops):
            reads_glb = any(nd.buf.is_global for i in members for nd in exec_steps[i].inputs)
            writes_glb = any(exec_steps[i].out.buf.is_global for i in members)
            sync_glb = 0
            if (reads_glb or writes_glb) and glb_written_since:
                sync_glb = 1
                barriers += 1
                glb_written_since = set()
            if writes_glb:
                glb_written_since.add(k)
            if op_kind[k]:
                op_words.append(self._encode_gather([exec_steps[i] for i in members]))
                n_gather += 1
            elif isinstance(exec_steps[members[0]], ReduceStep):
                op_words.append(self._encode_reduce(exec_steps[members[0]], sync_glb))
                n_reduce += 1
            elif getattr(exec_steps[members[0]], "fused", False):
                op_words.append(self._encode_fused(exec_steps[members[0]], sync_glb))
                n_fused += 1
            else:
                op_words.append(self._encode_step(exec_steps[members[0]], sync_glb))

        bytes_ops = sum(8 * (sum(nd.n for nd in st.inputs) + st.out.n) for st in exec_steps)
        n_ops = len(op_words)
        gates_off = PLAN_HDR
        steps_off = gates_off + GATE_WORDS * len(gates)
        steps_words = sum(len(w) for w in op_words)
        words = [len(gates), n_ops, gates_off, steps_off, _i32(result.buf.loc),
                 lds.peak, _i32(glb.peak & 0xFFFFFFFF), len(par_index), steps_words,
                 _i32(glb.peak >> 32), 0, len(free_ids)]
        for g in gates:
            words += g
        for w in op_words:
            words += w
        return Plan(words=np.asarray(words, dtype=np.int32), n_steps=len(steps), n_ops=n_ops,
                    n_gates=len(gates), treewidth=treewidth, max_rank=max_rank,
                    lds_need=lds.peak, glb_need=glb.peak, n_sliced=len(par_index),
                    bytes_step_model=total_bytes, flops=total_flops, n_barriers=barriers,
                    n_gather=n_gather, n_fused=n_fused, n_reduce=n_reduce, step_sizes=step_sizes,
                    steps_words=steps_words,
                    dag=dag, bytes_ops=bytes_ops)

    def _gather_ok(self, st, lanes, lim):
        return self.gather and st.out.n <= lim and st.gatherable(lanes, lim)

    def _schedule_first_fit(self, steps, lanes, lim):
        """Steps in order; a gatherable step joins the first gather op after its
        inputs with room, any other step opens a new op at the end."""
        ops, op_kind, op_fill = [], [], []
        op_of = [0] * len(steps)
        for i, st in enumerate(steps):
            ready = max([op_of[nd.step] + 1 for nd in st.inputs if nd.step >= 0] + [0])
            if self.window is not None:
                ready = max(ready, len(ops) - self.window)
            if self._gather_ok(st, lanes, lim):
                k = ready
                while k < len(ops) and not (op_kind[k] and op_fill[k] + st.out.n <= lanes
                                            and len(ops[k]) < GATHER_MAX_STEPS):
                    k += 1
                if k == len(ops):
                    ops.append([])
                    op_kind.append(True)
                    op_fill.append(0)
                ops[k].append(i)
                op_fill[k] += st.out.n
            else:
                k = len(ops)
                ops.append([i])
                op_kind.append(False)
                op_fill.append(st.out.n)
            op_of[i] = k
        return ops, op_kind

    def _schedule_levels(self, steps, lanes, lim, alap=False):
        """Level scheduling: every step gets its dependency level (as soon as
        possible, or as late as possible with ``alap``); the gatherable steps of
        a level are packed into gather ops (first fit by decreasing size, at most
        ``lanes`` outputs and GATHER_MAX_STEPS steps each), the others get one
        step op each.  The op count approaches the dependency depth."""
        n = len(steps)
        prods = [[nd.step for nd in st.inputs if nd.step >= 0] for st in steps]
        level = [0] * n
        for i in range(n):
            level[i] = 1 + max([level[j] for j in prods[i]] + [-1])
        if alap:
            last = max(level) if n else 0
            cons = [[] for _ in range(n)]
            for i in range(n):
                for j in prods[i]:
                    cons[j].append(i)
            for i in reversed(range(n)):
                level[i] = min([level[c] for c in cons[i]] + [last + 1]) - 1
        by_level = {}
        for i in range(n):
            by_level.setdefault(level[i], []).append(i)
        ops, op_kind = [], []
        for lv in sorted(by_level):
            members = by_level[lv]
            bins = []
            for i in sorted((i for i in members if self._gather_ok(steps[i], lanes, lim)),
                            key=lambda i: -steps[i].out.n):
                m = steps[i].out.n
                for b in bins:
                    if b[0] + m <= lanes and len(b[1]) < GATHER_MAX_STEPS:
                        b[0] += m
                        b[1].append(i)
                        break
                else:
                    bins.append([m, [i]])
            for _, b in bins:
                ops.append(sorted(b))
                op_kind.append(True)
            for i in members:
                if not self._gather_ok(steps[i], lanes, lim):
                    ops.append([i])
                    op_kind.append(False)
        return ops, op_kind

    def _schedule_list(self, steps, lanes, lim):
        """List scheduling by height (longest path to the result): each op is
        built from the ready steps, highest first; a gather op takes ready
        gatherable steps while it has room, a step that cannot be gathered gets
        an op of its own when it is the highest ready step."""
        n = len(steps)
        prods = [[nd.step for nd in st.inputs if nd.step >= 0] for st in steps]
        cons = [[] for _ in range(n)]
        for i in range(n):
            for j in set(prods[i]):
                cons[j].append(i)
        gok = [self._gather_ok(st, lanes, lim) for st in steps]
        height = [0] * n
        for i in reversed(range(n)):
            height[i] = (1 if gok[i] else 2) + max([height[c] for c in cons[i]] + [0])
        waiting = [len(set(pr)) for pr in prods]
        ready = [i for i in range(n) if waiting[i] == 0]
        ops, op_kind = [], []
        done = 0
        while done < n:
            ready.sort(key=lambda i: (-height[i], i))
            top = ready[0]
            if not gok[top]:
                chosen = [top]
                kind = False
            else:
                chosen, fill = [], 0
                for i in ready:
                    if gok[i] and fill + steps[i].out.n <= lanes and len(chosen) < GATHER_MAX_STEPS:
                        chosen.append(i)
                        fill += steps[i].out.n
                kind = True
            ops.append(sorted(chosen))
            op_kind.append(kind)
            chosen_set = set(chosen)
            ready = [i for i in ready if i not in chosen_set]
            done += len(chosen)
            for i in chosen:
                for c in cons[i]:
                    waiting[c] -= 1
                    if waiting[c] == 0:
                        ready.append(c)
        return ops, op_kind

    @staticmethod
    def _encode_fused(fs, sync_glb):
        """Fused step record: header, per input (loc, mask over union bits),
        then the offset tables: LO[o_lo][j] for the low output bits and
        X[x][j] for every assignment of the summed variables."""
        U = fs.union
        rank_u = len(U)
        pos = {v: rank_u - 1 - a for a, v in enumerate(U)}
        out_set = set(fs.out.dev)
        omask = 0
        for v in fs.out.dev:
            omask |= 1 << pos[v]
        smask = 0
        for v in U:
            if v not in out_set:
                smask |= 1 << pos[v]
        k = rank_u - len(out_set)
        rank_out = len(out_set)
        lo_bits = min(8, rank_out)
        n_in = len(fs.inputs)
        masks = []
        for nd in fs.inputs:
            m = 0
            for v in nd.dev:
                m |= 1 << pos[v]
            masks.append(m)
        words = [0, rank_u | (rank_out << 8) | (k << 16) | (lo_bits << 24), _i32(fs.out.buf.loc), 0,
                 _i32(omask), _i32(smask), 0, 0]
        for nd, m in zip(fs.inputs, masks):
            words += [_i32(nd.buf.loc), _i32(m), 0, 0]
        for o in range(1 << lo_bits):
            u = _pdep(o, omask)
            words += [_pext(u, m) for m in masks]
        for x in range(1 << k):
            u = _pdep(x, smask)
            words += [_pext(u, m) for m in masks]
        chain = OP_CHAIN if getattr(fs, "chain", False) else 0
        words[0] = OP_FUSED | chain | n_in | (sync_glb << 11) | (len(words) << 16)
        return words

    @staticmethod
    def _encode_reduce(rs, sync_glb):
        """Reduction op: w0 = OP_REDUCE | rank_out << 4 | sync << 11 | len << 16,
        then k (summed variables), the output and the input locations."""
        assert not rs.out.buf.is_global and not rs.inputs[0].buf.is_global
        w0 = OP_REDUCE | (rs.rank_out << 4) | (sync_glb << 11) | (REDUCE_HDR << 16)
        return [w0, rs.k, _i32(rs.out.buf.loc), _i32(rs.inputs[0].buf.loc)]

    @staticmethod
    def _encode_step(st, sync_glb):
        rec_len = STEP_HDR + IN_WORDS * len(st.inputs)
        all_lds = not st.out.buf.is_global and not any(nd.buf.is_global for nd in st.inputs)
        w0 = len(st.inputs) | (st.rank_out << 4) | (st.sum_flag << 9) | (sync_glb << 11) | \
            ((1 if all_lds else 0) << 12) | (rec_len << 16)
        words = [w0, 0, _i32(st.out.buf.loc)]
        for nd, low, top in zip(st.inputs, st.lows, st.tops):
            words += [_i32(nd.buf.loc), top, _i32(low), _i32(_runs_word(mask_runs(low)))]
        return words

    def _encode_gather(self, members):
        """Gather op: a 4-word header, then a lane table of ``n_act`` entries of
        NI + 1 words (NI = the members' largest input count), entry l at word
        l * (NI + 1) of the table.  Lane l of the op computes one output element
        of one member: word 0 of its entry holds the output element, word 1 + j
        the LDS elements of input j at the summed variable's 0 (low half) and 1
        (high half).  A member without a sum reads
        element LOC_ZERO as its first input's high half, so its second product
        is zero; missing inputs read LOC_ONE (1 + 0i)."""
        ni = max(len(st.inputs) for st in members)
        one = LOC_ONE | (LOC_ONE << 16)
        cols = [[] for _ in range(ni + 1)]
        for st in members:
            n_in, n = len(st.inputs), st.out.n
            base = st.out.buf.loc
            cols[0].extend(range(base, base + n))
            for j in range(ni):
                if j < n_in:
                    loc = st.inputs[j].buf.loc
                    offs = _pext_table(st.lows[j], n)
                    assert 0 <= loc and loc + offs[-1] + (st.tops[j] if st.sum_flag else 0) < (1 << 16)
                    if st.sum_flag:
                        hi = (loc + st.tops[j]) << 16
                        cols[1 + j].extend(loc + t + ((t << 16) + hi) for t in offs)
                    else:
                        hi = (LOC_ZERO if j == 0 else LOC_ONE) << 16
                        cols[1 + j].extend(loc + t + hi for t in offs)
                else:
                    cols[1 + j].extend([one] * n)
        n_act = len(cols[0])
        assert 1 <= n_act <= self.lanes
        length = GATHER_HDR + (ni + 1) * n_act
        words = [OP_GATHER | len(members) | (length << 16), n_act, GATHER_HDR, ni]
        words += [_i32(v) for entry in zip(*cols) for v in entry]
        return words

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
            if self.bake_angles:
                terms = ()
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
        node = Node(tuple(kept), tuple(int(v) for v in kept if v.size == 2))
        return node, entry, n
