"""Numpy interpreter of the executor's plan programs (test helper, CPU only).

Executes a compiled plan exactly as ``exec_kernel`` in
``qtensor_amd/csrc/qtensor_amd.hip`` defines it -- gate build from FP64 angles,
pext-run offsets, LDS/global locations, summed-variable offsets -- so the plan
compiler can be checked against the oracle without a GPU.
"""
import struct

import numpy as np

from qtensor_amd import plan as P


def _f32(bits):
    return struct.unpack("<f", struct.pack("<i", int(bits)))[0]


def _f64(lo, hi):
    return struct.unpack("<d", struct.pack("<ii", int(lo), int(hi)))[0]


def natural_gate(kind, a):
    re = np.zeros(8, dtype=np.complex128)
    if kind == 1:
        re[0], re[1] = 1, np.exp(1j * np.pi * a)
    elif kind in (2, 4):
        c, s, g = np.cos(np.pi * a / 2), np.sin(np.pi * a / 2), np.exp(1j * np.pi * a / 2)
        m = [[c, -1j * s], [-1j * s, c]] if kind == 2 else [[c, -s], [s, c]]
        re[:4] = g * np.asarray(m).reshape(-1)
    elif kind == 3:
        p, m = np.exp(1j * np.pi * a / 2), np.exp(-1j * np.pi * a / 2)
        re[:4] = [m, p, p, m]
    elif kind == 5:
        re[:4] = np.array([1, 1, 1, -1]) / np.sqrt(2)
    elif kind == 6:
        re[[0, 3, 5, 6]] = 1
    elif kind == 7:
        re[:2] = [1, -1]
    elif kind == 8:
        re[[0, 3]] = 1
    else:
        raise ValueError(kind)
    return re


def pext(x, m):
    r, bit = 0, 1
    while m:
        low = m & -m
        if x & low:
            r |= bit
        bit <<= 1
        m &= m - 1
    return r


def pext_vec(o, mask):
    o = np.atleast_1d(np.asarray(o, dtype=np.int64))
    r = np.zeros_like(o)
    bit = 0
    for b in range(32):
        if (mask >> b) & 1:
            r |= ((o >> b) & 1) << bit
            bit += 1
    return r


def _runs_offset(o, rw, mask):
    nr = rw >> 30
    if nr == 3:
        return pext_vec(o, mask)
    off = np.zeros_like(np.atleast_1d(o))
    for r in range(nr):
        word = (rw >> (15 * r)) & 0x7FFF
        src, wd, dst = word & 31, (word >> 5) & 31, (word >> 10) & 31
        off = off | (((np.atleast_1d(o) >> src) & ((1 << wd) - 1)) << dst)
    return off


def run_plan(words, params, slice_bits=0, dtype=np.complex128, lanes=64):
    """Execute a plan; ``lanes`` must be the item size the plan was compiled for."""
    w = [int(x) for x in np.asarray(words)]
    n_gates, n_ops, goff, soff, result_loc, lds_need = w[:6]
    glb_need = (w[6] & 0xFFFFFFFF) | ((w[9] & 0xFFFFFFFF) << 32)
    ops_words = w[8]
    lds = np.full(max(lds_need, P.GATE_BASE), np.nan + 0j, dtype=dtype)
    glb = np.full(max(glb_need, 1), np.nan + 0j, dtype=dtype)
    lds[P.LOC_ONE] = 1
    lds[P.LOC_ZERO] = 0
    # wide plans (no LDS) keep their gate tensors in the global arena
    gate_buf = glb if lds_need == 0 else lds
    for g in range(n_gates):
        e = w[goff + g * P.GATE_WORDS: goff + (g + 1) * P.GATE_WORDS]
        kind, dst, n, smap = e[0], e[1], e[2], e[3] & 0xFFFFFFFF
        alpha = _f64(e[8], e[9])
        if e[4] >= 0:
            alpha += _f32(e[5]) * params[e[4]]
        if e[6] >= 0:
            alpha += _f32(e[7]) * params[e[6]]
        shift = 0
        nsl = e[11] & 0xFF
        if nsl > 0:
            shift += ((slice_bits >> ((e[11] >> 8) & 0xFF)) & 1) * e[12]
        if nsl > 1:
            shift += ((slice_bits >> ((e[11] >> 16) & 0xFF)) & 1) * e[13]
        nat = natural_gate(kind, alpha)
        for k in range(n):
            gate_buf[dst + k] = nat[(((smap >> (4 * k)) & 15) + shift) & 7]
    pos = soff
    for _ in range(n_ops):
        w0 = w[pos] & 0xFFFFFFFF
        length = w0 >> 16
        assert pos + length <= soff + ops_words
        if w0 & P.OP_FUSED:
            n_in = w0 & 15
            meta = w[pos + 1] & 0xFFFFFFFF
            rank_u, rank_out, k, lo_bits = meta & 255, (meta >> 8) & 255, (meta >> 16) & 255, meta >> 24
            out_loc = w[pos + 2] & 0xFFFFFFFF
            omask, smask = w[pos + 4] & 0xFFFFFFFF, w[pos + 5] & 0xFFFFFFFF
            assert bin(omask).count("1") == rank_out and bin(smask).count("1") == k
            assert omask & smask == 0 and bin(omask | smask).count("1") == rank_u
            assert 1 <= n_in <= P.FUSED_MAX_IN and lo_bits == min(8, rank_out)
            ins = [(w[pos + P.FUSED_HDR + 4 * j] & 0xFFFFFFFF, w[pos + P.FUSED_HDR + 4 * j + 1] & 0xFFFFFFFF)
                   for j in range(n_in)]
            tab = pos + P.FUSED_HDR + 4 * n_in
            lo_tab = np.array(w[tab: tab + (n_in << lo_bits)], dtype=np.int64).reshape(-1, n_in)
            x_tab = np.array(w[tab + (n_in << lo_bits): tab + (n_in << lo_bits) + (n_in << k)],
                             dtype=np.int64).reshape(-1, n_in)
            assert pos + length == tab + (n_in << lo_bits) + (n_in << k)
            if w0 & P.OP_CHAIN:
                # the device's factorisation: sum over y of D times (sum over x
                # of A B), x = the (at most CHAIN_MAX_X lowest) summed variables D
                # lacks; the value is the same sum as any fused op's
                not_d = smask & ~ins[2][1]
                kx = min(bin(not_d).count("1"), P.CHAIN_MAX_X)
                assert n_in == 3 and rank_out >= 8 and kx >= 1 and k - kx <= P.CHAIN_MAX_Y
                assert bin(ins[2][1]).count("1") <= P.CHAIN_MAX_D and rank_u <= 31
            else:
                assert rank_u <= 30
            o = np.arange(1 << rank_out, dtype=np.int64)
            x = np.arange(1 << k, dtype=np.int64)
            # the device splits o into its low lo_bits (table) and the rest
            # (computed); both are checked against the direct pext/pdep form
            uo = np.array([P._pdep(int(v), omask) for v in o], dtype=np.int64)
            ux = np.array([P._pdep(int(v), smask) for v in x], dtype=np.int64)
            acc = np.zeros(1 << rank_out, dtype=dtype)
            offs_o, offs_x = [], []
            for j, (loc, m) in enumerate(ins):
                ao = pext_vec(uo, m)
                bx = pext_vec(ux, m)
                assert (lo_tab[:, j] == ao[:1 << lo_bits]).all() and (x_tab[:, j] == bx).all()
                hi = pext_vec(np.array([P._pdep(int(v) << lo_bits, omask) for v in (o >> lo_bits)],
                                       dtype=np.int64), m)
                assert (hi + lo_tab[o & ((1 << lo_bits) - 1), j] == ao).all()
                offs_o.append(ao)
                offs_x.append(bx)
            for xi in range(1 << k):
                prod = np.ones(1 << rank_out, dtype=dtype)
                for j, (loc, m) in enumerate(ins):
                    buf = glb if loc & P.GLOBAL else lds
                    prod = prod * buf[(loc & ~P.GLOBAL) + offs_o[j] + offs_x[j][xi]]
                acc = acc + prod
            buf = glb if out_loc & P.GLOBAL else lds
            buf[(out_loc & ~P.GLOBAL) + o] = acc
        elif w0 & P.OP_REDUCE:
            rank_out, k = (w0 >> 4) & 31, w[pos + 1]
            out_loc, in_loc = w[pos + 2], w[pos + 3]
            assert length == P.REDUCE_HDR and (1 << rank_out) <= P.REDUCE_MAX_OUT and k >= 1
            n_out = 1 << rank_out
            block = lds[in_loc: in_loc + (n_out << k)].reshape(1 << k, n_out)
            lds[out_loc: out_loc + n_out] = block.sum(axis=0)
        elif w0 & P.OP_GATHER:
            n_steps, n_act, tab, ni = w0 & 0x7FFF, w[pos + 1], w[pos + 2], w[pos + 3]
            assert 1 <= n_steps <= n_act <= lanes and tab == P.GATHER_HDR
            assert 1 <= ni <= P.MAX_IN and length == P.GATHER_HDR + (ni + 1) * n_act
            table = [x & 0xFFFFFFFF for x in w[pos + tab: pos + tab + (ni + 1) * n_act]]
            cols = [table[c::ni + 1] for c in range(ni + 1)]      # entry l = words l*(NI+1) ...
            results = []
            for lane in range(n_act):
                x, y = 1 + 0j, 1 + 0j
                for j in range(ni):
                    e = cols[1 + j][lane]
                    x *= lds[e & 0xFFFF]
                    y *= lds[e >> 16]
                results.append((cols[0][lane], x + y))
            for addr, v in results:        # all reads of an op precede its writes
                lds[addr] = v
        else:
            out_loc = w[pos + 2] & 0xFFFFFFFF
            n_in, rank_out, flags = w0 & 15, (w0 >> 4) & 31, (w0 >> 9) & 1
            assert length == P.STEP_HDR + P.IN_WORDS * n_in
            assert 1 <= n_in <= P.MAX_IN, "the device executor takes at most MAX_IN inputs per step"
            n_sum = 2 if flags else 1
            n_out = 1 << rank_out
            o = np.arange(n_out, dtype=np.int64)
            acc = np.zeros(n_out, dtype=dtype)
            for sv in range(n_sum):
                prod = np.ones(n_out, dtype=dtype)
                for j in range(n_in):
                    d = w[pos + P.STEP_HDR + j * P.IN_WORDS: pos + P.STEP_HDR + (j + 1) * P.IN_WORDS]
                    loc, top, mask, rw = d[0] & 0xFFFFFFFF, d[1], d[2] & 0xFFFFFFFF, d[3] & 0xFFFFFFFF
                    off = _runs_offset(o, rw, mask)
                    if sv:
                        off = off + top
                    buf = glb if loc & P.GLOBAL else lds
                    prod = prod * buf[(loc & ~P.GLOBAL) + off]
                acc = acc + prod
            buf = glb if out_loc & P.GLOBAL else lds
            buf[(out_loc & ~P.GLOBAL) + o] = acc
        pos += length
    buf = glb if result_loc & P.GLOBAL else lds
    loc = (result_loc & 0xFFFFFFFF) & ~P.GLOBAL
    if w[11]:            # open variables: the 2^rank result tensor
        return np.array(buf[loc: loc + (1 << w[11])], dtype=np.complex128)
    return complex(buf[loc])
