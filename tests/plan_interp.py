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


def run_plan(words, params, slice_bits=0, dtype=np.complex128):
    w = [int(x) for x in np.asarray(words)]
    n_gates, n_steps, goff, soff, result_loc, lds_need, glb_need = w[:7]
    lds = np.full(max(lds_need, 1), np.nan + 0j, dtype=dtype)
    glb = np.full(max(glb_need, 1), np.nan + 0j, dtype=dtype)
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
            lds[dst + k] = nat[(((smap >> (4 * k)) & 15) + shift) & 7]
    pos = soff
    for _ in range(n_steps):
        n_in, rank_out, flags, out_loc = w[pos:pos + 4]
        out_loc &= 0xFFFFFFFF
        descs = []
        for j in range(n_in):
            d = w[pos + 4 + j * P.IN_WORDS: pos + 4 + (j + 1) * P.IN_WORDS]
            descs.append(((d[0] & 0xFFFFFFFF), d[1], d[2], d[3] & 0xFFFFFFFF, d[4:8]))
        n_sum = 2 if flags & 1 else 1
        n_out = 1 << rank_out
        o = np.arange(n_out, dtype=np.int64)
        acc = np.zeros(n_out, dtype=dtype)
        for sv in range(n_sum):
            prod = np.ones(n_out, dtype=dtype)
            for loc, top, nr, mask, runs in descs:
                if nr > P.MAX_RUNS:
                    off = np.array([pext(int(x), mask) for x in o], dtype=np.int64)
                else:
                    off = np.zeros(n_out, dtype=np.int64)
                    for r in range(nr):
                        word = runs[r]
                        src, wd, dst = word & 31, (word >> 5) & 63, (word >> 11) & 31
                        off |= ((o >> src) & ((1 << wd) - 1)) << dst
                if sv:
                    off = off + top
                buf = glb if loc & P.GLOBAL else lds
                prod = prod * buf[(loc & ~P.GLOBAL) + off]
            acc = acc + prod
        buf = glb if out_loc & P.GLOBAL else lds
        buf[(out_loc & ~P.GLOBAL) + o] = acc
        pos += 4 + n_in * P.IN_WORDS
    buf = glb if result_loc & P.GLOBAL else lds
    return complex(buf[(result_loc & 0xFFFFFFFF) & ~P.GLOBAL])
