"""Plain-Python restatement of the plan compiler's interval allocator, the
checker for the native ``qb_place_intervals``: each buffer, in placement
order, takes the lowest offset >= base that is free over its whole op
interval among the buffers already placed (first fit over the overlapping
ones sorted by offset, ties in placement order)."""


def place_order(sizes, starts, ends, order, base):
    placed = []          # (off, end, start, last) in placement order
    offsets = [0] * len(sizes)
    peak = base
    for idx in order:
        n, s, e = int(sizes[idx]), int(starts[idx]), int(ends[idx])
        live = sorted((p for p in placed if p[2] <= e and s <= p[3]), key=lambda p: p[0])
        off = base
        for a, z, _, _ in live:
            if off + n <= a:
                break
            if z > off:
                off = z
        offsets[idx] = off
        placed.append((off, off + n, s, e))
        peak = max(peak, off + n)
    return offsets, peak
