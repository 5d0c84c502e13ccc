"""CPU models of the gfx950 kernels' algorithms, lane by lane (numpy), for tests.

They restate what gpu-compressor_amd/csrc/kernels.hip does per wavefront so that the window /
strip logic can be checked against the oracle on the CPU: a kernel-logic mistake shows up here
without a GPU.  They are not used by the library.
"""
import numpy as np

WAVE = 64
SHORT = 64  # GC_ZSTD_SHORT


def hash4(d, p):
    v = int(d[p]) | int(d[p + 1]) << 8 | int(d[p + 2]) << 16 | int(d[p + 3]) << 24
    return ((v * 2654435761) & 0xFFFFFFFF) >> 18


def cands(d):
    """cand(p) = last q < p with q + 3 < n and H(q) == H(p), else None (the table semantics)."""
    n = len(d)
    last = {}
    out = [None] * n
    for p in range(n):
        if p + 3 < n:
            h = hash4(d, p)
            out[p] = last.get(h)
            last[h] = p
    return out


def common(d, a, b, start, limit):
    k = start
    while k < limit and d[a + k] == d[b + k]:
        k += 1
    return min(k, limit)


def v3(d, x):
    n = len(d)
    return tuple(int(d[x + i]) if x + i < n else 0 for i in range(3))


def zstd_info(d, strip=8):
    """gc_zstd_info_kernel: info word per position (len << 16 | (p - cand) & 0xFFFF: the low half of a
    verified word is the match offset), strips walked backwards."""
    n = len(d)
    cand = cands(d)
    info = [0] * n
    for p0 in range(0, n, strip):
        if p0 + 3 >= n:
            break
        c = []
        for j in range(strip):
            p = p0 + j
            cj = cand[p] if (p < n and p + 3 < n) else None
            c.append(0 if cj is None else cj)
        next_len = 0
        out = [0] * strip
        for j in range(strip - 1, -1, -1):
            p = p0 + j
            ln = 0
            if p + 3 < n and p > 0 and v3(d, c[j]) == v3(d, p):
                lim = min(n - p, SHORT)
                if j + 1 < strip and next_len and c[j + 1] == c[j] + 1:
                    ln = min(next_len + 1, lim)
                else:
                    ln = common(d, c[j], p, 3, lim)
            next_len = ln
            out[j] = ln << 16 | ((p - c[j]) & 0xFFFF)
        for j in range(strip):
            if p0 + j + 3 < n:
                info[p0 + j] = out[j]
    return info


def zstd_parse(d, info, depth=16, max_seq=None):
    """gc_zstd_parse_kernel: window of 64 lanes + 64 ahead, slid after each sequence."""
    n = len(d)
    if max_seq is None:
        max_seq = 65536 // 3
    rep = [1, 4, 8]
    pos = lit_start = 0
    seqs = []

    def peek(x):
        if x + 3 < n:
            return info[x], v3(d, x)
        return 0, (0, 0, 0)

    def eq(x, r, w):
        return x + 3 < n and r <= x and v3(d, x - r) == w

    def flags(x, w):
        if x + 3 >= n:
            return 0
        return (1 if eq(x, rep[0], w) else 0) | (2 if eq(x, rep[1], w) else 0) | (4 if eq(x, rep[2], w) else 0)

    e = [0] * WAVE
    w = [(0, 0, 0)] * WAVE
    f = [0] * WAVE
    ae = [0] * WAVE
    aw = [(0, 0, 0)] * WAVE
    if n > 3:
        for i in range(WAVE):
            e[i], w[i] = peek(i)
            f[i] = flags(i, w[i])
            ae[i], aw[i] = peek(WAVE + i)
    while pos + 3 < n and len(seqs) < max_seq:
        m = [i for i in range(WAVE) if (e[i] >> 16) or f[i]]
        if not m:
            pos += WAVE
            e, w = ae[:], aw[:]
            f = [flags(pos + i, w[i]) for i in range(WAVE)]
            for i in range(WAVE):
                ae[i], aw[i] = peek(pos + WAVE + i)
            continue
        k = m[0]
        head = pos + k
        ek = e[k]
        if ek >> 16:
            off = ek & 0xFFFF
            ln = ek >> 16
            if ln >= SHORT:
                ln = common(d, head - off, head, SHORT, n - head)
        else:
            fk = f[k]
            off = rep[0] if fk & 1 else (rep[1] if fk & 2 else rep[2])
            ln = common(d, head - off, head, 3, n - head)
        if depth > 1 and head + 4 < n:
            e2 = e[k + 1] if k + 1 < WAVE else ae[0]
            if e2 >> 16:
                off2 = e2 & 0xFFFF
                l2 = e2 >> 16
                if l2 >= SHORT:
                    l2 = common(d, head + 1 - off2, head + 1, SHORT, n - head - 1)
                if l2 > ln + 1:
                    head += 1
                    ln = l2
                    off = off2
        seqs.append((head - lit_start, ln - 3, off))
        o0, o1, o2 = rep
        rep = [off, o0, o1]
        shift = head + ln - pos
        pos = head + ln
        lit_start = pos
        if pos + 3 >= n:
            break
        if shift >= 2 * WAVE:
            for i in range(WAVE):
                e[i], w[i] = peek(pos + i)
                f[i] = flags(pos + i, w[i])
                ae[i], aw[i] = peek(pos + WAVE + i)
            continue
        ne, nw, nf, nae, naw = [0] * WAVE, [None] * WAVE, [0] * WAVE, [0] * WAVE, [None] * WAVE
        for i in range(WAVE):
            fr = i + shift
            fl = fr & (WAVE - 1)
            stay = fr < WAVE
            x = pos + i
            if stay:
                ne[i], nw[i] = e[fl], w[fl]
            elif fr < 2 * WAVE:
                ne[i], nw[i] = ae[fl], aw[fl]
            else:
                ne[i], nw[i] = peek(x)
            if stay:
                sf = f[fl]
                if off == o0:
                    b0 = sf & 1
                elif off == o1:
                    b0 = (sf >> 1) & 1
                elif off == o2:
                    b0 = (sf >> 2) & 1
                else:
                    b0 = 1 if eq(x, off, nw[i]) else 0
                nf[i] = b0 | ((sf & 3) << 1)
            else:
                nf[i] = flags(x, nw[i])
            if shift < WAVE and stay:
                nae[i], naw[i] = ae[fl], aw[fl]
            else:
                nae[i], naw[i] = peek(pos + WAVE + i)
        e, w, f, ae, aw = ne, nw, nf, nae, naw
    return seqs


def zstd_parse_fixed(d, info, depth=16, max_seq=None, scan=False):
    """gc_zstd_parse_kernel, aligned-window form: lanes hold positions wbase + i (wbase a multiple
    of 64) and the next window ahead; lanes before pos are masked out of the search.  With `scan`,
    a second window in a row with no head starts the look-ahead: 256 positions per step, 4 per lane,
    to the first window that may hold a head (a lane whose 4 positions straddle a repeat offset
    counts as a hit)."""
    n = len(d)
    lim = n - 3
    if max_seq is None:
        max_seq = 65536 // 3
    rep = [1, 4, 8]
    pos = lit_start = 0
    seqs = []

    def peek(x):
        if x + 3 < n:
            return info[x], v3(d, x)
        return 0, (0, 0, 0)

    def eq(x, r, w):
        return x + 3 < n and r <= x and v3(d, x - r) == w

    def flags(x, w):
        return (1 if eq(x, rep[0], w) else 0) | (2 if eq(x, rep[1], w) else 0) | (4 if eq(x, rep[2], w) else 0)

    def look_ahead(sb):
        while sb < lim:
            hits = []
            for lane in range(WAVE):
                x0 = sb + 4 * lane
                hit = False
                if x0 < lim:
                    for x in range(x0, min(x0 + 4, lim)):
                        if info[x] >> 16 or any(x0 >= r and v3(d, x - r) == v3(d, x) for r in rep):
                            hit = True
                    if any(x0 < r <= x0 + 3 for r in rep):
                        hit = True
                hits.append(hit)
            if any(hits):
                return (sb + 4 * hits.index(True)) & ~(WAVE - 1)
            sb += 4 * WAVE
        return lim

    slid = False
    wbase = 0
    e, w = zip(*[peek(i) for i in range(WAVE)]) if n > 3 else ([0] * WAVE, [(0, 0, 0)] * WAVE)
    e, w = list(e), list(w)
    f = [flags(i, w[i]) for i in range(WAVE)]
    ae, aw = map(list, zip(*[peek(WAVE + i) for i in range(WAVE)]))
    while pos + 3 < n and len(seqs) < max_seq:
        lo = pos - wbase
        m = [i for i in range(lo, WAVE) if (e[i] >> 16) or f[i]]
        if not m:
            if scan and slid:
                slid = False
                nb = look_ahead(wbase + WAVE)
                if nb >= lim:
                    break
                wbase = pos = nb
                e, w = map(list, zip(*[peek(nb + i) for i in range(WAVE)]))
            else:
                slid = True
                wbase += WAVE
                pos = wbase
                e, w = ae, aw
            f = [flags(wbase + i, w[i]) for i in range(WAVE)]
            ae, aw = map(list, zip(*[peek(wbase + WAVE + i) for i in range(WAVE)]))
            continue
        slid = False
        k = m[0]
        head = wbase + k
        ek = e[k]
        if ek >> 16:
            off = ek & 0xFFFF
            ln = ek >> 16
            if ln >= SHORT:
                ln = common(d, head - off, head, SHORT, n - head)
        else:
            fk = f[k]
            off = rep[0] if fk & 1 else (rep[1] if fk & 2 else rep[2])
            ln = common(d, head - off, head, 3, n - head)
        if depth > 1 and head + 4 < n:
            e2 = e[k + 1] if k + 1 < WAVE else ae[0]
            if e2 >> 16:
                off2 = e2 & 0xFFFF
                l2 = e2 >> 16
                if l2 >= SHORT:
                    l2 = common(d, head + 1 - off2, head + 1, SHORT, n - head - 1)
                if l2 > ln + 1:
                    head += 1
                    ln = l2
                    off = off2
        seqs.append((head - lit_start, ln - 3, off))
        o0, o1, o2 = rep
        rep = [off, o0, o1]
        pos = head + ln
        lit_start = pos
        if pos + 3 >= n:
            break
        if pos >= wbase + WAVE:
            nb = pos & ~(WAVE - 1)
            if nb == wbase + WAVE:
                e, w = ae, aw
            else:
                e, w = map(list, zip(*[peek(nb + i) for i in range(WAVE)]))
            wbase = nb
            f = [flags(wbase + i, w[i]) for i in range(WAVE)]
            ae, aw = map(list, zip(*[peek(wbase + WAVE + i) for i in range(WAVE)]))
        else:
            nf = []
            for i in range(WAVE):
                sf = f[i]
                if off == o0:
                    b0 = sf & 1
                elif off == o1:
                    b0 = (sf >> 1) & 1
                elif off == o2:
                    b0 = (sf >> 2) & 1
                else:
                    b0 = 1 if eq(wbase + i, off, w[i]) else 0
                nf.append(b0 | ((sf & 3) << 1))
            f = nf
    return seqs

