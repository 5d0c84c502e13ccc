"""CPU model of the segmented Zstd parse (gc_zstd_parse_seg_kernel): the serial parse of
zstd_compress.cu:230-338 split into segments that start from a guessed state and are fixed up.

The parse's future from a decision point depends only on the position and the three repeat offsets
(the hash candidate of a position is a pure function of the sub-block).  So:
  * every segment j walks from a start state W_j (first: its first byte and offsets {1,4,8}; segment
    0's is the true start) up to its end e_j: it makes every decision at a position below e_j and stops
    with Y_j = (the end of its last sequence when that passes e_j, else e_j; its offsets).  A walk with
    no sequence passes its start's offsets through: Y_j = (max(W_j's position, e_j), W_j's offsets);
    such a segment is "transparent";
  * each fix-up round first derives every segment's expected start S_j from the segments before it:
    the end state of the nearest earlier segment k that is not transparent (or segment 0), moved on to
    e_{j-1} -- what it would be if the transparent ones between stayed so for the new offsets.  Every
    segment with W_j != S_j walks again from S_j, and stops as soon as it reaches one of its earlier
    walk's states (that walk's start, or the end of one of its sequences with the same offsets): from
    there on the earlier records are the true ones and the earlier Y_j stands.  A walk that reaches
    e_j without meeting them replaces the earlier walk;
  * once every S_j == W_j the walks are consistent and, from segment 0's true start, true.  When the
    segments before j are final, S_j is j's true start, so j is final after its next fix-up: there are
    at most (segments - 1) rounds.  A rare sequence in incompressible data, whose offsets every later
    segment must carry (there is no sequence there to meet an earlier walk at), costs one round, not one
    round per segment.
The records that remain, in order, are the serial parse's sequences; the sequence cap keeps the first
max_seq of them (the serial parse stops emitting there).  Used by tests/test_segment_model.py."""
import numpy as np

HASH_BITS = 14


def candidates(src: np.ndarray) -> np.ndarray:
    """cand(p): the last q < p with q + 3 < n and the same hash, else 0 (the empty slot)."""
    n = src.size
    cand = np.zeros(n, np.int64)
    table = [0] * (1 << HASH_BITS)
    for p in range(max(n - 3, 0)):
        h = ((int(src[p]) | int(src[p + 1]) << 8 | int(src[p + 2]) << 16 | int(src[p + 3]) << 24)
             * 2654435761 % (1 << 32)) >> (32 - HASH_BITS)
        cand[p] = table[h]
        table[h] = p
    return cand


def _ext(src, a, b, n):
    k = 3
    while b + k < n and a + k < n and src[a + k] == src[b + k] and k < 131074:
        k += 1
    return k


def states(x, recs):
    """A walk's decision states: its start, then after each record (position, offsets)."""
    out = [x]
    rep = list(x[1])
    for (h, ln, off) in recs:
        rep = [off, rep[0], rep[1]]
        out.append((h + ln, tuple(rep)))
    return out


def walk(src, cand, start, reps, stop, depth=16, earlier=None):
    """Decisions at positions in [start, min(stop, n - 3)).  Returns (records [(head, len, off)], Y,
    t, positions walked): with `earlier` (a dict of another walk's states -> record index), the walk
    stops at the first state of its own that is one of them, and t is that index (else None)."""
    n = src.size
    rep = list(reps)
    pos = start
    last_end = start
    recs = []
    if earlier is not None and (pos, tuple(rep)) in earlier:
        return recs, None, earlier[(pos, tuple(rep))], 0
    walked = 0
    while pos + 3 < n and pos < stop:
        c = int(cand[pos])
        off, ln = pos - c, 0
        if 0 < off <= pos and src[c] == src[pos] and src[c + 1] == src[pos + 1] and src[c + 2] == src[pos + 2]:
            ln = _ext(src, c, pos, n)
        r = 0
        while ln < 3 and r < 3:
            if rep[r] <= pos:
                rp = pos - rep[r]
                if src[rp] == src[pos] and src[rp + 1] == src[pos + 1] and src[rp + 2] == src[pos + 2]:
                    ln, off = _ext(src, rp, pos, n), rep[r]
            r += 1
        head = pos
        if ln >= 3 and depth > 1 and pos + 4 < n:
            c2 = int(cand[pos + 1])
            o2 = pos + 1 - c2
            if 0 < o2 <= pos + 1 and src[c2] == src[pos + 1] and src[c2 + 1] == src[pos + 2] \
                    and src[c2 + 2] == src[pos + 3]:
                l2 = _ext(src, c2, pos + 1, n)
                if l2 > ln + 1:
                    head, ln, off = pos + 1, l2, o2
        if ln >= 3:
            recs.append((head, ln, off))
            rep = [off, rep[0], rep[1]]
            walked += head + ln - pos
            pos = head + ln
            last_end = pos
            if earlier is not None and (pos, tuple(rep)) in earlier:
                return recs, None, earlier[(pos, tuple(rep))], walked
        else:
            pos += 1
            walked += 1
    y = (last_end if last_end >= stop else stop, tuple(rep))
    return recs, y, None, walked


def bounds(n, segments):
    """The kernel's segments: 64-aligned, ceil(n / segments) rounded up to 64 bytes each."""
    per = -(-n // segments)  # ceil(n / segments)
    seg = -(-per // 64) * 64
    starts = [min(j * seg, n) for j in range(segments)]
    return starts, [min(s + seg, n) for s in starts]


def segmented(src, segments=16, depth=16, max_seq=None, stats=None):
    """(records of the whole sub-block, fix-up rounds taken)."""
    cand = candidates(src)
    starts, ends = bounds(src.size, segments)
    S = len(ends)
    x = [(starts[j], (1, 4, 8)) for j in range(S)]  # the state the records in spec[j] start from
    spec, y = [], []
    for j in range(S):
        r, yy, _, _ = walk(src, cand, x[j][0], x[j][1], ends[j], depth)
        spec.append(r)
        y.append(yy)
    w = list(x)  # the state the segment's current records start from
    fix, tail = [[] for _ in range(S)], [0] * S
    rounds = 0
    while True:
        trans = [len(fix[j]) + len(spec[j]) - tail[j] == 0 for j in range(S)]
        expect = [w[0]]
        for j in range(1, S):
            k = j - 1
            while k > 0 and trans[k]:
                k -= 1
            expect.append((max(y[k][0], starts[j]), y[k][1]))
        need = [j > 0 and expect[j] != w[j] for j in range(S)]
        if not any(need):
            break
        assert rounds < S - 1
        for j in range(1, S):
            if not need[j]:
                continue
            earlier = {}
            for t, st in enumerate(states(x[j], spec[j])):
                earlier.setdefault(st, t)
            r, yy, t, wk = walk(src, cand, expect[j][0], expect[j][1], ends[j], depth, earlier)
            if stats is not None:
                stats.append((rounds, j, wk, t is not None))
            w[j] = expect[j]
            if t is not None:
                fix[j], tail[j] = r, t
            else:
                spec[j], fix[j], tail[j], x[j], y[j] = r, [], 0, expect[j], yy
        rounds += 1
    out = []
    for j in range(S):
        out += fix[j] + spec[j][tail[j]:]
    if max_seq is not None:
        out = out[:max_seq]
    return out, rounds
