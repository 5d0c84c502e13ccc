"""Pins of the CPU oracle (oracle/gc_oracle.c) against what the reference's algorithm fixes.

The reference ships no tests, fixtures or golden vectors for this path, and its kernels cannot be
built here (CUDA only).  The oracle is therefore checked against:
  * hand-derived known answers (hash values, repeat-offset path, lazy step and its strict rule,
    sequence cap), each worked out from zstd_compress.cu / lzma2_match_find.cu in the comments;
  * closed forms (all-zero and periodic inputs, tiny inputs);
  * an independent brute-force formulation on small inputs (candidate = last earlier position
    with the same hash, found by a linear scan; then the parse / per-position rule);
  * invariants on corpus data (reconstruction, match verification and maximality).
"""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

SUB = 65536
MASK24 = 0xFFFFFF


# ---------------------------------------------------------------------------------------------
# hash (zstd_compress.cu:228, lzma2_match_find.cu:92): (LE32 * 2654435761) mod 2^32 >> 18
# ---------------------------------------------------------------------------------------------
def test_hash_known_answers():
    # 0 * anything = 0
    assert oracle.hash4(b"\x00\x00\x00\x00") == 0
    # LE32 = 1: 2654435761 >> 18 = floor(2654435761 / 262144) = 10125 (262144 * 10125 = 2654208000)
    assert oracle.hash4(b"\x01\x00\x00\x00") == 10125
    # LE32 = 2^24: 2654435761 = 0x9E3779B1, times 2^24 mod 2^32 = 0xB1000000, >> 18 = 0x2C40 = 11328
    assert oracle.hash4(b"\x00\x00\x00\x01") == 11328


# ---------------------------------------------------------------------------------------------
# Zstd closed forms and known answers
# ---------------------------------------------------------------------------------------------
def zstd_ref(data, depth=16, max_seq=None, chunk=SUB):
    seq, sc, lc, lit = oracle.zstd_match_find(np.asarray(data, np.uint8), chunk, depth, max_seq, literals=True)
    return seq, sc, lc, lit


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4])
def test_zstd_tiny_inputs_are_all_literals(n):
    # the loop runs while pos + 3 < n: for n <= 4 only position 0 is a head, and it cannot match
    # (offset pos - 0 = 0 fails `match_offset > 0`, every repeat offset exceeds pos)
    seq, sc, lc, lit = zstd_ref(np.arange(n, dtype=np.uint8))
    assert sc.tolist() == ([0] if n else []) and lc.tolist() == ([n] if n else [])


@pytest.mark.parametrize("n", [5, 6, 100, 4096, SUB])
@pytest.mark.parametrize("depth", [1, 16, 64])
def test_zstd_all_zero(n, depth):
    # pos 0: no candidate (offset 0), reps 1,4,8 > 0.  pos 1: the empty slot reads as position 0,
    # offset 1, and the match runs to the end: length n-1.  Lazy at pos 2 gives n-2, not > n.
    seq, sc, lc, lit = zstd_ref(np.zeros(n, np.uint8), depth)
    assert sc[0] == 1 and lc[0] == 1
    assert seq[0, 0].tolist() == [1, n - 1 - 3, 1]


@pytest.mark.parametrize("k", [9, 13, 64, 200])
def test_zstd_periodic(k):
    # period k of distinct bytes: no 3-byte string repeats within a period and repeat offsets
    # 1,4,8 < k see different bytes, so positions < k are literals.  At pos k the hash slot
    # holds position 0 (no other position of the period shares H(0): checked below), offset k,
    # and the match runs to the end.
    n = 3000
    data = datagen.periodic(n, k, k)
    hs = [oracle.hash4(bytes(data[q:q + 4])) for q in range(k)]
    assert hs.count(hs[0]) == 1, "precondition: H(0) unique within the period"
    seq, sc, lc, _ = zstd_ref(data)
    assert sc[0] == 1 and seq[0, 0].tolist() == [k, n - k - 3, k] and lc[0] == k


def _distinct_hashes(data, allowed):
    """Precondition of a hand-derived case: 4-grams with equal hash are equal 4-grams, except the
    listed position pairs."""
    seen = {}
    for q in range(len(data) - 3):
        h = oracle.hash4(bytes(data[q:q + 4]))
        if h in seen and bytes(data[seen[h]:seen[h] + 4]) != bytes(data[q:q + 4]):
            raise AssertionError(f"hash collision {seen[h]} / {q}")
        seen.setdefault(h, q)


def _fill(n, start):
    return bytes(range(start, start + n))


def test_zstd_repeat_offset_known_answer():
    # 0..15  "ABCDEFGHIJKLMNOP"   16..31 "abcdefghijklmnop"   32..39 "abcdefgh"  40 'X'
    # 41..43 "jkl"   44.. "QRSTUVWYZ" (n = 53)
    # pos 32: hash slot of "abcd" holds 16 -> offset 16, common run "abcdefgh" = 8, then
    #         src[40]='X' != src[24]='i'.  Lazy at 33 finds 17 with 7, not > 9.
    #         -> {ll 32, ml 8, off 16}; reps become (16, 1, 4).
    # pos 40: "Xjkl" never seen (slot empty -> position 0, "ABC" != "Xjk"); reps 16,1,4 see 'i','h','e'.
    # pos 41: "jklQ" never seen; rep0 = 16 -> src[25..27] = "jkl": 3 bytes, then 'm' != 'Q'.
    #         Lazy at 42: "klQR" never seen.  -> {ll 1, ml 3, off 16} (a repeat-offset match).
    # pos 44..49: uppercase from Q on never seen and no rep hits -> literals 44..52 (9 bytes).
    data = np.frombuffer(b"ABCDEFGHIJKLMNOP" + b"abcdefghijklmnop" + b"abcdefgh" + b"X" + b"jkl" + b"QRSTUVWYZ",
                         np.uint8)
    assert data.size == 53
    _distinct_hashes(data, [])
    seq, sc, lc, lit = zstd_ref(data)
    assert sc[0] == 2
    assert seq[0, :2].tolist() == [[32, 5, 16], [1, 0, 16]]
    assert lc[0] == 32 + 1 + 9
    assert bytes(lit[:lc[0]]) == bytes(data[:32]) + b"X" + bytes(data[44:])


def _lazy_case(a_run):
    # F1 (8) | A = a_run (from "bcdefghijk") | F2 (8) | B = "abcdZ" | F3 (8) | "abcdefghijk" | F4 (6)
    f = iter(range(0x80, 0x100))
    F = lambda k: bytes(next(f) for _ in range(k))  # noqa: E731 - unique filler bytes
    parts = [F(8), a_run, F(8), b"abcdZ", F(8)]
    head = sum(map(len, parts))
    data = b"".join(parts) + b"abcdefghijk" + F(6)
    return np.frombuffer(data, np.uint8), head, len(parts[0])


def test_zstd_lazy_step_taken():
    # pos p ("abcd..."): hash slot of "abcd" holds B (26): common "abcd" = 4 ('Z' != 'e').
    # Lazy at p+1: slot of "bcde" holds A (8): common "bcdefghijk" = 10 > 4 + 1 -> the match moves
    # to p+1: {ll p+1, ml 10, off p+1-8}.  Everything else is literal.
    data, p, a = _lazy_case(b"bcdefghijk")
    _distinct_hashes(data, [])
    seq, sc, lc, _ = zstd_ref(data, 16)
    assert sc[0] == 1 and seq[0, 0].tolist() == [p + 1, 10 - 3, p + 1 - a]
    assert lc[0] == data.size - 10
    # depth 1 (level 3): no lazy step: {ll p, ml 4, off p-26}; then at p+4 "efgh" finds A+3 with
    # "efghijk" = 7 and offset p+4-(a+3); rep0 was p-26, so this is a new offset.
    seq, sc, lc, _ = zstd_ref(data, 1)
    assert sc[0] == 2
    assert seq[0, :2].tolist() == [[p, 1, p - 26], [0, 4, (p + 4) - (a + 3)]]


def test_zstd_lazy_rule_is_strict():
    # A = "bcdef" + filler: the lazy candidate gives 5 = 4 + 1, not > 4 + 1, so the match stays at
    # p with length 4 and offset p-26.  Then pos p+4 ("efgh"): never seen as a whole 4-gram; rep0
    # (p-26) sees 'Z', rep1 (1) sees 'd', rep2 (4) sees 'a': all literal to the end.
    data, p, a = _lazy_case(b"bcdef" + bytes([0x7F, 0x7E, 0x7D, 0x7C, 0x7B]))
    _distinct_hashes(data, [])
    seq, sc, lc, _ = zstd_ref(data, 16)
    assert sc[0] == 1 and seq[0, 0].tolist() == [p, 1, p - 26]


def test_zstd_sequence_cap():
    # zstd_compress.cu:309: once max_sequences_per_chunk sequences are out, later matches are not
    # emitted and everything after the last emitted match is literal.
    data = datagen.text(SUB, 3)
    full, fsc, _, _ = zstd_ref(data)
    for cap in (1, 2, 7, 100):
        seq, sc, lc, _ = zstd_ref(data, 16, cap)
        assert sc[0] == cap
        np.testing.assert_array_equal(seq[0, :cap], full[0, :cap])
        covered = int((seq[0, :cap, 1] + 3).sum())
        assert lc[0] == SUB - covered


# ---------------------------------------------------------------------------------------------
# brute force: candidate by linear scan, then the parse, in plain Python
# ---------------------------------------------------------------------------------------------
def _h(d, q):
    return ((int(d[q]) | int(d[q + 1]) << 8 | int(d[q + 2]) << 16 | int(d[q + 3]) << 24) * 2654435761 % 2**32) >> 18


def brute_cand(d, p, n):
    """max{q < p : q + 3 < n, H(q) = H(p)} or None (linear scan)."""
    hp = _h(d, p)
    for q in range(p - 1, -1, -1):
        if q + 3 < n and _h(d, q) == hp:
            return q
    return None


def common(d, a, b, limit):
    k = 0
    while k < limit and d[a + k] == d[b + k]:
        k += 1
    return k


def brute_zstd(d, depth):
    n = len(d)
    rep = [1, 4, 8]
    pos, lit_start, out = 0, 0, []
    while pos + 3 < n:
        c = brute_cand(d, pos, n)
        c = 0 if c is None else c
        L, off = 0, 0
        if pos > 0 and common(d, c, pos, 3) == 3:
            L, off = common(d, c, pos, n - pos), pos - c
        if L < 3:
            for r in rep:
                if r <= pos and common(d, pos - r, pos, 3) == 3:
                    L, off = common(d, pos - r, pos, n - pos), r
                    break
        if L >= 3 and depth > 1 and pos + 4 < n:
            c2 = brute_cand(d, pos + 1, n)
            c2 = 0 if c2 is None else c2
            if common(d, c2, pos + 1, 3) == 3:
                L2 = common(d, c2, pos + 1, n - pos - 1)
                if L2 > L + 1:
                    pos, L, off = pos + 1, L2, pos + 1 - c2
        if L >= 3:
            out.append([pos - lit_start, L - 3, off])
            rep = [off, rep[0], rep[1]]
            pos += L
            lit_start = pos
        else:
            pos += 1
    return out


def brute_lzma2(d):
    n = len(d)
    out = []
    for p in range(n):
        m = 0
        if p + 1 < n:
            maxlen = min(n - p, 273)
            if p + 3 < n:
                c = brute_cand(d, p, n)
                if c is not None:
                    L = common(d, c, p, maxlen)
                    if L >= 2:
                        m = (p - c) << 16 | L
            if m == 0 and p > 0 and d[p] == d[p - 1] and d[p + 1] == d[p]:
                m = 1 << 16 | common(d, p - 1, p, maxlen)
        out.append(m)
    return out


def small_inputs():
    rng = np.random.default_rng(7)
    cases = []
    for i in range(24):
        n = int(rng.integers(1, 700))
        alpha = int(rng.choice([2, 3, 4, 16, 256]))
        cases.append(rng.integers(0, alpha, n, dtype=np.uint8))
    cases.append(datagen.text(900, 1))
    cases.append(datagen.structured(900, 2))
    cases.append(datagen.backrefs(900, 3))
    return cases


@pytest.mark.parametrize("i", range(27))
@pytest.mark.parametrize("depth", [1, 16])
def test_zstd_brute_force(i, depth):
    d = small_inputs()[i]
    seq, sc, _, _ = zstd_ref(d, depth)
    got = seq[0, : sc[0]].tolist() if d.size else []
    assert got == brute_zstd(d.tolist(), depth)


@pytest.mark.parametrize("i", range(27))
def test_lzma2_brute_force(i):
    d = small_inputs()[i]
    assert oracle.lzma2_packed(d).tolist() == brute_lzma2(d.tolist())


# ---------------------------------------------------------------------------------------------
# LZMA2 closed forms
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 300, SUB])
def test_lzma2_all_zero(n):
    # pos 0: empty slot, no run (pos > 0 fails).  pos p >= 1 with p+3 < n: slot holds p-1:
    # distance 1, length min(273, n-p).  p = n-3, n-2: no hash; run check gives min(273, n-p).
    # p = n-1: nothing (p + 1 < n fails).
    want = [0] * n
    for p in range(1, n - 1):
        want[p] = 1 << 16 | min(273, n - p)
    assert oracle.lzma2_packed(np.zeros(n, np.uint8)).tolist() == want


def test_lzma2_reference_layout_and_tail_counts():
    data = datagen.mix(SUB + 100, 5)
    m, c = oracle.lzma2_match_find(data, SUB, 8, 32)
    assert m.shape == (2 * SUB, 8, 2)
    assert not c[SUB + 100:].any()            # tail counts zeroed (lzma2_match_find.cu:188-190)
    assert not m[:, 1:, :].any()              # one probe: never a second slot
    assert set(np.unique(c).tolist()) <= {0, 1}
    assert not m[c == 0].any()                # unused slots stay zero


def test_lzma2_depth_zero_keeps_only_runs():
    d = datagen.mix(SUB, 6)
    p0 = oracle.lzma2_packed(d, depth=0)
    assert set(np.unique(p0 >> 16).tolist()) <= {0, 1}
    p32 = oracle.lzma2_packed(d, depth=32)
    np.testing.assert_array_equal(p32, oracle.lzma2_packed(d, depth=48))  # depth > 0 is one probe


# ---------------------------------------------------------------------------------------------
# invariants on corpus data
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [1, 2])
def test_zstd_reconstructs_input(seed):
    data = gcompress.corpus(4 * SUB + 333, seed)
    seq, sc, lc, lit = zstd_ref(data)
    for s in range(sc.size):
        src = data[s * SUB:(s + 1) * SUB]
        out = bytearray()
        li = 0
        lits = lit[s * SUB: s * SUB + lc[s]]
        for ll, ml3, off in seq[s, : sc[s]].tolist():
            out += bytes(lits[li:li + ll])
            li += ll
            assert 1 <= off <= len(out)
            for _ in range(ml3 + 3):
                out.append(out[-off])
        out += bytes(lits[li:])
        assert bytes(out) == src.tobytes()


@pytest.mark.parametrize("seed", [1, 2])
def test_lzma2_matches_verify_and_are_maximal(seed):
    data = gcompress.corpus(2 * SUB, seed, (15, 25, 20, 40), 8192)
    pk = oracle.lzma2_packed(data)
    for s in range(2):
        d = data[s * SUB:(s + 1) * SUB]
        v = pk[s * SUB:(s + 1) * SUB]
        n = d.size
        for p in np.nonzero(v)[0].tolist():
            dist, L = int(v[p]) >> 16, int(v[p]) & 0xFFFF
            cap = min(273, n - p)
            assert 2 <= L <= cap and 1 <= dist <= p
            assert (d[p - dist:p - dist + L] == d[p:p + L]).all()
            assert L == cap or d[p + L] != d[p - dist + L]
