"""GPU results equal the CPU oracle bit for bit (both finders, several layouts)."""
import numpy as np
import pytest

import datagen
import gcompress
import oracle

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 4, 5, 7, 64, 255, 256, 257, 1000, 65535, 65536, 65537, 65536 * 3 + 12345]


def cases():
    out = []
    for n in SIZES:
        out.append((f"rand{n}", datagen.random_bytes(n, n)))
        out.append((f"zero{n}", np.zeros(n, np.uint8)))
    out.append(("text", datagen.text(4 * 65536 + 777, 1)))
    out.append(("backrefs", datagen.backrefs(4 * 65536 + 3, 2)))
    out.append(("structured", datagen.structured(4 * 65536 + 5, 3)))
    out.append(("mix", datagen.mix(8 * 65536 + 99, 4)))
    # copies of 64..160 bytes: the parse window slides past its ahead registers
    rng = np.random.default_rng(99)
    parts, total = [rng.integers(0, 256, 4096, dtype=np.uint8)], 4096
    while total < 3 * 65536:
        ln = int(rng.integers(64, 161))
        src = int(rng.integers(0, total - ln))
        parts.append(np.concatenate(parts)[src:src + ln])
        parts.append(rng.integers(0, 256, int(rng.integers(0, 40)), dtype=np.uint8))
        total = sum(p.size for p in parts)
    out.append(("longcopies", np.concatenate(parts)))
    for k in (1, 2, 3, 4, 7, 8, 9, 16, 63, 64, 65, 200):
        out.append((f"period{k}", datagen.periodic(65536 + 17, k, k)))
    # long literal stretches (the parse's look-ahead): random bytes with short repeats planted at
    # new offsets and then at the offsets those make current, some across a 4- or 256-position step
    # edge and in the last bytes of the input; and random bytes with a short text snippet every few KiB
    r = datagen.random_bytes(3 * 65536 + 999, 31).copy()
    at = 300
    for off, ln in [(77, 6), (77, 4), (1000, 5), (77, 3), (5000, 8), (1000, 4), (3, 9), (40000, 6), (5000, 3)] * 9:
        at += int(rng.integers(700, 9000))
        if at + ln >= r.size:
            break
        r[at:at + ln] = np.resize(r[at - off:at], ln) if off < ln else r[at - off:at - off + ln]
    r[-7:] = r[-7 - 64:-64]
    out.append(("planted", r))
    s = datagen.random_bytes(3 * 65536 + 31, 32).copy()
    snippet = datagen.text(40, 33)
    for at in range(1000, s.size - 40, 3001):
        s[at:at + 40] = snippet
    out.append(("sparsetext", s))
    return out


CASES = cases()


@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
def test_lzma2_packed_matches_oracle(gctx, name, data):
    print(f"[lzma2 packed] {name} n={data.size}", flush=True)
    got = gctx.lzma2_packed(data)
    want = oracle.lzma2_packed(data)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} positions differ, first {bad[:5]}: got {got[bad[:5]]} want {want[bad[:5]]}"


def _triage_cases():
    rng = np.random.default_rng(21)
    out = [(name, data) for name, data in CASES]
    # sub-blocks with and without matches side by side: random bytes have none, a zero run has
    out.append(("rand+zero+rand", np.concatenate([rng.integers(0, 256, 65536, dtype=np.uint8),
                                                  np.zeros(65536, np.uint8),
                                                  rng.integers(0, 256, 100000, dtype=np.uint8)])))
    # a single repeat of 5 bytes at the end of an otherwise match-free sub-block
    lone = rng.permutation(256).astype(np.uint8)
    lone = np.concatenate([lone, lone[:5]])
    out.append(("lone-pair", lone))
    return out


TRIAGE = _triage_cases()


@pytest.mark.parametrize("name,data", TRIAGE, ids=[c[0] for c in TRIAGE])
def test_lzma2_triage_matches_oracle(gctx, name, data):
    """Device-side triage: one byte per sub-block, 1 iff some position has a match (the oracle's lists)."""
    print(f"[lzma2 triage] {name} n={data.size}", flush=True)
    got = gctx.lzma2_triage(data)
    want = oracle.lzma2_packed(data)
    nsub = -(-data.size // 65536)
    want_any = np.array([int(np.any(want[s * 65536:(s + 1) * 65536] != 0)) for s in range(nsub)], np.uint8)
    np.testing.assert_array_equal(got, want_any)


@pytest.mark.parametrize("level", [1, 2, 3])
@pytest.mark.parametrize("name,data", CASES, ids=[c[0] for c in CASES])
def test_zstd_sequences_match_oracle(gctx, name, data, level):
    print(f"[zstd seq] {name} n={data.size} level={level}", flush=True)
    depth = {1: 16, 2: 64}.get(level, 1)
    seqs, counts = gctx.zstd_sequences(data, level)
    oseq, osc, _, _ = oracle.zstd_match_find(data, depth=depth)
    np.testing.assert_array_equal(counts, osc)
    want = oracle.zstd_as_zstd_sequences(oseq, osc)
    for s in range(counts.size):
        got = seqs[s, : counts[s]]
        if not np.array_equal(got, want[s]):
            d = np.nonzero((got != want[s]).any(axis=1))[0][0]
            raise AssertionError(f"sub-block {s} seq {d}: got {got[d]} want {want[s][d]}")


def test_gpu_find_matches_reference_layout(gctx):
    data = datagen.mix(3 * 65536 + 500, 7)
    md = gcompress.gpu_find_matches(data, 0)
    om, oc = oracle.lzma2_match_find(data, 65536, 8, 32)
    assert md.num_sub_blocks == 4 and md.sub_block_size == 65536 and md.max_matches == 8
    np.testing.assert_array_equal(md.match_counts, oc)
    np.testing.assert_array_equal(md.match_distances, om[:, :, 0].reshape(-1))
    np.testing.assert_array_equal(md.match_lengths, om[:, :, 1].reshape(-1))


def _dev_copy(ctx, a):
    p = ctx.alloc(a.nbytes + 64)
    ctx.htod(p, a)
    return p


@pytest.mark.parametrize("sub", [65536, 4096, 1000, 7])
def test_lzma2_kernel_abi(gctx, sub):
    """gc_lzma2_match_find_dev: the reference kernel's argument list and output layout."""
    data = datagen.mix(2 * 65536 + 4321, 11)
    n = data.size
    nsub = -(-n // sub)
    mm = 8
    d_in = _dev_copy(gctx, data)
    d_m = gctx.alloc(nsub * sub * mm * 8)
    d_c = gctx.alloc(nsub * sub * 4)
    gctx.zero(d_m, nsub * sub * mm * 8)
    gctx.zero(d_c, nsub * sub * 4)
    rc = gcompress.lib().gc_lzma2_match_find_dev(gctx.h, d_in, d_m, d_c, sub, nsub, n, mm, 32)
    assert rc == 0, gcompress.lib().gc_last_error()
    m = np.zeros((nsub * sub, mm, 2), np.uint32)
    c = np.zeros(nsub * sub, np.uint32)
    gctx.dtoh(m, d_m)
    gctx.dtoh(c, d_c)
    for p in (d_in, d_m, d_c):
        gctx.free(p)
    om, oc = oracle.lzma2_match_find(data, sub, mm, 32)
    np.testing.assert_array_equal(c, oc)
    np.testing.assert_array_equal(m, om)


@pytest.mark.parametrize("fmt", [0, 1])
@pytest.mark.parametrize("chunk", [65536, 4096, 999])
def test_zstd_kernel_abi(gctx, fmt, chunk):
    """gc_zstd_match_find_dev: reference argument list, both record layouts, literals."""
    data = datagen.mix(2 * 65536 + 4321, 12)
    n = data.size
    nsub = -(-n // chunk)
    max_seq = chunk // 3
    rec = 16 if fmt == 0 else 12
    d_in = _dev_copy(gctx, data)
    d_s = gctx.alloc(nsub * max_seq * rec)
    d_l = gctx.alloc(nsub * chunk)
    d_sc = gctx.alloc(nsub * 4)
    d_lc = gctx.alloc(nsub * 4)
    gctx.zero(d_s, nsub * max_seq * rec)
    rc = gcompress.lib().gc_zstd_match_find_dev(gctx.h, d_in, d_s, fmt, d_l, d_sc, d_lc, chunk, nsub, n, 16, max_seq)
    assert rc == 0, gcompress.lib().gc_last_error()
    s = np.zeros((nsub, max_seq, rec // 4), np.uint32)
    lit = np.zeros(nsub * chunk, np.uint8)
    sc = np.zeros(nsub, np.uint32)
    lc = np.zeros(nsub, np.uint32)
    gctx.dtoh(s, d_s)
    gctx.dtoh(lit, d_l)
    gctx.dtoh(sc, d_sc)
    gctx.dtoh(lc, d_lc)
    for p in (d_in, d_s, d_l, d_sc, d_lc):
        gctx.free(p)
    oseq, osc, olc, olit = oracle.zstd_match_find(data, chunk, 16, max_seq, literals=True)
    np.testing.assert_array_equal(sc, osc)
    np.testing.assert_array_equal(lc, olc)
    for b in range(nsub):
        k = osc[b]
        want = oseq[b, :k] if fmt == 1 else np.stack([oseq[b, :k, 2], oseq[b, :k, 0], oseq[b, :k, 1] + 3,
                                                     np.zeros(k, np.uint32)], 1)
        np.testing.assert_array_equal(s[b, :k], want)
        np.testing.assert_array_equal(lit[b * chunk: b * chunk + olc[b]], olit[b * chunk: b * chunk + olc[b]])
