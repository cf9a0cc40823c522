"""Hand-derived cases that pin the oracle's per-read rules no reference vector
covers: the vote table's first-min replacement when full and its count <= 1
guard (linear_gpu_locate.mojo:903-973), KEEP's stable top 16 with ties
(:975-1004), the extra pass-0 vote at vote_occ for reads the vote_fast pass
leaves unmapped (:2247-2300), the FW/RC merge's nm-then-MAPQ tie rule
(:2101-2113), and the Wang hash (gpu-common/src/gpu_kernels.mojo:128-137)
through its inverse.

Each case builds a reference by hand (random ACGT background with planted
copies of a read's seed k-mers), so the postings of every seed, their order,
and hence the vote table, KEEP picks, extension and MAPQ can be worked out on
paper; the expected records below are those worked results, and each case
says which plausible mistake would change them.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

ACGT = b"ACGT"
L = orc.lib()


def _rand(rng, n: int) -> bytearray:
    return bytearray(np.frombuffer(ACGT, np.uint8)[rng.integers(0, 4, n)].tobytes())


def _other(b: int) -> int:
    """A base different from b (the next one in A, C, G, T order)."""
    return ACGT[(ACGT.index(b) + 1) % 4]


def _mutate(s: bytes, *at: int) -> bytes:
    out = bytearray(s)
    for j in at:
        out[j] = _other(out[j])
    return bytes(out)


def _revcomp(s: bytes) -> bytes:
    return s.translate(bytes.maketrans(b"ACGT", b"TGCA"))[::-1]


def _index(contigs):
    seq = np.frombuffer(b"".join(bytes(c) for c in contigs), np.uint8).copy()
    off = np.zeros(len(contigs) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in contigs])
    km, of, po = orc.build_dense(seq, off, 21)
    return orc.OracleIndex(km, of, po, seq, off)


def _map(ix, read: bytes, prm):
    b = np.frombuffer(read, np.uint8).reshape(1, -1)
    mg, fw, rc, cnt = orc.map_batch(ix, b, np.array([len(read)], np.uint32), prm)
    one = lambda d: {k: int(v[0]) for k, v in d.items()}  # noqa: E731
    return one(mg), one(fw), one(rc), cnt


def _rec(r):
    return (r["cid"], r["pos"], r["flag"], r["nm"], r["mapq"], r["sl"], r["sr"])


# Stride 21 with k 21 gives a 63-base read three disjoint seeds: A = [0, 21),
# B = [21, 42), C = [42, 63), at query offsets 0, 21 and 42.
STRIDE21 = dict(stride=21)

# A scenario is (contigs, cases, counters): cases are (read, params kwargs,
# merged record (cid, pos, flag, nm, mapq, sl, sr), fw record or None, rc
# record or None), records worked out by hand; counters are the oracle's
# expected vote-table replacement attempts (or None).


def scenario_first_min_evicts_single_vote():
    """Seeds A, B, C each have 256 postings (occ ties break by slot: A, B, C).
    A: the true locus T (lowest position) then 255 decoys -> T at entry 0, the
    table full, every count 1. B: 255 decoys before T, then T: each decoy
    replaces the first minimum, entry 0 (count 1), so T is evicted by the
    first of them and put back, with count 1, by B's own posting of T. C: T
    first (count 2), then 255 decoys that churn entry 1. T leads with 2 votes
    and is exact: MAPQ 20 (fewer than 3 votes). Had the table replaced the
    last minimum, kept no newcomer when full, or given newcomers min + 1
    votes, T would reach 3 votes (MAPQ 40). Replacements attempted: 256 in B,
    255 in C."""
    rng = np.random.default_rng(101)
    R = bytes(_rand(rng, 63))
    ref = _rand(rng, 70_000)
    T = 20_000
    for i in range(255):
        ref[100 + 60 * i:100 + 60 * i + 21] = R[21:42]       # B decoys, before T
        ref[30_000 + 60 * i:30_000 + 60 * i + 21] = R[0:21]  # A decoys, after T
        ref[50_000 + 60 * i:50_000 + 60 * i + 21] = R[42:63]  # C decoys, after T
    ref[T:T + 63] = R
    return [ref], [(R, STRIDE21, (0, T, 0, 0, 20, 0, 0), None, None)], 256 + 255


def scenario_full_table_keeps_two_vote_entries():
    """A and B share 256 loci (T first, then 255 copies of R[0:42] whose next
    21 bases all mismatch R[42:63]): after B every entry has 2 votes. C's 256
    postings (copies of R[42:63] elsewhere; T's copy has a mismatch at 52) find
    the table full with minimum 2 > 1, so none replaces anything. KEEP takes T
    (entry 0) first; T extends with nm 1; the 255 decoys fail (21 tail
    mismatches, 13 after the 8-base clip, budget 6). T with 2 votes -> MAPQ 20;
    the RC pass (nm 1) finds nothing. Without the count <= 1 guard C's first
    posting would evict T (entry 0), no KEEP pick would extend, and the low-occ
    rescue would report T with 4 votes (MAPQ 40)."""
    rng = np.random.default_rng(102)
    R = bytes(_rand(rng, 63))
    tail = bytes(_other(b) for b in R[42:63])
    ref = _rand(rng, 50_000)
    T = 1000
    ref[T:T + 63] = _mutate(R, 52)
    for i in range(255):
        p = 2000 + 80 * i
        ref[p:p + 63] = R[0:42] + tail
    for i in range(256):
        p = 30_000 + 60 * i
        ref[p:p + 21] = R[42:63]
    unm = (-1, 0, 4, 0, 0, 0, 0)
    return [ref], [(R, STRIDE21, (0, T, 0, 1, 20, 0, 0), (0, T, 0, 1, 20, 0, 0), unm)], 256


def scenario_keep_stable_top16():
    """Twenty loci hold R with mismatches in B's and C's seeds, so only seed A
    votes and all twenty tie at 1 vote, in posting order. KEEP takes the first
    16 (strict '>' insertion, stable). Loci 1-16 have nm 3, loci 17-20 nm 2:
    the extension sees only 1-16 and reports the first, nm 3, MAPQ 20. A KEEP
    of more than 16, or one that lets later ties pass earlier ones, would
    report a locus from 17-20 (nm 2)."""
    rng = np.random.default_rng(103)
    R = bytes(_rand(rng, 63))
    ref = _rand(rng, 5000)
    for i in range(1, 21):
        p = 1000 + 100 * i
        ref[p:p + 63] = _mutate(R, 30, 50, 55) if i <= 16 else _mutate(R, 30, 50)
    return [ref], [(R, STRIDE21, (0, 1100, 0, 3, 20, 0, 0), None, None)], None


def scenario_vote_fast_extra_pass():
    """vote_fast 8 < vote_occ 256. Read R: every seed has 10 postings (T + 9
    decoys each): the vote_fast pass votes nothing and its rescue (occ <= 8)
    tries nothing, so the read is unmapped; the extra pass at vote_occ gives T
    3 votes, exact -> MAPQ 40 (as with vote_fast = vote_occ). Without that
    pass the high-occ tally (occ > 256 only) would leave the read unmapped.
    Read Q (seeds of its own elsewhere in the contig): seed A has 2 postings
    (T2, then a decoy), B and C 10 each: the vote_fast pass votes with A
    alone, maps T2 exactly with 1 vote -> MAPQ 20, and no further pass runs;
    a pass-0 cap of vote_occ would give T2 3 votes (MAPQ 40)."""
    rng = np.random.default_rng(104)
    R, Q = bytes(_rand(rng, 63)), bytes(_rand(rng, 63))
    ref = _rand(rng, 40_000)
    ref[1000:1063] = R
    ref[21_000:21_063] = Q
    ref[24_000:24_021] = Q[0:21]
    for i in range(9):
        for s, base in enumerate((5000, 10_000, 15_000)):
            ref[base + 50 * i:base + 50 * i + 21] = R[21 * s:21 * s + 21]
        for s, base in ((1, 30_000), (2, 35_000)):
            ref[base + 50 * i:base + 50 * i + 21] = Q[21 * s:21 * s + 21]
    fast = dict(stride=21, vote_occ=256, vote_fast=8)
    same = dict(stride=21, vote_occ=256, vote_fast=256)
    return [ref], [(R, fast, (0, 1000, 0, 0, 40, 0, 0), None, None),
                   (R, same, (0, 1000, 0, 0, 40, 0, 0), None, None),
                   (Q, fast, (0, 21_000, 0, 0, 20, 0, 0), None, None)], None


def scenario_fw_rc_merge():
    """A 40-base read has seeds at offsets 0, 5, 10, 15 (k 21, stride 5). A
    mismatch at 3 removes one seed's vote (3 votes, MAPQ 40); at 8, two (2
    votes, MAPQ 20); at 38, none. Five reads S1..S5; contig 0 holds each
    forward (or not) with its FW mismatches, contig 1 its reverse complement
    with its RC mismatches (flag 16). Merge: RC when FW is unmapped, or RC has
    fewer mismatches, or as many and a higher MAPQ; FW otherwise."""
    rng = np.random.default_rng(105)
    plan = [
        # fw mismatches, rc mismatches, expected merged (cid, pos, flag, nm, mapq)
        ((8,), (3,), "rc"),      # equal nm, RC's MAPQ higher -> RC
        ((3,), (8,), "fw"),      # equal nm, FW's MAPQ higher -> FW
        ((3,), (3,), "fw"),      # equal nm and MAPQ -> FW stays
        ((3, 38), (8,), "rc"),   # RC has fewer mismatches -> RC
        (None, (8,), "rc"),      # FW unmapped -> RC
    ]
    c0, c1 = _rand(rng, 6000), _rand(rng, 6000)
    cases = []
    for i, (fm, rm, pick) in enumerate(plan):
        S = bytes(_rand(rng, 40))
        p0, p1 = 500 + 1000 * i, 700 + 1000 * i
        if fm is not None:
            c0[p0:p0 + 40] = _mutate(S, *fm)
        c1[p1:p1 + 40] = _mutate(_revcomp(S), *rm)
        fw = (-1, 0, 4, 0, 0, 0, 0) if fm is None else (0, p0, 0, len(fm), 20 if 8 in fm else 40, 0, 0)
        rc = (1, p1, 16, 1, 40 if rm == (3,) else 20, 0, 0)
        cases.append((S, {}, rc if pick == "rc" else fw, fw, rc))
    return [c0, c1], cases, None


SCENARIOS = {
    "first_min_evicts_single_vote": scenario_first_min_evicts_single_vote,
    "full_table_keeps_two_vote_entries": scenario_full_table_keeps_two_vote_entries,
    "keep_stable_top16": scenario_keep_stable_top16,
    "vote_fast_extra_pass": scenario_vote_fast_extra_pass,
    "fw_rc_merge": scenario_fw_rc_merge,
}


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_oracle_hand_derived_scenario(name):
    contigs, cases, overflow = SCENARIOS[name]()
    ix = _index(contigs)
    total = 0
    for read, kw, want, want_fw, want_rc in cases:
        mg, fw, rc, cnt = _map(ix, read, orc.params(**kw))
        assert _rec(mg) == want, (name, read, fw, rc)
        if want_fw is not None:
            assert _rec(fw) == want_fw
        if want_rc is not None:
            assert _rec(rc) == want_rc
        total += cnt.overflow
    if overflow is not None:
        assert total == overflow


M64 = (1 << 64) - 1


def _unxorshift(y: int, s: int) -> int:
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x & M64


def wang_unhash(h: int) -> int:
    """Inverse of Thomas Wang's 64-bit mix, undone step by step from its
    published form: multiply by (2^31 + 1), xorshift 28, multiply by 21,
    xorshift 14, multiply by 265, xorshift 24, key * (2^21 - 1) - 1."""
    x = h
    x = (x * pow((1 << 31) + 1, -1, 1 << 64)) & M64
    x = _unxorshift(x, 28)
    x = (x * pow(21, -1, 1 << 64)) & M64
    x = _unxorshift(x, 14)
    x = (x * pow(265, -1, 1 << 64)) & M64
    x = _unxorshift(x, 24)
    x = ((x + 1) * pow((1 << 21) - 1, -1, 1 << 64)) & M64
    return x


def test_wang_hash_round_trips_through_its_inverse():
    rng = np.random.default_rng(106)
    keys = [int(k) for k in rng.integers(0, 2**63, 2000, dtype=np.uint64)] + [0, 1, 2**42 - 1, M64]
    for k in keys:
        h = L.or_wang_hash(k)
        assert wang_unhash(h) == k
    # the inverse is not blind: a mix with one shift changed does not invert
    def mixed_wrong(key):
        key = ((~key) + (key << 21)) & M64
        key ^= key >> 23  # 24 in the published mix
        key = (key + (key << 3) + (key << 8)) & M64
        key ^= key >> 14
        key = (key + (key << 2) + (key << 4)) & M64
        key ^= key >> 28
        return (key + (key << 31)) & M64
    assert sum(wang_unhash(mixed_wrong(k)) == k for k in keys[:200]) < 5
