"""The `-device cpu` host loop (fq2bam-meth/src/linear_mapper.mojo:209-311,
extend_read_with_seeds at linear_extend.mojo:143-211, vote_dense_seeds at
linear_index.mojo:530-578) in the oracle (or_map_forward), on hand-built reads
whose records follow from the rules:

- every k-mer position i of a read votes as if it sat at query offset 5 i
  (seed stride 5), so a read found once in a random reference gets its true
  start's vote from k-mer 0 alone; k-mer i votes for start - 4 i, a key of its
  own, and the first key to reach one vote (k-mer 0's first posting) wins with
  one vote: MAPQ 20;
- a read whose first k-mers are not in the reference therefore lands on a
  wrong start (k-mer i's start - 4 i) and stays unmapped;
- the verify compares upper-cased bytes (N against N matches), clips leading
  and trailing mismatch runs of at most max_soft only when the count exceeds
  max_diff, and wants 32 aligned bases.

The vote itself is pinned by a second formulation on reads from tandem
repeats (many keys, many votes): the winner is the key, among those with the
largest count M, whose M-th vote comes first, and MAPQ follows M."""

from __future__ import annotations

from bisect import bisect_left

import numpy as np
import pytest

from oracle import oracle as orc

ACGT = b"ACGT"


def _rand(rng, n):
    return bytearray(np.frombuffer(ACGT, np.uint8)[rng.integers(0, 4, n)].tobytes())


def _mut(s: bytes, *at) -> bytes:
    b = bytearray(s)
    for j in at:
        b[j] = ACGT[(ACGT.index(b[j]) + 1) % 4]
    return bytes(b)


class OracleBackend:
    """Index and host loop through the oracle (test_gpu_forward.py swaps in the GPU)."""

    def index(self, ref: bytes, k: int = 21):
        seq = np.frombuffer(bytes(ref), np.uint8).copy()
        off = np.array([0, len(seq)], np.uint64)
        km, of, po = orc.build_dense(seq, off, k)
        ix = orc.OracleIndex(km, of, po, seq, off)
        ix.k = k
        return ix

    def forward(self, ix, rows, lens, **kw):
        return orc.map_forward(ix, rows, lens, k=ix.k, **kw)


@pytest.fixture(scope="module")
def backend():
    return OracleBackend()


def _rows(reads):
    L = max(max(len(r) for r in reads), 1)
    rows = np.full((len(reads), L), ord("N"), np.uint8)
    for i, r in enumerate(reads):
        rows[i, :len(r)] = np.frombuffer(r, np.uint8)
    return rows, np.array([len(r) for r in reads], np.uint32)


def _one(be, ix, read: bytes, **kw):
    rows, lens = _rows([read])
    out = be.forward(ix, rows, lens, **kw)
    return tuple(int(out[f][0]) for f in ("cid", "pos", "mapq", "flag", "sl", "sr", "nm"))


UNMAPPED = (-1, 0, 0, 4, 0, 0, 0)


@pytest.fixture(scope="module")
def world():
    rng = np.random.default_rng(11)
    return _rand(rng, 20_000)


def test_unique_read_maps_with_one_vote(world, backend):
    ix = backend.index(world)
    assert _one(backend, ix, bytes(world[5000:5150])) == (0, 5000, 20, 0, 0, 0, 0)


def test_mismatches_are_counted_and_capped(world, backend):
    ix = backend.index(world)
    read = bytes(world[5000:5150])
    assert _one(backend, ix, _mut(read, 40, 80, 120)) == (0, 5000, 20, 0, 0, 0, 3)
    six = (30, 45, 60, 75, 90, 105)
    assert _one(backend, ix, _mut(read, *six)) == (0, 5000, 20, 0, 0, 0, 6)
    # seven inside the read: the first and last bases match, so nothing is
    # clipped and the recount is still 7 > 6
    assert _one(backend, ix, _mut(read, *six, 120)) == UNMAPPED


def test_a_broken_first_kmer_leaves_the_read_unmapped(world, backend):
    """Base 10 changed: k-mers 0..10 hold it and are not in the reference;
    k-mer 11 (at 5011) votes first, for 5011 - 55 = 4956, whose window is not
    the read."""
    ix = backend.index(world)
    assert _one(backend, ix, _mut(bytes(world[5000:5150]), 10)) == UNMAPPED


def test_trailing_mismatches_are_clipped(world, backend):
    ix = backend.index(world)
    read = bytes(world[5000:5150])
    # 8 mismatches at the end: 8 > 6, so the trailing run (8) is clipped
    assert _one(backend, ix, _mut(read, *range(142, 150))) == (0, 5000, 20, 0, 0, 8, 0)
    # 9: the clip stops at max_soft 8, base 141 stays and counts
    assert _one(backend, ix, _mut(read, *range(141, 150))) == (0, 5000, 20, 0, 0, 8, 1)
    # max_soft 4: 4 clipped, 4 left > 2 = max_diff
    assert _one(backend, ix, _mut(read, *range(142, 150)), max_soft=4, max_diff=2) == UNMAPPED
    assert _one(backend, ix, _mut(read, *range(142, 150)), max_soft=4, max_diff=4) == (0, 5000, 20, 0, 0, 4, 4)


def test_short_reads_need_32_aligned_bases(world, backend):
    ix = backend.index(world)
    assert _one(backend, ix, bytes(world[5000:5031])) == UNMAPPED       # 31 bases
    assert _one(backend, ix, bytes(world[5000:5032])) == (0, 5000, 20, 0, 0, 0, 0)
    assert _one(backend, ix, bytes(world[5000:5020])) == UNMAPPED       # shorter than k


def test_window_past_the_contig_end(world, backend):
    ix = backend.index(world)
    assert _one(backend, ix, bytes(world[19850:20000])) == (0, 19850, 20, 0, 0, 0, 0)
    tail = bytes(_rand(np.random.default_rng(12), 50))
    assert _one(backend, ix, bytes(world[19900:20000]) + tail) == UNMAPPED  # 19900 + 150 > 20000


def test_n_against_n_matches(world, backend):
    ref = bytearray(world)
    ref[5100] = ord("N")
    ix = backend.index(ref)
    read = bytearray(world[5000:5150])
    assert _one(backend, ix, bytes(read)) == (0, 5000, 20, 0, 0, 0, 1)   # reference N vs a base
    read[100] = ord("N")
    assert _one(backend, ix, bytes(read)) == (0, 5000, 20, 0, 0, 0, 0)   # N vs N
    read[100] = ord("n")
    assert _one(backend, ix, bytes(read)) == (0, 5000, 20, 0, 0, 0, 0)   # upper-cased first


def test_conversion_of_the_read_and_its_case(world, backend):
    """A C2T-converted reference has no C; the read is given unconverted (with
    C / c) and converted by the loop: for C2T it maps exactly, without the
    conversion its first k-mers (holding a C) are absent and it does not."""
    rng = np.random.default_rng(13)
    orig = bytearray(world)
    conv = bytes(orig).replace(b"C", b"T")
    ix = backend.index(conv)
    read = bytearray(orig[5000:5150])
    assert b"C" in bytes(read[:21])
    for j in rng.choice(150, 40, replace=False):
        read[j] = ord(chr(read[j]).lower())
    assert _one(backend, ix, bytes(read), bs_r1=1) == (0, 5000, 20, 0, 0, 0, 0)
    assert _one(backend, ix, bytes(read), bs_r1=0) == UNMAPPED
    # G2A: a G2A reference, the read's G / g become A
    ga = bytes(orig).replace(b"G", b"A")
    ix2 = backend.index(ga)
    read2 = bytes(orig[7000:7150]).lower()
    assert _one(backend, ix2, read2, bs_r1=2) == (0, 7000, 20, 0, 0, 0, 0)
    # rows [0, n_r1) take bs_r1, the rest bs_r2
    rows, lens = _rows([bytes(read), read2])
    out = backend.forward(ix, rows, lens, n_r1=1, bs_r1=1, bs_r2=2)
    assert int(out["cid"][0]) == 0 and int(out["pos"][0]) == 5000
    assert int(out["cid"][1]) == -1  # the G2A read against the C2T reference


def _vote_by_occurrences(ix, conv: bytes, k: int, stride: int, max_occ: int):
    """The vote as a batch: every (key, order) vote of the read in processing
    order; among the keys with the largest count M, the winner is the one
    whose M-th vote comes first."""
    kmers = ix.kmers.tolist()
    votes = {}
    t = 0
    for i in range(len(conv) - k + 1):
        mer = conv[i:i + k]
        if any(c not in ACGT for c in mer):
            continue
        key = 0
        for c in mer:
            key = key * 4 + ACGT.index(c)
        j = bisect_left(kmers, key)
        if j == len(kmers) or kmers[j] != key:
            continue
        a, b = int(ix.offsets[j]), int(ix.offsets[j + 1])
        if b - a > max_occ:
            continue
        for p in range(a, b):
            cid, pos = int(ix.postings[2 * p]), int(ix.postings[2 * p + 1])
            if pos >= i * stride:
                votes.setdefault((cid, pos - i * stride), []).append(t)
                t += 1
    if not votes:
        return None, 0
    M = max(len(v) for v in votes.values())
    key = min((v[M - 1], kk) for kk, v in votes.items() if len(v) == M)[1]
    return key, M


def test_vote_is_the_earliest_max_count_occurrence(backend):
    """Reads from a reference with tandem repeats (units of 2..6 bases, so a
    k-mer has many postings and keys many votes) and from short periodic
    stretches, k = 11: every mapped read's start (pos - left clip) and MAPQ
    match the vote written as above, over all three MAPQ bands."""
    rng = np.random.default_rng(14)
    ref = _rand(rng, 6000)
    for b0, unit in ((400, 4), (1500, 5), (2600, 3), (3700, 6), (4800, 2)):
        u = bytes(_rand(rng, unit))
        blk = (u * (400 // unit + 1))[:400]
        ref[b0:b0 + 400] = blk
    # short 4-periodic prefixes (14..39 bases) before unique sequence: a read
    # starting there collects a few votes for its true start (MAPQ 40 too)
    ref += _rand(rng, 4000)
    planted = []
    for m, lb in enumerate(range(14, 40)):
        s0 = 6100 + 140 * m
        ref[s0:s0 + lb] = (bytes(_rand(rng, 4)) * 20)[:lb]
        planted.append(s0)
    ref = bytes(ref)
    ix = backend.index(ref, k=11)
    oix = OracleBackend().index(ref, k=11)
    reads = [ref[s:s + 100] for s in list(range(300, 5400, 37)) + planted]
    rows, lens = _rows(reads)
    out = backend.forward(ix, rows, lens)
    bands = set()
    n_mapped = 0
    for i, r in enumerate(reads):
        key, M = _vote_by_occurrences(oix, r, 11, 5, 16384)
        if int(out["cid"][i]) < 0:
            continue
        n_mapped += 1
        start = int(out["pos"][i]) - int(out["sl"][i])
        assert (int(out["cid"][i]), start) == key, f"read {i}"
        want_q = 60 if M >= 5 else 40 if M >= 3 else 20
        assert int(out["mapq"][i]) == want_q, f"read {i}: M {M}"
        bands.add(want_q)
    assert n_mapped > 0.5 * len(reads)
    assert bands == {20, 40, 60}
