"""Pin the oracle (oracle/oracle.c) before trusting it as the GPU checker.

Each test checks the oracle against something that does not restate its own
code: the reference's known-answer tests (cited), brute force on small inputs,
closed forms, or hand-derived outcomes of specific branches. CPU only.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from mialign import synth
from oracle import oracle as orc

L = orc.lib()
ACGT = b"ACGT"


def conv(s: bytes, mode: int) -> bytes:
    a = np.frombuffer(s, dtype=np.uint8).copy()
    L.or_bs_convert(a.ctypes.data, a.size, mode)
    return a.tobytes()


# ---------------------------------------------------------------- A1 / A2
def test_conversion_reference_known_answers():
    # fq2bam-meth/tests/test_fq2bam_meth.py:34-42 (convert_fastq C2T / G2A)
    assert conv(b"ACGT", 1) == b"ATGT"
    assert conv(b"ACGT", 2) == b"ACAT"
    # fq2bam-meth/tests/test_linear_fastq.mojo:57-73 (arena conversion)
    assert conv(b"ACGTACGT", 1) == b"ATGTATGT"
    assert conv(b"GGTAGGTA", 2) == b"AATAAATA"
    assert conv(b"CCCC", 1) == b"TTTT"
    # _bs_convert keeps case (maketrans "Cc" -> "Tt"), linear_gpu_locate.mojo:63-66
    assert conv(b"acgtN", 1) == b"atgtN"


def test_fasta_conversion_known_answer():
    # fq2bam-meth/tests/test_fq2bam_meth.py:25-31
    a = np.frombuffer(b"ACGTCacgt", dtype=np.uint8).copy()
    L.or_convert_fasta_c2t(a.ctypes.data, a.size)
    assert a.tobytes() == b"ATGTTaTgt"


def test_code_table_closed_form():
    for b in range(256):
        want = {65: 0, 97: 0, 67: 1, 99: 1, 71: 2, 103: 2, 84: 3, 116: 3}.get(b, 255)
        assert L.or_code_of(b) == want


def test_wang_hash_is_a_bijection_on_a_sample():
    # Wang's 64-bit mix is a composition of invertible steps: no collisions.
    rng = np.random.default_rng(0)
    keys = np.unique(rng.integers(0, 2**62, size=200_000, dtype=np.uint64))
    hs = {L.or_wang_hash(int(k)) for k in keys[:50_000]}
    assert len(hs) == min(50_000, keys.size)
    # small keys also spread (avalanche): low bits of consecutive keys differ
    low = {L.or_wang_hash(k) & 0xFFFF for k in range(4096)}
    assert len(low) > 3500


# ---------------------------------------------------------------- A3 / A4
def brute_key(s: bytes) -> int | None:
    key = 0
    for c in s:
        if c not in ACGT:
            return None
        key = key * 4 + ACGT.index(c)
    return key


def test_encode_strided_brute_force():
    rng = np.random.default_rng(1)
    n, max_len, k, stride = 40, 60, 7, 5
    bases = np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.choice(5, size=(n, max_len), p=[.24, .24, .24, .24, .04])]
    lens = rng.integers(0, max_len + 1, size=n).astype(np.uint32)
    codes = np.empty_like(bases)
    L.or_pack_bases(bases.ctypes.data, codes.ctypes.data, bases.size)
    S = int(L.or_slots_per(max_len, k, stride))
    keys = np.empty(n * S, np.uint64)
    qoffs = np.empty(n * S, np.uint32)
    L.or_encode_strided(codes.ctypes.data, lens.ctypes.data, n, max_len, k, stride, S,
                        keys.ctypes.data, qoffs.ctypes.data)
    for r in range(n):
        for s in range(S):
            q = s * stride
            assert qoffs[r * S + s] == q
            want = brute_key(bases[r, q : q + k].tobytes()) if q + k <= lens[r] else None
            assert keys[r * S + s] == (0xFFFFFFFFFFFFFFFF if want is None else want)


def brute_index(seq: bytes, off, k):
    table: dict = {}
    for c in range(len(off) - 1):
        a, b = int(off[c]), int(off[c + 1])
        for p in range(0, b - a - k + 1):
            key = brute_key(seq[a + p : a + p + k])
            if key is not None:
                table.setdefault(key, []).append((c, p))
    return table


def test_build_dense_equals_brute_force_scan():
    rng = np.random.default_rng(2)
    raw = np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.choice(5, size=3000, p=[.3, .2, .2, .28, .02])]
    seq = raw.tobytes()
    off = np.array([0, 1000, 1005, 2500, 3000], dtype=np.uint64)  # includes a contig shorter than k
    k = 8
    km, of, po = orc.build_dense(raw, off, k)
    table = brute_index(seq, off, k)
    assert list(km) == sorted(table)
    for i, key in enumerate(km):
        got = [(int(po[2 * j]), int(po[2 * j + 1])) for j in range(int(of[i]), int(of[i + 1]))]
        assert got == table[int(key)]  # (contig, pos) ascending within a key
    assert of[-1] == po.size // 2


def test_locate_bsearch_brute_force():
    rng = np.random.default_rng(3)
    raw = np.frombuffer(ACGT, dtype=np.uint8)[rng.integers(0, 2, size=4000)]  # A/C only: repeats
    off = np.array([0, 4000], dtype=np.uint64)
    k = 6
    km, of, po = orc.build_dense(raw, off, k)
    table = brute_index(raw.tobytes(), off, k)
    queries = np.array(list(table)[:50] + [12345, 0xFFFFFFFFFFFFFFFF, 4**6 - 1], dtype=np.uint64)
    st = np.empty(queries.size, np.uint64)
    en = np.empty(queries.size, np.uint64)
    max_occ = 80
    L.or_locate_bsearch(queries.ctypes.data, queries.size, km.ctypes.data, of.ctypes.data, km.size,
                        max_occ, st.ctypes.data, en.ctypes.data)
    for q, a, b in zip(queries, st, en):
        hits = table.get(int(q), [])
        if not hits or len(hits) > max_occ:
            assert a == 0 and b == 0
        else:
            assert b - a == len(hits)
            got = [(int(po[2 * j]), int(po[2 * j + 1])) for j in range(int(a), int(b))]
            assert got == hits


# ------------------------------------------------------- A6-A8 invariants
@pytest.fixture(scope="module")
def uniq():
    ref = synth.gen_reference(31, 2, 400_000)
    dual = synth.dual_convert(ref)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    return ref, dual, orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)


def revcomp(s: bytes) -> bytes:
    return s.translate(bytes.maketrans(b"ACGTN", b"TGCAN"))[::-1]


def test_error_free_reads_map_to_their_origin(uniq):
    ref, dual, ix = uniq
    reads = synth.gen_reads(32, ref, 3000, p_err=0.0, dup_frac=0.0)
    mg, fw, rc, _ = orc.map_batch(ix, orc.convert_reads(reads.bases, 1), reads.lens)
    want_cid = 2 * reads.contig + reads.strand  # OT -> f copy forward, OB -> r copy reverse
    assert np.array_equal(mg["cid"], want_cid)
    assert np.array_equal(mg["pos"], reads.pos)
    assert np.array_equal(mg["flag"], np.where(reads.strand == 0, 0, 16))
    assert not mg["nm"].any() and not mg["sl"].any() and not mg["sr"].any()
    # 26 unique seeds: the vote stops after the 4th agreeing seed -> votes 4 -> MAPQ 40
    assert np.all(mg["mapq"] == 40)


def test_reported_nm_equals_recount(uniq):
    ref, dual, ix = uniq
    reads = synth.gen_reads(33, ref, 3000, p_err=0.02)
    convd = orc.convert_reads(reads.bases, 1)
    mg, _, _, _ = orc.map_batch(ix, convd, reads.lens)
    seq = dual.seq.tobytes()
    for i in range(reads.n):
        if mg["cid"][i] < 0:
            continue
        Lr = int(reads.lens[i])
        q = convd[i, :Lr].tobytes()
        if mg["flag"][i] == 16:
            q = revcomp(q)
        sl, sr = int(mg["sl"][i]), int(mg["sr"][i])
        g = int(dual.contig_off[mg["cid"][i]]) + int(mg["pos"][i])  # first aligned base
        win = seq[g : g + Lr - sl - sr]
        nm = sum(a != b for a, b in zip(q[sl : Lr - sr], win))
        assert nm == mg["nm"][i], i
        assert mg["mapq"][i] in (20, 40, 60)
        if mg["nm"][i] > 2:
            assert mg["mapq"][i] == 20


def test_recount_nm_checker(uniq):
    """or_recount_nm (the checker used at full size) equals the Python loop
    above and every reported nm; a record moved by one base no longer does."""
    ref, dual, ix = uniq
    reads = synth.gen_reads(34, ref, 1500, 60, 150, p_err=0.02, max_len=150)
    convd = orc.convert_reads(reads.bases, 1)
    mg, _, _, _ = orc.map_batch(ix, convd, reads.lens)
    got = orc.recount_nm(dual.seq, dual.contig_off, convd, reads.lens, mg)
    seq = dual.seq.tobytes()
    for i in range(0, reads.n, 7):
        if mg["cid"][i] < 0:
            assert got[i] == -1
            continue
        Lr = int(reads.lens[i])
        q = convd[i, :Lr].tobytes()
        if mg["flag"][i] == 16:
            q = revcomp(q)
        sl, sr = int(mg["sl"][i]), int(mg["sr"][i])
        g = int(dual.contig_off[mg["cid"][i]]) + int(mg["pos"][i])
        assert got[i] == sum(a != b for a, b in zip(q[sl : Lr - sr], seq[g : g + Lr - sl - sr])), i
    m = mg["cid"] >= 0
    assert m.mean() > 0.9 and np.array_equal(got[m], mg["nm"][m].astype(np.int32))
    moved = dict(mg)
    moved["pos"] = mg["pos"] + 1
    off = orc.recount_nm(dual.seq, dual.contig_off, convd, reads.lens, moved)
    assert (off[m] != got[m]).mean() > 0.9


def _tiny_world(seq: bytes, k=21):
    ref = synth.Reference(np.frombuffer(seq, dtype=np.uint8).copy(),
                          np.array([0, len(seq)], dtype=np.uint64), ["chr1"])
    dual = synth.dual_convert(ref)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, k)
    return ref, dual, orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)


def _map_one(ix, read: bytes, prm=None):
    b = np.frombuffer(read, dtype=np.uint8).reshape(1, -1)
    mg, fw, rc, _ = orc.map_batch(ix, b, np.array([len(read)], np.uint32), prm)
    return {k: int(v[0]) for k, v in mg.items()}


def _converted_ref(seed=41, n=5000):
    rng = np.random.default_rng(seed)
    return bytes(rng.choice(np.frombuffer(b"AGT", dtype=np.uint8), size=n))  # C-free: conversion no-op


def test_branch_softclip_left_hand_derived():
    seq = _converted_ref()
    _, _, ix = _tiny_world(seq)
    true = 1200
    read = bytearray(seq[true : true + 150])
    for j in range(8):  # first 8 bases mismatch (A<->T swap keeps the read C-free)
        read[j] = {65: 84, 84: 65, 71: 84}[read[j]]
    r = _map_one(ix, bytes(read))
    # nm_all 8 > budget 6 -> soft clip: sl = 8 (clip_cap 8), recount 0 -> accept
    assert (r["cid"], r["pos"], r["sl"], r["sr"], r["nm"], r["flag"]) == (0, true + 8, 8, 0, 0, 0)
    assert r["mapq"] == 40  # 4 agreeing seeds end the vote


def test_branch_three_mismatches_caps_mapq():
    seq = _converted_ref(42)
    _, _, ix = _tiny_world(seq)
    read = bytearray(seq[300:450])
    for j in (30, 75, 120):
        read[j] = 71 if read[j] != 71 else 65
    r = _map_one(ix, bytes(read))
    assert (r["cid"], r["pos"], r["nm"], r["mapq"]) == (0, 300, 3, 20)


def test_branch_short_read_low_occ_rescue():
    # L=20 < 24: extend rejects (alen < 24), the rescue accepts an exact hit
    # from a seed with occ <= 64 -> votes 5 -> MAPQ 60 (k=11 so 2 seeds exist).
    seq = _converted_ref(43, 20000)
    _, _, ix = _tiny_world(seq, k=11)
    for start in (700, 5000, 12345):
        read = seq[start : start + 20]
        r = _map_one(ix, read, orc.params(k=11))
        assert (r["cid"], r["pos"], r["nm"], r["mapq"], r["flag"]) == (0, start, 0, 60, 0)


def test_branch_high_occ_exact_tally():
    # A 100 bp segment copied 300 times: every seed has occ 300 > vote_occ, so
    # pass 0 finds nothing and pass 1 tallies exact loci. The tally keeps the
    # first 48 loci in posting order, all tie at 16 votes with equal min occ:
    # the first (lowest) copy wins with votes 1 -> MAPQ 20.
    rng = np.random.default_rng(44)
    seg = bytes(rng.choice(np.frombuffer(b"AGT", dtype=np.uint8), size=100))
    parts, starts, pos = [], [], 0
    for i in range(300):
        spacer = bytes(rng.choice(np.frombuffer(b"AGT", dtype=np.uint8), size=37))
        parts += [spacer, seg]
        starts.append(pos + 37)
        pos += 137
    seq = b"".join(parts)
    _, _, ix = _tiny_world(seq)
    r = _map_one(ix, seg)
    assert (r["cid"], r["pos"], r["nm"], r["mapq"], r["flag"]) == (0, starts[0], 0, 20, 0)


def test_ob_read_maps_reverse_on_r_copy():
    rng = np.random.default_rng(45)
    top = bytes(rng.choice(np.frombuffer(ACGT, dtype=np.uint8), size=4000))
    _, dual, ix = _tiny_world(top)
    # OB template: reverse complement of the top strand; bisulfite C->T on it
    tmpl = revcomp(top[1000:1150]).replace(b"C", b"T")
    r = _map_one(ix, tmpl)
    assert (r["cid"], r["pos"], r["flag"], r["nm"]) == (1, 1000, 16, 0)


# --------------------------------------------------------------- A10
def test_sort_key_formulas_hand_derived():
    assert L.or_unclipped5(0, 100, 2, 0, 150) == 98
    assert L.or_unclipped5(16, 100, 2, 3, 150) == 100 + 145 + 2 - 1
    assert L.or_unclipped5(4, 100, 2, 3, 150) == 0
    assert L.or_unclipped5(0, 100, 100, 60, 150) == 100  # clips cover the read: ignored
    assert L.or_coord_key(3, 77, 9) == (3 << 32) | 77
    assert L.or_coord_key(-1, 77, 9) == (0xFFFFFFFF << 32) | 9
    assert L.or_pack_end(0, 98, 0) == (1 << 48) | (98 + 2**30)
    assert L.or_pack_end(2, 5, 1) == (3 << 48) | (1 << 47) | (5 + 2**30)
    assert L.or_pack_end(0, -2**31, 0) == (1 << 48)  # negative clamps to 0


def test_sort_keys_se_records():
    rec = {f: np.zeros(3, dtype=np.int32 if f in ("cid", "flag") else np.uint32) for f in orc.REC_FIELDS}
    rec["cid"][:] = [1, 2, -1]
    rec["pos"][:] = [100, 50, 0]
    rec["flag"][:] = [16, 0, 4]
    rec["sl"][:] = [2, 0, 0]
    rec["sr"][:] = [3, 0, 0]
    rec["nm"][:] = [3, 300, 0]
    rec["mapq"][:] = [60, 20, 0]
    lens = np.array([150, 150, 150], np.uint32)
    tid = np.array([0, 0, 1, 1], np.int32)  # f0 r0 f1 r1
    coord, dhi, dlo, score, pair = orc.sort_keys_se(rec, lens, tid, pair_base=10)
    assert list(coord) == [(0 << 32) | 100, (1 << 32) | 50, (0xFFFFFFFF << 32) | 2]
    assert dhi[0] == (1 << 48) | (1 << 47) | (246 + 2**30)
    assert dhi[1] == (2 << 48) | (50 + 2**30)
    assert dhi[2] == 0xFFFFFFFFFFFFFFFF and dlo[2] == 2 and dlo[0] == 0
    assert list(score) == [(60 << 8) | 252, (20 << 8) | 0, 255]
    assert list(pair) == [10, 11, 12]
    # a target table shorter than the contig count: contig 2 has no target, so
    # its record keys as unmapped (rid1 < len(rid_to_tid), linear_gpu_fm.mojo:764-765)
    coord, dhi, dlo, _, _ = orc.sort_keys_se(rec, lens, tid[:2], pair_base=10)
    assert coord[1] == (0xFFFFFFFF << 32) | 1 and dhi[1] == 0xFFFFFFFFFFFFFFFF and dlo[1] == 1
    assert coord[0] == (0 << 32) | 100 and dhi[0] == (1 << 48) | (1 << 47) | (246 + 2**30)


# --------------------------------------------------------- A11 / A12
def test_order_run_perms_known_answer():
    # fq2bam-meth/tests/test_bam_emit.py:134-146
    coord = np.array([50, 10, 40, 20, 30], np.uint64)
    dhi = np.array([9, 1, 1, 2, 8], np.uint64)
    dlo = np.array([0, 5, 3, 0, 0], np.uint64)
    z = np.zeros(5, np.uint32)
    d, perm, dperm = orc.sort_markdup(coord, dhi, dlo, z, np.arange(5, dtype=np.uint32))
    assert list(perm) == [1, 3, 4, 2, 0]
    assert list(dperm) == [2, 1, 3, 4, 0]


def test_unmapped_sort_last_known_answer():
    # test_bam_emit.py:187-210: (tid,pos) (2,50),(0,10),(-1,0),(0,5) -> [3,1,0,2]
    coord = np.array([L.or_coord_key(2, 50, 0), L.or_coord_key(0, 10, 1),
                      L.or_coord_key(-1, 0, 2), L.or_coord_key(0, 5, 3)], np.uint64)
    z64 = np.zeros(4, np.uint64)
    z32 = np.zeros(4, np.uint32)
    _, perm, _ = orc.sort_markdup(coord, z64, z64, z32, z32, do_markdup=False)
    assert list(perm) == [3, 1, 0, 2]


def test_merge_markdup_marks_exactly_one_known_answer():
    # test_bam_emit.py:213-265 (two runs merged): records
    # (coord, dhi, dlo, score, pair): (100,1,1,100,0) (300,9,9,1,2) (200,1,1,50,1) (400,8,8,1,3)
    coord = np.array([100, 300, 200, 400], np.uint64)
    dhi = np.array([1, 9, 1, 8], np.uint64)
    dlo = np.array([1, 9, 1, 8], np.uint64)
    score = np.array([100, 1, 50, 1], np.uint32)
    pair = np.array([0, 2, 1, 3], np.uint32)
    d, perm, _ = orc.sort_markdup(coord, dhi, dlo, score, pair)
    assert list(d) == [0, 0, 1, 0]
    assert list(perm) == [0, 2, 1, 3]


def test_pair_id_uniqueness_known_answer():
    # test_bam_emit.py:268-308
    one = np.ones(2, np.uint64)
    d, _, _ = orc.sort_markdup(one, one, one, np.array([100, 50], np.uint32), np.array([0, 500], np.uint32))
    assert list(d) == [0, 1]
    d, _, _ = orc.sort_markdup(one, one, one, np.array([100, 50], np.uint32), np.array([0, 0], np.uint32))
    assert list(d) == [0, 0]


def test_markdup_brute_force_groupby():
    rng = np.random.default_rng(5)
    n = 3000
    dhi = rng.integers(0, 40, n).astype(np.uint64)
    dhi[rng.random(n) < 0.1] = 0xFFFFFFFFFFFFFFFF
    dlo = rng.integers(0, 3, n).astype(np.uint64)
    score = rng.integers(0, 5, n).astype(np.uint32)
    pair = rng.permutation(n).astype(np.uint32)
    pair[:20] = pair[20:40]  # some shared pair ids (PE mates)
    coord = rng.integers(0, 100, n).astype(np.uint64)
    d, perm, dperm = orc.sort_markdup(coord, dhi, dlo, score, pair)
    want = np.zeros(n, np.uint32)
    order = sorted(range(n), key=lambda i: (int(dhi[i]), int(dlo[i]), i))
    for key, grp in itertools.groupby(order, key=lambda i: (int(dhi[i]), int(dlo[i]))):
        grp = list(grp)
        if key[0] == 0xFFFFFFFFFFFFFFFF:
            continue
        best = max(grp, key=lambda i: (int(score[i]), -int(pair[i])))
        for i in grp:
            want[i] = int(pair[i] != pair[best])
    assert np.array_equal(d, want)
    assert list(dperm) == order
    assert list(perm) == sorted(range(n), key=lambda i: (int(coord[i]), i))
