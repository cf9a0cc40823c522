"""The single-indel branch (indel_cap 1..4) on the CPU: pin the oracle's
restatement of linear_gpu_locate.mojo:1015-1024, 1142-1235 and 1408-1511 with
reads carrying a planted insertion or deletion, checked by an independent
recount along the reported CIGAR, and pin the host CIGAR text of indel hits
(_hit_from_gpu, :217-284) with hand-worked cases."""

from __future__ import annotations

import numpy as np
import pytest

from mialign import linear_gpu_locate as lgl
from mialign import synth
from oracle import oracle as orc

CODE = {ord(c): i for i, c in enumerate("ACGT")}


def test_hit_cigar_hand_cases():
    names = ["chr1_f"]
    seq = b"A" * 100

    def cig(sl, sr, op, at, ln):
        return lgl.hit_from_gpu(names, "q", seq, b"", 0, 10, 60, 0, sl, sr, op, at, ln).cigar

    assert cig(0, 0, 1, 40, 2) == "40M2I58M"
    assert cig(0, 0, 2, 40, 3) == "40M3D60M"
    assert cig(0, 0, 1, 0, 2) == "2I98M"
    assert cig(0, 0, 2, 100, 1) == "100M1D"
    assert cig(0, 0, 1, 98, 2) == "98M2I"
    assert cig(5, 3, 1, 10, 4) == "5S10M4I78M3S"
    assert cig(0, 0, 0, 7, 7) == "100M"
    assert cig(0, 0, 1, 101, 1) == "100M"  # indel_at past the middle: no indel
    h = lgl.hit_from_gpu(names, "q", seq, b"", 0, 10, 60, 0, 0, 0, 1, 99, 2)  # right < 0
    assert (h.flag, h.contig, h.cigar) == (4, "*", "*")


def recount_cigar(ref: bytes, pos: int, q: bytes, iop: int, iat: int, ilen: int) -> int:
    """Mismatches (non-ACGT counts as a mismatch) plus indel length along the
    alignment an indel hit reports (no soft clips on indel hits)."""
    L = len(q)
    nm = ilen if iop else 0
    for t in range(L):
        if iop == 1 and iat <= t < iat + ilen:
            continue
        r = pos + t + (ilen if iop == 2 and t >= iat else 0) - (ilen if iop == 1 and t >= iat + ilen else 0)
        a, b = CODE.get(ref[r], 9), CODE.get(q[t], 9)
        nm += a > 3 or b > 3 or a != b
    return nm


@pytest.fixture(scope="module")
def genome():
    ref = synth.gen_reference(11, 2, 300_000)
    dual = synth.dual_convert(ref)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    return dual, oix


def planted_reads(dual, seed: int, n: int, L: int = 150):
    """Reads cut from the first converted contig with one planted indel each:
    (rows, lens, truth (start, kind, at, dlen))."""
    rng = np.random.default_rng(seed)
    c0 = dual.seq[dual.contig_off[0]:dual.contig_off[1]].tobytes()
    rows = np.full((n, L), ord("N"), np.uint8)
    truth = []
    for i in range(n):
        s = int(rng.integers(1000, len(c0) - L - 1000))
        d = int(rng.integers(1, 5))
        a = int(rng.integers(30, L - 30))
        if i % 2 == 0:  # deletion of d reference bases after a query bases
            r = c0[s:s + a] + c0[s + a + d:s + L + d]
            kind = 2
        else:  # insertion of d query bases
            ins = bytes(rng.choice(np.frombuffer(b"ACT", np.uint8), d))
            r = c0[s:s + a] + ins + c0[s + a:s + L - d]
            kind = 1
        rows[i] = np.frombuffer(r, np.uint8)
        truth.append((s, kind, a, d))
    return rows, np.full(n, L, np.uint32), truth


def test_oracle_recovers_planted_indels(genome):
    dual, oix = genome
    rows, lens, truth = planted_reads(dual, 5, 120)
    mg, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=4), nthreads=8)
    ref = dual.seq.tobytes()
    found = 0
    for i, (s, kind, a, d) in enumerate(truth):
        if mg["cid"][i] < 0 or mg["iop"][i] == 0:
            continue
        found += 1
        assert mg["flag"][i] == 0 and mg["cid"][i] == 0 and mg["sl"][i] == 0 and mg["sr"][i] == 0, i
        assert (mg["iop"][i], mg["ilen"][i]) == (kind, d), (i, truth[i])
        assert mg["mapq"][i] == 20 and mg["iat"][i] % 2 == 0
        assert abs(int(mg["pos"][i]) - s) <= 0, (i, truth[i], mg["pos"][i])
        q = rows[i].tobytes()
        off = int(dual.contig_off[0])
        nm = recount_cigar(ref, off + int(mg["pos"][i]), q, int(mg["iop"][i]), int(mg["iat"][i]),
                           int(mg["ilen"][i]))
        assert nm == mg["nm"][i] and nm <= 6, (i, nm, mg["nm"][i])
    assert found >= 100, found
    # the gapless contract maps none of them with an indel
    g0, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=0), nthreads=8)
    assert (g0["iop"] == 0).all() and (g0["cid"] < 0).sum() + (g0["nm"] > 0).sum() >= 100


def test_recount_catches_a_swapped_operation(genome):
    """The checker is not blind: reading a deletion as an insertion (or the
    reverse) changes the count."""
    dual, oix = genome
    rows, lens, truth = planted_reads(dual, 6, 40)
    mg, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=4), nthreads=8)
    ref = dual.seq.tobytes()
    off = int(dual.contig_off[0])
    differ = 0
    for i in range(40):
        if mg["iop"][i] == 0:
            continue
        args = (ref, off + int(mg["pos"][i]), rows[i].tobytes())
        ok = recount_cigar(*args, int(mg["iop"][i]), int(mg["iat"][i]), int(mg["ilen"][i]))
        bad = recount_cigar(*args, 3 - int(mg["iop"][i]), int(mg["iat"][i]), int(mg["ilen"][i]))
        differ += bad != ok
    assert differ >= 30


def test_exact_reads_map_alike_at_every_cap(genome):
    """Error-free reads map to their start with nm 0 and no indel at every
    indel_cap (0..4): the wider shift window (clamp(indel_cap, 2, 6)) and the
    indel candidates never displace an exact gapless hit."""
    dual, oix = genome
    c0 = dual.seq[dual.contig_off[0]:dual.contig_off[1]].tobytes()
    rng = np.random.default_rng(9)
    rows, starts = [], []
    for _ in range(40):
        s = int(rng.integers(1000, len(c0) - 2000))
        starts.append(s)
        rows.append(np.frombuffer(c0[s:s + 150], np.uint8))
    starts = np.array(starts, np.uint32)
    rows = np.stack(rows)
    lens = np.full(40, 150, np.uint32)
    for cap in (0, 1, 2, 3, 4):
        mg, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=cap), nthreads=8)
        assert (mg["cid"] == 0).all() and (mg["nm"] == 0).all() and (mg["iop"] == 0).all(), cap
        assert (mg["pos"] == starts).all(), cap
