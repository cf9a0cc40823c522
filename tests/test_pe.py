"""Paired-end emit (CPU): pin the oracle's restatement of the PE branch of the
reference's batch emit (linear_gpu_fm.mojo:826-977: flags, mate fields, signed
tlen, the pair's shared duplicate keys) with pairs worked by hand, with the
parity path's own pairing (pair_hits, linear_extend.mojo:263-294) on random
pairs, and with decoded records whose mate fields must name their mate."""

from __future__ import annotations

import numpy as np
import pytest

import bam_decode as bd
from mialign import linear_gpu_locate as lgl
from oracle import oracle as orc
from test_bam import expected_sam, random_se_batch

TID = np.array([0, 1, 0], np.int32)  # index contigs 0 and 2 share target 0 (the f / r strands)
E30 = 1 << 30


def pe_of(pairs, pair_base=0):
    """pairs: [((cid, pos, mapq, flag, sl, sr, nm, len) R1, (...) R2), ...]"""
    rows = [p[0] for p in pairs] + [p[1] for p in pairs]
    cols = list(zip(*rows))
    rec = {f: np.array(cols[k], dt) for k, (f, dt) in enumerate(
        (("cid", np.int32), ("pos", np.uint32), ("mapq", np.uint32), ("flag", np.int32), ("sl", np.uint32),
         ("sr", np.uint32), ("nm", np.uint32)))}
    return orc.pe_fields(rec, np.array(cols[7], np.uint32), TID, pair_base)


def end(tid, u5, strand):
    return ((tid + 1) << 48) | (strand << 47) | (u5 + E30)


def test_hand_worked_pairs():
    pairs = [
        ((0, 100, 60, 0, 0, 0, 1, 50), (2, 300, 40, 16, 0, 0, 2, 60)),     # one target, R1 leftmost
        ((0, 500, 20, 16, 2, 3, 0, 50), (0, 200, 60, 0, 4, 0, 0, 40)),     # one target, R2 leftmost
        ((1, 10, 60, 0, 0, 0, 0, 30), (0, 20, 60, 16, 0, 0, 0, 30)),       # two targets
        ((0, 1000, 60, 0, 0, 0, 0, 30), (-1, 7777, 0, 4, 0, 0, 0, 30)),    # R2 unmapped
        ((-1, 0, 0, 4, 0, 0, 0, 30), (2, 40, 60, 16, 0, 0, 300, 30)),      # R1 unmapped
        ((-1, 0, 0, 4, 0, 0, 0, 30), (-1, 0, 0, 4, 0, 0, 0, 30)),          # neither
        ((3, 55, 60, 0, 0, 0, 0, 30), (1, 66, 60, 0, 0, 0, 0, 30)),        # cid 3 has no target
    ]
    f = pe_of(pairs, pair_base=10)
    assert f["flag"].tolist() == [67, 147, 83, 131, 65, 145, 73, 133, 69, 153, 77, 141, 69, 137]
    assert f["tid"].tolist() == [0, 0, 0, 0, 1, 0, 0, -1, -1, 0, -1, -1, -1, 1]
    assert f["ntid"].tolist() == [0, 0, 0, 0, 0, 1, -1, 0, 0, -1, -1, -1, 1, -1]
    assert f["npos"].tolist() == [300, 100, 200, 500, 20, 10, -1, 1000, 40, -1, -1, -1, 66, -1]
    assert f["tlen"].tolist() == [260, -260, -340, 340, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0]
    assert f["row"].tolist() == [0, 7, 1, 8, 2, 9, 3, 10, 4, 11, 5, 12, 6, 13]
    e = [
        (end(0, 100, 0), end(0, 300 + 60 - 1, 1)),
        (end(0, 200 - 4, 0), end(0, 500 + 45 + 2 - 1, 1)),
        (end(0, 20 + 30 - 1, 1), end(1, 10, 0)),
        (end(0, 1000, 0), 0),
        (end(0, 40 + 30 - 1, 1), 0),
        (2 ** 64 - 1, 10),
        (end(1, 66, 0), 0),
    ]
    assert list(zip(f["dhi"].tolist()[::2], f["dlo"].tolist()[::2])) == e
    assert (f["dhi"][::2] == f["dhi"][1::2]).all() and (f["dlo"][::2] == f["dlo"][1::2]).all()
    assert f["score"].tolist()[::2] == [(100 << 8) | 252, (80 << 8) | 255, (120 << 8) | 255, (60 << 8) | 255,
                                        (60 << 8) | 0, 255, (120 << 8) | 255]
    assert f["pair"].tolist() == [5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11]
    unm = 0xFFFFFFFF << 32
    assert f["coord"].tolist() == [100, 300, 500, 200, (1 << 32) | 10, 20, 1000, unm | 7, unm | 8, 40,
                                   unm | 10, unm | 11, unm | 12, (1 << 32) | 66]


def random_pairs(seed, n_pairs):
    batch = random_se_batch(seed, 2 * n_pairs, max_len=150, n_cid=3)
    rec, tid_of_cid = batch[0], batch[1]
    rec["cid"][rec["cid"] >= 3] = 2  # mostly targets, some pairs on one target
    rng = np.random.default_rng(seed)
    same = rng.random(n_pairs) < 0.5
    rec["cid"][n_pairs:][same] = rec["cid"][:n_pairs][same]
    rec["flag"][:] = np.where(rec["cid"] < 0, 4, rec["flag"] & 16)  # the mapper's flags: 4 iff no contig
    rec["pos"][:] = rng.integers(0, 5000, 2 * n_pairs)
    lens = batch[6]
    for j in range(0, n_pairs, 9):  # duplicate pairs: copies of an earlier pair's ends
        k = int(rng.integers(0, j + 1))
        for a, b in ((j, k), (n_pairs + j, n_pairs + k)):
            for f in ("cid", "pos", "flag", "sl", "sr"):
                rec[f][a] = rec[f][b]
            lens[a] = lens[b]
    return batch


@pytest.mark.parametrize("seed", [3, 4])
def test_against_the_parity_pairing(seed):
    """Flags, mate target / position and |tlen| equal the SAM path's pairing of
    the same hits; the sign of tlen follows position order in the BAM path."""
    n_pairs = 300
    rec, tid_of_cid, _names, _blob, _off, _seqs, lens, *_ = random_pairs(seed, n_pairs)
    ref_names = ["chr1", "chr2", "chrX"]
    f = orc.pe_fields(rec, lens, tid_of_cid)
    for p in range(n_pairs):
        hits = []
        for r in (p, n_pairs + p):
            c = int(rec["cid"][r])
            t = int(tid_of_cid[c]) if 0 <= c < 3 else -1
            hits.append(lgl.LinearHit("q", int(rec["flag"][r]) | (4 if t < 0 else 0),
                                      ref_names[t] if t >= 0 else "*", int(rec["pos"][r]) + 1 if t >= 0 else 0,
                                      0, "*", "A" * int(lens[r])))
        a, b = lgl.pair_hits(*hits)
        for k, h in enumerate((a, b)):
            i = 2 * p + k
            assert int(f["flag"][i]) == h.flag, (p, k)
            nt = int(f["ntid"][i])
            if nt < 0:
                assert (h.rnext, h.pnext) == ("*", 0), (p, k)
            else:
                assert h.rnext == ("=" if nt == int(f["tid"][i]) else ref_names[nt]), (p, k)
                assert h.pnext == int(f["npos"][i]) + 1, (p, k)
            assert abs(int(f["tlen"][i])) == abs(h.tlen), (p, k)
    both = f["flag"][::2] & 2 != 0
    assert both.any() and (~both).any()


@pytest.mark.parametrize("seed", [5, 6])
def test_decoded_pe_records(seed):
    n_pairs = 250
    batch = random_pairs(seed, n_pairs)
    rec, tid_of_cid, names, blob, name_off, seqs, lens, quals, qual_lens = batch
    ref_names = ["chr1", "chr2", "chrX"]
    f = orc.pe_fields(rec, lens, tid_of_cid)
    is_dup, perm, _ = orc.sort_markdup(f["coord"], f["dhi"], f["dlo"], f["score"], f["pair"])
    assert (is_dup[::2] == is_dup[1::2]).all() and is_dup.any()  # a pair is marked as a whole
    stream = orc.bam_pack_pe(rec, tid_of_cid, blob, name_off, seqs, lens, quals, b"rg7", perm=perm,
                             is_dup=is_dup, apply_dup=True, qual_lens=qual_lens)
    got = bd.decode_records(stream)
    assert len(got) == 2 * n_pairs
    by_emit = {}
    for r, i in zip(got, perm):
        i = int(i)
        row = i // 2 + (n_pairs if i & 1 else 0)
        e = expected_sam(row, rec, tid_of_cid, names, seqs, lens, quals, qual_lens, b"rg7", ref_names)
        s = bd.to_sam_fields(r, ref_names)
        assert s[0] == e["qname"] and s[2] == e["rname"] and int(s[3]) == e["pos"] and s[5] == e["cigar"], i
        assert s[9] == e["seq"] and r["bin"] == e["bin"] and r["tags"]["NM"] == ("C", e["nm"]), i
        assert r["flag"] & (128 if i & 1 else 64) and r["flag"] & 1, i
        assert bool(r["flag"] & 1024) == bool(is_dup[i]), i
        by_emit[i] = r
    for p in range(n_pairs):
        a, b = by_emit[2 * p], by_emit[2 * p + 1]
        for x, y in ((a, b), (b, a)):
            if y["tid"] >= 0:
                assert (x["next_tid"], x["next_pos"]) == (y["tid"], y["pos"]), p
                assert not x["flag"] & 8, p
            else:
                assert (x["next_tid"], x["next_pos"]) == (-1, -1) and x["flag"] & 8, p
        assert a["tlen"] == -b["tlen"]
        if a["tid"] >= 0 and a["tid"] == b["tid"]:
            assert a["flag"] & 2 and b["flag"] & 2 and a["tlen"] != 0
            left = a if a["pos"] <= b["pos"] else b
            assert left["tlen"] > 0 or (a["pos"] == b["pos"] and a["tlen"] > 0)
        else:
            assert a["tlen"] == 0 and not a["flag"] & 2


def test_planted_mistake_is_caught():
    """A pairing that signs tlen by mate order instead of position order (the
    SAM path's rule) fails the hand-worked case."""
    f = pe_of([((0, 500, 20, 16, 2, 3, 0, 50), (0, 200, 60, 0, 4, 0, 0, 40))])
    wrong_sign = [abs(int(f["tlen"][0])), -abs(int(f["tlen"][0]))]
    assert f["tlen"].tolist() != wrong_sign
