"""Host-side logic (CPU): C-ABI exports, CLI parsing, FASTQ/SAM formatting,
pairing rules, dense-v1 pack files. Expected values come from the reference's
rules (cited) worked by hand."""

from __future__ import annotations

import io
import re
from pathlib import Path

import numpy as np
import pytest

from mialign import _lib
from mialign import linear_gpu_locate as lgl
from mialign import mojo_linear_pack as pack
from mialign.linear_mapper import get_kv_value

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "mialign.h").read_text()
    declared = set(re.findall(r"\b(ma_[a-z0-9_]+)\s*\(", header))
    assert len(declared) > 40
    lib = _lib.lib()  # loads without a GPU; no compute call is made
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == declared
    p = _lib.default_params()
    assert (p.k, p.stride, p.max_occ, p.vote_occ, p.vote_fast, p.max_diff, p.max_soft, p.max_indel,
            p.do_rc) == (21, 5, 16384, 256, 256, 6, 8, 0, 1)


def test_get_kv_value():
    # methylgrapher/src/utility.mojo:151-165
    args = ["-ref", "a.fa", "-fq1", "r1.fq", "-k", "21", "-device"]
    assert get_kv_value(args, "ref", "") == "a.fa"
    assert get_kv_value(args, "k", "15") == "21"
    assert get_kv_value(args, "device", "auto") == "auto"  # key without value
    assert get_kv_value(args, "fq2", "") == ""


def test_fastq_name_rules_and_conversion():
    # _read_one (linear_gpu_locate.mojo:73-96): name cut at space and '/'
    r = lgl.read_one(io.BytesIO(b"@read7/1 extra words\nACGTCacgt\n+\n\n"))
    assert r.name == "read7" and r.qual == "*"
    assert lgl.bs_convert(r.seq, "C2T") == b"ATGTTatgt"
    assert lgl.bs_convert(b"GGTAGGTA", "G2A") == b"AATAAATA"
    assert lgl.read_one(io.BytesIO(b"")) is None


def test_pe_batch_r2_short_raises():
    fq1 = io.BytesIO(b"@a\nACGT\n+\nIIII\n@b\nACGT\n+\nIIII\n")
    fq2 = io.BytesIO(b"@a\nACGT\n+\nIIII\n")
    with pytest.raises(_lib.MialignError):
        lgl.read_pe_batch(fq1, fq2, True, 10, "C2T", "G2A")


def test_pack_batch_pads_with_n_and_k():
    rows, lens, ml = lgl.pack_batch([b"ACG", b"ACGTACGT"], 21)
    assert ml == 21 and rows.shape == (2, 21)
    assert rows[0, :3].tobytes() == b"ACG" and (rows[0, 3:] == 78).all()
    assert list(lens) == [3, 8]


def test_hit_from_gpu_cigar_seq_and_contig():
    names = ["fchr1", "rchr1"]
    h = lgl.hit_from_gpu(names, "q", b"TTTTACGA", b"CCTTACGA", 1, 99, 60, 16, 2, 1)
    assert (h.flag, h.contig, h.pos, h.mapq, h.cigar) == (16, "chr1", 100, 60, "2S5M1S")
    assert h.seq == "TCGTAAGG"  # original read, reverse-complemented for flag 16
    u = lgl.hit_from_gpu(names, "q", b"ACGT", b"ACGT", -1, 0, 0, 4, 0, 0)
    assert (u.flag, u.contig, u.pos, u.cigar, u.seq) == (4, "*", 0, "*", "ACGT")
    z = lgl.hit_from_gpu(names, "q", b"AC", b"AC", 0, 5, 20, 0, 1, 1)  # nothing between clips
    assert z.flag == 4
    assert lgl.canonical_contig("f") == "f" and lgl.canonical_contig("chrX") == "chrX"


def test_pair_hits_rules():
    # linear_extend.mojo:263-294
    a = lgl.LinearHit("p", 0, "chr1", 100, 60, "10M", "A" * 10)
    b = lgl.LinearHit("p", 16, "chr1", 150, 60, "12M", "A" * 12)
    a, b = lgl.pair_hits(a, b)
    assert (a.flag, b.flag) == (1 | 64 | 2, 16 | 1 | 128 | 2)
    assert (a.rnext, a.pnext, a.tlen, b.pnext, b.tlen) == ("=", 150, 62, 100, -62)
    c = lgl.LinearHit("p", 0, "chr2", 5, 40, "10M", "A" * 10)
    d = lgl.LinearHit("p", 4, "*", 0, 0, "*", "A" * 10)
    c, d = lgl.pair_hits(c, d)
    assert (c.flag, d.flag) == (1 | 64 | 8, 4 | 1 | 128)
    assert (d.rnext, d.pnext, c.rnext) == ("chr2", 5, "*")


def test_sam_header_dedupes_canonical_names(monkeypatch):
    monkeypatch.setenv("METHYLGRAPHER_RG_ID", "rgX")
    h = lgl.sam_header(["fchr1", "rchr1", "fchr2", "rchr2"], [10, 10, 20, 20]).splitlines()
    assert h[0] == "@HD\tVN:1.6\tSO:unsorted"
    assert h[1:3] == ["@SQ\tSN:chr1\tLN:10", "@SQ\tSN:chr2\tLN:20"]
    assert h[3].startswith("@RG\tID:rgX\tSM:sample")


def test_env_knobs_quirk(monkeypatch):
    # _env_int maps values < 1 back to the default (linear_gpu_locate.mojo:139-145)
    monkeypatch.setenv("METHYLGRAPHER_LINEAR_MAX_DIFF", "0")
    monkeypatch.setenv("METHYLGRAPHER_LINEAR_MAX_SOFT", "3")
    monkeypatch.setenv("METHYLGRAPHER_LINEAR_RC", "no")
    p = lgl.knobs_from_env(21, 16384)
    assert (p.max_diff, p.max_soft, p.do_rc, p.max_indel) == (6, 3, 0, 0)


def test_dense_pack_files_roundtrip_and_cache_build_off(tmp_path, monkeypatch):
    km = np.array([3, 9], np.uint64)
    of = np.array([0, 2, 3], np.uint64)
    po = np.array([0, 1, 0, 5, 1, 2], np.uint32)
    seq = np.frombuffer(b"ACGTACGTAC", dtype=np.uint8)
    off = np.array([0, 6, 10], np.uint64)
    out = tmp_path / "p"
    pack.write_dense_dir(out, 4, km, of, po, seq, off, ["fa", "ra"])
    assert pack.pack_is_complete(out)
    meta, k2, o2, p2, s2, c2 = pack.load_dense_dir(out)
    assert meta["format"] == "dense-v1" and meta["n_keys"] == 2 and meta["n_postings"] == 3
    assert list(k2) == [3, 9] and list(p2) == list(po) and bytes(s2) == b"ACGTACGTAC"
    # METHYLGRAPHER_LINEAR_CACHE_BUILD=0 fails instead of building (mojo_linear_pack.py:122-127)
    monkeypatch.setenv("METHYLGRAPHER_LINEAR_CACHE_BUILD", "0")
    with pytest.raises(RuntimeError):
        pack.ensure_dense_pack(c2t_fasta=tmp_path / "x.fa", cache_dir=tmp_path / "missing", k=4)
    # a complete pack is returned as is
    assert pack.ensure_dense_pack(c2t_fasta=tmp_path / "x.fa", cache_dir=out, k=4) == out


def test_read_fasta_contigs(tmp_path):
    fa = tmp_path / "r.fa"
    fa.write_text(">chr1 desc\nacgt\nNNac\n>chr2\nTTTT\n")
    c = pack.read_fasta_contigs(fa)
    assert c == [("chr1", b"ACGTNNAC"), ("chr2", b"TTTT")]
