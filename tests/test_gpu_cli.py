"""End to end through the drop-in entry point: FASTQ + converted FASTA ->
dense-v1 pack (built on the GPU) -> map -> SAM, via the reference-shaped CLI."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from mialign import linear_mapper, synth
from mialign import linear_gpu_locate as lgl
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "fq2bam_fixture"


def sam_records(path: Path):
    return [l.split("\t") for l in path.read_text().splitlines() if not l.startswith("@")]


def test_reference_toy_fixture_pe(tmp_path):
    print("test_reference_toy_fixture_pe", flush=True)
    out = tmp_path / "toy.sam"
    rc = linear_mapper.main(["-ref", str(GOLD / "ref.fa"), "-fq1", str(GOLD / "R1.fastq"), "-fq2",
                             str(GOLD / "R2.fastq"), "-out_sam", str(out), "-device", "amd", "-k", "8",
                             "-cache_dir", str(tmp_path / "pack"), "-bs_r1", "C2T", "-bs_r2", "G2A"])
    assert rc == 0
    recs = sam_records(out)
    assert len(recs) == 2
    assert sum(1 for r in recs if r[2] != "*") >= 1  # test_fq2bam_meth.py:163-187
    assert all(int(r[1]) & 1 for r in recs)


def test_cli_sam_matches_oracle_records(tmp_path):
    print("test_cli_sam_matches_oracle_records", flush=True)
    ref = synth.gen_reference(41, 2, 300_000)
    dual = synth.dual_convert(ref)
    fa = tmp_path / "ref.c2t.fa"
    with fa.open("w") as fh:
        for i, nm in enumerate(dual.names):
            a, b = int(dual.contig_off[i]), int(dual.contig_off[i + 1])
            fh.write(f">{nm}\n{dual.seq[a:b].tobytes().decode()}\n")
    reads = synth.gen_reads(42, ref, 2000, 60, 150, max_len=150)
    fq = tmp_path / "r1.fq"
    with fq.open("w") as fh:
        for i in range(reads.n):
            s = reads.bases[i, : reads.lens[i]].tobytes().decode()
            fh.write(f"@r{i} x\n{s}\n+\n{'I' * len(s)}\n")
    out = tmp_path / "out.sam"
    rc = linear_mapper.main(["-ref", str(fa), "-fq1", str(fq), "-out_sam", str(out), "-k", "21",
                             "-cache_dir", str(tmp_path / "pack"), "-bs_r1", "C2T"])
    assert rc == 0
    recs = sam_records(out)
    assert len(recs) == reads.n
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases, 1), reads.lens)
    for i, r in enumerate(recs):
        orig = reads.bases[i, : reads.lens[i]].tobytes()
        conv = orig.replace(b"C", b"T")
        h = lgl.hit_from_gpu(dual.names, f"r{i}", conv, orig, int(want["cid"][i]), int(want["pos"][i]),
                             int(want["mapq"][i]), int(want["flag"][i]), int(want["sl"][i]),
                             int(want["sr"][i]))
        assert r[:11] == [h.query_name, str(h.flag), h.contig, str(h.pos), str(h.mapq), h.cigar, "*",
                          "0", "0", h.seq, "I" * len(orig)], i
        assert r[11] == "RG:Z:mojo1"
    assert sum(1 for r in recs if r[2] != "*") > 0.99 * reads.n


def test_cli_bam_honours_env_knobs(tmp_path, monkeypatch, capfd):
    """`-out_sam x.bam` maps with the METHYLGRAPHER_LINEAR_* knobs, as the SAM
    path does: with MAX_DIFF 3 and MAX_SOFT 4 the BAM stream equals the
    oracle's pack of the oracle's map under those knobs (sorted, duplicates
    flagged); MAX_INDEL 2 is reported and mapped gapless (BAM records carry
    slS mM srS CIGARs only)."""
    print("test_cli_bam_honours_env_knobs", flush=True)
    import bam_decode as bd

    ref = synth.gen_reference(43, 2, 300_000)
    dual = synth.dual_convert(ref)
    fa = tmp_path / "ref.c2t.fa"
    with fa.open("w") as fh:
        for i, nm in enumerate(dual.names):
            a, b = int(dual.contig_off[i]), int(dual.contig_off[i + 1])
            fh.write(f">{nm}\n{dual.seq[a:b].tobytes().decode()}\n")
    reads = synth.gen_reads(44, ref, 4000, 60, 150, max_len=150, p_err=0.02, dup_frac=0.1)
    blob, name_off, qual = synth.gen_names_quals(45, reads.lens, 150)
    fq = tmp_path / "r1.fq"
    fq.write_bytes(synth.write_fastq(reads.bases, qual, reads.lens, blob, name_off).tobytes())
    for var, val in (("MAX_DIFF", "3"), ("MAX_SOFT", "4"), ("MAX_INDEL", "2")):
        monkeypatch.setenv(f"METHYLGRAPHER_LINEAR_{var}", val)
    out = tmp_path / "out.bam"
    rc = linear_mapper.main(["-ref", str(fa), "-fq1", str(fq), "-out_sam", str(out), "-k", "21",
                             "-cache_dir", str(tmp_path / "pack"), "-bs_r1", "C2T"])
    assert rc == 0
    assert "MAX_INDEL=2 ignored for BAM output" in capfd.readouterr().err
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    conv = orc.convert_reads(reads.bases, 1)
    want, _, _, _ = orc.map_batch(oix, conv, reads.lens, orc.params(max_diff=3, max_soft=4), nthreads=8)
    dflt, _, _, _ = orc.map_batch(oix, conv, reads.lens, nthreads=8)
    # the knobs change records of this workload, so the comparison can tell them apart
    assert any(not np.array_equal(want[f], dflt[f]) for f in ("cid", "nm", "sl", "sr", "mapq"))
    tid = synth.tid_of_cid(dual.names)
    d, p, _ = orc.sort_markdup(*orc.sort_keys_se(want, reads.lens, tid))
    stream = orc.bam_pack_se(want, tid, blob, name_off, reads.bases, reads.lens, qual, b"mojo1", perm=p,
                             is_dup=d, apply_dup=True)
    data = b"".join(pl for _, pl in bd.bgzf_members(out.read_bytes()))
    assert data.startswith(b"BAM\x01") and data.endswith(stream)


def test_cli_device_cpu_host_loop(tmp_path):
    """`-device cpu`: the reference's forward-only host loop (linear_mapper.mojo:
    209-311) — SAM header with the pack's own contig names, eleven fields per
    line (no tags), SEQ the converted read, records equal the oracle's
    or_map_forward; pairs get the pair_hits fields (linear_extend.mojo:263-294):
    0x1 and 0x40 / 0x80, 0x2 with '=' mates, PNEXT and TLEN |pos2 - pos1| +
    len(R2) when both map to one contig, else the other mate's contig and
    position, 0x8 on the mate of an unmapped read. (Forward only: in a
    directional library R2 is the reverse complement and never maps with its
    R1, so the pairs here are two forward reads each.)"""
    print("test_cli_device_cpu_host_loop", flush=True)
    ref = synth.gen_reference(51, 2, 300_000)
    dual = synth.dual_convert(ref)
    fa = tmp_path / "ref.c2t.fa"
    with fa.open("w") as fh:
        for i, nm in enumerate(dual.names):
            a, b = int(dual.contig_off[i]), int(dual.contig_off[i + 1])
            fh.write(f">{nm}\n{dual.seq[a:b].tobytes().decode()}\n")
    # R2: reads of their own (either strand, 120 bp), converted C2T like R1, so
    # that both mates of a pair can map forward: to one contig (a proper pair)
    # or to two
    r1 = synth.gen_reads(52, ref, 1500, 150, 150)
    r2 = synth.gen_reads(53, ref, 1500, 120, 120)
    for nm, rd in (("R1.fq", r1), ("R2.fq", r2)):
        with (tmp_path / nm).open("w") as fh:
            for i in range(rd.n):
                s = rd.bases[i, : rd.lens[i]].tobytes().decode()
                fh.write(f"@p{i}/{nm[1]} x\n{s}\n+\n{'F' * len(s)}\n")
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    w1 = orc.map_forward(oix, r1.bases, r1.lens, k=21, bs_r1=1)
    w2 = orc.map_forward(oix, r2.bases, r2.lens, k=21, bs_r1=1)

    def expect(w, rd, i, mode):
        s = rd.bases[i, : rd.lens[i]].tobytes().replace(b"C" if mode == 1 else b"G", b"T" if mode == 1 else b"A")
        cid = int(w["cid"][i])
        if cid < 0:
            return [4, "*", 0, 0, "*", s.decode()]
        sl, sr = int(w["sl"][i]), int(w["sr"][i])
        cig = (f"{sl}S" if sl else "") + f"{len(s) - sl - sr}M" + (f"{sr}S" if sr else "")
        return [0, dual.names[cid], int(w["pos"][i]) + 1, int(w["mapq"][i]), cig, s.decode()]

    # single end
    out = tmp_path / "se.sam"
    rc = linear_mapper.main(["-ref", str(fa), "-fq1", str(tmp_path / "R1.fq"), "-out_sam", str(out), "-k", "21",
                             "-device", "cpu", "-cache_dir", str(tmp_path / "pack"), "-bs_r1", "C2T"])
    assert rc == 0
    text = out.read_text()
    sq = [l.split("\t")[1][3:] for l in text.splitlines() if l.startswith("@SQ")]
    assert sq == list(dual.names)
    recs = sam_records(out)
    assert len(recs) == r1.n and all(len(r) == 11 for r in recs)
    for i, r in enumerate(recs):
        e = expect(w1, r1, i, 1)
        assert r[0] == f"p{i}"
        assert [int(r[1]), r[2], int(r[3]), int(r[4]), r[5], r[9]] == e, f"read {i}"
        assert r[6:9] == ["*", "0", "0"] and r[10] == "F" * int(r1.lens[i])
    assert sum(1 for r in recs if r[2] != "*") > 0.3 * r1.n
    # paired
    out = tmp_path / "pe.sam"
    rc = linear_mapper.main(["-ref", str(fa), "-fq1", str(tmp_path / "R1.fq"), "-fq2", str(tmp_path / "R2.fq"),
                             "-out_sam", str(out), "-k", "21", "-device", "cpu", "-cache_dir", str(tmp_path / "pack"),
                             "-bs_r1", "C2T", "-bs_r2", "C2T"])
    assert rc == 0
    recs = sam_records(out)
    assert len(recs) == 2 * r1.n
    n_proper = n_both = 0
    for i in range(r1.n):
        a, b = recs[2 * i], recs[2 * i + 1]
        ea, eb = expect(w1, r1, i, 1), expect(w2, r2, i, 1)
        fa_, fb_ = ea[0] | 1 | 64, eb[0] | 1 | 128
        if ea[1] != "*" and eb[1] != "*" and ea[1] == eb[1]:
            fa_ |= 2
            fb_ |= 2
            tlen = abs(eb[2] - ea[2]) + int(r2.lens[i])
            mates = (["=", str(eb[2]), str(tlen)], ["=", str(ea[2]), str(-tlen)])
            n_proper += 1
        else:
            n_both += ea[1] != "*" and eb[1] != "*"
            if ea[1] == "*":
                fb_ |= 8
            if eb[1] == "*":
                fa_ |= 8
            mates = ([eb[1] if eb[1] != "*" else "*", str(eb[2]) if eb[1] != "*" else "0", "0"],
                     [ea[1] if ea[1] != "*" else "*", str(ea[2]) if ea[1] != "*" else "0", "0"])
        assert [int(a[1])] + [a[2], int(a[3]), int(a[4]), a[5], a[9]] == [fa_] + ea[1:], f"pair {i} R1"
        assert [int(b[1])] + [b[2], int(b[3]), int(b[4]), b[5], b[9]] == [fb_] + eb[1:], f"pair {i} R2"
        assert a[6:9] == mates[0] and b[6:9] == mates[1], f"pair {i} mate fields"
    assert n_proper > 0.05 * r1.n and n_both > 0.05 * r1.n
