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
