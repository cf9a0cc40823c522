"""The dense-v1 pack the reference's own builder wrote (tests/golden/
dense_v1_pack, made by tests/golden/make_dense_v1_fixture.py from its
ref.c2t.fa): the oracle's dense build reproduces its kmers / offsets /
postings bytes; this package's FASTA reader gives its contig names, bases and
offsets; this package's pack writer reproduces every file byte for byte
(meta.json up to the source path it records)."""

import json
from pathlib import Path

import numpy as np

from mialign import mojo_linear_pack as pack
from oracle import oracle as orc

GOLD = Path(__file__).resolve().parent / "golden" / "dense_v1_pack"


def _arrays():
    kmers = np.fromfile(GOLD / "kmers.bin", "<u8")
    offsets = np.fromfile(GOLD / "offsets.bin", "<u8")
    postings = np.fromfile(GOLD / "postings.bin", "<u4")
    seq = np.fromfile(GOLD / "sequences.bin", np.uint8)
    coff = np.fromfile(GOLD / "contig_offsets.bin", "<u8")
    return kmers, offsets, postings, seq, coff


def test_oracle_dense_build_equals_reference_pack():
    kmers, offsets, postings, seq, coff = _arrays()
    meta = json.loads((GOLD / "meta.json").read_text())
    km, of, po = orc.build_dense(seq, coff, int(meta["k"]))
    assert np.array_equal(km, kmers) and np.array_equal(of, offsets) and np.array_equal(po, postings)
    assert meta["n_keys"] == kmers.size and meta["n_postings"] * 2 == postings.size


def test_fasta_reader_matches_reference_pack():
    kmers, offsets, postings, seq, coff = _arrays()
    contigs = pack.read_fasta_contigs(GOLD / "ref.c2t.fa")
    names, s2, o2 = pack.contigs_to_arrays(contigs)
    meta = json.loads((GOLD / "meta.json").read_text())
    assert names == meta["contigs"] == (GOLD / "contigs.txt").read_text().split()
    assert np.array_equal(s2, seq) and np.array_equal(o2, coff)
    assert b"acgt" not in seq.tobytes() and b"N" in seq.tobytes()  # upper-cased, N run kept


def test_pack_writer_reproduces_reference_files(tmp_path):
    kmers, offsets, postings, seq, coff = _arrays()
    meta = json.loads((GOLD / "meta.json").read_text())
    out = tmp_path / "pack"
    mine = pack.write_dense_dir(out, int(meta["k"]), kmers, offsets, postings, seq, coff, meta["contigs"],
                                meta["c2t_fasta"])
    for f in ("kmers.bin", "offsets.bin", "postings.bin", "sequences.bin", "contig_offsets.bin", "contigs.txt",
              "meta.json"):
        assert (out / f).read_bytes() == (GOLD / f).read_bytes(), f
    assert mine == meta
    assert pack.pack_is_complete(GOLD)
