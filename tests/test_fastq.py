"""FASTQ ingest (CPU): pin the oracle's restatement of the reference's bulk
FASTQ reader against the reference's known answers, its FASTQ fixture, the host
reader of the parity path on random FASTQ text, and hand-cut edge cases."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from mialign import linear_gpu_locate as lgl
from oracle import oracle as orc

FIX = Path(__file__).resolve().parent / "golden" / "fq2bam_fixture"
L = orc.lib()


def names_of(r) -> list[bytes]:
    nm, off = r["names"].tobytes(), r["name_off"]
    return [nm[off[i]:off[i + 1]] for i in range(r["n"])]


def test_reference_known_answers():
    # fq2bam-meth/tests/test_linear_fastq.mojo:30-87
    r1 = orc.fastq_parse(b"@read1 extra/1\nACGTACGT\n+\nIIIIIIII\n@read2\nCCCC\n+\nJJJJ\n", 8, 1)
    r2 = orc.fastq_parse(b"@read1 extra/2\nGGTAGGTA\n+\nKKKKKKKK\n@read2\nGGGG\n+\nLLLL\n", 8, 2)
    assert r1["n"] == 2 and r1["max_len"] == 8
    assert names_of(r1) == [b"read1", b"read2"] and len(names_of(r1)[0]) == 5
    assert r1["rows"][0][1] == ord("C") and r1["conv"][0][1] == ord("T")
    assert r2["rows"][0][0] == ord("G") and r2["conv"][0][0] == ord("A")
    assert r1["conv"][1][0] == ord("T") and r1["lens"][1] == 4
    dest = np.concatenate([r1["conv"], r2["conv"]]).ravel()  # fq_pack_bases: R1 rows, then R2
    assert list(r1["lens"]) == [8, 4] and dest[1] == ord("T") and dest[12] == ord("N") and dest[16] == ord("A")
    assert r1["qual"][0].tobytes() == b"IIIIIIII" and list(r1["qual_lens"]) == [8, 4]


def test_reference_fixture():
    # fq2bam-meth/tests/test_fastq_batch.py:14-24 (R1 is A/T only, so C2T leaves it unchanged)
    r = orc.fastq_parse((FIX / "R1.fastq").read_bytes(), 256, 1)
    assert r["n"] >= 1 and names_of(r)[0] == b"read1" and r["status"] == 0
    assert np.array_equal(r["rows"], r["conv"])
    r2 = orc.fastq_parse((FIX / "R2.fastq").read_bytes(), 256, 2)
    assert r2["n"] == r["n"]


def test_conversion_table():
    # _cvt (linear_fastq.mojo:37-46): C/c -> 'T' for C2T, G/g -> 'A' for G2A, all else unchanged
    for b in range(256):
        assert L.or_fq_cvt(b, 0) == b
        assert L.or_fq_cvt(b, 1) == (ord("T") if chr(b) in "Cc" else b)
        assert L.or_fq_cvt(b, 2) == (ord("A") if chr(b) in "Gg" else b)


ALPHA = np.frombuffer(b"ACGTNacgtn", np.uint8)


def random_fastq(seed: int, n: int, max_len: int = 250, crlf: float = 0.2, tail_newline: bool = True) -> bytes:
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        L = int(rng.integers(0, max_len + 1))
        seq = ALPHA[rng.integers(0, ALPHA.size, L)].tobytes()
        qual = rng.integers(35, 75, L).astype(np.uint8).tobytes()
        kind = int(rng.integers(0, 4))
        name = [b"@r%d" % i, b"@r%d extra words" % i, b"@r%d/1" % i, b"r%d" % i][kind]
        nl = b"\r\n" if rng.random() < crlf else b"\n"
        out.append(name + nl + seq + nl + b"+" + nl + qual + nl)
    data = b"".join(out)
    return data if tail_newline else data.rstrip(b"\r\n")


def test_against_the_host_reader():
    """The parity path's host reader (read_one, linear_gpu_locate.mojo:73-96)
    yields the same names, bases and qualities on random FASTQ text."""
    import io

    data = random_fastq(3, 500)
    r = orc.fastq_parse(data, 250, 0)
    fh = io.BytesIO(data)
    host = []
    while True:
        rec = lgl.read_one(fh)
        if rec is None:
            break
        host.append(rec)
    assert r["n"] == len(host) == 500 and r["status"] == 0 and r["consumed"] == len(data)
    for i, h in enumerate(host):
        L = int(r["lens"][i])
        assert r["rows"][i][:L].tobytes() == h.seq and names_of(r)[i] == h.name.encode()
        # the host reader writes an empty quality line as "*" (its SAM QUAL)
        assert r["qual"][i][:int(r["qual_lens"][i])].tobytes() == (b"" if h.qual == "*" else h.qual.encode())
        assert (r["rows"][i][L:] == ord("N")).all()


def test_edge_cases_worked_by_hand():
    # no final newline: the last line still counts
    r = orc.fastq_parse(b"@a\nAC\n+\nII", 4)
    assert r["n"] == 1 and r["consumed"] == 10 and r["qual"][0][:2].tobytes() == b"II"
    # a final "\r" alone is not a line -> the record lacks its quality line
    assert orc.fastq_parse(b"@a\nAC\n+\n\r", 4)["status"] == 3
    # truncated records
    assert orc.fastq_parse(b"@a\n", 4)["status"] == 1
    assert orc.fastq_parse(b"@a\nAC\n", 4)["status"] == 2
    assert orc.fastq_parse(b"@a\nAC\n+\n", 4)["status"] == 3
    # CRLF stripped from every line; an empty line is a (nameless) header
    r = orc.fastq_parse(b"@x y\r\nACG\r\n+\r\nIII\r\n\nT\n+\nI\n", 4, 1)
    assert r["n"] == 2 and names_of(r) == [b"x", b""] and list(r["lens"]) == [3, 1]
    assert r["conv"][0][:3].tobytes() == b"ATG"
    # without `final`, an unterminated or incomplete tail waits for more input
    r = orc.fastq_parse(b"@a\nAC\n+\nII\n@b\nA", 4, final=False)
    assert r["n"] == 1 and r["consumed"] == 11 and r["status"] == 0
    r = orc.fastq_parse(b"@a\nAC\n+\nII", 4, final=False)
    assert r["n"] == 0 and r["consumed"] == 0
    # max_records stops the batch; the rest is not consumed
    d = random_fastq(5, 10)
    r = orc.fastq_parse(d, 250, max_records=4)
    assert r["n"] == 4 and d[r["consumed"]:].lstrip(b"@").startswith(b"r4")
    # a read longer than stride stops the batch before it (status 4)
    r = orc.fastq_parse(b"@a\nAC\n+\nII\n@b\nACGTA\n+\nIIIII\n@c\nA\n+\nI\n", 4)
    assert r["n"] == 1 and r["status"] == 4 and r["consumed"] == 11
    # names past names_cap stop the batch (status 5)
    r = orc.fastq_parse(b"@aaaa\nA\n+\nI\n@bbbb\nA\n+\nI\n", 4, names_cap=6)
    assert r["n"] == 1 and r["status"] == 5 and names_of(r) == [b"aaaa"]


@pytest.mark.parametrize("split_seed", [1, 2])
def test_streaming_in_pieces_equals_one_pass(split_seed):
    """Parsing without `final` piece by piece, carrying the unconsumed tail,
    gives the records of one final pass."""
    data = random_fastq(7, 300, crlf=0.5, tail_newline=False)
    whole = orc.fastq_parse(data, 250, 2)
    rng = np.random.default_rng(split_seed)
    cuts = sorted(rng.integers(1, len(data), 6).tolist()) + [len(data)]
    got_rows, got_names, carry, pos = [], [], b"", 0
    for c in cuts:
        carry += data[pos:c]
        pos = c
        r = orc.fastq_parse(carry, 250, 2, final=(c == len(data)))
        got_rows += [row.tobytes() for row in r["conv"]]
        got_names += names_of(r)
        carry = carry[r["consumed"]:]
    assert got_rows == [row.tobytes() for row in whole["conv"]] and got_names == names_of(whole)
