"""GPU parity of the FASTQ ingest: ma_fastq_parse_dev (through the C ABI)
produces the oracle's rows, converted rows, qualities, lengths, names and
batch info byte for byte, whole or in streamed pieces, and the streaming
reader's batches over files (plain and gzip, SE and PE) equal one oracle pass."""

from __future__ import annotations

import gzip
import sys

import numpy as np
import pytest

import mialign
from mialign import linear_fastq as lf
from oracle import oracle as orc
from test_fastq import random_fastq

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def compare(data: bytes, stride: int, mode: int, final: bool = True, max_records=None, names_cap=None):
    want = orc.fastq_parse(data, stride, mode, final=final, max_records=max_records, names_cap=names_cap)
    mr = max_records if max_records is not None else max(1, data.count(b"\n") // 4 + 1)
    info, got = lf.parse_fastq_bytes(data, stride, {0: "", 1: "C2T", 2: "G2A"}[mode], final=final, max_records=mr,
                                     names_cap=names_cap if names_cap is not None else max(1, len(data)))
    assert (info.records, info.consumed, info.max_len, info.status, info.name_bytes) == (
        want["n"], want["consumed"], want["max_len"], want["status"], want["name_bytes"]), (info, want["status"])
    n = want["n"]
    for k in ("rows", "conv", "qual", "lens", "qual_lens"):
        assert np.array_equal(got[k][:n], want[k]), k
    assert np.array_equal(got["name_off"][:n + 1], want["name_off"])
    assert got["names"].tobytes() == want["names"].tobytes()
    return want


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_parse_matches_oracle(mode):
    say(f"test_parse_matches_oracle mode={mode}: 20k records")
    data = random_fastq(21 + mode, 20_000, crlf=0.3)
    want = compare(data, 256, mode)
    assert want["n"] == 20_000 and want["status"] == 0
    compare(data.rstrip(b"\r\n"), 256, mode)  # no final newline
    compare(data, 250, mode, max_records=7777)  # a partial batch


def test_parse_edge_cases_match_oracle():
    say("test_parse_edge_cases_match_oracle")
    cases = [b"", b"@a\nAC\n+\nII", b"@a\nAC\n+\n\r", b"@a\n", b"@a\nAC\n", b"@a\nAC\n+\n",
             b"@x y\r\nACG\r\n+\r\nIII\r\n\nT\n+\nI\n", b"@a\nAC\n+\nII\n@b\nACGTA\n+\nIIIII\n@c\nA\n+\nI\n",
             b"\n\n\n\n", b"@r/1\nacgtn\n+\n!!!!!\n"]
    for c in cases:
        compare(c, 4, 1)
        compare(c, 4, 2, final=False)
    compare(b"@aaaa\nA\n+\nI\n@bbbb\nA\n+\nI\n", 4, 0, names_cap=6)


def test_parse_in_pieces_matches_oracle():
    say("test_parse_in_pieces_matches_oracle")
    data = random_fastq(31, 3000, crlf=0.5, tail_newline=False)
    rng = np.random.default_rng(3)
    cuts = sorted(rng.integers(1, len(data), 9).tolist()) + [len(data)]
    carry, pos, total = b"", 0, 0
    for c in cuts:
        carry += data[pos:c]
        pos = c
        w = compare(carry, 250, 1, final=(c == len(data)))
        total += w["n"]
        carry = carry[w["consumed"]:]
    assert total == 3000


@pytest.mark.parametrize("gz", ["plain", "gzip", "bgzf"])
def test_reader_over_files(tmp_path, gz):
    """The streaming reader's batches (small batches and chunks, so tails are
    carried across buffers) equal one oracle pass over each file, plain,
    gzip or BGZF (members inflated in parallel); R2 shorter than R1 raises the
    reference's error."""
    say(f"test_reader_over_files gz={gz}")
    d1 = random_fastq(41, 5000, crlf=0.1)
    d2 = random_fastq(42, 5000, crlf=0.1)
    sfx = "fq" if gz == "plain" else "fq.gz"
    p1, p2 = tmp_path / f"r1.{sfx}", tmp_path / f"r2.{sfx}"
    for p, d in ((p1, d1), (p2, d2)):
        if gz == "gzip":
            with gzip.open(p, "wb") as fh:
                fh.write(d)
        elif gz == "bgzf":
            p.write_bytes(orc.bgzf_stream(d, 1))
        else:
            p.write_bytes(d)
    w1, w2 = orc.fastq_parse(d1, 256, 1), orc.fastq_parse(d2, 256, 2)
    rd = lf.DeviceFastqReader(str(p1), str(p2), "C2T", "G2A", batch_reads=1234, chunk_bytes=100_000)
    rows1, rows2, names1, quals1 = [], [], [], []
    while True:
        b1, b2 = rd.read_batch()
        if b1.n == 0:
            break
        h1, h2 = b1.download(), b2.download()
        assert b2.n == b1.n
        rows1.append(h1["conv"])
        rows2.append(h2["conv"])
        quals1.append(h1["qual"])
        nm, off = h1["names"].tobytes(), h1["name_off"]
        names1 += [nm[off[i]:off[i + 1]] for i in range(b1.n)]
    rd.close()
    assert np.array_equal(np.concatenate(rows1), w1["conv"]) and np.array_equal(np.concatenate(rows2), w2["conv"])
    assert np.array_equal(np.concatenate(quals1), w1["qual"])
    wn, wo = w1["names"].tobytes(), w1["name_off"]
    assert names1 == [wn[wo[i]:wo[i + 1]] for i in range(w1["n"])]
    # R2 ends early
    p3 = tmp_path / "r2short.fq"
    p3.write_bytes(random_fastq(43, 100))
    rd = lf.DeviceFastqReader(str(p1), str(p3), "C2T", "G2A", batch_reads=1000, chunk_bytes=100_000)
    with pytest.raises(mialign.MialignError, match="R2 ended early"):
        while rd.read_batch()[0].n:
            pass
    rd.close()


def test_parse_unaligned_buffers_match_oracle():
    """Text and output rows at odd byte offsets (the C ABI takes any pointers):
    base rows, converted rows and quality rows each with another alignment, so
    the word writes fall back to bytes at every row edge and the converted rows
    take their own pass."""
    from mialign._lib import check, lib, ptr

    say("test_parse_unaligned_buffers_match_oracle")
    data = random_fastq(31, 3000, crlf=0.2)
    stride, mode = 151, 1
    want = orc.fastq_parse(data, stride, mode)
    n = want["n"]
    mr = n + 1
    text = mialign.DeviceArray((len(data) + 32,), np.uint8)
    host_text = np.frombuffer(data, np.uint8).copy()  # kept alive across the copy
    check(lib().ma_memcpy_h2d(text.ptr + 3, ptr(host_text), len(data), None), "h2d")
    size = mr * stride + 16
    rows, conv, qual = (mialign.DeviceArray((size,), np.uint8) for _ in range(3))
    offs = {"rows": 1, "conv": 2, "qual": 7}
    lens = mialign.DeviceArray((mr,), np.uint32)
    qlens = mialign.DeviceArray((mr,), np.uint32)
    names = mialign.DeviceArray((len(data),), np.uint8)
    name_off = mialign.DeviceArray((mr + 1,), np.int64)
    info = mialign.DeviceArray((8,), np.uint64)
    sb = int(lib().ma_fastq_scratch_bytes(len(data), mr))
    scratch = mialign.DeviceArray((sb,), np.uint8)
    check(lib().ma_fastq_parse_dev(None, text.ptr + 3, len(data), 1, mr, stride, mode, rows.ptr + offs["rows"],
                                   conv.ptr + offs["conv"], qual.ptr + offs["qual"], lens.ptr, qlens.ptr, names.ptr,
                                   len(data), name_off.ptr, info.ptr, scratch.ptr, sb), "ma_fastq_parse_dev")
    got = {}
    for k, arr in (("rows", rows), ("conv", conv), ("qual", qual)):
        host = arr.download()
        got[k] = host[offs[k]:offs[k] + n * stride].reshape(n, stride)
    assert int(info.download()[0]) == n
    for k in ("rows", "conv", "qual"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(lens.download()[:n], want["lens"])
    assert np.array_equal(name_off.download()[:n + 1], want["name_off"])
    assert names.download()[: int(want["name_off"][-1])].tobytes() == want["names"].tobytes()
