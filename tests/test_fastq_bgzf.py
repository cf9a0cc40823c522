"""Compressed FASTQ input on the host: BGZF (bgzip output) members inflated in
parallel by ma_bgzf_inflate, other gzip by zlib's gz reader (ma_gz_*), both
behind the reader's file interface. Expected bytes
come from Python's gzip module (multi-member gzip), independent of the native
inflate; members come from the oracle's BGZF writer and the native compressor."""

from __future__ import annotations

import gzip

import numpy as np
import pytest

from mialign import MialignError, bam_emit
from mialign import linear_fastq as lf
from oracle import oracle as orc


def fastq_text(seed: int, n: int) -> bytes:
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGTN", np.uint8)
    qual = np.frombuffer(b"FF:,#", np.uint8)
    out = []
    for i in range(n):
        L = int(rng.integers(50, 251))
        out.append(b"@read%d/1\n" % i + acgt[rng.integers(0, 5, L)].tobytes() + b"\n+\n"
                   + qual[rng.integers(0, 5, L)].tobytes() + b"\n")
    return b"".join(out)


def read_all(f, piece: int) -> bytes:
    buf = bytearray(piece)
    mv = memoryview(buf)
    parts = []
    while True:
        k = f.readinto(mv)
        if not k:
            return b"".join(parts)
        parts.append(bytes(mv[:k]))


@pytest.mark.parametrize("writer,threads", [("oracle", 1), ("oracle", 3), ("native", 4)])
def test_bgzf_fastq_reads_back(tmp_path, writer, threads):
    data = fastq_text(7, 9000)  # about 3 MB: 45+ members, several refills of a 128 KiB chunk
    z = orc.bgzf_stream(data, 1) if writer == "oracle" else bam_emit.bgzf_compress(data, 1, 4)
    p = tmp_path / "r.fq.gz"
    p.write_bytes(z)
    assert gzip.decompress(z) == data  # the members are well-formed gzip
    assert lf._is_bgzf(str(p))
    f = lf._BgzfFile(str(p), chunk=1 << 17, threads=threads)
    assert read_all(f, 100_003) == data
    f.close()
    s = lf._open_stream(str(p))
    assert isinstance(s, lf._BgzfFile) and read_all(s, 1 << 20) == data
    s.close()


def test_plain_gzip_reads_back_natively(tmp_path):
    """Single- and multi-member gzip (not BGZF) through zlib's gz reader; a
    flipped CRC32, a cut file and a file that is not gzip raise."""
    data = fastq_text(8, 3000)
    z = gzip.compress(data[:400_000]) + gzip.compress(data[400_000:])  # two members
    p = tmp_path / "r.fq.gz"
    p.write_bytes(z)
    assert not lf._is_bgzf(str(p))
    s = lf._open_stream(str(p))
    assert isinstance(s, lf._GzFile) and read_all(s, 65_537) == data
    s.close()
    one = gzip.compress(data)
    bad = bytearray(one)
    bad[-6] ^= 0xFF  # CRC32
    p.write_bytes(bytes(bad))
    with pytest.raises(MialignError, match="ma_gz_read"):
        read_all(lf._GzFile(str(p)), 1 << 16)
    p.write_bytes(one[: len(one) // 2])
    with pytest.raises(MialignError, match="ma_gz_read"):
        read_all(lf._GzFile(str(p)), 1 << 16)
    p.write_bytes(data[:5000])
    with pytest.raises(MialignError, match="not a gzip file"):
        read_all(lf._GzFile(str(p)), 1 << 16)


def test_bgzf_errors(tmp_path):
    data = fastq_text(9, 2000)
    z = bytearray(orc.bgzf_stream(data, 1, eof=False))
    first = (z[16] | z[17] << 8) + 1  # BSIZE + 1 = the first member's bytes
    # a flipped CRC32 byte in the second member
    second = first + (z[first + 16] | z[first + 17] << 8) + 1
    bad = bytearray(z)
    bad[second - 8] ^= 0xFF
    p = tmp_path / "crc.fq.gz"
    p.write_bytes(bytes(bad))
    with pytest.raises(MialignError, match="corrupt BGZF member"):
        read_all(lf._BgzfFile(str(p), chunk=1 << 17), 1 << 16)
    # an ISIZE beyond a BGZF member's 64 KiB, rejected before any buffer is sized by it
    bad = bytearray(z)
    bad[second - 1] = 0x7F
    p = tmp_path / "isize.fq.gz"
    p.write_bytes(bytes(bad))
    with pytest.raises(MialignError, match="ISIZE"):
        read_all(lf._BgzfFile(str(p), chunk=1 << 17), 1 << 16)
    # the file cut inside a member
    p = tmp_path / "cut.fq.gz"
    p.write_bytes(bytes(z[: second + 100]))
    with pytest.raises(MialignError, match="truncated BGZF member"):
        read_all(lf._BgzfFile(str(p), chunk=1 << 17), 1 << 16)
    # a plain gzip member after the BGZF ones
    p = tmp_path / "mixed.fq.gz"
    p.write_bytes(bytes(z) + gzip.compress(b"@x\nA\n+\nF\n"))
    with pytest.raises(MialignError, match="not a BGZF member"):
        read_all(lf._BgzfFile(str(p), chunk=1 << 17), 1 << 16)


def test_bgzf_inflate_sizes_only_pass():
    """dst null: the whole members at the front and their inflated size; a
    partial member at the end is left for the next call."""
    data = fastq_text(10, 1500)
    z = orc.bgzf_stream(data, 1, eof=False)
    import ctypes as C

    from mialign._lib import lib

    used, outn = C.c_int64(), C.c_int64()
    a = np.frombuffer(z, np.uint8)
    assert lib().ma_bgzf_inflate(a.ctypes.data, a.size, 2, None, 0, C.byref(used), C.byref(outn)) == 0
    assert used.value == a.size and outn.value == len(data)
    assert lib().ma_bgzf_inflate(a.ctypes.data, a.size - 5, 2, None, 0, C.byref(used), C.byref(outn)) == 0
    last = a.size - 5
    assert 0 < used.value < last and outn.value < len(data) and outn.value % 65280 == 0
