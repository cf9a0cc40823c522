"""FASTQ ingest on the device.

Mirrors the reference's bulk FASTQ path (fq2bam-meth/src/linear_fastq.mojo:
fq_open / fq_read_batch / fq_pack_bases / fq_close, :411-470; its Python twin
FastqPairReader, fq2bam-meth/python/fastq_batch.py:230-317): the host only
moves bytes (file chunks, gzip inflated on the host like the reference's pigz
pipe), and ma_fastq_parse_dev splits the lines, cuts the names, converts and
packs the rows in HBM, so a batch is ready for map_batch_dev and the BAM pack
without a host pass over the reads.

Semantics are the Mojo reader's: C2T maps C and c to 'T' and G2A maps G and g
to 'A' (upper case; the Python reader keeps case), rows are padded with 'N'.
"""

from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import MialignError, check, lib, ptr
from .core import DeviceArray, Stream

BS_CODES = {"": 0, "none": 0, "C2T": 1, "G2A": 2}
FQ_ERRORS = {
    1: "truncated FASTQ (missing SEQ)",
    2: "truncated FASTQ (missing +)",
    3: "truncated FASTQ (missing QUAL)",
}


@dataclass
class ParseInfo:
    records: int
    consumed: int
    max_len: int
    status: int
    name_bytes: int


class DeviceFastqBatch:
    """One parsed batch resident in HBM (layout of ma_fastq_parse_dev)."""

    def __init__(self, max_records: int, stride: int, names_cap: int, device: int = 0, with_conv: bool = True,
                 with_qual: bool = True):
        self.max_records, self.stride, self.names_cap = int(max_records), int(stride), int(names_cap)
        r = max(self.max_records, 1)
        self.rows = DeviceArray((r, self.stride), np.uint8, device)
        self.conv = DeviceArray((r, self.stride), np.uint8, device) if with_conv else None
        self.qual = DeviceArray((r, self.stride), np.uint8, device) if with_qual else None
        self.lens = DeviceArray((r,), np.uint32, device)
        self.qual_lens = DeviceArray((r,), np.uint32, device)
        self.names = DeviceArray((max(self.names_cap, 1),), np.uint8, device)
        self.name_off = DeviceArray((r + 1,), np.int64, device)
        self.info = DeviceArray((8,), np.uint64, device)
        self.n = 0
        self.max_len = 0

    def download(self) -> dict:
        """Host copies of the batch's first n records (for tests and tools)."""
        n = self.n
        out = {"rows": self.rows.download()[:n], "lens": self.lens.download()[:n],
               "qual_lens": self.qual_lens.download()[:n], "name_off": self.name_off.download()[:n + 1]}
        out["names"] = self.names.download()[: int(out["name_off"][-1]) if n else 0]
        if self.conv is not None:
            out["conv"] = self.conv.download()[:n]
        if self.qual is not None:
            out["qual"] = self.qual.download()[:n]
        return out


class FastqParser:
    """ma_fastq_parse_dev over device buffers of up to `max_bytes` bytes."""

    def __init__(self, max_bytes: int, max_records: int, device: int = 0):
        self.max_bytes, self.max_records = int(max_bytes), int(max_records)
        sb = int(lib().ma_fastq_scratch_bytes(self.max_bytes, self.max_records))
        self.scratch = DeviceArray((sb,), np.uint8, device)
        self.scratch_bytes = sb

    def parse(self, stream, d_buf_ptr: int, nbytes: int, final: bool, batch: DeviceFastqBatch, bs_mode: int,
              max_records: Optional[int] = None) -> ParseInfo:
        mr = min(batch.max_records, self.max_records if max_records is None else int(max_records))
        if nbytes > self.max_bytes:
            raise MialignError("FastqParser: buffer larger than its capacity")
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_fastq_parse_dev(h, d_buf_ptr if nbytes else None, int(nbytes), int(bool(final)), mr,
                                       batch.stride, int(bs_mode), batch.rows.ptr,
                                       batch.conv.ptr if batch.conv is not None else None,
                                       batch.qual.ptr if batch.qual is not None else None, batch.lens.ptr,
                                       batch.qual_lens.ptr, batch.names.ptr, batch.names_cap, batch.name_off.ptr,
                                       batch.info.ptr, self.scratch.ptr, self.scratch_bytes), "ma_fastq_parse_dev")
        info = np.zeros(8, np.uint64)
        check(lib().ma_memcpy_d2h(ptr(info), batch.info.ptr, 64, h), "ma_memcpy_d2h")
        if h is not None:
            check(lib().ma_stream_sync(h), "ma_stream_sync")
        pi = ParseInfo(int(info[0]), int(info[1]), int(info[2]), int(info[3]), int(info[4]))
        batch.n, batch.max_len = pi.records, pi.max_len
        return pi


def _is_bgzf(path: str) -> bool:
    """The file starts with a BGZF member (bgzip output): a gzip header with
    FEXTRA whose first extra subfield is BC with two bytes."""
    with open(path, "rb") as fh:
        h = fh.read(16)
    return len(h) == 16 and h[:4] == b"\x1f\x8b\x08\x04" and h[12:16] == b"BC\x02\x00"


class _BgzfFile:
    """A BGZF-compressed file read as its inflated bytes (readinto, like a
    file). The members carry their sizes, so a chunk's whole members are
    inflated at once by `threads` native workers (ma_bgzf_inflate), each
    member's ISIZE and CRC32 checked; other gzip files go through _GzFile on
    one thread."""

    def __init__(self, path: str, chunk: int = 32 << 20, threads: Optional[int] = None):
        self.path = path
        self.fh = open(path, "rb", buffering=0)
        self.src = np.empty(max(int(chunk), 1 << 17), np.uint8)  # >= two 64 KiB members
        self.have = 0  # compressed bytes at the front of src not inflated yet
        self.out = np.empty(0, np.uint8)
        self.pos = self.end = 0
        self.eof = False
        self.threads = int(threads) if threads else max(1, min(16, os.cpu_count() or 1))

    def readinto(self, mv) -> int:
        if self.pos == self.end:
            used, outn = C.c_int64(), C.c_int64()
            while True:  # the next whole members, inflated straight into mv when they fit
                if not self.eof and self.have < self.src.size:
                    k = self.fh.readinto(memoryview(self.src)[self.have:])
                    if k:
                        self.have += k
                    else:
                        self.eof = True
                if self.have == 0 and self.eof:
                    return 0
                check(lib().ma_bgzf_inflate(ptr(self.src), self.have, self.threads, None, 0, C.byref(used),
                                            C.byref(outn)), "ma_bgzf_inflate")
                if used.value == 0:
                    if self.eof:
                        raise MialignError(f"truncated BGZF member at the end of {self.path}")
                    continue
                n = outn.value
                direct = 0 < n <= len(mv)
                if direct:
                    dst = np.frombuffer(mv, np.uint8)
                else:
                    if self.out.size < n:
                        self.out = np.empty(n, np.uint8)
                    dst = self.out
                check(lib().ma_bgzf_inflate(ptr(self.src), used.value, self.threads, dst.ctypes.data if n else None,
                                            dst.size, C.byref(used), C.byref(outn)), "ma_bgzf_inflate")
                rest = self.have - used.value
                self.src[:rest] = self.src[used.value:self.have].copy()
                self.have = rest
                if direct:
                    return n
                if n:
                    self.pos, self.end = 0, n
                    break
        k = min(len(mv), self.end - self.pos)
        mv[:k] = memoryview(self.out)[self.pos:self.pos + k]
        self.pos += k
        return k

    def close(self):
        self.fh.close()


class _GzFile:
    """A gzip file of any framing read as its inflated bytes (readinto, like a
    file) by zlib's gz reader in native code (ma_gz_*): every member in turn,
    each CRC32 and ISIZE checked, on one thread."""

    def __init__(self, path: str):
        self.path = path
        self.h = C.c_void_p()
        check(lib().ma_gz_open(path.encode(), C.byref(self.h)), "ma_gz_open")

    def readinto(self, mv) -> int:
        got = C.c_int64()
        a = np.frombuffer(mv, np.uint8)
        check(lib().ma_gz_read(self.h, a.ctypes.data if a.size else None, a.size, C.byref(got)),
              f"ma_gz_read ({self.path})")
        return got.value

    def close(self):
        if self.h:
            lib().ma_gz_close(self.h)
            self.h = C.c_void_p()


def _open_stream(path: str):
    low = path.lower()
    if low.endswith(".gz") or low.endswith(".gzip"):
        return _BgzfFile(path) if _is_bgzf(path) else _GzFile(path)
    return open(path, "rb", buffering=0)


class _Side:
    """One FASTQ file. A reader thread reads the file's next chunk into a host
    buffer while the device parses (and the caller maps) the chunk before it;
    each chunk is copied into HBM after the unparsed tail of the last parse,
    which stays on the device (moved to the front of the other device buffer).
    Host buffers are plain pages: pinning buffers this size costs more than the
    copies it would speed up in a one-pass run."""

    def __init__(self, path: str, bs: str, chunk: int, device: int, stream: Stream):
        if bs not in BS_CODES:
            raise MialignError(f"unknown conversion mode {bs!r} (C2T, G2A or none)")
        self.fh = _open_stream(path)
        self.mode = BS_CODES[bs]
        self.chunk = int(chunk)
        self.device = device
        self.stream = stream
        self.filled = 0  # unparsed bytes at the front of dbuf
        self.eof = False
        self.host = [np.empty(self.chunk, np.uint8), np.empty(self.chunk, np.uint8)]
        self.dbufs = [DeviceArray((2 * self.chunk + 16,), np.uint8, device) for _ in range(2)]
        self.cur = 0
        self._slot = 0
        self._pool = ThreadPoolExecutor(max_workers=1)
        self._next = self._pool.submit(self._read, 0, self.chunk)

    @property
    def dbuf(self) -> DeviceArray:
        return self.dbufs[self.cur]

    def _read(self, slot: int, want: int):
        """Reader thread: up to `want` bytes of the file into host buffer `slot`;
        returns (count, the file ended)."""
        if self.host[slot].size < want:
            self.host[slot] = np.empty(want, np.uint8)
        mv = memoryview(self.host[slot])[:want]
        got = 0
        while got < want:
            k = self.fh.readinto(mv[got:])
            if not k:
                return got, True
            got += k
        return got, False

    def _grow(self, nbytes: int) -> None:
        """Both device buffers to nbytes, the buffered bytes kept."""
        dbufs = [DeviceArray((nbytes,), np.uint8, self.device) for _ in range(2)]
        if self.filled:
            check(lib().ma_memcpy_d2d(dbufs[0].ptr, self.dbuf.ptr, self.filled, self.stream.handle),
                  "ma_memcpy_d2d")
            self.stream.sync()
        self.dbufs, self.cur = dbufs, 0

    def fill(self) -> int:
        """Take read-ahead chunks until `chunk` bytes are buffered in HBM (or the
        file ends); returns the count."""
        while not self.eof and self.filled < self.chunk:
            got, self.eof = self._next.result()
            slot = self._slot
            if not self.eof:  # the next chunk is read while this one is copied, parsed and mapped
                self._slot ^= 1
                self._next = self._pool.submit(self._read, self._slot, self.chunk)
            if self.filled + got + 16 > self.dbuf.nbytes:
                self._grow(2 * (self.filled + got) + 16)
            if got:
                check(lib().ma_memcpy_h2d(self.dbuf.ptr + self.filled, self.host[slot].ctypes.data, got,
                                          self.stream.handle), "ma_memcpy_h2d")
                self.stream.sync()  # the reader thread refills this host buffer two chunks on
            self.filled += got
        return self.filled

    def consume(self, k: int) -> None:
        """Drop the first k bytes (the records just parsed): the rest moves to
        the front of the other device buffer."""
        rest = self.filled - k
        if rest > 0:
            check(lib().ma_memcpy_d2d(self.dbufs[self.cur ^ 1].ptr, self.dbuf.ptr + k, rest, self.stream.handle),
                  "ma_memcpy_d2d")
            self.cur ^= 1
        self.filled = max(rest, 0)

    def close(self):
        try:
            self._next.result()  # the reader thread is done with the file
        except Exception:
            pass
        self._pool.shutdown(wait=True)
        self.fh.close()


class DeviceFastqReader:
    """Streaming SE/PE FASTQ reader whose batches land in HBM, parsed on the
    device. read_batch() returns (batch R1, batch R2 or None), each with n
    records (R2 has as many as R1, else "paired FASTQ length mismatch (R2 ended
    early)" as in the reference)."""

    def __init__(self, path1: str, path2: str = "", bs_r1: str = "", bs_r2: str = "", batch_reads: int = 262144,
                 stride: int = 256, chunk_bytes: int = 0, device: int = 0, with_conv: bool = True):
        self.stride = int(stride)
        if chunk_bytes <= 0:  # a batch's worth of text at ~1.75 bytes per base per read
            chunk_bytes = max(1 << 20, int(batch_reads) * (2 * self.stride + 96))
        self.batch_reads = int(batch_reads)
        self.stream = Stream(device)
        self.sides = [_Side(path1, bs_r1, chunk_bytes, device, self.stream)]
        if path2:
            self.sides.append(_Side(path2, bs_r2, chunk_bytes, device, self.stream))
        self.parser = FastqParser(2 * chunk_bytes + 16, self.batch_reads, device)
        names_cap = 2 * chunk_bytes
        self.batches = [DeviceFastqBatch(self.batch_reads, self.stride, names_cap, device, with_conv=with_conv)
                        for _ in self.sides]

    def _parse_side(self, k: int, max_records: int) -> ParseInfo:
        """Parse up to max_records records of side k, reading more input until
        that many are complete or the file ends."""
        side, batch = self.sides[k], self.batches[k]
        while True:
            nbuf = side.fill()
            if nbuf > self.parser.max_bytes:
                self.parser = FastqParser(2 * nbuf, self.batch_reads, side.device)
            if nbuf > batch.names_cap:  # names never take more bytes than the buffer
                batch.names_cap = 2 * nbuf
                batch.names = DeviceArray((batch.names_cap,), np.uint8, side.device)
            info = self.parser.parse(self.stream, side.dbuf.ptr, nbuf, side.eof, batch, side.mode, max_records)
            if info.status in FQ_ERRORS:
                raise MialignError(FQ_ERRORS[info.status])
            if info.status == 4:
                raise MialignError(f"FASTQ read longer than {self.stride} bases")
            if info.status == 5:
                raise MialignError("FASTQ names larger than the batch's name buffer")
            if info.records >= max_records or side.eof:
                return info
            side.chunk *= 2  # fewer complete records than asked: read on

    def read_batch(self):
        """(R1 batch, R2 batch or None) of n records each, n = 0 at the end."""
        info1 = self._parse_side(0, self.batch_reads)
        n = info1.records
        if len(self.sides) == 2:
            if n > 0:
                info2 = self._parse_side(1, n)
                if info2.records < n:
                    raise MialignError("paired FASTQ length mismatch (R2 ended early)")
                self.sides[1].consume(info2.consumed)
            else:
                self.batches[1].n = 0
        self.sides[0].consume(info1.consumed)
        return self.batches[0], (self.batches[1] if len(self.sides) == 2 else None)

    def close(self):
        for s in self.sides:
            s.close()


def parse_fastq_bytes(data: bytes, stride: int, bs: str = "", final: bool = True, max_records: Optional[int] = None,
                      names_cap: Optional[int] = None, device: int = 0) -> tuple[ParseInfo, dict]:
    """Parse one in-memory FASTQ buffer on the device; returns (info, host copies)."""
    mr = max_records if max_records is not None else max(1, data.count(b"\n") // 4 + 1)
    nc = names_cap if names_cap is not None else max(1, len(data))
    st = Stream(device)
    d = DeviceArray((max(len(data), 1) + 16,), np.uint8, device)
    if data:
        check(lib().ma_memcpy_h2d(d.ptr, data, len(data), None), "ma_memcpy_h2d")
    batch = DeviceFastqBatch(mr, stride, nc, device)
    info = FastqParser(len(data), mr, device).parse(st, d.ptr, len(data), final, batch, BS_CODES[bs])
    return info, batch.download()


__all__ = ["DeviceFastqBatch", "DeviceFastqReader", "FastqParser", "ParseInfo", "parse_fastq_bytes"]
