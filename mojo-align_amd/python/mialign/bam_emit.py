"""BAM emit: records packed on the device in coordinate order with the
duplicate flag, BGZF members compressed by native worker threads.

Mirrors the reference's BAM writer module (fq2bam-meth/python/bam_emit.py:
BamWriter header :318-356, write_raw :477-481, write_bytes :483-487, close;
_reg2bin :34-48 and _pack_seq :52-61 as small host helpers) and the fm
driver's native-BAM path (fq2bam-meth/src/linear_gpu_fm.mojo: _open_bam_writer
:91-125, the SE branch of _emit_bam_from_batch :755-823, _pack_bam_aln
:497-665, _gather_perm_to_writer :177-225). The records themselves are built by
ma_bam_pack_se_dev / ma_bam_pack_dev (one wave per record, output order = the
sort's permutation, 0x400 applied in the same pass); the host only frames and
writes BGZF (ma_bgzf_*).
"""

from __future__ import annotations

import ctypes as C
import os
import struct
from typing import Optional, Sequence

import numpy as np

from ._lib import BAM_FIELDS, BamFields, BamReads, MialignError, check, lib, ptr
from .core import DeviceArray, DeviceRecs, Stream

BAM_MAGIC = b"BAM\x01"
BGZF_BLOCK = 65280
RG_MAX = 240
_SEQ_NT16 = [15] * 256
for _i, _c in enumerate(b"=ACMGRSVTWYHKDBN"):
    _SEQ_NT16[_c] = _i
    _SEQ_NT16[_c + 32] = _i


def _reg2bin(beg: int, end: int) -> int:
    """SAM-specification bin of [beg, end) (bam_emit.py:34-48)."""
    if end <= beg:
        end = beg + 1
    end -= 1
    for shift, first in ((14, 4681), (17, 585), (20, 73), (23, 9), (26, 1)):
        if beg >> shift == end >> shift:
            return first + (beg >> shift)
    return 0


def _pack_seq(seq: bytes) -> bytes:
    """4-bit BAM SEQ of an ASCII sequence (bam_emit.py:52-61)."""
    out = bytearray((len(seq) + 1) // 2)
    for i, b in enumerate(seq):
        out[i >> 1] |= _SEQ_NT16[b] << (4 * (1 - (i & 1)))
    return bytes(out)


def bam_threads() -> int:
    """METHYLGRAPHER_BAM_THREADS (default 32, at least 1), bam_emit.py:335-340."""
    try:
        t = int(os.environ.get("METHYLGRAPHER_BAM_THREADS", "32"))
    except ValueError:
        t = 32
    return max(1, t)


def bam_header(sq_names: Sequence[str], sq_lens: Sequence[int], rg_id: str = "mojo1", sm: str = "sample",
               lb: str = "lib1", pl: str = "ILLUMINA", pg_vn: str = "0.1.0-mojo-fm",
               sort_order: str = "unsorted") -> bytes:
    """BAM magic, SAM header text (@HD/@SQ/@RG/@PG) and the reference list
    (bam_emit.py:318-356)."""
    if len(sq_names) != len(sq_lens):
        raise ValueError("sq_names/sq_lens length mismatch")
    so = sort_order if sort_order in {"unsorted", "coordinate", "queryname", "unknown"} else "unsorted"
    hd = [f"@HD\tVN:1.6\tSO:{so}"]
    hd += [f"@SQ\tSN:{n}\tLN:{ln}" for n, ln in zip(sq_names, sq_lens)]
    hd.append(f"@RG\tID:{rg_id}\tSM:{sm}\tLB:{lb}\tPL:{pl}\tPU:{rg_id}")
    hd.append(f"@PG\tID:MojoFq2bamMeth\tPN:MojoFq2bamMeth\tVN:{pg_vn}")
    text = ("\n".join(hd) + "\n").encode("ascii")
    blob = bytearray(BAM_MAGIC) + struct.pack("<i", len(text)) + text + struct.pack("<i", len(sq_names))
    for n, ln in zip(sq_names, sq_lens):
        nb = n.encode("ascii") + b"\x00"
        blob += struct.pack("<i", len(nb)) + nb + struct.pack("<i", int(ln))
    return bytes(blob)


class BamWriter:
    """Stream BAM to `path` (BGZF): the header at open, raw record bytes from
    host memory after it, the EOF member at close (bam_emit.py:309-489)."""

    def __init__(self, path: str, sq_names: Sequence[str], sq_lens: Sequence[int], rg_id: str = "mojo1",
                 sm: str = "sample", lb: str = "lib1", pl: str = "ILLUMINA", pg_vn: str = "0.1.0-mojo-fm",
                 level: int = 1, sort_order: str = "unsorted", threads: Optional[int] = None) -> None:
        head = bam_header(sq_names, sq_lens, rg_id, sm, lb, pl, pg_vn, sort_order)
        self.rg = rg_id.encode("ascii")
        h = C.c_void_p()
        check(lib().ma_bgzf_open(os.fsencode(path), max(0, min(int(level), 9)),
                                 bam_threads() if threads is None else max(1, int(threads)), C.byref(h)),
              "ma_bgzf_open")
        self._h = h.value
        self.write_bytes(head)

    def write_raw(self, addr: int, n: int) -> None:
        """Append n bytes at host address addr."""
        if n > 0:
            check(lib().ma_bgzf_write(self._h, addr, int(n)), "ma_bgzf_write")

    def write_bytes(self, data) -> None:
        if data:
            a = np.frombuffer(data, dtype=np.uint8)
            self.write_raw(a.ctypes.data, a.size)

    def close(self) -> int:
        """Flush and close; returns the file's size in bytes."""
        if not self._h:
            return 0
        out = C.c_int64()
        h, self._h = self._h, None
        check(lib().ma_bgzf_close(h, C.byref(out)), "ma_bgzf_close")
        return out.value

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self.close()
        except Exception:
            pass


def bgzf_compress(data, level: int = 1, threads: Optional[int] = None, eof: bool = True) -> bytes:
    """BGZF bytes of a host buffer (65280-byte members, optional EOF member)."""
    a = np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    a = np.ascontiguousarray(a, dtype=np.uint8)
    out = np.empty(int(lib().ma_bgzf_bound(a.size)), np.uint8)
    n = C.c_int64()
    check(lib().ma_bgzf_compress(ptr(a) if a.size else None, a.size, int(level),
                                 bam_threads() if threads is None else int(threads), int(bool(eof)), ptr(out),
                                 out.size, C.byref(n)), "ma_bgzf_compress")
    return out[: n.value].tobytes()


class DeviceBgzf:
    """BGZF of device streams of up to `max_n` bytes, compressed on the GPU
    (ma_bgzf_compress_dev); the members land in a device buffer."""

    def __init__(self, max_n: int, device: int = 0):
        self.max_n = int(max_n)
        self.cap = int(lib().ma_bgzf_bound(self.max_n))
        self.d_dst = DeviceArray((max(self.cap, 4),), np.uint8, device)
        sb = int(lib().ma_bgzf_dev_scratch_bytes(self.max_n))
        self.d_scratch = DeviceArray((sb,), np.uint8, device)
        self.scratch_bytes = sb
        self.d_total = DeviceArray((1,), np.uint64, device)

    def compress(self, stream, d_src_ptr: int, n: int, level: int = 1, eof: bool = True) -> None:
        """Enqueue the compression of n bytes at device address d_src_ptr."""
        if n > self.max_n:
            raise MialignError("DeviceBgzf: stream longer than its capacity")
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_bgzf_compress_dev(h, d_src_ptr, int(n), int(level), int(bool(eof)), self.d_dst.ptr,
                                         self.cap, self.d_total.ptr, self.d_scratch.ptr, self.scratch_bytes),
              "ma_bgzf_compress_dev")

    def total(self, stream=None) -> int:
        t = np.zeros(1, np.uint64)
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_memcpy_d2h(ptr(t), self.d_total.ptr, 8, h), "ma_memcpy_d2h")
        if h is not None:
            check(lib().ma_stream_sync(h), "ma_stream_sync")
        return int(t[0])

    def download(self, stream=None) -> bytes:
        total = self.total(stream)
        out = np.empty(max(total, 1), np.uint8)
        h = stream.handle if isinstance(stream, Stream) else stream
        if total:
            check(lib().ma_memcpy_d2h(ptr(out), self.d_dst.ptr, total, h), "ma_memcpy_d2h")
            if h is not None:
                check(lib().ma_stream_sync(h), "ma_stream_sync")
        return out[:total].tobytes()


class DeviceReads:
    """A batch's names, original bases and qualities resident on the device,
    in the layout ma_bam_reads describes (rows of `stride` bytes)."""

    def __init__(self, names_blob: np.ndarray, name_off: np.ndarray, seq_rows: np.ndarray, lens: np.ndarray,
                 qual_rows: Optional[np.ndarray] = None, qual_lens: Optional[np.ndarray] = None,
                 device: int = 0, stream=None):
        n = int(lens.size)
        name_off = np.ascontiguousarray(name_off, dtype=np.int64)
        if name_off.size != n + 1:
            raise MialignError("name_off must hold n + 1 offsets")
        blob = np.ascontiguousarray(names_blob, dtype=np.uint8)
        if blob.size == 0:
            blob = np.zeros(1, np.uint8)
        seq_rows = np.ascontiguousarray(seq_rows, dtype=np.uint8)
        self.n = n
        self.stride = int(seq_rows.shape[1]) if seq_rows.ndim == 2 else 0
        lens = np.asarray(lens)
        if seq_rows.ndim != 2 or seq_rows.shape[0] != n or (n and int(lens.max()) > self.stride):
            raise MialignError("base rows must be [n x stride] with every length <= stride")
        if name_off[0] != 0 or np.any(np.diff(name_off) < 0) or int(name_off[-1]) > blob.size:
            raise MialignError("name offsets must rise from 0 and stay inside the names")
        self.d_names = DeviceArray.from_host(blob, device, stream)
        self.d_name_off = DeviceArray.from_host(name_off, device, stream)
        self.d_seq = DeviceArray.from_host(seq_rows, device, stream)
        self.d_lens = DeviceArray.from_host(np.ascontiguousarray(lens, dtype=np.uint32), device, stream)
        self.d_qual = None
        self.d_qual_lens = None
        if qual_rows is not None:
            qual_rows = np.ascontiguousarray(qual_rows, dtype=np.uint8)
            if qual_rows.shape != seq_rows.shape:
                raise MialignError("quality rows must have the shape of the base rows")
            self.d_qual = DeviceArray.from_host(qual_rows, device, stream)
            if qual_lens is not None:
                self.d_qual_lens = DeviceArray.from_host(np.ascontiguousarray(qual_lens, dtype=np.uint32),
                                                         device, stream)
        self.name_bytes = int(name_off[-1])
        self.struct = BamReads(self.d_names.ptr, self.d_name_off.ptr, self.d_seq.ptr,
                               self.d_qual.ptr if self.d_qual else None, self.d_lens.ptr,
                               self.d_qual_lens.ptr if self.d_qual_lens else None, self.stride)

    @classmethod
    def on_device(cls, names: DeviceArray, name_off: DeviceArray, seq: DeviceArray, lens: DeviceArray, n: int,
                  stride: int, name_bytes: int, qual: Optional[DeviceArray] = None,
                  qual_lens: Optional[DeviceArray] = None) -> "DeviceReads":
        """Wrap reads already in HBM (the layout ma_bam_reads describes)."""
        self = cls.__new__(cls)
        self.n, self.stride, self.name_bytes = int(n), int(stride), int(name_bytes)
        self.d_names, self.d_name_off, self.d_seq, self.d_lens = names, name_off, seq, lens
        self.d_qual, self.d_qual_lens = qual, qual_lens
        self.struct = BamReads(names.ptr, name_off.ptr, seq.ptr, qual.ptr if qual is not None else None, lens.ptr,
                               qual_lens.ptr if qual_lens is not None else None, self.stride)
        return self

    @classmethod
    def from_names(cls, names: Sequence[bytes], seq_rows, lens, qual_rows=None, qual_lens=None, device: int = 0,
                   stream=None) -> "DeviceReads":
        blob, off = names_blob(names)
        return cls(blob, off, seq_rows, lens, qual_rows, qual_lens, device, stream)

    def out_bound(self, rg_n: int) -> int:
        """Bytes the batch's records take at most (every read at the row width)."""
        return self.name_bytes + self.n * (36 + 1 + 12 + 3 + 1 + 4 + rg_n + self.stride + (self.stride + 1) // 2)


def names_blob(names: Sequence[bytes]):
    """Names back to back and their offsets [n+1] (the ma_bam_reads layout)."""
    off = np.zeros(len(names) + 1, dtype=np.int64)
    if names:
        off[1:] = np.cumsum([len(x) for x in names])
    return np.frombuffer(b"".join(names) or b"\0", dtype=np.uint8), off


class BamPacker:
    """Device buffers for packing batches of up to `cap_reads` records."""

    def __init__(self, cap_reads: int, out_cap: int, device: int = 0):
        self.cap_reads = int(cap_reads)
        self.out_cap = int(out_cap)
        self.d_out = DeviceArray((max(self.out_cap, 4),), np.uint8, device)
        self.d_rec_off = DeviceArray((self.cap_reads + 1,), np.uint64, device)
        sb = int(lib().ma_bam_scratch_bytes(self.cap_reads))
        self.d_scratch = DeviceArray((sb,), np.uint8, device)
        self.scratch_bytes = sb

    def pack_se(self, stream, n: int, recs: DeviceRecs, d_tid_of_cid: DeviceArray, n_tid: int,
                reads: DeviceReads, rg: bytes, d_perm: Optional[DeviceArray] = None,
                d_is_dup: Optional[DeviceArray] = None, apply_dup: bool = False) -> None:
        """Enqueue the pack of n SE records (ma_bam_pack_se_dev)."""
        if n > self.cap_reads:
            raise MialignError("BamPacker: batch larger than its capacity")
        if len(rg) > RG_MAX:
            raise MialignError(f"read-group id longer than {RG_MAX} bytes")
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_bam_pack_se_dev(h, n, C.byref(recs.struct), d_tid_of_cid.ptr, int(n_tid),
                                       C.byref(reads.struct), rg, len(rg), d_perm.ptr if d_perm else None,
                                       d_is_dup.ptr if d_is_dup else None, int(bool(apply_dup)), self.d_out.ptr,
                                       self.out_cap, self.d_rec_off.ptr, self.d_scratch.ptr, self.scratch_bytes),
              "ma_bam_pack_se_dev")

    def pack_pe(self, stream, n_pairs: int, recs, d_tid_of_cid: DeviceArray, n_tid: int, reads: DeviceReads,
                rg: bytes, d_perm: Optional[DeviceArray] = None, d_is_dup: Optional[DeviceArray] = None,
                apply_dup: bool = False) -> None:
        """Enqueue the pack of n_pairs pairs, 2 n_pairs records (ma_bam_pack_pe_dev):
        records and reads R1 block then R2 block, records out R1, R2 per pair."""
        if 2 * n_pairs > self.cap_reads:
            raise MialignError("BamPacker: batch larger than its capacity")
        if len(rg) > RG_MAX:
            raise MialignError(f"read-group id longer than {RG_MAX} bytes")
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_bam_pack_pe_dev(h, n_pairs, C.byref(recs.struct), d_tid_of_cid.ptr, int(n_tid),
                                       C.byref(reads.struct), rg, len(rg), d_perm.ptr if d_perm else None,
                                       d_is_dup.ptr if d_is_dup else None, int(bool(apply_dup)), self.d_out.ptr,
                                       self.out_cap, self.d_rec_off.ptr, self.d_scratch.ptr, self.scratch_bytes),
              "ma_bam_pack_pe_dev")

    def pack_fields(self, stream, n: int, fields: dict, reads: DeviceReads, rg: bytes,
                    d_perm: Optional[DeviceArray] = None, d_is_dup: Optional[DeviceArray] = None,
                    apply_dup: bool = False) -> None:
        """Enqueue the pack of n records from explicit device int32 field
        arrays (ma_bam_pack_dev); fields maps names of BAM_FIELDS to DeviceArray."""
        if n > self.cap_reads:
            raise MialignError("BamPacker: batch larger than its capacity")
        st = BamFields(*[fields[f].ptr if fields.get(f) is not None else None for f in BAM_FIELDS])
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_bam_pack_dev(h, n, C.byref(st), C.byref(reads.struct), rg, len(rg),
                                    d_perm.ptr if d_perm else None, d_is_dup.ptr if d_is_dup else None,
                                    int(bool(apply_dup)), self.d_out.ptr, self.out_cap, self.d_rec_off.ptr,
                                    self.d_scratch.ptr, self.scratch_bytes), "ma_bam_pack_dev")

    def total(self, n: int, stream=None) -> int:
        """Stream length of the last pack (synchronises `stream`); raises the
        reference's overflow error when it did not fit."""
        t = np.zeros(1, np.uint64)
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_memcpy_d2h(ptr(t), self.d_rec_off.ptr + 8 * n, 8, h), "ma_memcpy_d2h")
        if h is not None:
            check(lib().ma_stream_sync(h), "ma_stream_sync")
        total = int(t[0])
        if total > self.out_cap:
            raise MialignError("BAM batch buffer overflow")
        return total

    def download(self, n: int, stream=None) -> bytes:
        total = self.total(n, stream)
        out = np.empty(max(total, 1), np.uint8)
        if total:
            h = stream.handle if isinstance(stream, Stream) else stream
            check(lib().ma_memcpy_d2h(ptr(out), self.d_out.ptr, total, h), "ma_memcpy_d2h")
            if h is not None:
                check(lib().ma_stream_sync(h), "ma_stream_sync")
        return out[:total].tobytes()


def tid_table(contig_names: Sequence[str]):
    """Canonical (f/r-stripped) @SQ names in first-seen order and the contig ->
    target table (_rid_to_tid_table / _open_bam_writer, linear_gpu_fm.mojo:91-125, 423-435)."""
    from .linear_gpu_locate import canonical_contig

    seen: dict[str, int] = {}
    names: list[str] = []
    table = np.empty(len(contig_names), dtype=np.int32)
    first_idx: list[int] = []
    for ci, nm in enumerate(contig_names):
        c = canonical_contig(nm)
        if c not in seen:
            seen[c] = len(names)
            names.append(c)
            first_idx.append(ci)
        table[ci] = seen[c]
    return names, first_idx, table


__all__ = [
    "BamPacker",
    "DeviceBgzf",
    "BamWriter",
    "DeviceReads",
    "bam_header",
    "bgzf_compress",
    "names_blob",
    "tid_table",
]
