"""FASTQ (SE or PE) -> coordinate-sorted, duplicate-marked BAM with every stage on the GPU.

The end-to-end product of the reference's native-BAM path
(fq2bam-meth/src/linear_gpu_fm.mojo: FASTQ arena -> GPU map -> _emit_bam_from_batch
keys -> gpu_sort_markdup -> _gather_perm_to_writer -> BamWriter, :2140-3290),
with this build's science (the hash-seeded parity mapper) and every step on
the device: ma_fastq_parse_dev, ma_map_batch_dev, ma_sort_keys_se_dev,
ma_sort_markdup_dev, ma_bam_pack_se_dev and ma_bgzf_compress_dev (or the host
zlib writer, whose members are the reference's byte for byte). The reads,
records and keys of the whole run stay in HBM (an MI355X holds ~700M reads of
150 bp this way). Pairs follow the PE branch of the same emit (:826-977):
ma_sort_keys_pe_dev and ma_bam_pack_pe_dev.
"""

from __future__ import annotations

import ctypes as C
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from ._lib import BS_MODES, MA_ENC_ASCII, REC_FIELDS, MialignError, Recs, check, default_params, lib
from .bam_emit import BamPacker, BamWriter, DeviceBgzf, DeviceReads, bam_header, bgzf_compress, tid_table
from .core import DeviceArray, DeviceRecs, Event, Stream, Workspace, map_batch_dev
from .linear_fastq import DeviceFastqReader
from .serve import pinned_empty


class _Grow:
    """A device buffer that doubles when an append would not fit."""

    def __init__(self, itemsize: int, cap: int, device: int):
        self.itemsize, self.device = itemsize, device
        self.cap = max(int(cap), 1)
        self.buf = DeviceArray((self.cap * itemsize,), np.uint8, device)
        self.n = 0

    def append_dev(self, src_ptr: int, count: int, stream) -> None:
        if self.n + count > self.cap:
            new_cap = max(self.cap * 2, self.n + count)
            nb = DeviceArray((new_cap * self.itemsize,), np.uint8, self.device)
            if self.n:
                check(lib().ma_memcpy_d2d(nb.ptr, self.buf.ptr, self.n * self.itemsize, stream.handle), "d2d")
            stream.sync()
            self.buf, self.cap = nb, new_cap
        if count:
            check(lib().ma_memcpy_d2d(self.buf.ptr + self.n * self.itemsize, src_ptr, count * self.itemsize,
                                      stream.handle), "d2d")
        self.n += count

    def view(self, dtype) -> DeviceArray:
        """A non-owning DeviceArray over the first n items."""
        v = DeviceArray.__new__(DeviceArray)
        v.shape, v.dtype = (self.n * self.itemsize // np.dtype(dtype).itemsize,), np.dtype(dtype)
        v.nbytes, v.device, v.ptr = self.n * self.itemsize, self.device, self.buf.ptr
        v.free = lambda: None
        return v


class _Acc:
    """The reads and merged records of one FASTQ side, appended batch by batch in HBM."""

    FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")

    def __init__(self, batch_reads: int, stride: int, device: int):
        self.rows, self.quals = _Grow(stride, batch_reads, device), _Grow(stride, batch_reads, device)
        self.lens, self.qlens = _Grow(4, batch_reads, device), _Grow(4, batch_reads, device)
        self.names = _Grow(1, 32 * batch_reads, device)
        self.fields = {f: _Grow(4, batch_reads, device) for f in self.FIELDS}
        self.name_off = [np.zeros(1, np.int64)]

    def append(self, b, recs_b: DeviceRecs, st) -> None:
        n = b.n
        base = self.names.n
        self.rows.append_dev(b.rows.ptr, n, st)
        self.quals.append_dev(b.qual.ptr, n, st)
        self.lens.append_dev(b.lens.ptr, n, st)
        self.qlens.append_dev(b.qual_lens.ptr, n, st)
        off = b.name_off.download()[: n + 1]
        self.names.append_dev(b.names.ptr, int(off[n]), st)
        self.name_off.append(off[1:] + base)
        for f, g in self.fields.items():
            g.append_dev(recs_b.field_ptr(f), n, st)

    def extend(self, other: "_Acc", st) -> None:
        """Append another side's reads and records after these (R2 after R1)."""
        base = self.names.n
        for a, b in ((self.rows, other.rows), (self.quals, other.quals), (self.lens, other.lens),
                     (self.qlens, other.qlens), (self.names, other.names)):
            a.append_dev(b.buf.ptr, b.n, st)
        for f in self.FIELDS:
            self.fields[f].append_dev(other.fields[f].buf.ptr, other.fields[f].n, st)
        self.name_off.append(np.concatenate(other.name_off)[1:] + base)


def fastq_to_bam(index, fq1: str, out_bam: str, bs_r1: str = "C2T", *, fq2: str = "", bs_r2: str = "G2A",
                 batch_reads: int = 1 << 20, stride: int = 256, do_markdup: bool = True, device_bgzf: bool = True,
                 level: int = 1, rg_id: str = "mojo1", sm: str = "sample", lb: str = "lib1", pl: str = "ILLUMINA",
                 pg_vn: str = "0.1.0-mojo-fm", device: int = 0, params=None, rescue_window: int = 512) -> dict:
    """FASTQ (single-end, or paired with fq2) -> BAM in coordinate order, 0x400
    on duplicates. `index` is a LinearIndex with its GPU index loaded. Pairs are
    mapped R1 with bs_r1 and R2 with bs_r2, emitted R1, R2 per pair with the PE
    flags and mate fields, and marked as pairs. `params` are the mapping knobs
    (default: the reference defaults for the index's k); the BAM records carry
    slS mM srS CIGARs only, as the reference's native-BAM records do, so
    params.max_indel must be 0. Pairs with one mate unmapped get the native
    BAM driver's mate rescue within +-rescue_window bases of the mapped mate
    (0 = off; ma_mate_rescue_dev). Returns counts and stage times."""
    paired = bool(fq2)
    gix = index.gpu
    prm = default_params(k=gix.k) if params is None else params
    if prm.k != gix.k:
        raise MialignError(f"params.k {prm.k} differs from the index k {gix.k}")
    if prm.max_indel != 0:
        raise MialignError("BAM output carries gapless records only (slS mM srS): params.max_indel must be 0")
    for m in (bs_r1, bs_r2) if paired else (bs_r1,):
        if m not in BS_MODES:
            raise MialignError(f"unknown conversion mode {m!r} (C2T, G2A or none)")
    if stride > 256:
        raise MialignError("reads longer than 256 bases are not supported")
    st = Stream(device)
    t0 = time.perf_counter()
    reader = DeviceFastqReader(fq1, fq2, bs_r1, bs_r2 if paired else "", batch_reads=batch_reads, stride=stride,
                               device=device, with_conv=False)
    ws = Workspace(gix, batch_reads, stride)
    sides = [_Acc(batch_reads, stride, device) for _ in range(2 if paired else 1)]
    recs_b = [DeviceRecs(batch_reads, device) for _ in sides]
    modes = [BS_MODES[bs_r1]] + ([BS_MODES[bs_r2]] if paired else [])
    t_map = 0.0
    ev0, ev1 = Event(), Event()
    try:
        while True:
            batches = reader.read_batch()
            n = batches[0].n
            if n == 0:
                break
            ev0.record(st)
            for k in range(len(sides)):
                b = batches[k]
                max_len = min(stride, max(b.max_len, gix.k))
                map_batch_dev(gix, ws, st, b.rows, stride, b.lens, n, max_len, recs_b[k], encoding=MA_ENC_ASCII,
                              bs_r1=modes[k], params=prm)
            ev1.record(st)
            for k in range(len(sides)):
                sides[k].append(batches[k], recs_b[k], st)
            st.sync()  # the reader reuses its batch buffers
            t_map += ev0.elapsed_ms(ev1)
    finally:
        reader.close()
    acc = sides[0]
    if paired:
        acc.extend(sides[1], st)
        st.sync()
        del sides[1:]
    N = acc.lens.n
    t_ingest = time.perf_counter() - t0
    if N == 0:
        raise MialignError("no reads in " + fq1)
    # sort keys, sort + markdup over the whole run
    t1 = time.perf_counter()
    sq_names, first, tid = tid_table(index.contig_names)
    sq_lens = [index.contig_lengths[i] for i in first]
    d_tid = DeviceArray.from_host(tid, device)
    recs = Recs()
    for f in REC_FIELDS:
        setattr(recs, f, acc.fields[f].buf.ptr if f in acc.fields else None)
    rescued = 0
    if paired and rescue_window > 0:
        before = int((acc.fields["cid"].view(np.int32).download() >= 0).sum())
        check(lib().ma_mate_rescue_dev(gix.handle, st.handle, acc.rows.buf.ptr, stride, acc.lens.buf.ptr, N // 2,
                                       stride, MA_ENC_ASCII, modes[0], modes[1], int(rescue_window), prm.max_diff,
                                       C.byref(recs)), "ma_mate_rescue_dev")
        rescued = int((acc.fields["cid"].view(np.int32).download(st.handle) >= 0).sum()) - before
    keys = {f: DeviceArray((N,), dt, device) for f, dt in (("coord", np.uint64), ("dhi", np.uint64),
                                                          ("dlo", np.uint64), ("score", np.uint32),
                                                          ("pair", np.uint32))}
    kargs = [keys[f].ptr for f in ("coord", "dhi", "dlo", "score", "pair")]
    if paired:
        check(lib().ma_sort_keys_pe_dev(st.handle, N // 2, C.byref(recs), acc.lens.buf.ptr, d_tid.ptr, tid.size, 0,
                                        *kargs), "ma_sort_keys_pe_dev")
    else:
        check(lib().ma_sort_keys_se_dev(st.handle, N, C.byref(recs), acc.lens.buf.ptr, d_tid.ptr, tid.size, 0,
                                        *kargs),
              "ma_sort_keys_se_dev")
    perm = DeviceArray((N,), np.uint32, device)
    is_dup = DeviceArray((N,), np.uint32, device)
    sb = int(lib().ma_sort_scratch_bytes(N))
    scratch = DeviceArray((sb,), np.uint8, device)
    check(lib().ma_sort_markdup_dev(st.handle, N, keys["coord"].ptr, keys["dhi"].ptr, keys["dlo"].ptr,
                                    keys["score"].ptr, keys["pair"].ptr, is_dup.ptr, perm.ptr, int(do_markdup), None,
                                    scratch.ptr, sb), "ma_sort_markdup_dev")
    del keys, scratch
    # BAM records in coordinate order, then BGZF
    offs = np.concatenate(acc.name_off)
    d_off = DeviceArray.from_host(offs, device)
    reads = DeviceReads.on_device(acc.names.view(np.uint8), d_off, acc.rows.view(np.uint8),
                                  acc.lens.view(np.uint32), N, stride, int(offs[-1]), acc.quals.view(np.uint8),
                                  acc.qlens.view(np.uint32))
    rg = rg_id.encode("ascii")
    packer = BamPacker(N, reads.out_bound(len(rg)), device)
    if paired:
        packer.pack_pe(st, N // 2, _recs_view(recs, N), d_tid, tid.size, reads, rg, perm, is_dup, bool(do_markdup))
    else:
        packer.pack_se(st, N, _recs_view(recs, N), d_tid, tid.size, reads, rg, perm, is_dup, bool(do_markdup))
    total = packer.total(N, st)
    t_sort_pack = time.perf_counter() - t1
    t2 = time.perf_counter()
    head = bam_header(sq_names, sq_lens, rg_id, sm, lb, pl, pg_vn, "coordinate")
    if device_bgzf:
        dz = DeviceBgzf(total, device)
        dz.compress(st, packer.d_out.ptr, total, level, True)
        zt = dz.total(st)
        with open(out_bam, "wb") as fh:
            fh.write(bgzf_compress(head, level, eof=False))
            fh.flush()
            _write_from_device(fh.fileno(), dz.d_dst.ptr, zt, fh.tell(), st)
        size = zt
    else:
        stream_bytes = packer.download(N, st)
        w = BamWriter(out_bam, sq_names, sq_lens, rg_id, sm, lb, pl, pg_vn, level, "coordinate")
        w.write_bytes(stream_bytes)
        size = w.close()
    t_bgzf = time.perf_counter() - t2
    dups = int(is_dup.download().sum()) if do_markdup else 0
    mapped = int((acc.fields["cid"].view(np.int32).download() >= 0).sum())
    return {"reads": N, "pairs": N // 2 if paired else 0, "mapped": mapped, "rescued": rescued, "duplicates": dups,
            "bam_stream_bytes": total, "bgzf_bytes": size, "ingest_and_map_s": round(t_ingest, 3),
            "map_kernel_ms": round(t_map, 3), "sort_and_pack_s": round(t_sort_pack, 3), "bgzf_s": round(t_bgzf, 3)}


def _pwrite_all(fd: int, buf: np.ndarray, n: int, off: int) -> None:
    mv = memoryview(buf)[:n]
    while len(mv):
        k = os.pwrite(fd, mv, off)
        mv, off = mv[k:], off + k


def _write_from_device(fd: int, d_src: int, nbytes: int, file_off: int, st: Stream, piece: int = 32 << 20,
                       ring: int = 3) -> None:
    """nbytes of device memory -> file fd at file_off: pieces copied into a ring
    of pinned buffers, each written out by a writer thread while the next
    pieces are copied."""
    bufs = [pinned_empty((min(piece, max(nbytes, 1)),), np.uint8) for _ in range(min(ring, max(1, -(-nbytes // piece))))]
    futs = [None] * len(bufs)
    with ThreadPoolExecutor(max_workers=len(bufs)) as pool:
        for i, o in enumerate(range(0, nbytes, piece)):
            k = i % len(bufs)
            if futs[k] is not None:
                futs[k].result()  # the buffer's last piece is written
            n = min(piece, nbytes - o)
            check(lib().ma_memcpy_d2h(bufs[k].ctypes.data, d_src + o, n, st.handle), "ma_memcpy_d2h")
            st.sync()
            futs[k] = pool.submit(_pwrite_all, fd, bufs[k], n, file_off + o)
        for f in futs:
            if f is not None:
                f.result()


class _recs_view:
    """The accumulated records as the DeviceRecs interface BamPacker takes."""

    def __init__(self, recs: Recs, n: int):
        self.struct, self.n = recs, n


__all__ = ["fastq_to_bam"]
