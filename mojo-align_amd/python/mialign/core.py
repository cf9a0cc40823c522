"""Python host layer over the C ABI: device memory, the resident index, batch
mapping, sort keys and sort + duplicate marking."""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._lib import (
    BuildOpts,
    MA_ENC_2BIT,
    MA_ENC_ASCII,
    REC_DTYPES,
    REC_FIELDS,
    MialignError,
    Params,
    Recs,
    check,
    default_params,
    host_recs,
    lib,
    ptr,
    recs_struct,
)


# ------------------------------------------------------------------ memory
class DeviceArray:
    """Owning device allocation with a numpy dtype/shape view description."""

    def __init__(self, shape, dtype, device: int = 0):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.device = device
        p = C.c_void_p()
        check(lib().ma_dev_alloc(device, max(self.nbytes, 1), C.byref(p)), "ma_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, device: int = 0, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than allocation")
        check(lib().ma_memcpy_h2d(self.ptr, ptr(a), a.nbytes, stream), "ma_memcpy_h2d")

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().ma_memcpy_d2h(ptr(out), self.ptr, self.nbytes, stream), "ma_memcpy_d2h")
        if stream is not None:
            sync(stream)
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().ma_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self, device: int = 0):
        p = C.c_void_p()
        check(lib().ma_stream_create(device, C.byref(p)), "ma_stream_create")
        self.handle = p.value

    def sync(self) -> None:
        check(lib().ma_stream_sync(self.handle), "ma_stream_sync")

    def __del__(self):
        try:
            if self.handle:
                lib().ma_stream_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """A stage chain captured from a stream and replayed as one HIP graph.

    with Graph.capture(stream) as g:   # the *_dev calls made here are recorded
        ...
    g.launch(stream)                   # replays them (same buffers, same shapes)
    """

    def __init__(self, handle: int, n_nodes: int):
        self.handle, self.n_nodes = handle, n_nodes

    @classmethod
    def capture(cls, stream):
        return _Capture(stream)

    def launch(self, stream) -> None:
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_graph_launch(self.handle, h), "ma_graph_launch")

    def __del__(self):
        try:
            if self.handle:
                lib().ma_graph_destroy(self.handle)
        except Exception:
            pass


class _Capture:
    def __init__(self, stream):
        self.h = stream.handle if isinstance(stream, Stream) else stream
        self.graph: Optional[Graph] = None

    def __enter__(self) -> "_Capture":
        check(lib().ma_graph_begin(self.h), "ma_graph_begin")
        return self

    def __exit__(self, et, ev, tb) -> None:
        x, nn = C.c_void_p(), C.c_int64(0)
        rc = lib().ma_graph_end(self.h, C.byref(x), C.byref(nn))
        if et is None:
            check(rc, "ma_graph_end")
            self.graph = Graph(x.value, int(nn.value))
        elif rc == 0 and x.value:
            lib().ma_graph_destroy(x.value)


def sync(stream) -> None:
    h = stream.handle if isinstance(stream, Stream) else stream
    check(lib().ma_stream_sync(h), "ma_stream_sync")


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().ma_event_create(C.byref(p)), "ma_event_create")
        self.handle = p.value

    def record(self, stream) -> None:
        h = stream.handle if isinstance(stream, Stream) else stream
        check(lib().ma_event_record(self.handle, h), "ma_event_record")

    def synchronize(self) -> None:
        check(lib().ma_event_sync(self.handle), "ma_event_sync")

    def elapsed_ms(self, end: "Event") -> float:
        v = C.c_float(0)
        check(lib().ma_event_elapsed_ms(self.handle, end.handle, C.byref(v)), "ma_event_elapsed_ms")
        return float(v.value)

    def __del__(self):
        try:
            if self.handle:
                lib().ma_event_destroy(self.handle)
        except Exception:
            pass


def device_info(device: int = 0) -> dict:
    name = C.create_string_buffer(256)
    tot, free = C.c_int64(0), C.c_int64(0)
    ncu, lds = C.c_int32(0), C.c_int32(0)
    check(
        lib().ma_device_info(device, name, 256, C.byref(tot), C.byref(free), C.byref(ncu), C.byref(lds)),
        "ma_device_info",
    )
    return {
        "name": name.value.decode(),
        "hbm_total": tot.value,
        "hbm_free": free.value,
        "n_cu": ncu.value,
        "lds_per_block": lds.value,
    }


# ------------------------------------------------------------------- index
class GpuIndex:
    """Resident dense-v1 index + 2-bit reference + hash table on one device."""

    def __init__(self, handle: int, device: int):
        self.handle = handle
        self.device = device
        s = self.stats()
        self.k = s["k"]
        self.n_keys = s["n_keys"]
        self.n_postings = s["n_postings"]
        self.n_bases = s["n_bases"]
        self.n_contigs = s["n_contigs"]

    @classmethod
    def open(cls, dense_v1_dir: str, device: int = 0) -> "GpuIndex":
        h = C.c_void_p()
        check(lib().ma_index_open(str(dense_v1_dir).encode(), device, C.byref(h)), "ma_index_open")
        return cls(h.value, device)

    @classmethod
    def build(cls, seq: np.ndarray, contig_off: np.ndarray, k: int, device: int = 0, *,
              keep_dense: bool = True, partition_bits: int = 10, chunk_windows: int = 0) -> "GpuIndex":
        """GPU dense-v1 build (one key-range chunk of windows at a time) + hash table."""
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        contig_off = np.ascontiguousarray(contig_off, dtype=np.uint64)
        opts = BuildOpts()
        lib().ma_build_opts_default(C.byref(opts))
        opts.keep_dense, opts.partition_bits, opts.chunk_windows = int(keep_dense), int(partition_bits), int(chunk_windows)
        h = C.c_void_p()
        check(
            lib().ma_index_build_opts(device, k, ptr(seq), seq.size, ptr(contig_off), contig_off.size - 1,
                                      C.byref(opts), C.byref(h)),
            "ma_index_build",
        )
        return cls(h.value, device)

    @classmethod
    def from_dense(cls, k, kmers, offsets, postings, seq, contig_off, device: int = 0) -> "GpuIndex":
        kmers = np.ascontiguousarray(kmers, dtype=np.uint64)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        postings = np.ascontiguousarray(postings, dtype=np.uint32)
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        contig_off = np.ascontiguousarray(contig_off, dtype=np.uint64)
        h = C.c_void_p()
        check(
            lib().ma_index_from_dense(device, k, ptr(kmers), kmers.size, ptr(offsets), ptr(postings),
                                      postings.size // 2, ptr(seq), seq.size, ptr(contig_off),
                                      contig_off.size - 1, C.byref(h)),
            "ma_index_from_dense",
        )
        return cls(h.value, device)

    def stats(self) -> dict:
        v = [C.c_int64(0) for _ in range(5)]
        k = C.c_int32(0)
        check(lib().ma_index_stats(self.handle, *[C.byref(x) for x in v], C.byref(k)), "ma_index_stats")
        return {
            "n_keys": v[0].value,
            "n_postings": v[1].value,
            "n_bases": v[2].value,
            "n_contigs": v[3].value,
            "device_bytes": v[4].value,
            "k": k.value,
        }

    def hash_stats(self) -> dict:
        nb, no, sbytes, inl = C.c_int64(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
        check(lib().ma_index_hash_stats(self.handle, C.byref(nb), C.byref(no), C.byref(sbytes), C.byref(inl)),
              "ma_index_hash_stats")
        return {"n_buckets": nb.value, "n_overflow": no.value, "slot_bytes": sbytes.value,
                "slots_per_bucket": 64 // sbytes.value, "inline_postings": bool(inl.value)}

    def export_dense(self):
        kmers = np.empty(self.n_keys, dtype=np.uint64)
        offsets = np.empty(self.n_keys + 1, dtype=np.uint64)
        postings = np.empty(2 * self.n_postings, dtype=np.uint32)
        check(lib().ma_index_export_dense(self.handle, ptr(kmers), ptr(offsets), ptr(postings)),
              "ma_index_export_dense")
        return kmers, offsets, postings

    def close(self) -> None:
        if self.handle:
            lib().ma_index_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------- map
def map_batch(
    index: GpuIndex,
    bases: np.ndarray,
    lens: np.ndarray,
    max_len: int,
    *,
    encoding: int = MA_ENC_ASCII,
    bs_r1: int = 0,
    bs_r2: int = 0,
    n_r1: Optional[int] = None,
    params: Optional[Params] = None,
    strands: bool = False,
):
    """Map one batch (host arrays). bases is [n, stride] uint8. Returns the
    merged records, plus FW and RC records when strands=True."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = lens.size
    if bases.ndim != 2 or bases.shape[0] != n:
        raise ValueError("bases must be [n_reads, stride]")
    stride = bases.shape[1]
    params = params or default_params(k=index.k)
    merged = host_recs(n)
    fw = host_recs(n) if strands else None
    rc = host_recs(n) if strands else None
    m_s, f_s, r_s = recs_struct(merged), recs_struct(fw), recs_struct(rc)
    check(
        lib().ma_map_batch(
            index.handle, ptr(bases), stride, ptr(lens), n, int(max_len), encoding, bs_r1, bs_r2,
            n if n_r1 is None else int(n_r1), C.byref(params), C.byref(m_s),
            C.byref(f_s) if f_s is not None else None, C.byref(r_s) if r_s is not None else None,
        ),
        "ma_map_batch",
    )
    if strands:
        return merged, fw, rc
    return merged


def map_forward(index: GpuIndex, bases: np.ndarray, lens: np.ndarray, max_len: int, *, bs_r1: int = 0,
                bs_r2: int = 0, n_r1: Optional[int] = None, params: Optional[Params] = None) -> dict:
    """The reference's `-device cpu` host loop (forward strand only) on the
    device, over ASCII rows [n, stride]: records as ma_map_forward."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = lens.size
    if bases.ndim != 2 or bases.shape[0] != n:
        raise ValueError("bases must be [n_reads, stride]")
    params = params or default_params(k=index.k)
    out = host_recs(n)
    st = recs_struct(out)
    check(lib().ma_map_forward(index.handle, ptr(bases), bases.shape[1], ptr(lens), n, int(max_len), bs_r1, bs_r2,
                               n if n_r1 is None else int(n_r1), C.byref(params), C.byref(st)), "ma_map_forward")
    return out


class DeviceRecs:
    """Merged-record arrays resident on the device (one allocation)."""

    def __init__(self, n: int, device: int = 0):
        self.n = n
        self.buf = DeviceArray((max(n, 1) * 4 * len(REC_FIELDS),), np.uint8, device)
        self.struct = Recs()
        for i, f in enumerate(REC_FIELDS):
            setattr(self.struct, f, self.buf.ptr + i * max(n, 1) * 4)

    def field_ptr(self, f: str) -> int:
        return getattr(self.struct, f)

    @classmethod
    def from_host(cls, recs: dict, n: int, device: int = 0, stream=None) -> "DeviceRecs":
        """Upload host record arrays (missing fields are zero)."""
        d = cls(n, device)
        raw = np.zeros(d.buf.nbytes, dtype=np.uint8)
        n1 = max(n, 1)
        for i, f in enumerate(REC_FIELDS):
            if f in recs:
                a = np.ascontiguousarray(recs[f][:n]).astype(REC_DTYPES[f], copy=False)
                raw[i * n1 * 4 : i * n1 * 4 + n * 4] = a.view(np.uint8)
        d.buf.upload(raw, stream)
        return d

    def download(self) -> dict[str, np.ndarray]:
        raw = self.buf.download()
        out = {}
        n1 = max(self.n, 1)
        for i, f in enumerate(REC_FIELDS):
            out[f] = raw[i * n1 * 4 : i * n1 * 4 + self.n * 4].view(REC_DTYPES[f]).copy()
        return out


class Workspace:
    def __init__(self, index: GpuIndex, max_reads: int, max_len: int, persistent: bool = True):
        h = C.c_void_p()
        check(lib().ma_workspace_create(index.handle, max_reads, max_len, C.byref(h)),
              "ma_workspace_create")
        self.handle = h.value
        check(lib().ma_workspace_set_persistent(self.handle, int(persistent)),
              "ma_workspace_set_persistent")

    def __del__(self):
        try:
            if self.handle:
                lib().ma_workspace_destroy(self.handle)
        except Exception:
            pass


def map_batch_dev(index: GpuIndex, ws: Workspace, stream, d_bases: DeviceArray, stride: int,
                  d_lens: DeviceArray, n: int, max_len: int, out: DeviceRecs, *,
                  encoding: int = MA_ENC_2BIT, bs_r1: int = 0, bs_r2: int = 0,
                  n_r1: Optional[int] = None, params: Optional[Params] = None) -> None:
    params = params or default_params(k=index.k)
    h = stream.handle if isinstance(stream, Stream) else stream
    check(
        lib().ma_map_batch_dev(index.handle, ws.handle, h, d_bases.ptr, stride, d_lens.ptr, n,
                               int(max_len), encoding, bs_r1, bs_r2, n if n_r1 is None else n_r1,
                               C.byref(params), C.byref(out.struct), None, None),
        "ma_map_batch_dev",
    )


def mate_rescue_dev(index: GpuIndex, stream, d_bases: DeviceArray, stride: int, d_lens: DeviceArray, n_pairs: int,
                    max_len: int, recs: DeviceRecs, *, encoding: int = MA_ENC_2BIT, bs_r1: int = 0, bs_r2: int = 0,
                    window: int = 512, max_diff: int = 6) -> None:
    """Mate rescue of a paired batch in place on its device records (R1 block
    then R2 block of n_pairs each); ma_mate_rescue_dev."""
    h = stream.handle if isinstance(stream, Stream) else stream
    check(lib().ma_mate_rescue_dev(index.handle, h, d_bases.ptr, stride, d_lens.ptr, n_pairs, int(max_len), encoding,
                                   bs_r1, bs_r2, int(window), int(max_diff), C.byref(recs.struct)),
          "ma_mate_rescue_dev")


def mate_rescue(index: GpuIndex, bases: np.ndarray, lens: np.ndarray, max_len: int, recs: dict, *,
                encoding: int = MA_ENC_ASCII, bs_r1: int = 0, bs_r2: int = 0, window: int = 512,
                max_diff: int = 6) -> dict:
    """Host-array form: the merged records of a paired batch with rescued
    mates filled in (a copy)."""
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = lens.size
    st = Stream(index.device)
    d_b = DeviceArray.from_host(np.ascontiguousarray(bases, dtype=np.uint8), index.device)
    d_l = DeviceArray.from_host(lens, index.device)
    d_r = DeviceRecs.from_host(recs, n, index.device)
    mate_rescue_dev(index, st, d_b, bases.shape[1], d_l, n // 2, max_len, d_r, encoding=encoding, bs_r1=bs_r1,
                    bs_r2=bs_r2, window=window, max_diff=max_diff)
    st.sync()
    return d_r.download()


# -------------------------------------------------------- sort + markdup
def sort_markdup(coord, dhi, dlo, score, pair, do_markdup: bool = True, device: int = 0):
    """gpu_sort_markdup contract on host arrays -> (is_dup, perm, dup_perm)."""
    coord = np.ascontiguousarray(coord, dtype=np.uint64)
    dhi = np.ascontiguousarray(dhi, dtype=np.uint64)
    dlo = np.ascontiguousarray(dlo, dtype=np.uint64)
    score = np.ascontiguousarray(score, dtype=np.uint32)
    pair = np.ascontiguousarray(pair, dtype=np.uint32)
    n = coord.size
    is_dup = np.zeros(n, dtype=np.uint32)
    perm = np.zeros(n, dtype=np.uint32)
    dup_perm = np.zeros(n, dtype=np.uint32)
    if n:
        check(
            lib().ma_sort_markdup(device, n, ptr(coord), ptr(dhi), ptr(dlo), ptr(score), ptr(pair),
                                  ptr(is_dup), ptr(perm), int(bool(do_markdup)), ptr(dup_perm)),
            "ma_sort_markdup",
        )
    return is_dup, perm, dup_perm


def seed_kmers(bases: np.ndarray, k: int, device: int = 0) -> np.ndarray:
    """Per-position Wang hash of MSB-first k-mers (0 = invalid), [n, max_len]."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    n, max_len = bases.shape
    out = np.zeros((n, max_len), dtype=np.uint64)
    check(lib().ma_seed_kmers(device, ptr(bases), n, max_len, k, ptr(out)), "ma_seed_kmers")
    return out


__all__ = [
    "DeviceArray",
    "DeviceRecs",
    "Event",
    "GpuIndex",
    "MialignError",
    "Stream",
    "Workspace",
    "device_info",
    "map_batch",
    "map_batch_dev",
    "seed_kmers",
    "sort_markdup",
    "sync",
]
