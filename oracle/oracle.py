"""ctypes wrapper of the oracle (oracle.c) — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; the product never imports it. See oracle.c for the reference
file:line each function restates and for which functions are pinned by what.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REC_FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm", "iop", "iat", "ilen")
REC_DTYPES = {"cid": np.int32, "flag": np.int32}

_lib = None


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run(
        ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", str(tmp),
         str(HERE / "oracle.c")],
        check=True,
    )
    os.replace(tmp, LIB)
    return LIB


class Index(C.Structure):
    _fields_ = [
        ("kmers", C.c_void_p),
        ("offsets", C.c_void_p),
        ("postings", C.c_void_p),
        ("n_keys", C.c_int64),
        ("n_postings", C.c_int64),
        ("ref_codes", C.c_void_p),
        ("contig_off", C.c_void_p),
        ("n_contigs", C.c_int64),
    ]


class Params(C.Structure):
    _fields_ = [(f, C.c_int32) for f in
                ("k", "stride", "max_occ", "vote_occ", "vote_fast", "max_diff", "max_soft",
                 "max_indel", "do_rc")]


class Recs(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in REC_FIELDS]


class Counters(C.Structure):
    _fields_ = [("probes", C.c_int64), ("postings", C.c_int64), ("windows", C.c_int64),
                ("reads_rc", C.c_int64), ("overflow", C.c_int64), ("tally", C.c_int64)]


def lib():
    global _lib
    if _lib is None:
        if not LIB.is_file():
            build()
        L = C.CDLL(str(LIB))
        P, I64 = C.c_void_p, C.c_int64
        L.or_code_of.restype = C.c_uint8
        L.or_code_of.argtypes = [C.c_uint8]
        L.or_pack_bases.argtypes = [P, P, I64]
        L.or_bs_convert.argtypes = [P, I64, C.c_int]
        L.or_convert_fasta_c2t.argtypes = [P, I64]
        L.or_wang_hash.restype = C.c_uint64
        L.or_wang_hash.argtypes = [C.c_uint64]
        L.or_encode_strided.argtypes = [P, P, I64, I64, I64, I64, I64, P, P]
        L.or_locate_bsearch.argtypes = [P, I64, P, P, I64, I64, P, P]
        L.or_count_windows.restype = I64
        L.or_count_windows.argtypes = [P, I64, C.c_int]
        L.or_build_dense.restype = C.c_int
        L.or_build_dense.argtypes = [P, P, I64, C.c_int, P, P, P, C.POINTER(I64), C.POINTER(I64)]
        L.or_vote_extend.argtypes = [C.POINTER(Index), P, P, I64, I64, I64, P, P, P, I64, I64, I64,
                                     I64, C.c_int, I64, P, C.c_int32, C.POINTER(Recs)]
        L.or_slots_per.restype = I64
        L.or_slots_per.argtypes = [I64, I64, I64]
        L.or_map_batch.restype = C.c_int
        L.or_map_batch.argtypes = [C.POINTER(Index), C.POINTER(Params), P, P, I64, I64,
                                   C.POINTER(Recs), C.POINTER(Recs), C.POINTER(Recs),
                                   C.POINTER(Counters), C.c_int]
        L.or_unclipped5.restype = I64
        L.or_unclipped5.argtypes = [I64, I64, I64, I64, I64]
        L.or_coord_key.restype = C.c_uint64
        L.or_coord_key.argtypes = [I64, I64, I64]
        L.or_pack_end.restype = C.c_uint64
        L.or_pack_end.argtypes = [I64, I64, I64]
        L.or_sort_keys_se.argtypes = [I64, P, P, P, P, P, P, P, P, P, I64, I64, P, P, P, P, P]
        L.or_mate_rescue.argtypes = [C.POINTER(Index), P, P, I64, I64, I64, I64, C.POINTER(Recs)]
        L.or_mate_rescue.restype = None
        L.or_sort_markdup.restype = C.c_int
        L.or_sort_markdup.argtypes = [I64, P, P, P, P, P, P, P, C.c_int, P]
        L.or_num_threads.restype = C.c_int
        L.or_reg2bin.restype = I64
        L.or_reg2bin.argtypes = [I64, I64]
        L.or_nt16.restype = C.c_uint8
        L.or_nt16.argtypes = [C.c_uint8]
        L.or_nt16_comp.restype = C.c_uint8
        L.or_nt16_comp.argtypes = [C.c_uint8]
        L.or_pack_bam_aln.restype = I64
        L.or_pack_bam_aln.argtypes = [P, I64, I64, P, I64, I64, I64, I64, I64, I64, I64, P, I64, P,
                                      I64, I64, I64, I64, I64, P, I64]
        L.or_fq_cvt.restype = C.c_uint8
        L.or_fq_cvt.argtypes = [C.c_uint8, C.c_int]
        L.or_fastq_parse.argtypes = [P, I64, C.c_int, I64, I64, C.c_int, P, P, P, P, P, P, I64, P, P]
        L.or_bam_pack_fields.restype = I64
        L.or_bam_pack_fields.argtypes = [I64] + [P] * 13 + [I64, P, P, P, P, I64, P, P, C.c_int, P, I64]
        L.or_bam_pack_se.restype = I64
        L.or_bam_pack_se.argtypes = [I64] + [P] * 8 + [I64, P, P, P, I64, P, P, P, P, I64, P, P,
                                                       C.c_int, P, I64]
        L.or_pe_fields.restype = None
        L.or_pe_fields.argtypes = [I64] + [P] * 9 + [I64, I64] + [P] * 16
        L.or_bam_pack_pe.restype = I64
        L.or_bam_pack_pe.argtypes = [I64] + [P] * 8 + [I64, P, P, P, I64, P, P, P, P, I64, P, P,
                                                       C.c_int, P, I64]
        L.or_recount_nm.argtypes = [P, P, I64, P, I64, P, I64, P, P, P, P, P, P]
        L.or_map_forward.restype = C.c_int
        L.or_map_forward.argtypes = [C.POINTER(Index), P, P, I64, P, I64, I64, C.c_int, C.c_int, C.c_int,
                                     C.c_int, I64, C.c_int, C.c_int, C.POINTER(Recs), C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def params(k=21, stride=5, max_occ=16384, vote_occ=256, vote_fast=None, max_diff=6, max_soft=8,
           max_indel=0, do_rc=1) -> Params:
    return Params(k, stride, max_occ, vote_occ, vote_occ if vote_fast is None else vote_fast,
                  max_diff, max_soft, max_indel, do_rc)


def new_recs(n: int) -> dict:
    return {f: np.zeros(n, dtype=REC_DTYPES.get(f, np.uint32)) for f in REC_FIELDS}


def recs_struct(d: dict) -> Recs:
    r = Recs()
    for f in REC_FIELDS:
        setattr(r, f, _p(d[f]))
    return r


def build_dense(seq: np.ndarray, contig_off: np.ndarray, k: int):
    """dense-v1 (kmers, offsets, postings) by stable sort (build_dense_pack)."""
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    contig_off = np.ascontiguousarray(contig_off, dtype=np.uint64)
    nc = contig_off.size - 1
    cap = int(lib().or_count_windows(_p(contig_off), nc, k))
    kmers = np.empty(max(cap, 1), dtype=np.uint64)
    offsets = np.empty(max(cap, 1) + 1, dtype=np.uint64)
    postings = np.empty(2 * max(cap, 1), dtype=np.uint32)
    nk, npost = C.c_int64(0), C.c_int64(0)
    rc = lib().or_build_dense(_p(seq), _p(contig_off), nc, k, _p(kmers), _p(offsets), _p(postings),
                              C.byref(nk), C.byref(npost))
    if rc != 0:
        raise MemoryError("or_build_dense failed")
    return kmers[: nk.value].copy(), offsets[: nk.value + 1].copy(), postings[: 2 * npost.value].copy()


class OracleIndex:
    """Holds the arrays an or_index points at."""

    def __init__(self, kmers, offsets, postings, seq, contig_off):
        self.kmers = np.ascontiguousarray(kmers, dtype=np.uint64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.postings = np.ascontiguousarray(postings, dtype=np.uint32)
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        self.seq = seq  # the pack's bytes (the host loop compares these)
        self.ref_codes = np.empty_like(seq)
        lib().or_pack_bases(_p(seq), _p(self.ref_codes), seq.size)
        self.contig_off = np.ascontiguousarray(contig_off, dtype=np.uint64)
        self.struct = Index(_p(self.kmers), _p(self.offsets), _p(self.postings), self.kmers.size,
                            self.postings.size // 2, _p(self.ref_codes), _p(self.contig_off),
                            self.contig_off.size - 1)


def recount_nm(seq, contig_off, conv_reads, lens, recs) -> np.ndarray:
    """Mismatches of each record at its reported locus (-1 unmapped, -2 out of contig)."""
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    contig_off = np.ascontiguousarray(contig_off, dtype=np.uint64)
    conv_reads = np.ascontiguousarray(conv_reads, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    f = {k: np.ascontiguousarray(recs[k]) for k in ("cid", "pos", "flag", "sl", "sr")}
    out = np.empty(lens.size, dtype=np.int32)
    lib().or_recount_nm(_p(seq), _p(contig_off), contig_off.size - 1, _p(conv_reads), conv_reads.shape[1], _p(lens),
                        lens.size, _p(f["cid"]), _p(f["pos"]), _p(f["flag"]), _p(f["sl"]), _p(f["sr"]), _p(out))
    return out


def convert_reads(bases: np.ndarray, mode: int) -> np.ndarray:
    out = np.ascontiguousarray(bases, dtype=np.uint8).copy()
    lib().or_bs_convert(_p(out), out.size, mode)
    return out


def map_batch(ix: OracleIndex, bases: np.ndarray, lens: np.ndarray, prm: Params | None = None,
              nthreads: int = 1):
    """Reads already in-silico converted, [n, max_len] ASCII padded 'N'.
    Returns (merged, fw, rc, counters)."""
    prm = prm or params()
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n, max_len = bases.shape
    if max_len < prm.k:
        pad = np.full((n, prm.k), ord("N"), dtype=np.uint8)
        pad[:, :max_len] = bases
        bases, max_len = pad, prm.k
    mg, fw, rc = new_recs(n), new_recs(n), new_recs(n)
    cnt = Counters()
    r = lib().or_map_batch(C.byref(ix.struct), C.byref(prm), _p(bases), _p(lens), n, max_len,
                           C.byref(recs_struct(fw)), C.byref(recs_struct(rc)),
                           C.byref(recs_struct(mg)), C.byref(cnt), nthreads)
    if r != 0:
        raise RuntimeError(f"or_map_batch failed ({r})")
    return mg, fw, rc, cnt


def mate_rescue(ix: OracleIndex, bases: np.ndarray, lens: np.ndarray, rec: dict, window: int = 512,
                max_diff: int = 6) -> dict:
    """Mate rescue of a paired batch (R1 block then R2 block of n_pairs each;
    reads already in-silico converted, [2 n_pairs, max_len] ASCII): a copy of
    the merged records with rescued mates filled in."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n, max_len = bases.shape
    codes = np.empty_like(bases)
    lib().or_pack_bases(_p(bases), _p(codes), bases.size)
    out = {f: np.ascontiguousarray(rec[f]).copy() for f in REC_FIELDS}
    lib().or_mate_rescue(C.byref(ix.struct), _p(codes), _p(lens), max_len, n // 2, window, max_diff,
                         C.byref(recs_struct(out)))
    return out


def map_forward(ix: OracleIndex, bases: np.ndarray, lens: np.ndarray, *, k: int, n_r1: int | None = None,
                bs_r1: int = 0, bs_r2: int = 0, stride: int = 5, max_occ: int = 16384, max_diff: int = 6,
                max_soft: int = 8, nthreads: int = 8) -> dict:
    """The reference's `-device cpu` host loop (or_map_forward) over raw ASCII
    rows [n, row bytes]: converts each read itself (bs_r1 for rows < n_r1,
    else bs_r2) and returns the records (forward strand only)."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n, row = bases.shape
    out = new_recs(n)
    r = lib().or_map_forward(C.byref(ix.struct), _p(ix.seq), _p(bases), row, _p(lens), n,
                             n if n_r1 is None else n_r1, bs_r1, bs_r2, k, stride, max_occ, max_diff, max_soft,
                             C.byref(recs_struct(out)), nthreads)
    if r != 0:
        raise RuntimeError(f"or_map_forward failed ({r})")
    return out


def sort_keys_se(rec: dict, lens: np.ndarray, tid_of_cid: np.ndarray, pair_base: int = 0):
    n = lens.size
    coord = np.empty(n, np.uint64)
    dhi = np.empty(n, np.uint64)
    dlo = np.empty(n, np.uint64)
    score = np.empty(n, np.uint32)
    pair = np.empty(n, np.uint32)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    tid = np.ascontiguousarray(tid_of_cid, dtype=np.int32)
    lib().or_sort_keys_se(n, _p(rec["cid"]), _p(rec["pos"]), _p(rec["flag"]), _p(rec["sl"]),
                          _p(rec["sr"]), _p(rec["nm"]), _p(rec["mapq"]), _p(lens), _p(tid), tid.size,
                          pair_base, _p(coord), _p(dhi), _p(dlo), _p(score), _p(pair))
    return coord, dhi, dlo, score, pair


def sort_markdup(coord, dhi, dlo, score, pair, do_markdup=True):
    n = len(coord)
    coord = np.ascontiguousarray(coord, dtype=np.uint64)
    dhi = np.ascontiguousarray(dhi, dtype=np.uint64)
    dlo = np.ascontiguousarray(dlo, dtype=np.uint64)
    score = np.ascontiguousarray(score, dtype=np.uint32)
    pair = np.ascontiguousarray(pair, dtype=np.uint32)
    is_dup = np.zeros(n, np.uint32)
    perm = np.zeros(n, np.uint32)
    dup_perm = np.zeros(n, np.uint32)
    r = lib().or_sort_markdup(n, _p(coord), _p(dhi), _p(dlo), _p(score), _p(pair), _p(is_dup),
                              _p(perm), int(bool(do_markdup)), _p(dup_perm))
    if r != 0:
        raise MemoryError("or_sort_markdup failed")
    return is_dup, perm, dup_perm


# ------------------------------------------------------------ BAM pack / BGZF

def bam_record_bound(lens: np.ndarray, name_off: np.ndarray, rg_n: int) -> int:
    lens = np.asarray(lens, dtype=np.int64)
    nl = np.diff(np.asarray(name_off, dtype=np.int64))
    return int((36 + nl + 1 + 12 + (lens + 1) // 2 + lens + 8 + rg_n).sum())


def bam_pack_se(rec: dict, tid_of_cid: np.ndarray, names: np.ndarray, name_off: np.ndarray,
                seq: np.ndarray, lens: np.ndarray, qual: np.ndarray | None, rg: bytes,
                perm: np.ndarray | None = None, is_dup: np.ndarray | None = None,
                apply_dup: bool = False, qual_lens: np.ndarray | None = None) -> bytes:
    """Uncompressed BAM records of single-end map records (or_bam_pack_se)."""
    n = int(lens.size)
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    stride = seq.shape[1] if seq.ndim == 2 else 0
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    names = np.ascontiguousarray(names, dtype=np.uint8)
    name_off = np.ascontiguousarray(name_off, dtype=np.int64)
    tid = np.ascontiguousarray(tid_of_cid, dtype=np.int32)
    q = None if qual is None else np.ascontiguousarray(qual, dtype=np.uint8)
    ql = None if qual_lens is None else np.ascontiguousarray(qual_lens, dtype=np.uint32)
    pm = None if perm is None else np.ascontiguousarray(perm, dtype=np.uint32)
    dp = None if is_dup is None else np.ascontiguousarray(is_dup, dtype=np.uint32)
    rgb = np.frombuffer(rg, dtype=np.uint8) if rg else np.zeros(1, np.uint8)
    cols = {f: np.ascontiguousarray(rec[f]) for f in ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")}
    cap = bam_record_bound(lens, name_off, len(rg))
    out = np.empty(max(cap, 1), np.uint8)
    g = lib().or_bam_pack_se(n, _p(cols["cid"]), _p(cols["pos"]), _p(cols["mapq"]), _p(cols["flag"]),
                             _p(cols["sl"]), _p(cols["sr"]), _p(cols["nm"]), _p(tid), tid.size,
                             _p(names), _p(name_off), _p(seq), stride, _p(lens), _p(q), _p(ql),
                             _p(rgb), len(rg), _p(pm), _p(dp), int(bool(apply_dup)), _p(out), cap)
    if g < 0:
        raise RuntimeError("or_bam_pack_se: BAM batch buffer overflow")
    return out[:g].tobytes()


PE_FIELDS = ("flag", "tid", "pos", "mapq", "sl", "sr", "nm", "ntid", "npos", "tlen", "row")


def pe_fields(rec: dict, lens: np.ndarray, tid_of_cid: np.ndarray, pair_base: int = 0) -> dict:
    """Paired-end fields and keys (or_pe_fields) of 2 np merged records, R1
    block then R2 block; output record 2p is pair p's R1, 2p + 1 its R2."""
    n = int(lens.size)
    assert n % 2 == 0
    out = {f: np.empty(n, np.int32) for f in PE_FIELDS}
    for f, dt in (("coord", np.uint64), ("dhi", np.uint64), ("dlo", np.uint64), ("score", np.uint32),
                  ("pair", np.uint32)):
        out[f] = np.empty(n, dt)
    cols = {f: np.ascontiguousarray(rec[f]) for f in ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")}
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    tid = np.ascontiguousarray(tid_of_cid, dtype=np.int32)
    lib().or_pe_fields(n // 2, _p(cols["cid"]), _p(cols["pos"]), _p(cols["mapq"]), _p(cols["flag"]),
                       _p(cols["sl"]), _p(cols["sr"]), _p(cols["nm"]), _p(lens), _p(tid), tid.size, pair_base,
                       *[_p(out[f]) for f in PE_FIELDS + ("coord", "dhi", "dlo", "score", "pair")])
    return out


def bam_pack_pe(rec: dict, tid_of_cid: np.ndarray, names: np.ndarray, name_off: np.ndarray,
                seq: np.ndarray, lens: np.ndarray, qual: np.ndarray | None, rg: bytes,
                perm: np.ndarray | None = None, is_dup: np.ndarray | None = None,
                apply_dup: bool = False, qual_lens: np.ndarray | None = None) -> bytes:
    """Uncompressed BAM records of paired-end map records (or_bam_pack_pe);
    reads and records R1 block then R2 block, records out as R1, R2 per pair."""
    n = int(lens.size)
    assert n % 2 == 0
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    stride = seq.shape[1] if seq.ndim == 2 else 0
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    names = np.ascontiguousarray(names, dtype=np.uint8)
    name_off = np.ascontiguousarray(name_off, dtype=np.int64)
    tid = np.ascontiguousarray(tid_of_cid, dtype=np.int32)
    q = None if qual is None else np.ascontiguousarray(qual, dtype=np.uint8)
    ql = None if qual_lens is None else np.ascontiguousarray(qual_lens, dtype=np.uint32)
    pm = None if perm is None else np.ascontiguousarray(perm, dtype=np.uint32)
    dp = None if is_dup is None else np.ascontiguousarray(is_dup, dtype=np.uint32)
    rgb = np.frombuffer(rg, dtype=np.uint8) if rg else np.zeros(1, np.uint8)
    cols = {f: np.ascontiguousarray(rec[f]) for f in ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")}
    cap = bam_record_bound(lens, name_off, len(rg))
    out = np.empty(max(cap, 1), np.uint8)
    g = lib().or_bam_pack_pe(n // 2, _p(cols["cid"]), _p(cols["pos"]), _p(cols["mapq"]), _p(cols["flag"]),
                             _p(cols["sl"]), _p(cols["sr"]), _p(cols["nm"]), _p(tid), tid.size,
                             _p(names), _p(name_off), _p(seq), stride, _p(lens), _p(q), _p(ql),
                             _p(rgb), len(rg), _p(pm), _p(dp), int(bool(apply_dup)), _p(out), cap)
    if g < 0:
        raise RuntimeError("or_bam_pack_pe: BAM batch buffer overflow")
    return out[:g].tobytes()


BGZF_MAX_BLOCK = 65280
BGZF_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def bgzf_block(chunk: bytes, level: int) -> bytes:
    """One BGZF member: gzip header with the BC extra field (BSIZE = member
    length - 1), raw deflate at `level`, CRC32 and ISIZE; follows _bgzf_block,
    fq2bam-meth/python/bam_emit.py:77-100 (zlib's raw deflate, wbits -15)."""
    import struct
    import zlib
    cobj = zlib.compressobj(level, zlib.DEFLATED, -15)
    payload = cobj.compress(chunk) + cobj.flush()
    head = struct.pack("<BBBBLBBHBBHH", 31, 139, 8, 4, 0, 0, 255, 6, 66, 67, 2, 25 + len(payload))
    return head + payload + struct.pack("<II", zlib.crc32(chunk) & 0xFFFFFFFF, len(chunk))


def bgzf_stream(data: bytes, level: int = 1, eof: bool = True) -> bytes:
    """A whole stream cut into consecutive 65280-byte members, as _BgzfWriter
    writes it (bam_emit.py:103-185), then the 28-byte EOF member."""
    out = [bgzf_block(data[i:i + BGZF_MAX_BLOCK], level) for i in range(0, len(data), BGZF_MAX_BLOCK)]
    return b"".join(out) + (BGZF_EOF if eof else b"")


# ------------------------------------------------------------------ FASTQ ingest

def fastq_parse(buf: bytes, stride: int, mode: int = 0, final: bool = True, max_records: int | None = None,
                names_cap: int | None = None) -> dict:
    """or_fastq_parse over a FASTQ byte buffer; returns the rows and the info fields."""
    a = np.frombuffer(buf, dtype=np.uint8) if buf else np.zeros(1, np.uint8)
    n = len(buf)
    cap = max_records if max_records is not None else max(1, buf.count(b"\n") // 4 + 1)
    names_cap = names_cap if names_cap is not None else max(1, n)
    rows = np.zeros((cap, stride), np.uint8)
    conv = np.zeros((cap, stride), np.uint8)
    qual = np.zeros((cap, stride), np.uint8)
    lens = np.zeros(cap, np.uint32)
    qlens = np.zeros(cap, np.uint32)
    names = np.zeros(names_cap, np.uint8)
    off = np.zeros(cap + 1, np.int64)
    info = np.zeros(5, np.int64)
    lib().or_fastq_parse(_p(a), n, int(final), cap, stride, mode, _p(rows), _p(conv), _p(qual), _p(lens), _p(qlens),
                         _p(names), names_cap, _p(off), _p(info))
    r = int(info[0])
    return {"n": r, "consumed": int(info[1]), "max_len": int(info[2]), "status": int(info[3]),
            "name_bytes": int(info[4]), "rows": rows[:r], "conv": conv[:r], "qual": qual[:r], "lens": lens[:r],
            "qual_lens": qlens[:r], "names": names[:int(info[4])], "name_off": off[:r + 1]}
