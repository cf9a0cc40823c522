"""Seeded synthetic WGBS workloads (references, bisulfite reads, sort records).

Thin numpy wrapper over libmasynth.so (csrc/synth.c). Each workload is a pure
function of its seed, so the GPU path and the oracle see identical bytes.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None


def _synth():
    global _lib
    if _lib is None:
        path = HERE / "libmasynth.so"
        if not path.is_file():
            raise RuntimeError(f"{path} missing: run __graft_entry__.build() first")
        L = C.CDLL(str(path))
        P, I64, D, U64 = C.c_void_p, C.c_int64, C.c_double, C.c_uint64
        L.ms_gen_reference.restype = C.c_int
        L.ms_gen_reference.argtypes = [U64, I64, I64, D, D, I64, I64, I64, D, P, P]
        L.ms_dual_convert.restype = None
        L.ms_dual_convert.argtypes = [P, P, I64, P, P]
        L.ms_gen_reads.restype = C.c_int
        L.ms_gen_reads.argtypes = [U64, P, P, I64, I64, C.c_int32, C.c_int32, D, D, D, D, I64, P, P,
                                   P, P, P, P]
        L.ms_pack_2bit.restype = I64
        L.ms_pack_2bit.argtypes = [P, P, I64, I64, I64, P]
        L.ms_gen_sort_records.restype = None
        L.ms_gen_sort_records.argtypes = [U64, I64, I64, I64, D, P, P, P]
        L.ms_write_fastq.restype = I64
        L.ms_write_fastq.argtypes = [I64, P, P, I64, P, P, P, P, I64, P]
        L.ms_gen_names_quals.restype = I64
        L.ms_gen_names_quals.argtypes = [U64, I64, P, I64, P, P, I64, P]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class Reference:
    seq: np.ndarray  # uint8 ASCII, all contigs concatenated
    contig_off: np.ndarray  # uint64 [n + 1]
    names: list

    @property
    def n_contigs(self) -> int:
        return self.contig_off.size - 1


def gen_reference(seed: int, n_contigs: int, total_len: int, *, cpg_keep: float = 0.16,
                  repeat_frac: float = 0.0, seg_families: int = 0, seg_copies: int = 0,
                  seg_len: int = 0, seg_div: float = 0.0) -> Reference:
    """Uniform ACGT with CpG depleted to ~cpg_keep/16 of dinucleotides (0.16 -> 1%)."""
    seq = np.empty(total_len, dtype=np.uint8)
    off = np.empty(n_contigs + 1, dtype=np.uint64)
    rc = _synth().ms_gen_reference(seed, n_contigs, total_len, cpg_keep, repeat_frac, seg_families,
                                   seg_copies, seg_len, seg_div, _p(seq), _p(off))
    if rc != 0:
        raise ValueError(f"ms_gen_reference failed ({rc})")
    return Reference(seq, off, [f"chr{i + 1}" for i in range(n_contigs)])


def dual_convert(ref: Reference) -> Reference:
    """bwameth dual copy: f<name> = C->T, r<name> = G->A, interleaved per contig."""
    out = np.empty(2 * ref.seq.size, dtype=np.uint8)
    off = np.empty(2 * ref.n_contigs + 1, dtype=np.uint64)
    _synth().ms_dual_convert(_p(ref.seq), _p(ref.contig_off), ref.n_contigs, _p(out), _p(off))
    names = []
    for n in ref.names:
        names += ["f" + n, "r" + n]
    return Reference(out, off, names)


@dataclass
class Reads:
    bases: np.ndarray  # uint8 [n, max_len] ASCII (pre in-silico conversion), 'N' padded
    lens: np.ndarray  # uint32 [n]
    contig: np.ndarray  # int32 truth: original contig
    pos: np.ndarray  # uint32 truth: template start on the top strand
    strand: np.ndarray  # uint8 truth: 0 OT, 1 OB
    dup_of: np.ndarray  # int32: -1 or the copied read

    @property
    def n(self) -> int:
        return self.lens.size

    @property
    def max_len(self) -> int:
        return self.bases.shape[1]


def gen_reads(seed: int, ref: Reference, n: int, len_min: int = 150, len_max: int = 150, *,
              p_conv: float = 0.99, p_meth: float = 0.7, p_err: float = 0.005,
              dup_frac: float = 0.05, max_len: int | None = None) -> Reads:
    max_len = max_len or len_max
    bases = np.empty((n, max_len), dtype=np.uint8)
    lens = np.empty(n, dtype=np.uint32)
    contig = np.empty(n, dtype=np.int32)
    pos = np.empty(n, dtype=np.uint32)
    strand = np.empty(n, dtype=np.uint8)
    dup_of = np.empty(n, dtype=np.int32)
    rc = _synth().ms_gen_reads(seed, _p(ref.seq), _p(ref.contig_off), ref.n_contigs, n, len_min,
                               len_max, p_conv, p_meth, p_err, dup_frac, max_len, _p(bases), _p(lens),
                               _p(contig), _p(pos), _p(strand), _p(dup_of))
    if rc != 0:
        raise ValueError(f"ms_gen_reads failed ({rc})")
    return Reads(bases, lens, contig, pos, strand, dup_of)


_COMP = np.arange(256, dtype=np.uint8)
_COMP[np.frombuffer(b"ACGTacgt", np.uint8)] = np.frombuffer(b"TGCAtgca", np.uint8)


def gen_pairs(seed: int, ref: Reference, n_pairs: int, frag_len: int = 300, len1: int = 150, len2: int = 150,
              **kw) -> tuple[Reads, Reads]:
    """Directional bisulfite pairs: fragments of frag_len bases drawn as
    gen_reads draws reads (conversion, errors, duplicate fragments); R1 = a
    fragment's first len1 bases, R2 = the reverse complement of its last len2.
    Truth fields describe the fragment."""
    if not (1 <= len1 <= frag_len and 1 <= len2 <= frag_len):
        raise ValueError("gen_pairs: mate lengths must lie in [1, frag_len]")
    fr = gen_reads(seed, ref, n_pairs, frag_len, frag_len, max_len=frag_len, **kw)
    m = max(len1, len2)
    b1 = np.full((n_pairs, m), ord("N"), np.uint8)
    b1[:, :len1] = fr.bases[:, :len1]
    b2 = np.full((n_pairs, m), ord("N"), np.uint8)
    b2[:, :len2] = _COMP[fr.bases[:, frag_len - len2:frag_len]][:, ::-1]
    r1 = Reads(b1, np.full(n_pairs, len1, np.uint32), fr.contig, fr.pos, fr.strand, fr.dup_of)
    r2 = Reads(b2, np.full(n_pairs, len2, np.uint32), fr.contig, fr.pos, fr.strand, fr.dup_of)
    return r1, r2


def pack_2bit(bases: np.ndarray, lens: np.ndarray, stride_bytes: int | None = None) -> np.ndarray:
    """[n, max_len] ASCII -> [n, stride] 2-bit (base 4j in bits 7-6 of byte j)."""
    n, max_len = bases.shape
    stride_bytes = stride_bytes or (max_len + 3) // 4
    out = np.empty((n, stride_bytes), dtype=np.uint8)
    bad = _synth().ms_pack_2bit(_p(np.ascontiguousarray(bases)), _p(np.ascontiguousarray(lens)), n,
                                max_len, stride_bytes, _p(out))
    if bad:
        raise ValueError(f"{bad} non-ACGT bases cannot be 2-bit packed")
    return out


def gen_sort_records(seed: int, n: int, n_contigs: int = 24, genome_len: int = 3_100_000_000,
                     dup_frac: float = 0.12):
    """Sort-stress records: key = contig<<40 | pos<<1 | strand, u32 id, u8 nm."""
    key = np.empty(n, dtype=np.uint64)
    rid = np.empty(n, dtype=np.uint32)
    nm = np.empty(n, dtype=np.uint8)
    _synth().ms_gen_sort_records(seed, n, n_contigs, genome_len, dup_frac, _p(key), _p(rid), _p(nm))
    return key, rid, nm


def tid_of_cid(names: list) -> np.ndarray:
    """Canonical target id per index contig: f/r prefixes stripped, first-seen
    order (_canonical_contig + _write_sam_header, linear_gpu_locate.mojo:156-208)."""
    seen: dict = {}
    out = np.empty(len(names), dtype=np.int32)
    for i, n in enumerate(names):
        c = n[1:] if len(n) >= 2 and n[0] in "fr" else n
        if c not in seen:
            seen[c] = len(seen)
        out[i] = seen[c]
    return out


def gen_names_quals(seed: int, lens: np.ndarray, max_len: int):
    """Illumina-style read names (blob + offsets [n+1]) and phred+33 quality
    rows [n, max_len] for n reads (seeded)."""
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    n = lens.size
    cap = 48 * max(n, 1)
    names = np.empty(cap, dtype=np.uint8)
    off = np.empty(n + 1, dtype=np.int64)
    qual = np.empty((n, max_len), dtype=np.uint8)
    got = _synth().ms_gen_names_quals(seed, n, _p(lens), max_len, _p(qual), _p(names), cap, _p(off))
    if got < 0:
        raise ValueError("ms_gen_names_quals: name buffer too small")
    return names[:max(int(got), 1)], off, qual


def write_fastq(bases: np.ndarray, qual: np.ndarray, lens: np.ndarray, names: np.ndarray,
                name_off: np.ndarray) -> np.ndarray:
    """FASTQ text (uint8 array) of the reads: "@name", bases, "+", qualities."""
    bases = np.ascontiguousarray(bases, dtype=np.uint8)
    qual = np.ascontiguousarray(qual, dtype=np.uint8)
    lens = np.ascontiguousarray(lens, dtype=np.uint32)
    name_off = np.ascontiguousarray(name_off, dtype=np.int64)
    n, stride = bases.shape
    cap = int(n * 6 + 2 * int(lens.astype(np.int64).sum()) + int(name_off[-1]))
    out = np.empty(max(cap, 1), dtype=np.uint8)
    off = np.empty(n + 1, dtype=np.int64)
    got = _synth().ms_write_fastq(n, _p(bases), _p(qual), stride, _p(lens), _p(np.ascontiguousarray(names)),
                                  _p(name_off), _p(out), cap, _p(off))
    if got < 0:
        raise ValueError("ms_write_fastq: buffer too small")
    return out[:got]
