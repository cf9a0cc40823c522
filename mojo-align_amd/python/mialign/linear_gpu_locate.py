"""map_fastq_dense_gpu_locate: FASTQ -> GPU seed/extend -> SAM.

Same entry point, read-batch interface and SAM output as the reference's
parity engine (fq2bam-meth/src/linear_gpu_locate.mojo:291-2911): FASTQ
records are read in batches (R1 then R2, in-silico converted with bs_r1 /
bs_r2), packed into [n_seq x max_len] rows padded with 'N', mapped on the
GPU (every pass, the need_rc test and the FW/RC merge happen on the device
in one kernel launch), and written as SAM with the original (pre-conversion)
SEQ, reverse-complemented for flag 16, contig names with the bwameth f/r
prefix stripped, and an RG:Z tag.
"""

from __future__ import annotations

import gzip
import os
import sys
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from ._lib import BS_MODES, MA_ENC_ASCII, MialignError, default_params
from .core import map_batch

_C2T = bytes.maketrans(b"Cc", b"Tt")
_G2A = bytes.maketrans(b"Gg", b"Aa")
_RC_TABLE = bytes({65: 84, 84: 65, 67: 71, 71: 67}.get(b, 78) for b in range(256))
MAX_READ = 256


def env_int(name: str, default: int) -> int:
    """Integer knob; values < 1 fall back to the default (linear_gpu_locate.mojo:139-145)."""
    n = int(os.environ.get(name, str(default)))
    return default if n < 1 else n


def env_str(name: str, default: str) -> str:
    raw = os.environ.get(name, default)
    return default if raw == "" else raw


def bs_convert(seq: bytes, mode: str) -> bytes:
    """In-silico conversion: C2T (Cc->Tt) or G2A (Gg->Aa) (:58-70)."""
    if mode == "C2T":
        return seq.translate(_C2T)
    if mode == "G2A":
        return seq.translate(_G2A)
    return seq


def reverse_complement(seq: bytes) -> bytes:
    """A<->T, C<->G, anything else -> N (methylgrapher/src/utility.mojo:28-48)."""
    return seq.translate(_RC_TABLE)[::-1]


def canonical_contig(name: str) -> str:
    """Strip the bwameth f/r prefix (:156-162)."""
    if len(name) >= 2 and name[0] in "fr":
        return name[1:]
    return name


@dataclass
class FastqRec:
    name: str
    seq: bytes  # converted (aligned)
    qual: str = "*"
    original_seq: bytes = b""


def _open_fastq(path: str):
    low = path.lower()
    if low.endswith(".gz") or low.endswith(".gzip"):
        return gzip.open(path, "rb")
    return open(path, "rb")


def _strip_nl(b: bytes) -> bytes:
    return b.rstrip(b"\r\n")


def read_one(fh) -> Optional[FastqRec]:
    """One FASTQ record; name is the first word before any '/' (:73-96)."""
    n = fh.readline()
    if not n:
        return None
    s = _strip_nl(fh.readline())
    fh.readline()
    q = _strip_nl(fh.readline()).decode()
    name = _strip_nl(n).decode()
    if name.startswith("@"):
        name = name[1:]
    name = name.split(" ")[0].split("/")[0]
    return FastqRec(name, s, q if q else "*", s)


def read_pe_batch(fh1, fh2, paired: bool, batch_size: int, bs_r1: str, bs_r2: str):
    """Up to batch_size SE/PE records; raises when R2 ends early (:98-137)."""
    b1: List[FastqRec] = []
    b2: List[FastqRec] = []
    while len(b1) < batch_size:
        r1 = read_one(fh1)
        if r1 is None:
            break
        r1.seq = bs_convert(r1.seq, bs_r1)
        b1.append(r1)
        if paired:
            r2 = read_one(fh2)
            if r2 is None:
                raise MialignError("paired FASTQ length mismatch (R2 ended early)")
            r2.seq = bs_convert(r2.seq, bs_r2)
            b2.append(r2)
    return b1, b2


def pack_batch(seqs: List[bytes], k: int) -> Tuple[np.ndarray, np.ndarray, int]:
    """[n_seq x max_len] rows padded with 'N' (78), max_len >= k (:1838-1876)."""
    lens = np.array([len(s) for s in seqs], dtype=np.uint32)
    max_len = max(int(lens.max()) if lens.size else 0, k)
    if max_len > MAX_READ:
        raise MialignError(f"reads longer than {MAX_READ} bases are not supported (got {max_len})")
    rows = np.full((len(seqs), max_len), 78, dtype=np.uint8)
    for i, s in enumerate(seqs):
        if s:
            rows[i, : len(s)] = np.frombuffer(s, dtype=np.uint8)
    return rows, lens, max_len


@dataclass
class LinearHit:
    """SAM record fields (fq2bam-meth/src/linear_extend.mojo:11-48)."""

    query_name: str
    flag: int
    contig: str
    pos: int  # 1-based
    mapq: int
    cigar: str
    seq: str
    qual: str = "*"
    rnext: str = "*"
    pnext: int = 0
    tlen: int = 0


def hit_from_gpu(names: List[str], name: str, align_seq: bytes, original_seq: bytes, cid: int,
                 start0: int, mapq: int, flag: int, soft_l: int, soft_r: int, indel_op: int = 0,
                 indel_at: int = 0, indel_len: int = 0) -> LinearHit:
    """CIGAR slS mM srS, or with the hit's single indel inside the middle m
    bases (insertion: atM lenI (m-at-len)M, unmapped if that is negative;
    deletion: atM lenD (m-at)M; zero-length M runs left out); SEQ = original
    read (RC for flag 16), canonical contig; unmapped when cid < 0 or nothing
    is left between the clips (:217-284)."""
    qlen = len(align_seq)
    emit = original_seq if original_seq else align_seq
    m = qlen - soft_l - soft_r
    unmapped = LinearHit(name, 4, "*", 0, 0, "*", emit.decode(), "*")
    if cid < 0 or m < 1:
        return unmapped
    mid = f"{m}M"
    if indel_op == 1 and indel_len > 0 and 0 <= indel_at <= m:
        right = m - indel_at - indel_len
        if right < 0:
            return unmapped
        mid = (f"{indel_at}M" if indel_at > 0 else "") + f"{indel_len}I" + (f"{right}M" if right > 0 else "")
    elif indel_op == 2 and indel_len > 0 and 0 <= indel_at <= m:
        right = m - indel_at
        mid = (f"{indel_at}M" if indel_at > 0 else "") + f"{indel_len}D" + (f"{right}M" if right > 0 else "")
    cigar = (f"{soft_l}S" if soft_l > 0 else "") + mid + (f"{soft_r}S" if soft_r > 0 else "")
    if flag & 16:
        emit = reverse_complement(emit)
    return LinearHit(name, flag, canonical_contig(names[cid]), start0 + 1, mapq, cigar, emit.decode(), "*")


def pair_hits(h1: LinearHit, h2: LinearHit) -> Tuple[LinearHit, LinearHit]:
    """PE flags and mate fields (linear_extend.mojo:263-294)."""
    a, b = h1, h2
    a.flag |= 1 | 64
    b.flag |= 1 | 128
    if a.contig != "*" and b.contig != "*" and a.contig == b.contig:
        a.flag |= 2
        b.flag |= 2
        a.rnext = b.rnext = "="
        a.pnext = b.pos
        b.pnext = a.pos
        tlen = abs(b.pos - a.pos) + len(b.seq)
        a.tlen = tlen
        b.tlen = -tlen
    else:
        if a.contig == "*":
            b.flag |= 8
        if b.contig == "*":
            a.flag |= 8
        if a.contig != "*":
            b.rnext = a.contig
            b.pnext = a.pos
        if b.contig != "*":
            a.rnext = b.contig
            a.pnext = b.pos
    return a, b


def hit_to_sam_line(h: LinearHit, rg_id: str) -> str:
    return (f"{h.query_name}\t{h.flag}\t{h.contig}\t{h.pos}\t{h.mapq}\t{h.cigar}\t{h.rnext}\t"
            f"{h.pnext}\t{h.tlen}\t{h.seq}\t{h.qual}\tRG:Z:{rg_id}")


def sam_header(names: List[str], lengths: List[int]) -> str:
    """@HD unsorted, @SQ canonical and de-duplicated, @RG from env, @PG (:165-208)."""
    out = ["@HD\tVN:1.6\tSO:unsorted"]
    seen = set()
    for n, ln in zip(names, lengths):
        c = canonical_contig(n)
        if c not in seen:
            seen.add(c)
            out.append(f"@SQ\tSN:{c}\tLN:{ln}")
    rg = env_str("METHYLGRAPHER_RG_ID", "mojo1")
    sm = env_str("METHYLGRAPHER_RG_SM", "sample")
    lb = env_str("METHYLGRAPHER_RG_LB", "lib1")
    pl = env_str("METHYLGRAPHER_RG_PL", "ILLUMINA")
    out.append(f"@RG\tID:{rg}\tSM:{sm}\tLB:{lb}\tPL:{pl}\tPU:{rg}")
    out.append("@PG\tID:MojoFq2bamMeth\tPN:MojoFq2bamMeth\tVN:0.1.0-mialign")
    return "\n".join(out) + "\n"


def knobs_from_env(k: int, max_occ: int):
    """METHYLGRAPHER_LINEAR_* knobs with the reference defaults (:1729-1751),
    except MAX_INDEL: unset (or below 1) keeps the extension gapless
    (indel_cap 0, this build's default); 1..4 turns on the single-indel branch
    at that cap, as the reference's default of 4 does."""
    vote_occ = env_int("METHYLGRAPHER_LINEAR_VOTE_OCC", 256)
    rc_raw = env_str("METHYLGRAPHER_LINEAR_RC", "1").lower()
    return default_params(
        k=k,
        stride=env_int("METHYLGRAPHER_LINEAR_SEED_STRIDE", 5),
        max_occ=max_occ,
        vote_occ=vote_occ,
        vote_fast=env_int("METHYLGRAPHER_LINEAR_VOTE_OCC_FAST", vote_occ),
        max_diff=env_int("METHYLGRAPHER_LINEAR_MAX_DIFF", 6),
        max_soft=env_int("METHYLGRAPHER_LINEAR_MAX_SOFT", 8),
        max_indel=env_int("METHYLGRAPHER_LINEAR_MAX_INDEL", 0),
        do_rc=0 if rc_raw in ("0", "false", "no") else 1,
    )


def map_fastq_dense_gpu_locate(index, fq1: str, out_sam: str, device: str = "auto", fq2: str = "",
                               bs_r1: str = "", bs_r2: str = "") -> int:
    """Map FASTQ (SE or PE) against a resident dense-v1 index; returns the number
    of mapped records (linear_gpu_locate.mojo:291-299)."""
    if not getattr(index, "dense", False) or index.gpu is None:
        raise MialignError("map_fastq_dense_gpu_locate requires a resident dense-v1 index")
    for mode in (bs_r1, bs_r2):
        if mode not in BS_MODES:
            raise MialignError(f"unknown conversion mode {mode!r} (C2T, G2A or none)")
    k = index.k
    batch_size = env_int("METHYLGRAPHER_LINEAR_READ_BATCH", 262144)
    prm = knobs_from_env(k, index.max_occ())
    rg = env_str("METHYLGRAPHER_RG_ID", "mojo1")
    paired = bool(fq2)
    fh1 = _open_fastq(fq1)
    fh2 = _open_fastq(fq2) if paired else None
    out = sys.stdout if out_sam == "-" else open(out_sam, "w")
    n_mapped = n_reads = 0
    t0 = time.perf_counter()
    try:
        out.write(sam_header(index.contig_names, index.contig_lengths))
        while True:
            b1, b2 = read_pe_batch(fh1, fh2, paired, batch_size, bs_r1, bs_r2)
            if not b1:
                break
            recs = b1 + b2
            rows, lens, max_len = pack_batch([r.seq for r in recs], k)
            n1 = len(b1)
            got = map_batch(index.gpu, rows, lens, max_len, encoding=MA_ENC_ASCII, params=prm,
                            n_r1=n1)
            hits = [
                hit_from_gpu(index.contig_names, r.name, r.seq, r.original_seq, int(got["cid"][i]),
                             int(got["pos"][i]), int(got["mapq"][i]), int(got["flag"][i]),
                             int(got["sl"][i]), int(got["sr"][i]), int(got["iop"][i]), int(got["iat"][i]),
                             int(got["ilen"][i]))
                for i, r in enumerate(recs)
            ]
            for h, r in zip(hits, recs):
                h.qual = r.qual
            lines = []
            if not paired:
                for h in hits:
                    n_mapped += h.contig != "*"
                    lines.append(hit_to_sam_line(h, rg))
                n_reads += n1
            else:
                for j in range(n1):
                    p1, p2 = pair_hits(hits[j], hits[n1 + j])
                    n_mapped += (p1.contig != "*") + (p2.contig != "*")
                    lines.append(hit_to_sam_line(p1, rg))
                    lines.append(hit_to_sam_line(p2, rg))
                n_reads += 2 * n1
            out.write("\n".join(lines) + "\n")
    finally:
        fh1.close()
        if fh2 is not None:
            fh2.close()
        if out is not sys.stdout:
            out.close()
    dt = time.perf_counter() - t0
    print(f"mialign GPU map: reads={n_reads} mapped={n_mapped} wall_s={dt:.2f}", flush=True)
    return n_mapped
