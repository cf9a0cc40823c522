"""The reference's `-device cpu` mapping loop, with its mapping on the GPU.

fq2bam-meth/src/linear_mapper.mojo:209-311 reads FASTQ in batches, converts
each read (_bs_convert, :44-61), maps it forward-only (extend_read_with_seeds,
linear_extend.mojo:143-211) and writes SAM lines with the pack's own contig
names (_write_sam_header, linear_mapper.mojo:99-117; hit_to_sam_line,
linear_extend.mojo:297-320), pairing mates with pair_hits (:263-294). Here the
per-read vote and verify run in ma_map_forward's kernel; this module keeps
the loop's batching, conversion, SAM text and pairing.

Batches are METHYLGRAPHER_LINEAR_READ_BATCH reads (default 262144 here, 4096
in the reference; the records do not depend on it).
"""

from __future__ import annotations

import os
from typing import List

import numpy as np

from ._lib import MialignError, default_params
from .core import map_forward
from .linear_gpu_locate import LinearHit, _open_fastq, env_int, pair_hits, read_one

MODES = {"": 0, "none": 0, "C2T": 1, "G2A": 2}
MAX_READ = 256


def host_loop_convert(seq: bytes, mode: str) -> bytes:
    """_bs_convert of the host loop (linear_mapper.mojo:44-61): for C2T a C or
    c becomes T, for G2A a G or g becomes A (upper case), other bytes kept."""
    if mode == "C2T":
        return seq.replace(b"c", b"T").replace(b"C", b"T")
    if mode == "G2A":
        return seq.replace(b"g", b"A").replace(b"G", b"A")
    return seq


def sam_header_host_loop(index) -> str:
    """@HD, one @SQ per pack contig under its own name, @PG (:99-117)."""
    out = ["@HD\tVN:1.6\tSO:unsorted"]
    for i in range(index.contig_count()):
        out.append(f"@SQ\tSN:{index.contig_name(i)}\tLN:{index.contig_length(i)}")
    out.append("@PG\tID:MojoFq2bamMeth\tPN:MojoFq2bamMeth\tVN:0.1.0-mialign")
    return "\n".join(out) + "\n"


def hit_line(h: LinearHit) -> str:
    """hit_to_sam_line (linear_extend.mojo:297-320): eleven fields, no tags."""
    return (f"{h.query_name}\t{h.flag}\t{h.contig}\t{h.pos}\t{h.mapq}\t{h.cigar}\t{h.rnext}\t"
            f"{h.pnext}\t{h.tlen}\t{h.seq}\t{h.qual}")


def _hits(index, recs: dict, names: List[str], seqs: List[bytes], quals: List[str]) -> List[LinearHit]:
    out = []
    for i, nm in enumerate(names):
        seq = seqs[i].decode(errors="replace")
        cid = int(recs["cid"][i])
        if cid < 0:
            out.append(LinearHit(nm, 4, "*", 0, 0, "*", seq, quals[i]))
            continue
        sl, sr = int(recs["sl"][i]), int(recs["sr"][i])
        alen = len(seqs[i]) - sl - sr
        cigar = (f"{sl}S" if sl else "") + f"{alen}M" + (f"{sr}S" if sr else "")
        out.append(LinearHit(nm, 0, index.contig_name(cid), int(recs["pos"][i]) + 1, int(recs["mapq"][i]), cigar,
                             seq, quals[i]))
    return out


def _map(index, seqs: List[bytes], prm) -> dict:
    lens = np.array([len(s) for s in seqs], np.uint32)
    max_len = max(int(lens.max()) if lens.size else 1, 1)
    if max_len > MAX_READ:
        raise MialignError(f"reads longer than {MAX_READ} bases are not supported (got {max_len})")
    rows = np.full((len(seqs), max_len), ord("N"), np.uint8)
    for i, s in enumerate(seqs):
        if s:
            rows[i, :len(s)] = np.frombuffer(s, np.uint8)
    # the reads are converted already (as the loop converts them): no conversion on the device
    return map_forward(index.gpu, rows, lens, max_len, params=prm)


def map_fastq_host_loop(index, fq1: str, out_sam: str, fq2: str = "", bs_r1: str = "", bs_r2: str = "") -> int:
    """SAM of SE or PE FASTQ through the host loop's science; returns the
    number of mapped records (linear_mapper.mojo:309)."""
    for mode in (bs_r1, bs_r2):
        if mode not in MODES:
            raise MialignError(f"unknown conversion mode {mode!r} (C2T, G2A or none)")
    if index.gpu is None:
        raise MialignError("the host loop needs a resident dense-v1 pack")
    # the loop's knobs: seed stride (_seed_stride, linear_extend.mojo:117-126), max_diff and
    # max_soft (:156-161), read as plain integers like the reference does
    prm = default_params(k=index.k, stride=env_int("METHYLGRAPHER_LINEAR_SEED_STRIDE", 5), max_occ=index.max_occ(),
                         max_diff=int(os.environ.get("METHYLGRAPHER_LINEAR_MAX_DIFF", "6")),
                         max_soft=int(os.environ.get("METHYLGRAPHER_LINEAR_MAX_SOFT", "8")))
    batch = env_int("METHYLGRAPHER_LINEAR_READ_BATCH", 262144)
    paired = bool(fq2)
    fh1 = _open_fastq(fq1)
    fh2 = _open_fastq(fq2) if paired else None
    n_mapped = n_reads = n_batches = 0
    print(f"MojoLinear map start batch={batch} max_occ={index.max_occ()} paired={paired}", flush=True)
    try:
        with open(out_sam, "w") as out:
            out.write(sam_header_host_loop(index))
            while True:
                b1, b2 = [], []
                while len(b1) < batch:
                    r1 = read_one(fh1)
                    if r1 is None:
                        break
                    b1.append(r1)
                    if paired:
                        r2 = read_one(fh2)
                        if r2 is None:
                            raise MialignError("paired FASTQ length mismatch (R2 ended early)")
                        b2.append(r2)
                if not b1:
                    break
                s1 = [host_loop_convert(r.original_seq, bs_r1) for r in b1]
                h1 = _hits(index, _map(index, s1, prm), [r.name for r in b1], s1, [r.qual for r in b1])
                if paired:
                    s2 = [host_loop_convert(r.original_seq, bs_r2) for r in b2]
                    h2 = _hits(index, _map(index, s2, prm), [r.name for r in b2], s2, [r.qual for r in b2])
                    lines = []
                    for a, b in zip(h1, h2):
                        p1, p2 = pair_hits(a, b)
                        n_mapped += (p1.contig != "*") + (p2.contig != "*")
                        lines.append(hit_line(p1))
                        lines.append(hit_line(p2))
                    n_reads += 2 * len(h1)
                else:
                    n_mapped += sum(h.contig != "*" for h in h1)
                    lines = [hit_line(h) for h in h1]
                    n_reads += len(h1)
                out.write("\n".join(lines) + "\n")
                n_batches += 1
                if n_batches == 1 or n_batches % 5 == 0:
                    print(f"MojoLinear progress batches={n_batches} reads={n_reads} mapped={n_mapped}", flush=True)
    finally:
        fh1.close()
        if fh2 is not None:
            fh2.close()
    print(f"wrote SAM -> {out_sam} mapped_records={n_mapped} reads={n_reads}", flush=True)
    return n_mapped
