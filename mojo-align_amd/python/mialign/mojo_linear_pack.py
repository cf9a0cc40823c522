"""dense-v1 k-mer pack: reader/writer and a GPU builder.

On-disk layout under a pack directory (the reference's format, so existing
packs load unchanged — fq2bam-meth/python/mojo_linear_pack.py:1-11):

  meta.json           format=dense-v1, k, n_keys, n_postings, contigs
  kmers.bin           uint64 LE, sorted unique MSB-first 2-bit k-mers
  offsets.bin         uint64 LE, n_keys + 1 CSR starts into postings
  postings.bin        uint32 LE (contig_id, pos0) pairs, ordered (key, contig, pos)
  sequences.bin       contig bases, upper-cased, no newlines
  contig_offsets.bin  uint64 LE, n_contigs + 1
  contigs.txt         one contig name per line

build_dense_pack() builds on the GPU (GpuIndex.build: one key per window,
stable radix sort, CSR by scan) instead of numpy's host mergesort; the bytes
are the same. ensure_dense_pack() keeps the reference's build-once contract
for concurrent workers (:98-157).
"""

from __future__ import annotations

import contextlib
import fcntl
import json
import os
import time
from pathlib import Path
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

PACK_FILES = ("meta.json", "kmers.bin", "offsets.bin", "postings.bin", "sequences.bin",
              "contig_offsets.bin")


def pack_is_complete(cache_dir: Path) -> bool:
    cache_dir = Path(cache_dir)
    return all((cache_dir / f).is_file() for f in PACK_FILES)


def read_fasta_contigs(path: Path) -> List[Tuple[str, bytes]]:
    """Contigs of a FASTA file as (name, bases), in file order. A record starts
    at a line beginning with '>'; its name is the header's first word
    ("unnamed" for a blank header); its bases are its lines with surrounding
    whitespace removed, joined and upper-cased. Lines before the first header
    belong to no record. (The contract of the reference's reader,
    fq2bam-meth/python/mojo_linear_pack.py:160-176.)"""
    records: List[Tuple[str, List[bytes]]] = []
    for line in Path(path).read_bytes().split(b"\n"):
        if line.startswith(b">"):
            words = line[1:].decode("ascii", errors="replace").split()
            records.append((words[0] if words else "unnamed", []))
        elif records:
            records[-1][1].append(line.strip())
    return [(name, b"".join(parts).upper()) for name, parts in records]


def contigs_to_arrays(contigs: Sequence[Tuple[str, bytes]]):
    names = [n for n, _ in contigs]
    off = np.zeros(len(contigs) + 1, dtype=np.uint64)
    total = 0
    for i, (_, s) in enumerate(contigs):
        off[i] = total
        total += len(s)
    off[-1] = total
    seq = np.frombuffer(b"".join(s for _, s in contigs), dtype=np.uint8).copy()
    return names, seq, off


def write_dense_dir(out_dir: Path, k: int, kmers: np.ndarray, offsets: np.ndarray,
                    postings: np.ndarray, seq: np.ndarray, contig_off: np.ndarray,
                    names: Sequence[str], source: str = "") -> dict:
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    (out_dir / "kmers.bin").write_bytes(np.ascontiguousarray(kmers, dtype="<u8").tobytes())
    (out_dir / "offsets.bin").write_bytes(np.ascontiguousarray(offsets, dtype="<u8").tobytes())
    (out_dir / "postings.bin").write_bytes(np.ascontiguousarray(postings, dtype="<u4").tobytes())
    (out_dir / "sequences.bin").write_bytes(np.ascontiguousarray(seq, dtype=np.uint8).tobytes())
    (out_dir / "contig_offsets.bin").write_bytes(np.ascontiguousarray(contig_off, dtype="<u8").tobytes())
    (out_dir / "contigs.txt").write_text("\n".join(names) + "\n", encoding="utf-8")
    meta = {
        "format": "dense-v1",
        "k": int(k),
        "n_keys": int(kmers.size),
        "n_postings": int(postings.size // 2),
        "contigs": list(names),
        "c2t_fasta": source,
        "endian": "little",
        "kmer_encoding": "2bit-ACGT",
        "posting": "u32_contig_id,u32_pos0",
        "n_contigs": len(names),
        "n_bases": int(contig_off[-1]),
        "sequences": "sequences.bin+contig_offsets.bin",
    }
    (out_dir / "meta.json").write_text(json.dumps(meta, indent=2) + "\n", encoding="utf-8")
    return meta


def read_meta(cache_dir: Path) -> dict:
    meta = json.loads((Path(cache_dir) / "meta.json").read_text(encoding="utf-8"))
    if meta.get("format") != "dense-v1":
        raise ValueError(f"{cache_dir}: not a dense-v1 pack")
    return meta


def build_dense_pack(c2t_fasta: Path, out_dir: Path, k: int = 15, device: int = 0) -> dict:
    """Build a dense-v1 pack for an (already converted) FASTA on the GPU."""
    from .core import GpuIndex

    if k < 1 or k > 31:
        raise ValueError(f"k must be in 1..31, got {k}")
    contigs = read_fasta_contigs(Path(c2t_fasta))
    if not contigs:
        raise RuntimeError(f"no contigs in {c2t_fasta}")
    names, seq, off = contigs_to_arrays(contigs)
    ix = GpuIndex.build(seq, off, k, device)
    kmers, offsets, postings = ix.export_dense()
    ix.close()
    if kmers.size == 0:
        raise RuntimeError("no valid k-mers (check alphabet / k)")
    return write_dense_dir(Path(out_dir), k, kmers, offsets, postings, seq, off, names,
                           str(c2t_fasta))


@contextlib.contextmanager
def _exclusive(lock_path: Path):
    """Hold an exclusive flock on lock_path (workers building the same pack
    queue here; flock, the lock the reference's workers take on the same file)."""
    fd = os.open(lock_path, os.O_RDWR | os.O_CREAT | os.O_APPEND, 0o644)
    try:
        fcntl.flock(fd, fcntl.LOCK_EX)
        yield fd
    finally:
        os.close(fd)  # closing the descriptor releases the lock


def ensure_dense_pack(*, c2t_fasta: Path, cache_dir: Path, k: int = 15,
                      log: Optional[Callable[[str], None]] = None) -> Path:
    """A complete dense-v1 pack at cache_dir, built from c2t_fasta when missing
    (the contract of fq2bam-meth/python/mojo_linear_pack.py:98-157): workers
    serialize on `<cache_dir>.lock` and whoever gets the lock first builds;
    the others find the pack complete once they hold it.
    METHYLGRAPHER_LINEAR_CACHE_BUILD=0 makes a missing pack an error."""
    cache_dir = Path(cache_dir)
    say = log or (lambda m: print(m, flush=True))
    if pack_is_complete(cache_dir):
        return cache_dir
    policy = os.environ.get("METHYLGRAPHER_LINEAR_CACHE_BUILD", "1").strip().lower()
    if policy in ("0", "false", "no", "off"):
        raise RuntimeError(f"dense-v1 pack missing at {cache_dir} (METHYLGRAPHER_LINEAR_CACHE_BUILD={policy})")
    if not Path(c2t_fasta).is_file():
        raise FileNotFoundError(f"C2T FASTA required to build pack: {c2t_fasta}")
    cache_dir.parent.mkdir(parents=True, exist_ok=True)
    started = time.time()
    with _exclusive(Path(f"{cache_dir}.lock")) as fd:
        say(f"mialign pack: lock held after {time.time() - started:.1f}s")
        if not pack_is_complete(cache_dir):  # nobody built it while we waited
            build_dense_pack(Path(c2t_fasta), cache_dir, k=k)
            os.write(fd, f"built k={k} t={time.time():.0f}\n".encode())
    return cache_dir


def load_dense_dir(cache_dir: Path):
    """(meta, kmers, offsets, postings, seq, contig_off) as numpy memmaps."""
    cache_dir = Path(cache_dir)
    meta = read_meta(cache_dir)
    mm = lambda f, dt: np.memmap(cache_dir / f, dtype=dt, mode="r")  # noqa: E731
    return (meta, mm("kmers.bin", "<u8"), mm("offsets.bin", "<u8"), mm("postings.bin", "<u4"),
            mm("sequences.bin", np.uint8), mm("contig_offsets.bin", "<u8"))
