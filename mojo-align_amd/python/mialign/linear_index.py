"""LinearIndex: a dense-v1 pack made resident on the GPU.

Mirrors the parts of LinearIndex the GPU mapper uses
(fq2bam-meth/src/linear_index.mojo:98-150, 217-303, 425-434): contig names and
lengths, k, key/posting counts, max_occ() from METHYLGRAPHER_LINEAR_MAX_OCC.
The tables themselves live in HBM behind ma_index_open (one hash-bucket probe
per seed instead of the reference's binary search over kmers.bin).
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import List

import numpy as np

from .core import GpuIndex
from .mojo_linear_pack import pack_is_complete, read_meta


class LinearIndex:
    def __init__(self, k: int = 15):
        self.k = k
        self.contig_names: List[str] = []
        self.contig_lengths: List[int] = []
        self.n_keys = 0
        self.n_postings = 0
        self.dense = False
        self.cache_dir = ""
        self.gpu: GpuIndex | None = None

    def load_cache(self, cache_dir: str, device: int = 0) -> bool:
        """Open a dense-v1 pack; False when the directory is not a complete pack."""
        d = Path(cache_dir)
        if not pack_is_complete(d):
            return False
        meta = read_meta(d)
        self.k = int(meta["k"])
        self.n_keys = int(meta["n_keys"])
        self.n_postings = int(meta["n_postings"])
        self.contig_names = [str(x) for x in meta["contigs"]]
        off = np.fromfile(d / "contig_offsets.bin", dtype="<u8")
        self.contig_lengths = [int(b - a) for a, b in zip(off[:-1], off[1:])]
        self.gpu = GpuIndex.open(str(d), device)
        self.dense = True
        self.cache_dir = str(d)
        return True

    @classmethod
    def from_gpu(cls, gpu: GpuIndex, names: List[str], lengths: List[int]) -> "LinearIndex":
        ix = cls(gpu.k)
        ix.gpu = gpu
        ix.contig_names = list(names)
        ix.contig_lengths = list(lengths)
        ix.n_keys = gpu.n_keys
        ix.n_postings = gpu.n_postings
        ix.dense = True
        return ix

    def contig_count(self) -> int:
        return len(self.contig_names)

    def contig_name(self, i: int) -> str:
        return self.contig_names[i]

    def contig_length(self, i: int) -> int:
        return self.contig_lengths[i]

    def total_bases(self) -> int:
        return int(sum(self.contig_lengths))

    def max_occ(self) -> int:
        """Locate cap; values < 1 fall back to 512 (linear_index.mojo:425-434)."""
        n = int(os.environ.get("METHYLGRAPHER_LINEAR_MAX_OCC", "16384"))
        return 512 if n < 1 else n

    def close(self) -> None:
        if self.gpu is not None:
            self.gpu.close()
            self.gpu = None
