"""The linear Align entry: map_fastq_to_sam and its `-key value` CLI.

Mirrors fq2bam-meth/src/linear_mapper.mojo:129-354. Same argv
(-ref -fq1 -out_sam [-fq2] [-device] [-k] [-cache_dir] [-bs_r1] [-bs_r2]
[-build_cache_only]) and environment (METHYLGRAPHER_LINEAR_ENGINE, ...), same
dense-v1 cache directory. Exit code 0 on success; any failure exits non-zero
(the orchestrator's signal to fall back, fq2bam_meth.py:808-838).

`-device cpu` runs the reference's forward-only host loop (its science, SAM
text and pairing; linear_mapper.mojo:209-311) with the per-read vote and
verify on the GPU (linear_host_loop.py); every other device value runs the
hash-seeded parity science on the GPU. The fm / speed engines are not part of
this build and fail with a message.
"""

from __future__ import annotations

import os
import sys
from pathlib import Path
from typing import List, Sequence

from ._lib import MialignError, device_count
from .linear_gpu_locate import map_fastq_dense_gpu_locate
from .linear_index import LinearIndex
from .mojo_linear_pack import ensure_dense_pack


def get_kv_value(args: Sequence[str], key: str, default: str) -> str:
    """Value after `-key` (methylgrapher/src/utility.mojo:151-165)."""
    for i, a in enumerate(args):
        if len(a) > 1 and a.startswith("-") and a[1:] == key:
            return args[i + 1] if i + 1 < len(args) else default
    return default


def map_fastq_to_sam(ref_fasta: str, fq1: str, out_sam: str, device: str = "auto", k: int = 15,
                     fq2: str = "", cache_dir: str = "", bs_r1: str = "", bs_r2: str = "",
                     build_cache_only: bool = False) -> int:
    engine = os.environ.get("METHYLGRAPHER_LINEAR_ENGINE", "parity").lower()
    if engine in ("fm", "bwa-mem", "bwamem", "speed", "fast"):
        raise MialignError(f"engine {engine!r} is not part of this build (parity only)")
    dev = device.lower()
    if device_count() < 1 and not build_cache_only:
        raise MialignError("no GPU visible (every device choice maps on the GPU here)")
    if not cache_dir:
        cache_dir = str(Path(ref_fasta).with_name(Path(ref_fasta).name + f".mojo_linear_k{k}"))
    ensure_dense_pack(c2t_fasta=Path(ref_fasta), cache_dir=Path(cache_dir), k=k)
    if build_cache_only:
        print(f"mialign cache ready -> {cache_dir}", flush=True)
        return 0
    index = LinearIndex(k)
    if not index.load_cache(cache_dir):
        raise MialignError(f"dense-v1 pack incomplete: {cache_dir}")
    try:
        print(f"mialign index contigs={index.contig_count()} bases={index.total_bases()} "
              f"postings={index.n_postings} device={device}", flush=True)
        if dev == "cpu":
            from .linear_host_loop import map_fastq_host_loop

            return map_fastq_host_loop(index, fq1, out_sam, fq2, bs_r1, bs_r2)
        if out_sam.lower().endswith(".bam"):
            # native BAM like the reference's fm driver (_path_is_bam, linear_gpu_fm.mojo:84-88):
            # coordinate-sorted, duplicates marked, every stage on the GPU; pairs as R1, R2
            from .fq2bam import fastq_to_bam
            from .linear_gpu_locate import env_int, env_str, knobs_from_env

            prm = knobs_from_env(k, index.max_occ())
            if prm.max_indel != 0:
                # BAM records are written with slS mM srS CIGARs only (the reference's
                # native-BAM records carry no indel); indel hits need SAM output
                print(f"mialign: METHYLGRAPHER_LINEAR_MAX_INDEL={prm.max_indel} ignored for BAM output "
                      "(gapless extension); write SAM for the single-indel branch", file=sys.stderr, flush=True)
                prm.max_indel = 0
            stats = fastq_to_bam(index, fq1, out_sam, bs_r1, fq2=fq2, bs_r2=bs_r2, params=prm,
                                 rescue_window=env_int("METHYLGRAPHER_FM_RESCUE_WIN", 512),
                                 level=env_int("METHYLGRAPHER_BAM_LEVEL", 1),
                                 device_bgzf=os.environ.get("MIALIGN_BGZF", "device").lower() != "host",
                                 rg_id=env_str("METHYLGRAPHER_RG_ID", "mojo1"),
                                 sm=env_str("METHYLGRAPHER_RG_SM", "sample"),
                                 lb=env_str("METHYLGRAPHER_RG_LB", "lib1"),
                                 pl=env_str("METHYLGRAPHER_RG_PL", "ILLUMINA"))
            print(f"mialign GPU fq2bam: {stats}", flush=True)
            return stats["mapped"]
        return map_fastq_dense_gpu_locate(index, fq1, out_sam, device, fq2, bs_r1, bs_r2)
    finally:
        index.close()


def run_mojo_linear_cli(args: List[str]) -> int:
    ref_fa = get_kv_value(args, "ref", "")
    fq1 = get_kv_value(args, "fq1", "")
    fq2 = get_kv_value(args, "fq2", "")
    out_sam = get_kv_value(args, "out_sam", "aligned.sam")
    device = get_kv_value(args, "device", "auto")
    k = int(get_kv_value(args, "k", "15"))
    cache_dir = get_kv_value(args, "cache_dir", "")
    bs_r1 = get_kv_value(args, "bs_r1", "")
    bs_r2 = get_kv_value(args, "bs_r2", "")
    build_cache_only = any(a in ("-build_cache_only", "--build_cache_only") for a in args)
    if not ref_fa:
        raise MialignError("MojoLinearMap requires -ref")
    if not build_cache_only and not fq1:
        raise MialignError("MojoLinearMap requires -fq1 (or pass -build_cache_only)")
    map_fastq_to_sam(ref_fa, fq1, out_sam, device, k, fq2, cache_dir, bs_r1, bs_r2, build_cache_only)
    return 0


def main(argv: List[str] | None = None) -> int:
    try:
        return run_mojo_linear_cli(list(sys.argv[1:] if argv is None else argv))
    except (MialignError, OSError, ValueError, RuntimeError) as e:
        print(f"mialign: {e}", file=sys.stderr, flush=True)
        return 1


if __name__ == "__main__":
    sys.exit(main())
