"""Writes tests/golden/dense_v1_pack/: a tiny dense-v1 pack built by the
reference's own pack builder (fq2bam-meth/python/mojo_linear_pack.py,
build_dense_pack), for the tiny bisulfite-converted FASTA this script writes
(tests/golden/dense_v1_pack/ref.c2t.fa: three contigs with lower-case bases,
an N run and IUPAC codes, two of them the f / r conversion copies of one
chromosome). Run in the build container, where the reference is readable:

    python tests/golden/make_dense_v1_fixture.py

It pins the pack's on-disk layout (meta.json, kmers / offsets / postings,
sequences.bin, contig_offsets.bin) that no other check covers. The files it
writes are the fixture; nothing of the reference's source is copied. The
tests read the fixture only, never the reference.
"""

from __future__ import annotations

import random
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "dense_v1_pack"
REFERENCE_PY = Path("/root/reference/fq2bam-meth/python")


def fasta_text() -> str:
    rng = random.Random(20261204)
    top = "".join(rng.choice("ACGT") for _ in range(700))
    top = top[:200] + "NNNNNNNNNN" + top[210:400] + "acgtac" + top[406:500] + "RYK" + top[503:]
    f = top.upper().replace("C", "T")  # C->T copy (the reference's f<name>)
    r = top.upper().replace("G", "A")  # G->A copy (r<name>)
    other = "".join(rng.choice("ACGT") for _ in range(333)).replace("C", "T")
    lines = []
    for name, seq in (("fchr1 converted", f), ("rchr1", r), ("fchr2", other.lower()[:120] + other[120:])):
        lines.append(f">{name}")
        lines += [seq[i:i + 60] for i in range(0, len(seq), 60)]
    return "\n".join(lines) + "\n"


def main() -> int:
    sys.path.insert(0, str(REFERENCE_PY))
    import mojo_linear_pack as reference_pack  # noqa: E402

    if OUT.exists():
        shutil.rmtree(OUT)
    OUT.mkdir(parents=True)
    fa = OUT / "ref.c2t.fa"
    fa.write_text(fasta_text())
    build = OUT / "_build"
    reference_pack.build_dense_pack(fa, build, k=11)
    for p in sorted(build.iterdir()):
        if p.is_file():
            shutil.copy2(p, OUT / p.name)
    shutil.rmtree(build)
    print("wrote", sorted(p.name for p in OUT.iterdir()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
