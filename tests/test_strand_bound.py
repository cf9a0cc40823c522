"""The bound behind the map kernel's strand order and strand skip, checked on
the oracle's per-strand records.

The kernel maps first the strand with fewer "broken" seeds -- seeds inside the
part of the read that no soft clip can remove (clip_cap = min(max_soft, L/4)
bases at either end) whose k-mer holds an N or is absent from the index -- and
maps the other strand only when it could still win the FW/RC merge
(linear_gpu_locate.mojo:2101-2113: the later strand wins only with nm <= the
earlier record's). That is exact only if every record of nm mismatches (an
indel's length counted in nm, linear_gpu_locate.mojo:1156, 1193, 1429) has at
most per_base * nm broken seeds, per_base = (k - 1) // stride + 1. Here the
bound is checked on every record the oracle makes for either orientation of
bisulfite reads with substitutions, N bases and planted indels, gapless and at
max_indel 4, and shown to be tight: records with exactly per_base * nm broken
seeds occur, so a smaller per_base would break the kernel's shortcut."""

from __future__ import annotations

import numpy as np
import pytest

from mialign import synth
from oracle import oracle as orc

K, STRIDE, MAX_SOFT = 21, 5, 8
PER_BASE = (K - 1) // STRIDE + 1
CODE = np.full(256, 255, np.uint8)
CODE[np.frombuffer(b"ACGT", np.uint8)] = np.arange(4, dtype=np.uint8)
COMP = np.arange(256, dtype=np.uint8)
COMP[np.frombuffer(b"ACGT", np.uint8)] = np.frombuffer(b"TGCA", np.uint8)


@pytest.fixture(scope="module")
def genome():
    ref = synth.gen_reference(23, 3, 400_000, repeat_frac=0.05)
    dual = synth.dual_convert(ref)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, K)
    return ref, dual, orc.OracleIndex(km, of, po, dual.seq, dual.contig_off), km


def broken_inner_seeds(row: np.ndarray, L: int, kmers: np.ndarray) -> int:
    clip = min(MAX_SOFT, L // 4)
    n = 0
    for q in range(0, L - K + 1, STRIDE):
        if q < clip or q + K > L - clip:
            continue
        c = CODE[row[q:q + K]]
        if (c > 3).any():
            n += 1
            continue
        key = 0
        for x in c:
            key = (key << 2) | int(x)
        j = np.searchsorted(kmers, np.uint64(key))
        n += not (j < kmers.size and kmers[j] == key)
    return n


def reads_under_test(ref, dual):
    """Converted reads (C->T) of both orientations' sources: bisulfite reads
    with 0.5% and 3% substitutions, some with N runs, and reads carrying one
    planted indel of 1..4 bases."""
    rows, lens = [], []
    for seed, p_err in ((31, 0.005), (32, 0.03)):
        r = synth.gen_reads(seed, ref, 600, 60, 150, p_err=p_err, max_len=150)
        rows.append(r.bases)
        lens.append(r.lens)
    b = np.concatenate(rows)
    ln = np.concatenate(lens)
    rng = np.random.default_rng(5)
    for i in rng.choice(b.shape[0], 150, replace=False):
        a = int(rng.integers(0, ln[i]))
        b[i, a:a + int(rng.integers(1, 4))] = ord("N")
    b[:, :][np.arange(150)[None, :] >= ln[:, None]] = ord("N")
    conv = orc.convert_reads(b, 0)
    c0 = dual.seq[dual.contig_off[0]:dual.contig_off[1]].tobytes()
    ind = np.full((300, 150), ord("N"), np.uint8)
    for i in range(300):
        s = int(rng.integers(1000, len(c0) - 1200))
        d = int(rng.integers(1, 5))
        a = int(rng.integers(25, 125))
        if i % 2:
            r = c0[s:s + a] + c0[s + a + d:s + 150 + d]
        else:
            r = c0[s:s + a] + bytes(rng.choice(np.frombuffer(b"ACT", np.uint8), d)) + c0[s + a:s + 150 - d]
        ind[i] = np.frombuffer(r, np.uint8)
    conv = np.concatenate([conv, ind])
    ln = np.concatenate([ln, np.full(300, 150, np.uint32)])
    return conv, ln


@pytest.mark.parametrize("max_indel", [0, 4])
def test_records_have_at_most_per_base_broken_seeds_per_mismatch(genome, max_indel):
    ref, dual, oix, kmers = genome
    conv, lens = reads_under_test(ref, dual)
    rcrow = np.full_like(conv, ord("N"))
    for i, L in enumerate(lens):
        rcrow[i, :L] = COMP[conv[i, :L]][::-1]
    prm = orc.params(max_indel=max_indel, do_rc=0)
    checked = tight = indel_recs = 0
    for rows in (conv, rcrow):  # each orientation mapped as the forward strand
        _, fw, _, _ = orc.map_batch(oix, rows, lens, prm, nthreads=8)
        for i in np.flatnonzero(fw["cid"] >= 0):
            L, nm = int(lens[i]), int(fw["nm"][i])
            nb = broken_inner_seeds(rows[i], L, kmers)
            assert nb <= PER_BASE * nm, (i, L, nm, nb, int(fw["iop"][i]), int(fw["ilen"][i]))
            checked += 1
            tight += nm > 0 and nb == PER_BASE * nm
            indel_recs += int(fw["iop"][i]) != 0
    assert checked >= 1000, checked
    assert tight >= 5, tight  # per_base - 1 would fail on these
    if max_indel:
        assert indel_recs >= 100, indel_recs
