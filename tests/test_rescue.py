"""Mate rescue in the oracle (fm_mate_rescue_kernel, fq2bam-meth/src/
linear_gpu_fm.mojo:1839-2011), on hand-built pairs whose results are worked
out from the rule: with exactly one mate mapped, the other is scanned gapless
within +-window (clamped to 32..1024) of the mate on the mate's contig, the
strand opposite the mate's first; the first least mismatch count within
budget (clamp(max_diff, 2, len / 2)) wins, a later strand only when strictly
better; with none, a single-base deletion at the best window, split points
every 4 bases; MAPQ 60 / 40 / 20 for nm 0 / <= 2 / more."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

ACGT = b"ACGT"


def _rand(rng, n):
    return bytearray(np.frombuffer(ACGT, np.uint8)[rng.integers(0, 4, n)].tobytes())


def _rc(s: bytes) -> bytes:
    return s.translate(bytes.maketrans(b"ACGT", b"TGCA"))[::-1]


def _mut(s: bytes, *at) -> bytes:
    b = bytearray(s)
    for j in at:
        b[j] = ACGT[(ACGT.index(b[j]) + 1) % 4]
    return bytes(b)


class OracleBackend:
    """Index and rescue through the oracle (test_gpu_rescue.py swaps in the GPU)."""

    def index(self, ref: bytes):
        seq = np.frombuffer(bytes(ref), np.uint8).copy()
        off = np.array([0, len(seq)], np.uint64)
        km, of, po = orc.build_dense(seq, off, 21)
        return orc.OracleIndex(km, of, po, seq, off)

    def rescue(self, ix, rows, lens, rec, window, max_diff):
        return orc.mate_rescue(ix, rows, lens, rec, window, max_diff)


@pytest.fixture(scope="module")
def backend():
    return OracleBackend()


def _pair(be, ix, r1: bytes, r2: bytes, mate=(0, 5000, 0), window=512, max_diff=6, r1_mapped=True):
    """R1 mapped at mate = (cid, pos, flag) (or R2 mapped there when
    r1_mapped is False), the other mate unmapped; returns the rescued record of
    the unmapped mate as (cid, pos, flag, nm, mapq, sl, sr)."""
    L = max(len(r1), len(r2))
    rows = np.full((2, L), ord("N"), np.uint8)
    rows[0, :len(r1)] = np.frombuffer(r1, np.uint8)
    rows[1, :len(r2)] = np.frombuffer(r2, np.uint8)
    lens = np.array([len(r1), len(r2)], np.uint32)
    rec = orc.new_recs(2)
    m, u = (0, 1) if r1_mapped else (1, 0)
    rec["cid"][:] = -1
    rec["flag"][:] = 4
    rec["cid"][m], rec["pos"][m], rec["flag"][m], rec["mapq"][m] = mate[0], mate[1], mate[2], 60
    out = be.rescue(ix, rows, lens, rec, window, max_diff)
    for f in rec:  # the mapped mate is never touched
        assert out[f][m] == rec[f][m]
    return tuple(int(out[f][u]) for f in ("cid", "pos", "flag", "nm", "mapq", "sl", "sr"))


@pytest.fixture(scope="module")
def world():
    rng = np.random.default_rng(7)
    ref = _rand(rng, 20_000)
    return rng, ref


def test_opposite_strand_first_exact(world, backend):
    rng, ref = world
    ix = backend.index(ref)
    r2 = _rc(bytes(ref[5200:5300]))  # the fragment's other end, reverse strand
    assert _pair(backend, ix, bytes(ref[5000:5100]), r2) == (0, 5200, 16, 0, 60, 0, 0)


def test_same_strand_when_opposite_has_none(world, backend):
    rng, ref = world
    ix = backend.index(ref)
    r2 = _mut(bytes(ref[4850:4950]), 30)
    assert _pair(backend, ix, bytes(ref[5000:5100]), r2) == (0, 4850, 0, 1, 40, 0, 0)


def test_later_strand_only_when_strictly_better(backend):
    rng = np.random.default_rng(8)
    ref = _rand(rng, 20_000)
    s = bytes(_rand(rng, 100))
    ref[5300:5400] = _rc(_mut(s, 10, 60))     # opposite strand (searched first): nm 2
    ref[4700:4800] = s                        # mate's strand: nm 0, strictly better
    ix = backend.index(ref)
    assert _pair(backend, ix, bytes(ref[5000:5100]), s) == (0, 4700, 0, 0, 60, 0, 0)
    ref[4700:4800] = _mut(s, 20, 70)          # now as good, not better: the first stays
    ix = backend.index(ref)
    assert _pair(backend, ix, bytes(ref[5000:5100]), s) == (0, 5300, 16, 2, 40, 0, 0)


def test_window_bounds_and_clamp(world, backend):
    rng, ref0 = world
    ref = bytearray(ref0)
    s = bytes(ref[5600:5700])
    ix = backend.index(ref)
    r1 = bytes(ref[5000:5100])
    unm = (-1, 0, 4, 0, 0, 0, 0)
    assert _pair(backend, ix, r1, _rc(s), window=512) == unm          # 600 bases away: outside +-512
    assert _pair(backend, ix, r1, _rc(s), window=600) == (0, 5600, 16, 0, 60, 0, 0)
    far = bytes(ref[6100:6200])
    assert _pair(backend, ix, r1, _rc(far), window=5000) == unm       # clamped to 1024: 1100 away
    near = bytes(ref[6000:6100])
    assert _pair(backend, ix, r1, _rc(near), window=5000) == (0, 6000, 16, 0, 60, 0, 0)
    assert _pair(backend, ix, r1, _rc(bytes(ref[5010:5030])), window=512) == unm  # shorter than 24


def test_rescue_from_r2_and_mapq_bands(world, backend):
    rng, ref = world
    ix = backend.index(ref)
    r1 = _rc(_mut(bytes(ref[5100:5200]), 5, 50, 90))  # nm 3 -> MAPQ 20
    got = _pair(backend, ix, r1, bytes(ref[5000:5100]), r1_mapped=False)
    assert got == (0, 5100, 16, 3, 20, 0, 0)


def test_single_base_deletion_on_the_four_base_grid(backend):
    """The reference has one base more than the read after read base 89 (so
    at 90): gapless at S the read's last 10 bases are shifted, 10 > budget 6
    mismatches (no adjacent equal bases there), every other window near 75;
    the deletion grid's best split is 88 (1 + read[88] vs ref[89] + read[89]
    vs the extra base = 3), where a grid of every base would find 90 (nm 1)."""
    rng = np.random.default_rng(9)
    while True:
        read = bytes(_rand(rng, 100))
        tail = read[86:]
        if all(tail[i] != tail[i + 1] for i in range(len(tail) - 1)):
            break
    extra = next(b for b in ACGT if b not in (read[89], read[90]))
    ref = _rand(rng, 20_000)
    S = 5300
    ref[S:S + 90] = read[:90]
    ref[S + 90] = extra
    ref[S + 91:S + 101] = read[90:]
    ix = backend.index(ref)
    got = _pair(backend, ix, bytes(ref[5000:5100]), read, mate=(0, 5000, 16))  # mate reverse: forward first
    assert got == (0, S, 0, 3, 20, 0, 0)


def test_pairs_with_both_or_no_mates_mapped_are_left_alone(world, backend):
    rng, ref = world
    ix = backend.index(ref)
    rows = np.frombuffer(bytes(ref[5000:5100]) + _rc(bytes(ref[5200:5300])), np.uint8).reshape(2, 100).copy()
    lens = np.array([100, 100], np.uint32)
    for cids in ((0, 0), (-1, -1)):
        rec = orc.new_recs(2)
        rec["cid"][:] = cids
        rec["flag"][:] = [0 if c >= 0 else 4 for c in cids]
        rec["pos"][:] = [5000, 0]
        out = backend.rescue(ix, rows, lens, rec, 512, 6)
        for f in rec:
            assert np.array_equal(out[f], rec[f])
