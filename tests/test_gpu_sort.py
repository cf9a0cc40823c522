"""Sort + markdup on the device against the oracle's stable sorts and markdup
scan (linear_gpu_sort.mojo:156-551), bit for bit: is_dup, perm and dup_perm.

Covers the sort's own decisions: 128-bit duplicate keys whose low word moves
with the high word (pairs' ordered ends), single-end keys whose low word is
implied by the high word (0, or the record index when unmapped) and is not
sorted, high bytes that only follow the unmapped sentinels (no pass of their
own) and a key breaking that pattern, sizes around the radix tiles (2048 / 4096 records), keys that are all
equal (no pass runs), one and two records, and the device entry point without
a dup_perm output, followed by its status check."""

import ctypes as C

import numpy as np
import pytest

import mialign
from mialign._lib import check, lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def say(msg):
    print(f"[test] {msg}", flush=True)


def _both(coord, dhi, dlo, score, pair, do_markdup=True):
    got = mialign.sort_markdup(coord, dhi, dlo, score, pair, do_markdup=do_markdup)
    want = orc.sort_markdup(coord, dhi, dlo, score, pair, do_markdup=do_markdup)
    for name, g, w in zip(("is_dup", "perm", "dup_perm"), got, want):
        bad = np.flatnonzero(g != w)
        assert bad.size == 0, f"{name}: {bad.size} differ (first {bad[:5]})"
    return want


def _pe_like(rng, n, span):
    """Pairs' keys: coord per record, (dhi, dlo) the ordered ends, shared by
    both records of a pair; some pairs without a target (~0, index)."""
    hi = rng.integers(0, span, n).astype(np.uint64) << np.uint64(17)
    lo = hi + (rng.integers(0, span, n).astype(np.uint64) << np.uint64(3))
    none = rng.random(n) < 0.05
    hi[none] = M64
    lo[none] = np.arange(n, dtype=np.uint64)[none] ^ np.uint64(1)  # not the index: lo is not implied
    coord = rng.integers(0, 1 << 40, n).astype(np.uint64)
    score = rng.integers(0, 1 << 16, n).astype(np.uint32)
    pair = (np.arange(n) // 2).astype(np.uint32)
    return coord, hi, lo, score, pair


@pytest.mark.parametrize("n", [1, 2, 3, 100, 2047, 2048, 2049, 4095, 4096, 4097, 100_000, 1_000_003])
def test_sort_markdup_128bit_keys(n):
    say(f"test_sort_markdup_128bit_keys n={n}")
    rng = np.random.default_rng(n)
    want = _both(*_pe_like(rng, n, max(2, n // 8)))
    if n >= 100:
        assert want[0].sum() > 0


@pytest.mark.parametrize("n", [5, 4097, 300_000])
def test_sort_markdup_se_keys_implied_low_word(n):
    say(f"test_sort_markdup_se_keys_implied_low_word n={n}")
    rng = np.random.default_rng(100 + n)
    dhi = (rng.integers(1, 25, n).astype(np.uint64) << np.uint64(48)) | (
        rng.integers(0, 2, n).astype(np.uint64) << np.uint64(47)) | rng.integers(0, max(4, n // 4), n).astype(np.uint64)
    dlo = np.zeros(n, np.uint64)
    un = rng.random(n) < 0.1
    dhi[un] = M64
    dlo[un] = np.arange(n, dtype=np.uint64)[un]
    coord = np.where(un, (np.uint64(0xFFFFFFFF) << np.uint64(32)) | np.arange(n, dtype=np.uint64),
                     rng.integers(0, 1 << 36, n).astype(np.uint64))
    score = rng.integers(0, 1 << 14, n).astype(np.uint32)
    pair = np.arange(n, dtype=np.uint32)
    _both(coord, dhi, dlo, score, pair)
    _both(coord, dhi, dlo, score, pair, do_markdup=False)


@pytest.mark.parametrize("break_one", [False, True])
def test_sort_markdup_sentinel_bytes_follow(break_one):
    """High bytes that are 0xFF exactly where the byte below them is, and 0
    elsewhere (the all-ones high words of unmapped records), need no pass of
    their own: the records are already in their order. One key breaking the
    pattern (byte 7 all ones over a byte 6 of 0) makes the pass run again."""
    say(f"test_sort_markdup_sentinel_bytes_follow break_one={break_one}")
    rng = np.random.default_rng(7 + int(break_one))
    n = 200_000
    idx = np.arange(n, dtype=np.uint64)
    coord = rng.integers(0, 1 << 32, n).astype(np.uint64) | (rng.integers(0, 200, n).astype(np.uint64) << np.uint64(32))
    dhi = (rng.integers(1, 200, n).astype(np.uint64) << np.uint64(48)) | (
        rng.integers(0, 2, n).astype(np.uint64) << np.uint64(47)) | rng.integers(0, 1 << 20, n).astype(np.uint64)
    dlo = np.zeros(n, np.uint64)
    un = rng.random(n) < 0.05
    coord[un] = (np.uint64(0xFFFFFFFF) << np.uint64(32)) | idx[un]
    dhi[un] = M64
    dlo[un] = idx[un]
    if break_one:
        coord[17] = (np.uint64(0xFF) << np.uint64(56)) | np.uint64(5)
        dhi[23] = (np.uint64(0xFF) << np.uint64(56)) | (np.uint64(3) << np.uint64(48)) | np.uint64(9)
        dlo[23] = np.uint64(0)
    score = rng.integers(0, 1 << 14, n).astype(np.uint32)
    pair = np.arange(n, dtype=np.uint32)
    _both(coord, dhi, dlo, score, pair)
    _both(coord, dhi, dlo, score, pair, do_markdup=False)


def test_sort_markdup_all_equal_keys():
    """One duplicate group of 2M records (every key equal): no radix pass runs,
    and the group's best is folded per wave, not per record."""
    say("test_sort_markdup_all_equal_keys")
    n = 2_000_000
    k = np.full(n, 12345, np.uint64)
    score = np.random.default_rng(3).integers(0, 8, n).astype(np.uint32)
    want = _both(k, k, np.zeros(n, np.uint64), score, np.arange(n, dtype=np.uint32))
    assert want[0].sum() == n - 1


def test_sort_markdup_dev_without_dup_perm():
    say("test_sort_markdup_dev_without_dup_perm")
    rng = np.random.default_rng(9)
    n = 250_000
    coord, dhi, dlo, score, pair = _pe_like(rng, n, n // 6)
    wd, wp, _ = orc.sort_markdup(coord, dhi, dlo, score, pair)
    st = mialign.Stream(0)
    dev = [mialign.DeviceArray.from_host(a) for a in (coord, dhi, dlo, score, pair)]
    d_dup = mialign.DeviceArray((n,), np.uint32)
    d_perm = mialign.DeviceArray((n,), np.uint32)
    sb = int(lib().ma_sort_scratch_bytes(n))
    scr = mialign.DeviceArray((sb,), np.uint8)
    for _ in range(2):  # the same scratch twice in a row
        check(lib().ma_sort_markdup_dev(st.handle, n, *[a.ptr for a in dev], d_dup.ptr, d_perm.ptr, 1, None,
                                        scr.ptr, sb), "ma_sort_markdup_dev")
        check(lib().ma_sort_dev_status(st.handle, n, scr.ptr, sb), "ma_sort_dev_status")
        assert np.array_equal(d_dup.download(), wd)
        assert np.array_equal(d_perm.download(), wp)
