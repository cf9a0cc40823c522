"""Single large batch at full size: a 100 Mbp reference in 10 contigs (seeded
ACGT, 1% CpG) converted both ways, 10M single-end 150 bp bisulfite reads as
2-bit packed bytes, the k=21 index resident in HBM. Every read is mapped in
one launch; the first 1M records equal the oracle's bit for bit, the device
sort keys of all 10M records equal the oracle's keys of the same records, and
the device coordinate order and duplicate flags of all 10M records equal the
oracle's stable sorts and markdup scan. The oracle's own records of the first
1M reads sort and mark the same on the device as in the oracle.
"""
import os

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import MA_BS_C2T, MA_ENC_2BIT
from mialign.linear_gpu_sort import sort_keys_se as gpu_sort_keys_se
from oracle import oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")
THREADS = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[test] {msg}", flush=True)


def first_diff(got, want, fields):
    for f in fields:
        bad = np.nonzero(got[f] != want[f])[0]
        if bad.size:
            i = int(bad[0])
            return f"field {f} differs on {bad.size} reads; first {i}: " + str(
                {g: (int(got[g][i]), int(want[g][i])) for g in fields})
    return None


def test_config2_full_batch_against_oracle():
    say("test_config2_full_batch_against_oracle: 100 Mbp x 2, 10M x 150 bp, one launch")
    ref = synth.gen_reference(1001, 10, 100_000_000)
    dual = synth.dual_convert(ref)
    tid = synth.tid_of_cid(dual.names)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    reads = synth.gen_reads(1002, ref, 10_000_000, 150, 150)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    say("reads generated; mapping 10M reads")
    got = mialign.map_batch(gix, packed, reads.lens, 150, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
    assert (got["cid"] >= 0).mean() > 0.99
    km, of, po = gix.export_dense()
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    ns = 1_000_000
    say(f"oracle on the first {ns} reads ({THREADS} threads)")
    want, _, _, cnt = orc.map_batch(oix, orc.convert_reads(reads.bases[:ns], 1), reads.lens[:ns],
                                    nthreads=THREADS)
    sub = {f: got[f][:ns] for f in FIELDS}
    msg = first_diff(sub, want, FIELDS)
    assert msg is None, msg
    assert cnt.reads_rc > 0.5 * ns  # the RC pass ran on most reads and its records were compared

    # sort + markdup of the oracle's records (1M): device == oracle
    keys = orc.sort_keys_se(want, reads.lens[:ns], tid)
    d, p, dp = mialign.sort_markdup(*keys)
    wd, wp, wdp = orc.sort_markdup(*keys)
    assert np.array_equal(p, wp) and np.array_equal(d, wd) and np.array_equal(dp, wdp)
    assert wd.sum() > 0.03 * ns  # 5% planted duplicates: the markdup scan has groups to resolve

    # all 10M records: device keys == oracle keys; device order and flags == oracle's
    say("keys, sort and markdup of all 10M records")
    gk = gpu_sort_keys_se(got, reads.lens, tid)
    ok = orc.sort_keys_se(got, reads.lens, tid)
    for name, a, b in zip(("coord", "dhi", "dlo", "score", "pair"), gk, ok):
        assert np.array_equal(a, b), f"device sort key {name} differs"
    d, p, _ = mialign.sort_markdup(*ok)
    wd, wp, _ = orc.sort_markdup(*ok)
    assert np.array_equal(p, wp), "coordinate order differs"
    assert np.array_equal(d, wd), "duplicate flags differ"
    say(f"10M records: {int(wd.sum())} duplicates, order and flags equal the oracle's")
    gix.close()


def test_sort_keys_truncated_target_table():
    """Index contigs outside the target table key as unmapped on the device, as
    in the oracle (the reference guards rid1 < len(rid_to_tid))."""
    say("test_sort_keys_truncated_target_table")
    rng = np.random.default_rng(5)
    n = 50_000
    rec = {
        "cid": rng.integers(-1, 8, n).astype(np.int32),
        "pos": rng.integers(0, 1 << 30, n).astype(np.uint32),
        "mapq": rng.choice([20, 40, 60], n).astype(np.uint32),
        "flag": rng.choice([0, 16], n).astype(np.int32),
        "sl": rng.integers(0, 8, n).astype(np.uint32),
        "sr": rng.integers(0, 8, n).astype(np.uint32),
        "nm": rng.integers(0, 7, n).astype(np.uint32),
    }
    lens = rng.integers(50, 251, n).astype(np.uint32)
    tid = np.array([0, 0, 1, 1, 2], np.int32)  # contigs 5..7 have no target
    gk = gpu_sort_keys_se(rec, lens, tid)
    ok = orc.sort_keys_se(rec, lens, tid)
    for name, a, b in zip(("coord", "dhi", "dlo", "score", "pair"), gk, ok):
        assert np.array_equal(a, b), f"device sort key {name} differs"
    outside = rec["cid"] >= tid.size
    assert outside.any() and (ok[1][outside] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
