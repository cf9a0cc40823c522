"""Mate rescue on the device (ma_mate_rescue_dev): the hand-derived cases of
test_rescue.py through the GPU kernel, and a mapped paired batch (30% of the
R2 mates with a mismatch in every 21-mer, so no seed finds them) rescued on
the device equal to the oracle's rescue of the same records, all fields."""

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import MA_BS_C2T, MA_BS_G2A, MA_ENC_ASCII, REC_FIELDS
from oracle import oracle as orc
from test_rescue import (  # noqa: F401  (the hand-derived cases, run here on the GPU)
    test_opposite_strand_first_exact,
    test_pairs_with_both_or_no_mates_mapped_are_left_alone,
    test_rescue_from_r2_and_mapq_bands,
    test_same_strand_when_opposite_has_none,
    test_later_strand_only_when_strictly_better,
    test_single_base_deletion_on_the_four_base_grid,
    test_window_bounds_and_clamp,
    world,
)

pytestmark = pytest.mark.gpu


class GpuBackend:
    def index(self, ref: bytes):
        seq = np.frombuffer(bytes(ref), np.uint8).copy()
        return mialign.GpuIndex.build(seq, np.array([0, len(seq)], np.uint64), 21)

    def rescue(self, ix, rows, lens, rec, window, max_diff):
        return mialign.mate_rescue(ix, rows, lens, rows.shape[1], rec, encoding=MA_ENC_ASCII, window=window,
                                   max_diff=max_diff)


@pytest.fixture(scope="module")
def backend():
    return GpuBackend()


def test_rescue_of_a_mapped_paired_batch_equals_oracle():
    print("test_rescue_of_a_mapped_paired_batch_equals_oracle", flush=True)
    ref = synth.gen_reference(61, 3, 2_000_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    n = 30_000
    r1, r2 = synth.gen_pairs(62, ref, n, 320, 150, 110)
    rng = np.random.default_rng(63)
    # R2 mates of 110 bases with a mismatch at 20, 40, 60, 80 and 100: every
    # 21-mer seed (offsets 0, 5, ..., 85) holds one, so no seed finds them
    hide = rng.random(n) < 0.3
    for i in np.flatnonzero(hide):
        for j in range(20, 101, 20):
            b = int(r2.bases[i, j])
            r2.bases[i, j] = {65: 84, 84: 65, 67: 65, 71: 84}.get(b, b)
    bases = np.concatenate([r1.bases, r2.bases])
    lens = np.concatenate([r1.lens, r2.lens])
    merged = mialign.map_batch(gix, bases, lens, 150, bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A, n_r1=n)
    assert (merged["cid"][n:][hide] < 0).mean() > 0.9
    got = mialign.mate_rescue(gix, bases, lens, 150, merged, bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A, max_diff=8)
    km, of, po = gix.export_dense()
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    conv = np.concatenate([orc.convert_reads(bases[:n], 1), orc.convert_reads(bases[n:], 2)])
    want = orc.mate_rescue(oix, conv, lens, merged, 512, 8)
    for f in REC_FIELDS:
        bad = np.flatnonzero(got[f] != want[f])
        assert bad.size == 0, f"field {f}: {bad.size} records differ (first {bad[:5]})"
    before = int((merged["cid"] >= 0).sum())
    after = int((got["cid"] >= 0).sum())
    print(f"mapped {before} -> {after} records after rescue", flush=True)
    assert after - before > 0.2 * n
    gix.close()
