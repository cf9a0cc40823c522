"""Launch knobs change launch shapes only: the map records, sort order and
duplicate flags are identical under every setting the autotuner can choose
(claim cap, persistent blocks per CU, chunks per wave, the per-length-class
split of uneven batches, the radix tile), and equal the oracle's."""

import numpy as np
import pytest

import mialign
from mialign import autotune, synth
from mialign._lib import MA_BS_C2T, MA_ENC_2BIT, REC_FIELDS
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")


def say(msg):
    print(f"[test] {msg}", flush=True)


@pytest.fixture(scope="module")
def world():
    ref = synth.gen_reference(91, 3, 2_000_000, repeat_frac=0.05)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    reads = synth.gen_reads(92, ref, 150_000, 50, 250, max_len=250)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    yield ref, dual, gix, reads, packed
    autotune.reset()
    gix.close()


SETTINGS = [
    {},
    {"map_class_split": 1},
    {"map_class_split": 1, "map_max_claim": 4, "map_blocks_per_cu": 4},
    {"map_max_claim": 64, "map_blocks_per_cu": 16, "map_chunks_per_wave": 4},
    {"map_class_split": 1, "map_max_claim": 1, "map_chunks_per_wave": 8},
]


@pytest.mark.parametrize("k", range(len(SETTINGS)))
def test_map_records_identical_under_knobs(world, k):
    ref, dual, gix, reads, packed = world
    say(f"test_map_records_identical_under_knobs {SETTINGS[k]}")
    autotune.reset()
    base = mialign.map_batch(gix, packed, reads.lens, 250, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
    autotune.apply(SETTINGS[k])
    try:
        got = mialign.map_batch(gix, packed, reads.lens, 250, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
        for f in REC_FIELDS:
            bad = np.flatnonzero(got[f] != base[f])
            assert bad.size == 0, f"{SETTINGS[k]}: field {f} differs on {bad.size} reads (first {bad[:5]})"
        if k == 1:
            km, of, po = gix.export_dense()
            oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
            ns = 20_000
            want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases[:ns], 1), reads.lens[:ns], nthreads=8)
            for f in FIELDS:
                assert np.array_equal(got[f][:ns], want[f]), f
    finally:
        autotune.reset()


@pytest.mark.parametrize("items", [8, 16])
def test_sort_identical_under_tile_size(world, items):
    say(f"test_sort_identical_under_tile_size items={items}")
    key, rid, nm = synth.gen_sort_records(93, 1_500_000)
    score = (255 - nm.astype(np.uint32)).astype(np.uint32)
    want = orc.sort_markdup(key, key, np.zeros_like(key), score, rid)
    autotune.apply({"sort_tile_items": items})
    try:
        got = mialign.sort_markdup(key, key, np.zeros_like(key), score, rid)
    finally:
        autotune.reset()
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
