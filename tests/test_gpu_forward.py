"""The `-device cpu` host loop on the device (ma_map_forward / _dev): the
hand-built reads of test_forward.py through the GPU kernel, and batches equal
to the oracle's or_map_forward on every field — 50-250 bp reads of a 2 Mbp
dual-converted reference with tandem repeats (reads in repeats vote for
thousands of keys, past any on-chip table), C2T and G2A halves, and one
workspace used for two batches in a row (the vote tables and their tag clock
carried over)."""

import ctypes as C

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import REC_FIELDS, check, default_params, lib, recs_struct
from oracle import oracle as orc
from test_forward import (  # noqa: F401  (the hand-built cases, run here on the GPU)
    OracleBackend,
    test_a_broken_first_kmer_leaves_the_read_unmapped,
    test_conversion_of_the_read_and_its_case,
    test_mismatches_are_counted_and_capped,
    test_n_against_n_matches,
    test_short_reads_need_32_aligned_bases,
    test_trailing_mismatches_are_clipped,
    test_unique_read_maps_with_one_vote,
    test_vote_is_the_earliest_max_count_occurrence,
    test_window_past_the_contig_end,
    world,
)

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[test] {msg}", flush=True)


class GpuBackend:
    def index(self, ref: bytes, k: int = 21):
        seq = np.frombuffer(bytes(ref), np.uint8).copy()
        return mialign.GpuIndex.build(seq, np.array([0, len(seq)], np.uint64), k)

    def forward(self, ix, rows, lens, n_r1=None, bs_r1=0, bs_r2=0, max_diff=6, max_soft=8):
        prm = default_params(k=ix.k, max_diff=max_diff, max_soft=max_soft)
        return mialign.map_forward(ix, rows, lens, rows.shape[1], bs_r1=bs_r1, bs_r2=bs_r2, n_r1=n_r1, params=prm)


@pytest.fixture(scope="module")
def backend():
    return GpuBackend()


@pytest.fixture(scope="module")
def genome():
    say("fixture: 2 Mbp reference, 3 contigs, 5% tandem repeats, dual-converted, k=21")
    ref = synth.gen_reference(81, 3, 2_000_000, repeat_frac=0.05)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    km, of, po = gix.export_dense()
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    yield ref, gix, oix
    gix.close()


def _equal(got, want, what):
    for f in REC_FIELDS:
        bad = np.flatnonzero(np.asarray(got[f]) != np.asarray(want[f]))
        assert bad.size == 0, f"{what}: {f} differs on {bad.size} reads (first {bad[:5]})"


def test_forward_batch_equals_oracle(genome):
    ref, gix, oix = genome
    say("test_forward_batch_equals_oracle: 60k reads of 50-250 bp, C2T then G2A halves")
    rd = synth.gen_reads(82, ref, 60_000, 50, 250, max_len=250)
    n_r1 = 30_000
    got = mialign.map_forward(gix, rd.bases, rd.lens, 250, bs_r1=1, bs_r2=2, n_r1=n_r1)
    want = orc.map_forward(oix, rd.bases, rd.lens, k=21, n_r1=n_r1, bs_r1=1, bs_r2=2)
    _equal(got, want, "GPU vs oracle")
    mapped = got["cid"] >= 0
    say(f"mapped {mapped.mean():.3f}; MAPQ 20/40/60: {[(got['mapq'][mapped] == q).sum() for q in (20, 40, 60)]}")
    assert mapped.mean() > 0.2
    assert (got["sr"][mapped] > 0).any() or (got["sl"][mapped] > 0).any()


def test_forward_repeat_reads_equal_oracle(genome):
    """Reads from the tandem-repeat blocks only (thousands of postings per
    k-mer, so the tables grow past any on-chip size)."""
    ref, gix, oix = genome
    say("test_forward_repeat_reads_equal_oracle")
    seq = ref.seq
    # windows whose bases are mostly one short unit: tandem repeat blocks
    starts = []
    rng = np.random.default_rng(83)
    for s in rng.integers(0, seq.size - 200, 400_000):
        w = seq[s:s + 40]
        if (w[6:] == w[:-6]).all() or (w[5:] == w[:-5]).all() or (w[4:] == w[:-4]).all():
            starts.append(int(s))
        if len(starts) >= 3000:
            break
    assert len(starts) > 500
    rows = np.stack([seq[s:s + 150] for s in starts]).astype(np.uint8)
    lens = np.full(len(starts), 150, np.uint32)
    got = mialign.map_forward(gix, rows, lens, 150, bs_r1=1)
    want = orc.map_forward(oix, rows, lens, k=21, bs_r1=1)
    _equal(got, want, "repeat reads GPU vs oracle")


def test_forward_dev_workspace_twice(genome):
    """ma_map_forward_dev twice on one workspace, the second batch smaller:
    both equal the oracle (stale table entries never count for a later read)."""
    ref, gix, oix = genome
    say("test_forward_dev_workspace_twice")
    rd = synth.gen_reads(84, ref, 20_000, 150, 150)
    ws = mialign.Workspace(gix, 0, 256)
    st = mialign.Stream(0)
    prm = default_params(k=21)
    for n in (20_000, 7_000):
        d_b = mialign.DeviceArray.from_host(np.ascontiguousarray(rd.bases[:n]))
        d_l = mialign.DeviceArray.from_host(np.ascontiguousarray(rd.lens[:n]))
        d_o = mialign.DeviceRecs(n)
        check(lib().ma_map_forward_dev(gix.handle, ws.handle, st.handle, d_b.ptr, rd.bases.shape[1], d_l.ptr, n, 150,
                                       1, 0, n, C.byref(prm), C.byref(d_o.struct)), "ma_map_forward_dev")
        st.sync()
        want = orc.map_forward(oix, rd.bases[:n], rd.lens[:n], k=21, bs_r1=1)
        _equal(d_o.download(), want, f"dev batch of {n}")
