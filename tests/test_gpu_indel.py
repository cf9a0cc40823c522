"""GPU parity of the single-indel branch (max_indel 1..4, the reference's
default 4 included): ma_map_batch's records, iop / iat / ilen included, equal
the oracle's on reads with a planted insertion or deletion on either strand and
on error-prone bisulfite reads of uneven length."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import MA_BS_C2T, REC_FIELDS, default_params
from oracle import oracle as orc
from test_indel import planted_reads

pytestmark = pytest.mark.gpu

COMP = np.arange(256, dtype=np.uint8)
COMP[np.frombuffer(b"ACGTN", np.uint8)] = np.frombuffer(b"TGCAN", np.uint8)


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def genome():
    ref = synth.gen_reference(11, 2, 300_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    return ref, dual, gix, orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)


def assert_all_fields_equal(got, want, what):
    for f in REC_FIELDS:
        g, w = np.asarray(got[f]), np.asarray(want[f])
        if not np.array_equal(g, w):
            i = int(np.flatnonzero(g != w)[0])
            raise AssertionError(f"{what}: {f}[{i}] got {g[i]} want {w[i]} "
                                 f"(got {[int(got[k][i]) for k in REC_FIELDS]}, "
                                 f"want {[int(want[k][i]) for k in REC_FIELDS]})")


@pytest.mark.parametrize("cap", [1, 2, 4])
def test_planted_indels_match_oracle(genome, cap):
    say(f"test_planted_indels_match_oracle cap={cap}: 600 reads, both strands")
    _ref, _dual, gix, oix = genome
    rows, lens, _truth = planted_reads(_dual, 5 + cap, 300)
    rows = np.concatenate([rows, COMP[rows][:, ::-1]])  # the same reads on the other strand
    lens = np.concatenate([lens, lens])
    got = mialign.map_batch(gix, rows, lens, 150, params=default_params(k=21, max_indel=cap))
    want, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=cap), nthreads=8)
    assert_all_fields_equal(got, want, f"cap {cap}")
    n_ind = int((want["iop"] != 0).sum())
    assert n_ind >= (300 if cap == 4 else 100), n_ind
    assert ((want["flag"] == 16) & (want["iop"] != 0)).any()


def test_noisy_reads_match_oracle(genome):
    """3% substitution errors on 50-150 bp C2T reads: most map gapless or
    soft-clipped, some fall to the indel candidates and the rescue."""
    say("test_noisy_reads_match_oracle: 6000 reads, p_err 0.03, cap 4")
    ref, _dual, gix, oix = genome
    reads = synth.gen_reads(21, ref, 6000, 50, 150, max_len=150, p_err=0.03)
    got = mialign.map_batch(gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T,
                            params=default_params(k=21, max_indel=4))
    want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases, 1), reads.lens, orc.params(max_indel=4),
                                  nthreads=8)
    assert_all_fields_equal(got, want, "noisy, cap 4")
    assert (want["iop"] != 0).sum() > 0 and (want["cid"] >= 0).mean() > 0.9


def test_indels_with_offsets_in_hbm():
    """600 contigs (the contig offsets in HBM, no inline postings: the other
    indel kernel): planted indels and noisy reads equal the oracle at cap 4."""
    say("test_indels_with_offsets_in_hbm: 300 contigs x 2, cap 4")
    ref = synth.gen_reference(12, 300, 750_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    assert not gix.hash_stats()["inline_postings"]
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    rows, lens, _ = planted_reads(dual, 31, 200)
    reads = synth.gen_reads(32, ref, 3000, 60, 150, max_len=150, p_err=0.03)
    prm = default_params(k=21, max_indel=4)
    got = mialign.map_batch(gix, rows, lens, 150, params=prm)
    want, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(max_indel=4), nthreads=8)
    assert_all_fields_equal(got, want, "planted, 600 contigs")
    assert (want["iop"] != 0).sum() >= 150
    got = mialign.map_batch(gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T, params=prm)
    want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases, 1), reads.lens, orc.params(max_indel=4),
                                  nthreads=8)
    assert_all_fields_equal(got, want, "noisy, 600 contigs")


def test_served_loop_with_indels(genome):
    """The served request loop with max_indel 4: requests of mixed sizes (one
    of planted indels), merged and cut into 1024-read launches, give the same
    records as each batch mapped alone, iop / iat / ilen included."""
    say("test_served_loop_with_indels: 6 requests, cap 4")
    from mialign import serve

    ref, dual, gix, _oix = genome
    rng = np.random.default_rng(17)
    batches = [planted_reads(dual, 40, 500)[:2]]
    for i, n in enumerate((1, 900, 3000, 64, 2500)):
        rd = synth.gen_reads(41 + i, ref, n, 50, 150, max_len=150, p_err=0.02)
        batches.append((orc.convert_reads(rd.bases, 1), rd.lens))
    reqs = [serve.Request(serve.pinned_copy(b), serve.pinned_copy(l), 150, 0,
                          arrival=float(rng.uniform(0, 0.01))) for b, l in batches]
    prm = default_params(k=21, max_indel=4)
    srv = serve.Server(gix, slots=3, hbm_budget=2 << 20, launch_reads=1024)
    st, outs = serve.run_request_loop(srv, reqs, params=prm, keep_outputs=True)
    srv.close()
    assert len(outs) == len(reqs)
    for i, (b, l) in enumerate(batches):
        alone = mialign.map_batch(gix, b, l, 150, params=prm)
        assert_all_fields_equal(outs[i], alone, f"request {i} served vs alone")
    assert (outs[0]["iop"] != 0).sum() >= 400
