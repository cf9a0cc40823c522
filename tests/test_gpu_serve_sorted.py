"""Per-sample sort and duplicate marking in the served loop
(ma_serve_submit_sorted). Samples' batches are served together: merged into
shared launches with other samples and cut across several launches. Each
sample's is_dup / perm must equal its own batch sorted alone by the oracle,
keys built from its own records (or_sort_keys_se, or or_pe_fields for paired
samples, then or_sort_markdup): the reference marks duplicates per sample,
never over a device's mix. Two samples with the same reads prove it: each
gets the same flags as the other, where a sort over the mix would mark one
of them all duplicate. A small budget makes samples wait for admission, and
the server's held device memory stays within it."""

import numpy as np
import pytest

import mialign
from mialign import serve, synth
from mialign._lib import MA_BS_C2T, MA_BS_G2A, REC_FIELDS
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[test] {msg}", flush=True)


def _with_dups(rng, bases, frac):
    """Copy rows over other rows so ~frac of the reads repeat another read."""
    n = bases.shape[0]
    dst = np.flatnonzero(rng.random(n) < frac)
    src = rng.integers(0, n, dst.size)
    out = bases.copy()
    out[dst] = bases[src]
    return out


class World:
    def __init__(self):
        self.ref = synth.gen_reference(71, 2, 1_500_000)
        self.dual = synth.dual_convert(self.ref)
        self.gix = mialign.GpuIndex.build(self.dual.seq, self.dual.contig_off, 21)
        self.tid = synth.tid_of_cid(self.dual.names)


@pytest.fixture(scope="module")
def world():
    say("fixture: 1.5 Mbp reference, 2 contigs, k=21")
    w = World()
    yield w
    w.gix.close()


def _samples(w):
    rng = np.random.default_rng(72)
    reqs = []
    for i, n in enumerate((3_000, 17_000, 9_000, 30_000, 5_000, 12_000, 21_000)):  # SE samples
        rd = synth.gen_reads(100 + i, w.ref, n)
        b = _with_dups(rng, rd.bases, 0.2)
        reqs.append(serve.Request(serve.pinned_copy(b), serve.pinned_copy(rd.lens), 150, 0, bs_r1=MA_BS_C2T,
                                  keys=1, markdup=(i != 2)))
    twin = reqs[1]  # the same reads as sample 1, as another sample
    reqs.append(serve.Request(twin.bases, twin.lens, 150, 0, bs_r1=MA_BS_C2T, keys=1))
    for i, npairs in enumerate((4_000, 11_000, 7_000)):  # PE samples: R1 rows, then R2 rows
        r1, r2 = synth.gen_pairs(200 + i, w.ref, npairs, 300, 150, 150)
        dst = np.flatnonzero(rng.random(npairs) < 0.2)
        src = rng.integers(0, npairs, dst.size)
        b1, b2 = r1.bases.copy(), r2.bases.copy()
        b1[dst], b2[dst] = r1.bases[src], r2.bases[src]
        reqs.append(serve.Request(serve.pinned_copy(np.concatenate([b1, b2])),
                                  serve.pinned_copy(np.concatenate([r1.lens, r2.lens])), 150, 0,
                                  bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A, n_r1=npairs, keys=2))
    t = 0.0
    for r in reqs:
        r.arrival = t
        t += 0.0003
    return reqs


def _oracle_sorted(w, r, rec):
    """The sample's batch sorted alone: keys from its own records, then the oracle's sort."""
    if r.keys == 1:
        keys = orc.sort_keys_se(rec, r.lens, w.tid)
    else:
        f = orc.pe_fields(rec, r.lens, w.tid)
        keys = tuple(f[k] for k in ("coord", "dhi", "dlo", "score", "pair"))
    is_dup, perm, _ = orc.sort_markdup(*keys, do_markdup=r.markdup)
    return is_dup, perm


def _check(w, reqs, outs):
    for i, r in enumerate(reqs):
        got = outs[i]
        want_dup, want_perm = _oracle_sorted(w, r, got)
        for name, g, wnt in (("is_dup", got["is_dup"], want_dup), ("perm", got["perm"], want_perm)):
            bad = np.flatnonzero(np.asarray(g)[:r.n] != wnt)
            assert bad.size == 0, f"sample {i} ({r.n} reads, keys {r.keys}): {name} differs on {bad.size} (first {bad[:5]})"
        if r.markdup:
            assert want_dup.sum() > 0.05 * r.n, f"sample {i}: too few duplicates to test ({want_dup.sum()})"
        else:
            assert want_dup.sum() == 0


@pytest.mark.parametrize("slots,launch_reads", [(3, 4096), (2, 65536)])
def test_served_samples_sorted_each_alone(world, slots, launch_reads):
    say(f"test_served_samples_sorted_each_alone slots={slots} launch_reads={launch_reads}")
    w = world
    reqs = _samples(w)
    srv = serve.Server(w.gix, slots=slots, launch_reads=launch_reads)
    srv.set_targets(w.tid)
    st, outs = serve.run_request_loop(srv, reqs, keep_outputs=True)
    say(f"served {len(reqs)} samples, {st.reads} reads: p50 {st.pct(50):.2f} ms, p99 {st.pct(99):.2f} ms")
    _check(w, reqs, outs)
    assert np.array_equal(outs[1]["is_dup"], outs[7]["is_dup"])  # twins: per sample, not per mix
    # records of a sample cut across launches equal the sample mapped alone
    for i in (3, 9):
        r = reqs[i]
        alone = mialign.map_batch(w.gix, r.bases, r.lens, 150, bs_r1=r.bs_r1, bs_r2=r.bs_r2,
                                  n_r1=r.n if r.n_r1 is None else r.n_r1)
        for f in REC_FIELDS:
            assert np.array_equal(outs[i][f], alone[f]), f"sample {i}: {f} served vs alone"
    srv.close()


def test_served_samples_wait_for_admission_within_budget(world):
    """A budget that holds the slot buffers and only a few samples' arenas:
    later samples are admitted as earlier ones complete, and the results are
    the same; the server never holds more than its budget."""
    say("test_served_samples_wait_for_admission_within_budget")
    w = world
    reqs = _samples(w)
    budget = 16 << 20
    srv = serve.Server(w.gix, slots=2, hbm_budget=budget, launch_reads=4096)
    srv.set_targets(w.tid)
    mem0 = srv.memory()
    assert 0 < mem0["held"] <= budget and mem0["budget"] == budget and mem0["free_hbm"] > 0
    st, outs = serve.run_request_loop(srv, reqs, keep_outputs=True)
    mem1 = srv.memory()
    say(f"held {mem0['held']} -> {mem1['held']} bytes of {budget}; free HBM {mem1['free_hbm']}")
    assert mem1["held"] <= budget
    _check(w, reqs, outs)
    srv.close()
