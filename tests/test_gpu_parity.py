"""GPU parity: every record the HIP path produces equals the oracle's, bit for
bit, on seeded synthetic inputs (the reference's CPU science restated in
oracle/oracle.c). Called through the C ABI (libmialign.so)."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import MA_BS_C2T, MA_BS_G2A, MA_ENC_2BIT
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def assert_recs_equal(got, want, what):
    for f in FIELDS:
        bad = np.nonzero(got[f] != want[f])[0]
        if bad.size:
            i = int(bad[0])
            detail = {g: (int(got[g][i]), int(want[g][i])) for g in FIELDS}
            raise AssertionError(
                f"{what}: field {f} differs on {bad.size} reads; first read {i} (got, want): {detail}"
            )


class World:
    def __init__(self, ref, k=21):
        self.ref = ref
        self.dual = synth.dual_convert(ref)
        self.k = k
        km, of, po = orc.build_dense(self.dual.seq, self.dual.contig_off, k)
        self.dense = (km, of, po)
        self.oix = orc.OracleIndex(km, of, po, self.dual.seq, self.dual.contig_off)
        self.gix = mialign.GpuIndex.build(self.dual.seq, self.dual.contig_off, k)
        self.tid = synth.tid_of_cid(self.dual.names)

    def oracle_map(self, reads, mode=1, prm=None):
        conv = orc.convert_reads(reads.bases, mode)
        return orc.map_batch(self.oix, conv, reads.lens, prm, nthreads=8)


@pytest.fixture(scope="module")
def w1():
    say("fixture w1: 1 Mbp reference, k=21")
    return World(synth.gen_reference(1, 1, 1_000_000))


@pytest.fixture(scope="module")
def wrep():
    say("fixture wrep: 2 contigs with diverged (300 x 400 bp) and exact (800 x 300 bp) duplicates")
    fam_a = synth.gen_reference(6, 1, 1_000_000, repeat_frac=0.05, seg_families=1, seg_copies=300,
                                seg_len=400, seg_div=0.02)
    fam_b = synth.gen_reference(7, 1, 1_000_000, seg_families=1, seg_copies=800, seg_len=300,
                                seg_div=0.0)
    seq = np.concatenate([fam_a.seq, fam_b.seq])
    off = np.array([0, fam_a.seq.size, seq.size], dtype=np.uint64)
    return World(synth.Reference(seq, off, ["chr1", "chr2"]))


def test_index_build_matches_oracle(w1):
    say("test_index_build_matches_oracle")
    gk, go, gp = w1.gix.export_dense()
    km, of, po = w1.dense
    assert np.array_equal(gk, km)
    assert np.array_equal(go, of)
    assert np.array_equal(gp, po)


def test_index_build_in_key_range_chunks(w1):
    """The build sorts one key-range chunk of windows at a time: many small
    chunks, a coarse 2-partition split and one chunk all give the oracle's
    dense-v1 bytes; a lean index (no resident kmers/offsets) maps the same."""
    say("test_index_build_in_key_range_chunks")
    km, of, po = w1.dense
    for bits, cw in ((10, 50_000), (3, 300_000), (1, 0)):
        g = mialign.GpuIndex.build(w1.dual.seq, w1.dual.contig_off, 21, partition_bits=bits, chunk_windows=cw)
        gk, go, gp = g.export_dense()
        assert np.array_equal(gk, km) and np.array_equal(go, of) and np.array_equal(gp, po), (bits, cw)
    lean = mialign.GpuIndex.build(w1.dual.seq, w1.dual.contig_off, 21, keep_dense=False, chunk_windows=100_000)
    assert lean.n_keys == km.size and lean.n_postings == po.size // 2
    with pytest.raises(mialign.MialignError):
        lean.export_dense()
    reads = synth.gen_reads(22, w1.ref, 5000)
    got = mialign.map_batch(lean, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, "lean chunked index vs oracle")


def test_index_build_with_n_runs():
    say("test_index_build_with_n_runs")
    ref = synth.gen_reference(3, 3, 300_000)
    seq = ref.seq.copy()
    for s in (1000, 150_000, 299_990):
        seq[s : s + 10] = ord("N")
    seq[50_000:50_003] = ord("R")  # IUPAC code: not ACGT
    ref.seq = seq
    dual = synth.dual_convert(ref)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    g = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    gk, go, gp = g.export_dense()
    assert np.array_equal(gk, km) and np.array_equal(go, of) and np.array_equal(gp, po)


@pytest.mark.parametrize("fmt", ["wide", "compact", "compact-plain"])
def test_hash_table_overflow_placements_bit_exact(w1, monkeypatch, fmt):
    """Tables built nearly full: 16-byte slots at load 0.8 (many keys placed by
    the cuckoo pass, lookups read both buckets at once) and compact 8-byte
    slots at ~0.95 (2^18 buckets for the 2M keys: many keys find both buckets
    full and go to the stash), with one-posting keys held inline or (plain)
    as posting starts. Lookups must still return every key's range: records
    equal the oracle's and those of the default table (compact, low load,
    first bucket first, inline postings)."""
    say(f"test_hash_table_overflow_placements_bit_exact[{fmt}]")
    monkeypatch.setenv("MIALIGN_HASH_FORMAT", fmt.split("-")[0])
    if fmt == "wide":
        monkeypatch.setenv("MIALIGN_HASH_LOAD", "0.8")
    else:
        monkeypatch.setenv("MIALIGN_HASH_BUCKET_BITS", "18")
    if fmt == "compact-plain":
        monkeypatch.setenv("MIALIGN_HASH_INLINE", "0")
    full_ix = mialign.GpuIndex.build(w1.dual.seq, w1.dual.contig_off, 21)
    for v in ("MIALIGN_HASH_FORMAT", "MIALIGN_HASH_LOAD", "MIALIGN_HASH_BUCKET_BITS", "MIALIGN_HASH_INLINE"):
        monkeypatch.delenv(v, raising=False)
    hs, hs0 = full_ix.hash_stats(), w1.gix.hash_stats()
    say(f"full table {hs}, default table {hs0}")
    assert hs["slot_bytes"] == (16 if fmt == "wide" else 8)
    assert hs["inline_postings"] == (fmt == "compact") and hs0["inline_postings"]
    load = full_ix.n_keys / (hs["n_buckets"] * hs["slots_per_bucket"])
    assert (0.72 < load <= 0.8) if fmt == "wide" else (0.9 < load <= 1.0), load
    assert hs["n_overflow"] > 1000, hs
    assert hs0["slot_bytes"] == 8 and w1.gix.n_keys <= hs0["n_buckets"] * 8 * 0.35
    reads = synth.gen_reads(21, w1.ref, 20_000)
    got = mialign.map_batch(full_ix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, f"full {fmt} table vs oracle")
    base = mialign.map_batch(w1.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T)
    assert_recs_equal(got, base, f"full {fmt} table vs the default table")


def test_map_config1_ascii_all_records(w1):
    say("test_map_config1_ascii_all_records: 10k x 150 bp, C2T")
    reads = synth.gen_reads(2, w1.ref, 10_000)
    got, gfw, grc = mialign.map_batch(w1.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T,
                                      strands=True)
    want, wfw, wrc, _ = w1.oracle_map(reads)
    assert_recs_equal(gfw, wfw, "FW records")
    assert_recs_equal(grc, wrc, "RC records")
    assert_recs_equal(got, want, "merged records")
    assert (got["cid"] >= 0).mean() > 0.99


def test_map_config1_2bit_equals_ascii(w1):
    say("test_map_config1_2bit_equals_ascii")
    reads = synth.gen_reads(3, w1.ref, 5000)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    got = mialign.map_batch(w1.gix, packed, reads.lens, 150, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, "2-bit input")


def test_map_uneven_lengths(w1):
    say("test_map_uneven_lengths: 50-250 bp")
    reads = synth.gen_reads(4, w1.ref, 8000, 50, 250)
    got = mialign.map_batch(w1.gix, reads.bases, reads.lens, 250, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, "uneven lengths")


def test_map_g2a_and_r1_r2_split(w1):
    say("test_map_g2a_and_r1_r2_split")
    reads = synth.gen_reads(8, w1.ref, 4000)
    n1 = 2500
    got = mialign.map_batch(w1.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A,
                            n_r1=n1)
    conv = np.concatenate([orc.convert_reads(reads.bases[:n1], 1),
                           orc.convert_reads(reads.bases[n1:], 2)])
    want, _, _, _ = orc.map_batch(w1.oix, conv, reads.lens, nthreads=8)
    assert_recs_equal(got, want, "R1 C2T / R2 G2A")


def test_map_many_contigs_offsets_in_hbm():
    """600 contigs (300 x 2 conversions, 2.5 kbp each): the contig offsets no
    longer fit the kernel's LDS cache, so the kernel reads them from HBM and
    the table holds every key's posting start (no inline postings). Reads
    near contig ends exercise the window bounds checks."""
    say("test_map_many_contigs_offsets_in_hbm: 300 contigs, 750 kbp")
    w = World(synth.gen_reference(12, 300, 750_000))
    assert w.dual.contig_off.size - 1 >= 256
    hs = w.gix.hash_stats()
    assert hs["slot_bytes"] == 8 and not hs["inline_postings"], hs
    reads = synth.gen_reads(13, w.ref, 6000, 60, 150)
    got, gfw, grc = mialign.map_batch(w.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T, strands=True)
    want, wfw, wrc, _ = w.oracle_map(reads)
    assert_recs_equal(gfw, wfw, "FW records (600 contigs)")
    assert_recs_equal(grc, wrc, "RC records (600 contigs)")
    assert_recs_equal(got, want, "merged records (600 contigs)")
    assert (got["cid"] >= 0).mean() > 0.5


def test_map_repeats_overflow_and_high_occ(wrep):
    say("test_map_repeats_overflow_and_high_occ")
    reads = synth.gen_reads(9, wrep.ref, 8000, 60, 200, max_len=200)
    got, gfw, grc = mialign.map_batch(wrep.gix, reads.bases, reads.lens, 200, bs_r1=MA_BS_C2T,
                                      strands=True)
    want, wfw, wrc, cnt = wrep.oracle_map(reads)
    assert_recs_equal(gfw, wfw, "FW records (repeats)")
    assert_recs_equal(grc, wrc, "RC records (repeats)")
    assert_recs_equal(got, want, "merged records (repeats)")
    # both rare branches ran: vote-table overflow replacement and the high-occ exact tally
    assert cnt.overflow > 0 and cnt.tally > 0, (cnt.overflow, cnt.tally)


def test_map_reads_with_n_and_short_reads(w1):
    say("test_map_reads_with_n_and_short_reads")
    reads = synth.gen_reads(10, w1.ref, 3000, 10, 150)
    b = reads.bases.copy()
    rng = np.random.default_rng(0)
    for i in rng.choice(reads.n, 600, replace=False):
        L = int(reads.lens[i])
        for j in rng.choice(L, min(L, 3), replace=False):
            b[i, j] = ord("N")
    reads.bases = b
    reads.lens[:5] = 0
    got = mialign.map_batch(w1.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, "reads with N / short reads")


def test_map_knobs(w1):
    say("test_map_knobs: max_diff 3, max_soft 4, vote_fast < vote_occ, no RC")
    reads = synth.gen_reads(13, w1.ref, 3000, p_err=0.02)
    for kw in ({"max_diff": 3, "max_soft": 4}, {"vote_fast": 2, "vote_occ": 256}, {"do_rc": 0}):
        p = mialign.default_params(**kw)
        op = orc.params(**{k: v for k, v in kw.items() if k != "vote_fast"},
                        vote_fast=kw.get("vote_fast"))
        got = mialign.map_batch(w1.gix, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T, params=p)
        want, _, _, _ = w1.oracle_map(reads, prm=op)
        assert_recs_equal(got, want, f"knobs {kw}")


def test_sort_markdup_parity_on_mapped_records(w1):
    say("test_sort_markdup_parity_on_mapped_records")
    reads = synth.gen_reads(14, w1.ref, 10_000)
    want, _, _, _ = w1.oracle_map(reads)
    keys = orc.sort_keys_se(want, reads.lens, w1.tid)
    d, p, dp = mialign.sort_markdup(*keys)
    wd, wp, wdp = orc.sort_markdup(*keys)
    assert np.array_equal(p, wp)
    assert np.array_equal(dp, wdp)
    assert np.array_equal(d, wd)
    assert d.sum() > 0


def test_sort_markdup_stress_records():
    say("test_sort_markdup_stress_records: 2M records")
    key, rid, nm = synth.gen_sort_records(21, 2_000_000)
    score = (255 - nm.astype(np.uint32)).astype(np.uint32)
    dlo = np.zeros_like(key)
    d, p, dp = mialign.sort_markdup(key, key, dlo, score, rid)
    wd, wp, wdp = orc.sort_markdup(key, key, dlo, score, rid)
    assert np.array_equal(p, wp)
    assert np.array_equal(dp, wdp)
    assert np.array_equal(d, wd)


def test_sort_markdup_full_size_properties():
    """200M sort-stress records (the full sort + markdup workload): perm is the
    stable key order, every duplicate group keeps exactly one record, the first
    of its highest score (pair = record index), checked with vectorised numpy."""
    say("test_sort_markdup_full_size_properties: 200M records")
    n = 200_000_000
    key, rid, nm = synth.gen_sort_records(23, n)
    score = (255 - nm.astype(np.uint32)).astype(np.uint32)
    del nm
    d, p, _ = mialign.sort_markdup(key, key, np.zeros(n, np.uint64), score, rid)
    k = key[p]
    assert (k[1:] >= k[:-1]).all(), "perm is not the key order"
    eq = k[1:] == k[:-1]
    assert (p[1:][eq] > p[:-1][eq]).all(), "perm is not stable"
    del key
    starts = np.flatnonzero(np.r_[True, ~eq])
    del eq, k
    s = score[p]
    gmax = np.maximum.reduceat(s, starts)
    gid = np.repeat(np.arange(starts.size), np.diff(np.r_[starts, n]))
    at_max = s == gmax[gid]
    del s, gmax
    pos = np.arange(n, dtype=np.int64)
    first_max = np.minimum.reduceat(np.where(at_max, pos, n), starts)
    del at_max
    keep = d[p] == 0
    assert np.array_equal(np.bincount(gid, weights=keep, minlength=starts.size), np.ones(starts.size)), \
        "a duplicate group does not keep exactly one record"
    assert np.array_equal(pos[keep], first_max), "the kept record is not the first of the best score"
    assert n - starts.size == int(d.sum()) > 0.05 * n


def test_sort_markdup_no_markdup_and_empty():
    say("test_sort_markdup_no_markdup_and_empty")
    key, rid, nm = synth.gen_sort_records(22, 10_000)
    d, p, dp = mialign.sort_markdup(key, key, np.zeros_like(key), rid, rid, do_markdup=False)
    assert not d.any()
    assert np.array_equal(dp, np.arange(key.size, dtype=np.uint32))
    assert np.array_equal(p, np.argsort(key, kind="stable").astype(np.uint32))
    e = np.zeros(0, np.uint64)
    d0, p0, _ = mialign.sort_markdup(e, e, e, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert d0.size == 0 and p0.size == 0


def test_seed_kmers_wang_hash():
    say("test_seed_kmers_wang_hash")
    rng = np.random.default_rng(1)
    bases = np.frombuffer(b"ACGTN", dtype=np.uint8)[rng.integers(0, 5, size=(50, 40))]
    bases = np.where(rng.random((50, 40)) < 0.9, bases, ord("A")).astype(np.uint8)
    got = mialign.seed_kmers(bases, 11)
    L = orc.lib()
    for r in range(50):
        for p in range(40):
            w = bases[r, p : p + 11]
            if p + 11 > 40 or any(c not in b"ACGT" for c in w.tobytes()):
                assert got[r, p] == 0
                continue
            key = 0
            for c in w.tobytes():
                key = (key << 2) | b"ACGT".index(c)
            assert got[r, p] == L.or_wang_hash(key)


def test_index_open_dense_v1_dir(w1, tmp_path):
    """A pack the reference's own builder wrote (tests/golden/dense_v1_pack)
    opens unchanged: its arrays come back byte for byte, the device build of
    its sequences equals them, and reads map through it as the oracle maps
    them on the same arrays; a pack this package writes opens the same way."""
    say("test_index_open_dense_v1_dir")
    from pathlib import Path

    from mialign import mojo_linear_pack as pack

    gold = Path(__file__).resolve().parent / "golden" / "dense_v1_pack"
    g = mialign.GpuIndex.open(str(gold))
    km = np.fromfile(gold / "kmers.bin", "<u8")
    of = np.fromfile(gold / "offsets.bin", "<u8")
    po = np.fromfile(gold / "postings.bin", "<u4")
    seq = np.fromfile(gold / "sequences.bin", np.uint8)
    coff = np.fromfile(gold / "contig_offsets.bin", "<u8")
    gk, go, gp = g.export_dense()
    assert np.array_equal(gk, km) and np.array_equal(go, of) and np.array_equal(gp, po)
    b = mialign.GpuIndex.build(seq, coff, 11)
    bk, bo, bp = b.export_dense()
    assert np.array_equal(bk, km) and np.array_equal(bo, of) and np.array_equal(bp, po)
    b.close()
    rng = np.random.default_rng(17)
    rows = np.full((400, 90), ord("N"), np.uint8)
    lens = rng.integers(40, 91, 400).astype(np.uint32)
    for i in range(400):
        c = int(rng.integers(0, coff.size - 1))
        a0, a1 = int(coff[c]), int(coff[c + 1])
        st = int(rng.integers(a0, a1 - int(lens[i])))
        rows[i, : lens[i]] = seq[st : st + int(lens[i])]
    prm = mialign.default_params(k=11)
    got = mialign.map_batch(g, rows, lens, 90, params=prm)
    oix = orc.OracleIndex(km, of, po, seq, coff)
    want, _, _, _ = orc.map_batch(oix, rows, lens, orc.params(k=11), nthreads=8)
    assert_recs_equal(got, want, "index opened from the reference's dense-v1 pack")
    assert (got["cid"] >= 0).mean() > 0.9
    g.close()
    out = tmp_path / "pack"
    pack.write_dense_dir(out, 21, *w1.dense, w1.dual.seq, w1.dual.contig_off, w1.dual.names)
    g = mialign.GpuIndex.open(str(out))
    assert g.n_keys == w1.dense[0].size
    reads = synth.gen_reads(15, w1.ref, 1000)
    got = mialign.map_batch(g, reads.bases, reads.lens, 150, bs_r1=MA_BS_C2T)
    want, _, _, _ = w1.oracle_map(reads)
    assert_recs_equal(got, want, "index opened from dense-v1 files")


def test_map_short_read_class(w1):
    say("test_map_short_read_class: 40-128 bp (4-word kernel)")
    reads = synth.gen_reads(16, w1.ref, 6000, 40, 128, max_len=128)
    got, gfw, grc = mialign.map_batch(w1.gix, reads.bases, reads.lens, 128, bs_r1=MA_BS_C2T,
                                      strands=True)
    want, wfw, wrc, _ = w1.oracle_map(reads)
    assert_recs_equal(gfw, wfw, "FW records (short)")
    assert_recs_equal(grc, wrc, "RC records (short)")
    assert_recs_equal(got, want, "merged records (short)")


@pytest.mark.parametrize("size_weight,launch_reads", [(2.0, 2048), (0.0, 131072)])
def test_served_loop_results_same_alone_and_in_the_mix(w1, size_weight, launch_reads):
    """Requests of mixed sizes arriving over 20 ms, merged and (with 2048-read
    launches) cut into pieces; one paired request (R1 C2T, R2 G2A) is split at
    its conversion boundary; some requests take records into one pinned block
    (one 2D copy per piece), others into separate pageable arrays (one copy per
    field). Every request's records equal the same batch mapped alone and the
    oracle's."""
    say(f"test_served_loop_results_same_alone_and_in_the_mix[{size_weight}-{launch_reads}]")
    from mialign import serve
    from mialign._lib import host_recs

    rng = np.random.default_rng(7)
    reqs, modes = [], []
    for i, n in enumerate((1, 700, 5000, 120, 3000, 64, 9000, 2500, 17, 4000, 6000)):
        rd = synth.gen_reads(100 + i, w1.ref, n, 50, 150, max_len=150)
        paired = i == 10
        n_r1 = 3500 if paired else None
        out = host_recs(n) if i % 3 == 1 else None
        reqs.append(serve.Request(serve.pinned_copy(rd.bases), serve.pinned_copy(rd.lens), 150, 0,
                                  bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A if paired else 0, n_r1=n_r1,
                                  arrival=float(rng.uniform(0, 0.02)), out=out))
        modes.append(n_r1)
    srv = serve.Server(w1.gix, slots=4, hbm_budget=4 << 20, launch_reads=launch_reads,
                       size_weight=size_weight)  # small budget: admission must wait
    st, outs = serve.run_request_loop(srv, reqs, keep_outputs=True)
    srv.close()
    assert len(outs) == len(reqs) and len(st.latencies_ms) == len(reqs)
    for i, r in enumerate(reqs):
        n_r1 = modes[i]
        alone = mialign.map_batch(w1.gix, r.bases, r.lens, 150, bs_r1=MA_BS_C2T,
                                  bs_r2=MA_BS_G2A if n_r1 is not None else 0, n_r1=n_r1)
        assert_recs_equal(outs[i], alone, f"request {i} served vs alone")
        if n_r1 is None:
            conv = orc.convert_reads(r.bases, 1)
        else:
            conv = np.concatenate([orc.convert_reads(r.bases[:n_r1], 1), orc.convert_reads(r.bases[n_r1:], 2)])
        want, _, _, _ = orc.map_batch(w1.oix, conv, r.lens, nthreads=8)
        assert_recs_equal(outs[i], want, f"request {i} served vs oracle")
