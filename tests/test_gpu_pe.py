"""GPU parity of the paired-end emit: the PE sort keys (ma_sort_keys_pe_dev)
and PE records (ma_bam_pack_pe_dev), through the C ABI, equal the oracle's
restatement of the reference's PE batch emit byte for byte, on random pairs
and after mapping R1 (C2T) and R2 (G2A) of synthetic fragments on the device,
sorting and marking duplicates."""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import bam_decode as bd
import mialign
from mialign import bam_emit, synth
from mialign._lib import MA_BS_C2T, MA_BS_G2A, check, lib
from mialign.core import DeviceArray, DeviceRecs, Stream
from oracle import oracle as orc
from test_pe import random_pairs

pytestmark = pytest.mark.gpu

KEYS = (("coord", np.uint64), ("dhi", np.uint64), ("dlo", np.uint64), ("score", np.uint32), ("pair", np.uint32))


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_keys_pe(rec, lens, tid_of_cid, pair_base=0):
    n = int(lens.size)
    st = Stream(0)
    d_recs = DeviceRecs.from_host(rec, n)
    d_lens = DeviceArray.from_host(np.ascontiguousarray(lens, dtype=np.uint32))
    d_tid = DeviceArray.from_host(np.ascontiguousarray(tid_of_cid, dtype=np.int32))
    out = {f: DeviceArray((max(n, 1),), dt) for f, dt in KEYS}
    check(lib().ma_sort_keys_pe_dev(st.handle, n // 2, C.byref(d_recs.struct), d_lens.ptr, d_tid.ptr,
                                    int(tid_of_cid.size), int(pair_base), *[out[f].ptr for f, _ in KEYS]),
          "ma_sort_keys_pe_dev")
    st.sync()
    return {f: out[f].download()[:n] for f, _ in KEYS}


def gpu_pack_pe(rec, tid_of_cid, blob, name_off, seqs, lens, quals, rg, perm=None, dup=None, apply_dup=False,
                qual_lens=None):
    n = int(lens.size)
    st = Stream(0)
    reads = bam_emit.DeviceReads(blob, name_off, seqs, lens, quals, qual_lens)
    d_recs = DeviceRecs.from_host(rec, n)
    d_tid = DeviceArray.from_host(np.ascontiguousarray(tid_of_cid, dtype=np.int32))
    d_perm = DeviceArray.from_host(np.ascontiguousarray(perm, dtype=np.uint32)) if perm is not None else None
    d_dup = DeviceArray.from_host(np.ascontiguousarray(dup, dtype=np.uint32)) if dup is not None else None
    pk = bam_emit.BamPacker(max(n, 1), reads.out_bound(len(rg)))
    pk.pack_pe(st, n // 2, d_recs, d_tid, tid_of_cid.size, reads, rg, d_perm, d_dup, apply_dup)
    return pk.download(n, st)


def assert_keys_equal(got, want):
    for f, _ in KEYS:
        if not np.array_equal(got[f], want[f]):
            i = int(np.flatnonzero(got[f] != want[f])[0])
            raise AssertionError(f"{f}[{i}]: got {got[f][i]} want {want[f][i]}")


@pytest.mark.parametrize("seed,n_pairs,pair_base", [(1, 1, 0), (2, 3000, 0), (3, 5000, 12345)])
def test_pe_keys_match_oracle(seed, n_pairs, pair_base):
    say(f"test_pe_keys_match_oracle seed={seed} pairs={n_pairs}")
    rec, tid, *_rest = random_pairs(seed, n_pairs)
    lens = _rest[4]
    want = orc.pe_fields(rec, lens, tid, pair_base)
    assert_keys_equal(gpu_keys_pe(rec, lens, tid, pair_base), want)


@pytest.mark.parametrize("seed,n_pairs", [(4, 1), (5, 2000)])
def test_pe_pack_matches_oracle(seed, n_pairs):
    say(f"test_pe_pack_matches_oracle seed={seed} pairs={n_pairs}")
    rec, tid, _names, blob, name_off, seqs, lens, quals, qual_lens = random_pairs(seed, n_pairs)
    f = orc.pe_fields(rec, lens, tid)
    dup, perm, _ = orc.sort_markdup(f["coord"], f["dhi"], f["dlo"], f["score"], f["pair"])
    rng = np.random.default_rng(seed)
    rand_perm = rng.permutation(2 * n_pairs).astype(np.uint32)
    for rg, p, apply_dup, with_qual in ((b"mojo1", None, False, True), (b"rg", perm, True, True),
                                        (b"", rand_perm, True, False)):
        q = quals if with_qual else None
        ql = qual_lens if with_qual else None
        want = orc.bam_pack_pe(rec, tid, blob, name_off, seqs, lens, q, rg, perm=p, is_dup=dup,
                               apply_dup=apply_dup, qual_lens=ql)
        got = gpu_pack_pe(rec, tid, blob, name_off, seqs, lens, q, rg, p, dup, apply_dup, ql)
        if got != want:
            i = next((k for k in range(min(len(got), len(want))) if got[k] != want[k]), min(len(got), len(want)))
            raise AssertionError(f"rg={rg!r} perm={p is not None} dup={apply_dup}: first differing byte {i} "
                                 f"of {len(want)} (got {len(got)})")


def test_pe_zero_pairs():
    say("test_pe_zero_pairs")
    rec, tid, _names, blob, _off, seqs, lens, quals, _ql = random_pairs(7, 1)
    empty = {k: v[:0] for k, v in rec.items()}
    got = gpu_keys_pe(empty, lens[:0], tid)
    assert all(v.size == 0 for v in got.values())
    assert gpu_pack_pe(empty, tid, blob[:1], np.zeros(1, np.int64), seqs[:0], lens[:0], None, b"x") == b""


def test_pe_after_map_and_sort_matches_oracle():
    """R1 (C2T) and R2 (G2A) of 12k synthetic fragments mapped in one device
    launch -> PE keys -> sort + markdup -> PE pack in coordinate order with
    0x400: equal to the oracle's map, keys, sort and pack; most pairs proper,
    duplicate fragments marked as pairs."""
    say("test_pe_after_map_and_sort_matches_oracle: 400 kbp reference, 12k pairs")
    ref = synth.gen_reference(81, 3, 400_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    r1, r2 = synth.gen_pairs(82, ref, 12_000, 300, 150, 120, dup_frac=0.1)
    n = r1.n
    bases = np.concatenate([r1.bases, r2.bases])
    lens = np.concatenate([r1.lens, r2.lens])
    got = mialign.map_batch(gix, bases, lens, 150, bs_r1=MA_BS_C2T, bs_r2=MA_BS_G2A, n_r1=n)
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    conv = np.concatenate([orc.convert_reads(r1.bases, 1), orc.convert_reads(r2.bases, 2)])
    want, _, _, _ = orc.map_batch(oix, conv, lens, nthreads=8)
    for f in ("cid", "pos", "mapq", "flag", "sl", "sr", "nm"):
        assert np.array_equal(got[f], want[f]), f
    tid = synth.tid_of_cid(dual.names)
    wf = orc.pe_fields(want, lens, tid)
    gk = gpu_keys_pe(got, lens, tid)
    assert_keys_equal(gk, wf)
    d, p, _ = mialign.sort_markdup(gk["coord"], gk["dhi"], gk["dlo"], gk["score"], gk["pair"])
    wd, wp, _ = orc.sort_markdup(wf["coord"], wf["dhi"], wf["dlo"], wf["score"], wf["pair"])
    assert np.array_equal(p, wp) and np.array_equal(d, wd)
    assert d.sum() > 0 and (d[::2] == d[1::2]).all()
    proper = (wf["flag"][::2] & 2) != 0
    assert proper.mean() > 0.7, proper.mean()
    blob, name_off, qual = synth.gen_names_quals(83, lens, 150)
    stream = orc.bam_pack_pe(want, tid, blob, name_off, bases, lens, qual, b"mojo1", perm=wp, is_dup=wd,
                             apply_dup=True)
    assert gpu_pack_pe(got, tid, blob, name_off, bases, lens, qual, b"mojo1", p, d, True) == stream
    recs = bd.decode_records(stream)
    coords = [(r["tid"] if r["tid"] >= 0 else 1 << 31, r["pos"]) for r in recs]
    assert coords == sorted(coords)
    # tlen = |pos2 - pos1| + len(R2): 300 when R1 starts the fragment (top strand), 150 + 120 when R2 does
    spans = [abs(r["tlen"]) for r in recs if r["flag"] & 2 and r["tlen"] > 0]
    assert spans and np.isin(spans, [270, 300]).mean() > 0.9


def test_fastq_to_bam_pe_end_to_end(tmp_path):
    """Paired FASTQ files -> device parse (R1, R2) -> map (C2T, G2A) -> PE keys
    -> sort + markdup -> PE pack -> device BGZF, in several batches: the BAM's
    records equal the oracle's PE pack of the oracle's map, gathered by the
    oracle's sort."""
    say("test_fastq_to_bam_pe_end_to_end: 9000 pairs in batches of 4000")
    from mialign.fq2bam import fastq_to_bam
    from mialign.linear_index import LinearIndex

    ref = synth.gen_reference(91, 3, 400_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    lengths = [int(dual.contig_off[i + 1] - dual.contig_off[i]) for i in range(len(dual.names))]
    index = LinearIndex.from_gpu(gix, dual.names, lengths)
    r1, r2 = synth.gen_pairs(92, ref, 9000, 280, 150, 130, dup_frac=0.1)
    files, blobs, offs, quals = [], [], [], []
    for k, r in enumerate((r1, r2)):
        blob, name_off, qual = synth.gen_names_quals(93 + k, r.lens, 150)
        path = tmp_path / f"r{k + 1}.fq"
        path.write_bytes(synth.write_fastq(r.bases, qual, r.lens, blob, name_off).tobytes())
        files.append(str(path))
        blobs.append(blob[:int(name_off[-1])])
        offs.append(name_off)
        quals.append(qual)
    out = tmp_path / "pe.bam"
    stats = fastq_to_bam(index, files[0], str(out), "C2T", fq2=files[1], bs_r2="G2A", batch_reads=4000)
    assert stats["reads"] == 18_000 and stats["pairs"] == 9000
    bases = np.concatenate([r1.bases, r2.bases])
    lens = np.concatenate([r1.lens, r2.lens])
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    conv = np.concatenate([orc.convert_reads(r1.bases, 1), orc.convert_reads(r2.bases, 2)])
    want, _, _, _ = orc.map_batch(oix, conv, lens, nthreads=8)
    want = orc.mate_rescue(oix, conv, lens, want, 512, 6)  # the native-BAM driver's mate rescue
    tid = synth.tid_of_cid(dual.names)
    wf = orc.pe_fields(want, lens, tid)
    d, p, _ = orc.sort_markdup(wf["coord"], wf["dhi"], wf["dlo"], wf["score"], wf["pair"])
    blob = np.concatenate(blobs)
    name_off = np.concatenate([offs[0], offs[1][1:] + offs[0][-1]])
    stream = orc.bam_pack_pe(want, tid, blob, name_off, bases, lens, np.concatenate(quals), b"mojo1", perm=p,
                             is_dup=d, apply_dup=True)
    data = b"".join(pl for _, pl in bd.bgzf_members(out.read_bytes()))
    assert data.startswith(b"BAM\x01") and data.endswith(stream)
    assert stats["duplicates"] == int(d.sum()) > 0
