"""GPU parity of the BAM record pack: the device stream (ma_bam_pack_se_dev /
ma_bam_pack_dev through the C ABI) equals the oracle's restatement of the
reference's writer byte for byte, in any output permutation, with and without
the duplicate flag; BGZF of the result decodes back to it."""

from __future__ import annotations

import gzip
import sys

import numpy as np
import pytest

import bam_decode as bd
import mialign
from mialign import bam_emit, synth
from mialign._lib import MA_BS_C2T, MA_ENC_2BIT
from mialign.core import DeviceArray, DeviceRecs, Stream
from oracle import oracle as orc
from test_bam import random_se_batch

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_pack_se(rec, tid_of_cid, blob, name_off, seqs, lens, quals, rg, perm=None, dup=None, apply_dup=False,
                qual_lens=None, out_cap=None):
    n = int(lens.size)
    st = Stream(0)
    reads = bam_emit.DeviceReads(blob, name_off, seqs, lens, quals, qual_lens)
    d_recs = DeviceRecs.from_host(rec, n)
    d_tid = DeviceArray.from_host(np.ascontiguousarray(tid_of_cid, dtype=np.int32))
    d_perm = DeviceArray.from_host(np.ascontiguousarray(perm, dtype=np.uint32)) if perm is not None else None
    d_dup = DeviceArray.from_host(np.ascontiguousarray(dup, dtype=np.uint32)) if dup is not None else None
    cap = reads.out_bound(len(rg)) if out_cap is None else out_cap
    pk = bam_emit.BamPacker(max(n, 1), cap)
    pk.pack_se(st, n, d_recs, d_tid, tid_of_cid.size, reads, rg, d_perm, d_dup, apply_dup)
    out = pk.download(n, st)
    off = np.empty(n + 1, np.uint64)
    mialign._lib.check(mialign._lib.lib().ma_memcpy_d2h(off.ctypes.data, pk.d_rec_off.ptr, 8 * (n + 1), None),
                       "ma_memcpy_d2h")
    return out, off


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 777), (3, 20_000)])
def test_bam_pack_se_matches_oracle(seed, n):
    say(f"test_bam_pack_se_matches_oracle seed={seed} n={n}")
    rec, tid, names, blob, name_off, seqs, lens, quals, qual_lens = random_se_batch(seed, n)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n).astype(np.uint32)
    dup = (rng.random(n) < 0.3).astype(np.uint32)
    for rg, use_perm, apply_dup, with_qual in ((b"mojo1", False, False, True), (b"", True, True, True),
                                               (b"X" * 240, True, False, False), (b"rg7", True, True, False)):
        q = quals if with_qual else None
        ql = qual_lens if with_qual else None
        p = perm if use_perm else None
        want = orc.bam_pack_se(rec, tid, blob, name_off, seqs, lens, q, rg, perm=p, is_dup=dup,
                               apply_dup=apply_dup, qual_lens=ql)
        got, off = gpu_pack_se(rec, tid, blob, name_off, seqs, lens, q, rg, p, dup, apply_dup, ql)
        if got != want:
            i = next(k for k in range(min(len(got), len(want))) if got[k] != want[k])
            raise AssertionError(f"rg={rg[:8]!r} perm={use_perm} dup={apply_dup}: first differing byte {i} "
                                 f"of {len(want)} (got {len(got)})")
        sizes = [r["block"] + 4 for r in bd.decode_records(want)]
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64))


def test_bam_pack_reads_longer_than_the_stage():
    """Reads over 256 bases (the per-record LDS stage) take the byte-by-byte
    path from HBM; mixed with staged ones, the stream still equals the oracle's."""
    say("test_bam_pack_reads_longer_than_the_stage")
    n = 600
    rec, tid, names, blob, name_off, seqs, lens, quals, qual_lens = random_se_batch(21, n, max_len=400)
    rng = np.random.default_rng(21)
    perm = rng.permutation(n).astype(np.uint32)
    dup = (rng.random(n) < 0.3).astype(np.uint32)
    assert (lens > 256).sum() > 100 and (lens <= 256).sum() > 100
    want = orc.bam_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"rg1", perm=perm, is_dup=dup,
                           apply_dup=True, qual_lens=qual_lens)
    got, _ = gpu_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"rg1", perm, dup, True, qual_lens)
    assert got == want


def test_bam_pack_fields_matches_oracle():
    say("test_bam_pack_fields_matches_oracle")
    n = 3000
    rec, _tid, names, blob, name_off, seqs, lens, quals, qual_lens = random_se_batch(9, n)
    rng = np.random.default_rng(9)
    f = {
        "flag": rng.choice([0, 16, 4, 99, 147, 83, 163, 0x400 | 16], n).astype(np.int32),
        "tid": rng.integers(-2, 5, n).astype(np.int32),
        "pos": rng.integers(-3, 1 << 29, n).astype(np.int32),
        "mapq": rng.integers(-5, 300, n).astype(np.int32),
        "sl": rng.integers(-3, 20, n).astype(np.int32),
        "sr": rng.integers(-3, 20, n).astype(np.int32),
        "nm": rng.integers(-3, 300, n).astype(np.int32),
        "ntid": rng.integers(-2, 5, n).astype(np.int32),
        "npos": rng.integers(0, 1 << 28, n).astype(np.int32),
        "tlen": rng.integers(-600, 600, n).astype(np.int32),
    }
    perm = rng.permutation(n).astype(np.uint32)
    dup = (rng.random(n) < 0.5).astype(np.uint32)
    L = orc.lib()
    cap = orc.bam_record_bound(lens, name_off, 5)
    want_buf = np.empty(cap, np.uint8)
    rgb = np.frombuffer(b"mojo1", np.uint8)
    cols = [np.ascontiguousarray(f[k]) for k in mialign._lib.BAM_FIELDS]
    g = L.or_bam_pack_fields(n, *[c.ctypes.data for c in cols], blob.ctypes.data, name_off.ctypes.data,
                             seqs.ctypes.data, seqs.shape[1], lens.ctypes.data, quals.ctypes.data,
                             qual_lens.ctypes.data, rgb.ctypes.data, 5, perm.ctypes.data, dup.ctypes.data, 1,
                             want_buf.ctypes.data, cap)
    assert g > 0
    want = want_buf[:g].tobytes()
    st = Stream(0)
    reads = bam_emit.DeviceReads(blob, name_off, seqs, lens, quals, qual_lens)
    dev = {k: DeviceArray.from_host(v) for k, v in f.items()}
    pk = bam_emit.BamPacker(n, reads.out_bound(5))
    pk.pack_fields(st, n, dev, reads, b"mojo1", DeviceArray.from_host(perm), DeviceArray.from_host(dup), True)
    assert pk.download(n, st) == want
    # without the mate fields: -1 / -1 / 0
    dev2 = dict(dev, ntid=None, npos=None, tlen=None)
    pk.pack_fields(st, n, dev2, reads, b"mojo1")
    got = bd.decode_records(pk.download(n, st))
    assert all((r["next_tid"], r["next_pos"], r["tlen"]) == (-1, -1, 0) for r in got)


def test_bam_pack_overflow_and_empty():
    say("test_bam_pack_overflow_and_empty")
    rec, tid, names, blob, name_off, seqs, lens, quals, qual_lens = random_se_batch(4, 500)
    want = orc.bam_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"mojo1", qual_lens=qual_lens)
    with pytest.raises(mialign.MialignError, match="BAM batch buffer overflow"):
        gpu_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"mojo1", qual_lens=qual_lens,
                    out_cap=len(want) - 1)
    got, _ = gpu_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"mojo1", qual_lens=qual_lens,
                         out_cap=len(want))
    assert got == want
    e = np.zeros(0, np.uint32)
    got, off = gpu_pack_se({k: v[:0] for k, v in rec.items()}, tid, blob[:1], np.zeros(1, np.int64),
                           seqs[:0], e, None, b"mojo1")
    assert got == b"" and off.tolist() == [0]


def test_bam_after_map_and_sort_matches_oracle(tmp_path):
    """Map -> SE keys -> sort + markdup -> pack in coordinate order with 0x400,
    all on the device; the stream equals the oracle's pack of the same records
    gathered by the oracle's own sort; the BAM file decodes back to it."""
    say("test_bam_after_map_and_sort_matches_oracle: 400 kbp reference, 30k reads")
    ref = synth.gen_reference(61, 3, 400_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    reads = synth.gen_reads(62, ref, 30_000, 60, 150, max_len=150, dup_frac=0.1)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    got = mialign.map_batch(gix, packed, reads.lens, 150, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
    names, first, tid = bam_emit.tid_table(dual.names)
    assert np.array_equal(tid, synth.tid_of_cid(dual.names))
    keys = orc.sort_keys_se(got, reads.lens, tid)
    d, p, _ = mialign.sort_markdup(*keys)
    wd, wp, _ = orc.sort_markdup(*keys)
    assert np.array_equal(p, wp) and np.array_equal(d, wd) and d.sum() > 0
    blob, name_off, qual = synth.gen_names_quals(63, reads.lens, 150)
    want = orc.bam_pack_se(got, tid, blob, name_off, reads.bases, reads.lens, qual, b"mojo1", perm=wp,
                           is_dup=wd, apply_dup=True)
    stream, _ = gpu_pack_se(got, tid, blob, name_off, reads.bases, reads.lens, qual, b"mojo1", p, d, True)
    assert stream == want
    recs = bd.decode_records(stream)
    coords = [(r["tid"] if r["tid"] >= 0 else 1 << 31, r["pos"]) for r in recs]
    assert coords == sorted(coords)
    assert sum(1 for r in recs if r["flag"] & 1024) == int(d.sum())
    path = tmp_path / "s.bam"
    lens_sq = [int(dual.contig_off[i + 1] - dual.contig_off[i]) for i in first]
    with bam_emit.BamWriter(str(path), names, lens_sq, sort_order="coordinate", threads=8) as w:
        w.write_bytes(stream)
    raw = path.read_bytes()
    members = bd.bgzf_members(raw)
    head = bam_emit.bam_header(names, lens_sq, sort_order="coordinate")
    assert b"".join(pl for _, pl in members) == head + stream
    assert raw == orc.bgzf_stream(head + stream, 1)


@pytest.mark.parametrize("level", [1, 0])
def test_device_bgzf_members_decode_to_the_stream(level):
    """GPU BGZF: every member's framing, BSIZE, CRC32 and ISIZE check out
    (zlib inflates each body), the members are the stream's consecutive
    65280-byte blocks, the EOF member closes the file; incompressible blocks
    fall back to stored ones. Sizes cover empty, one byte, block edges and
    the BAM stream of mapped records."""
    say(f"test_device_bgzf_members_decode_to_the_stream level={level}")
    rng = np.random.default_rng(11)
    rec, tid, names, blob, name_off, seqs, lens, quals, qual_lens = random_se_batch(12, 3000)
    bam = orc.bam_pack_se(rec, tid, blob, name_off, seqs, lens, quals, b"mojo1", qual_lens=qual_lens)
    cases = {
        "empty": b"", "one": b"A", "block": bytes(rng.integers(0, 40, 65280).astype(np.uint8)),
        "block+1": bytes(rng.integers(0, 40, 65281).astype(np.uint8)), "zeros": bytes(200_000),
        "random": bytes(rng.integers(0, 256, 150_000).astype(np.uint8)),
        "two symbols": bytes(rng.integers(0, 2, 70_000).astype(np.uint8)),
        "skewed": bytes(np.minimum(rng.geometric(0.5, 300_000), 255).astype(np.uint8)),
        "bam": bam,
    }
    st = Stream(0)
    dz = bam_emit.DeviceBgzf(max(len(v) for v in cases.values()))
    for name, data in cases.items():
        src = DeviceArray.from_host(np.frombuffer(data or b"\0", np.uint8))
        for eof in (True, False):
            dz.compress(st, src.ptr, len(data), level, eof)
            z = dz.download(st)
            members = bd.bgzf_members(z)
            want_n = (len(data) + 65279) // 65280
            assert len(members) == want_n + int(eof), name
            for k in range(want_n):
                assert members[k][1] == data[65280 * k:65280 * (k + 1)], (name, k)
            if eof:
                assert members[-1][0] == orc.BGZF_EOF
            assert gzip.decompress(z) == data if z else data == b""
            if level == 0:
                assert len(z) == sum(18 + len(p) + 5 + 8 for p in (data[i:i + 65280] for i in
                                     range(0, len(data), 65280))) + 28 * int(eof)
        say(f"  {name}: {len(data)} -> {len(z)} bytes")
    if level:
        # incompressible blocks are stored; a BAM stream compresses
        dz.compress(st, DeviceArray.from_host(np.frombuffer(cases["random"], np.uint8)).ptr, 150_000, level, False)
        assert dz.total(st) <= 150_000 + 3 * (18 + 5 + 8)
        dz.compress(st, DeviceArray.from_host(np.frombuffer(bam, np.uint8)).ptr, len(bam), level, False)
        assert dz.total(st) < 0.8 * len(bam)


@pytest.mark.parametrize("device_bgzf", [True, False])
def test_fastq_to_bam_end_to_end(tmp_path, device_bgzf):
    """FASTQ file -> device parse -> map -> keys -> sort + markdup -> pack ->
    BGZF: the BAM's records equal the oracle's pack of the oracle's map of the
    same reads, gathered by the oracle's sort; the header names the canonical
    contigs; host BGZF members equal zlib's."""
    say(f"test_fastq_to_bam_end_to_end device_bgzf={device_bgzf}")
    from mialign.fq2bam import fastq_to_bam
    from mialign.linear_index import LinearIndex

    ref = synth.gen_reference(71, 3, 400_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    lengths = [int(dual.contig_off[i + 1] - dual.contig_off[i]) for i in range(len(dual.names))]
    index = LinearIndex.from_gpu(gix, dual.names, lengths)
    reads = synth.gen_reads(72, ref, 20_000, 60, 150, max_len=150, dup_frac=0.1)
    blob, name_off, qual = synth.gen_names_quals(73, reads.lens, 150)
    fq = tmp_path / "r1.fq"
    fq.write_bytes(synth.write_fastq(reads.bases, qual, reads.lens, blob, name_off).tobytes())
    out = tmp_path / "out.bam"
    stats = fastq_to_bam(index, str(fq), str(out), "C2T", batch_reads=7000, device_bgzf=device_bgzf)
    assert stats["reads"] == 20_000
    # the oracle: map, keys, sort + markdup, pack
    km, of, po = orc.build_dense(dual.seq, dual.contig_off, 21)
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases, 1), reads.lens, nthreads=8)
    tid = synth.tid_of_cid(dual.names)
    d, p, _ = orc.sort_markdup(*orc.sort_keys_se(want, reads.lens, tid))
    stream = orc.bam_pack_se(want, tid, blob, name_off, reads.bases, reads.lens, qual, b"mojo1", perm=p,
                             is_dup=d, apply_dup=True)
    raw = out.read_bytes()
    data = b"".join(pl for _, pl in bd.bgzf_members(raw))
    assert data.startswith(b"BAM\x01")
    (lt,) = np.frombuffer(data[4:8], np.int32)
    text = data[8:8 + lt].decode()
    assert "@HD\tVN:1.6\tSO:coordinate" in text and "@RG\tID:mojo1" in text
    assert data.endswith(stream)
    assert stats["duplicates"] == int(d.sum()) and stats["mapped"] == int((want["cid"] >= 0).sum())
    if not device_bgzf:
        assert raw == orc.bgzf_stream(data, 1)
