"""BAM record pack + BGZF (CPU): pin the oracle's restatement of the
reference's BAM writer, then check the library's host side (header, BGZF
writer) against it. Expected values come from the reference's known-answer
tests (cited), the SAM/BAM specification (bam_decode.py decodes records
independently of any packer) and zlib itself."""

from __future__ import annotations

import gzip
import struct
import zlib

import numpy as np
import pytest

import bam_decode as bd
from mialign import bam_emit
from oracle import oracle as orc

L = orc.lib()
RC = bytes.maketrans(b"ACGTacgt", b"TGCAtgca")


def pack_one(name=b"r1", flag=0, tid=0, pos=10, mapq=60, sl=0, sr=0, seq=b"ACGT", qual=None, ntid=-1,
             npos=-1, tlen=0, nm=0, rg=b"mojo1") -> bytes:
    cap = 4096
    buf = np.zeros(cap, np.uint8)
    s = np.frombuffer(seq, np.uint8) if seq else np.zeros(1, np.uint8)
    q = None if qual is None else np.frombuffer(qual, np.uint8)
    nmb = np.frombuffer(name, np.uint8) if name else np.zeros(1, np.uint8)
    rgb = np.frombuffer(rg, np.uint8) if rg else np.zeros(1, np.uint8)
    off = L.or_pack_bam_aln(buf.ctypes.data, 0, cap, nmb.ctypes.data, len(name), flag, tid, pos, mapq, sl, sr,
                            s.ctypes.data, len(seq), None if q is None else q.ctypes.data,
                            0 if qual is None else len(qual), ntid, npos, tlen, nm, rgb.ctypes.data, len(rg))
    assert off > 0
    return buf[:off].tobytes()


# ------------------------------------------------------------- oracle pins
def test_reference_known_answers():
    # fq2bam-meth/tests/test_bam_emit.py:15-17
    assert L.or_reg2bin(0, 1) == ((1 << 15) - 1) // 7
    rec = bd.decode_records(pack_one(seq=b"ACGT"))[0]
    seq_off = 36 + rec["l_qname"] + 4 * rec["n_cigar"]
    assert pack_one(seq=b"ACGT")[seq_off:seq_off + 2] == bytes([0x12, 0x48])  # A=1 C=2 G=4 T=8
    # test_bam_emit.py:21-62 (samtools view of r1: flag 99, tid 0, pos0 10, 10M, RG, NM 0)
    blob = pack_one(name=b"r1", flag=99, tid=0, pos=10, mapq=60, seq=b"ACGTACGTAC", qual=b"IIIIIIIIII",
                    ntid=0, npos=50, tlen=50, nm=0)
    f = bd.to_sam_fields(bd.decode_records(blob)[0], ["chr1"])
    assert f[0] == "r1" and f[1] == "99" and f[2] == "chr1" and f[3] == "11"
    assert f[5] == "10M" and f[9] == "ACGTACGTAC"
    assert "RG:Z:mojo1" in f and "NM:i:0" in f
    assert f[6] == "=" and f[7] == "51" and f[8] == "50"


def test_reg2bin_against_bin_definition():
    rng = np.random.default_rng(5)
    cases = [(0, 1), (0, 0), (5, 3), (16383, 16384), (16383, 16385), (1 << 17, (1 << 17) + 1),
             ((1 << 26) - 1, (1 << 26) + 1), (0, 1 << 29)]
    for _ in range(3000):
        b = int(rng.integers(0, 1 << 29))
        e = b + int(rng.choice([1, 2, 150, 20000, 1 << 18, 1 << 22, 1 << 27]))
        cases.append((b, e))
    for b, e in cases:
        assert L.or_reg2bin(b, e) == bd.reg2bin_by_definition(b, e), (b, e)
        assert bam_emit._reg2bin(b, e) == bd.reg2bin_by_definition(b, e), (b, e)


def test_nt16_codes_follow_the_bam_alphabet():
    for b in range(256):
        c = chr(b)
        want = bd.SEQ_ALPHABET.index(c.upper()) if c in "ACGTacgt" else 15
        assert L.or_nt16(b) == want, b
        if c in "ACGTacgt":
            assert L.or_nt16_comp(L.or_nt16(b)) == L.or_nt16(ord(c.translate(str.maketrans("ACGTacgt", "TGCAtgca"))))
    assert L.or_nt16_comp(15) == 15


def expected_sam(i, rec, tid_of_cid, names, seqs, lens, quals, qual_lens, rg, ref_names):
    """SAM columns of one SE record worked from the writer's rules as the
    reference states them (flag | 0x41, tid lookup, clips swapped for flag 16,
    SEQ/QUAL reversed and SEQ complemented for flag 16)."""
    cid = int(rec["cid"][i])
    flag = int(rec["flag"][i]) | 0x41
    tid = int(tid_of_cid[cid]) if 0 <= cid < len(tid_of_cid) else -1
    if tid < 0:
        flag |= 4
    qlen = int(lens[i])
    seq = bytes(seqs[i][:qlen])
    sl, sr = int(rec["sl"][i]), int(rec["sr"][i])
    if sl + sr >= qlen:
        sl = sr = 0
    mid = qlen - sl - sr
    rev = bool(flag & 16)
    if rev:
        sl, sr = sr, sl
    mapped = tid >= 0 and not flag & 4
    cigar = "".join(f"{n}{op}" for n, op in ((sl, "S"), (mid, "M"), (sr, "S")) if n > 0) if mapped else ""
    s = seq.translate(RC)[::-1] if rev else seq
    s = "".join(ch if ch in "ACGT" else "N" for ch in s.decode("latin1").upper())
    q = None
    if quals is not None and int(qual_lens[i]) == qlen:
        qq = bytes(quals[i][:qlen])
        qq = qq[::-1] if rev else qq
        q = [x - 33 if x >= 33 else 255 for x in qq]
    return {
        "qname": names[i].decode(), "flag": flag, "rname": ref_names[tid] if mapped else "*",
        "pos": int(rec["pos"][i]) + 1 if mapped else 0, "mapq": min(255, int(rec["mapq"][i])),
        "cigar": cigar or "*", "seq": s or "*", "qual": q, "nm": min(255, int(rec["nm"][i])),
        "bin": bd.reg2bin_by_definition(int(rec["pos"][i]), int(rec["pos"][i]) + max(mid, 1)) if mapped else 4680,
    }


def random_se_batch(seed, n, max_len=250, n_cid=6):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, n).astype(np.uint32)
    lens[:4] = [0, 1, 2, max_len][:min(n, 4)]
    alphabet = np.frombuffer(b"ACGTACGTACGTacgtNnRY.", np.uint8)
    seqs = alphabet[rng.integers(0, alphabet.size, (n, max_len))]
    quals = rng.integers(30, 75, (n, max_len)).astype(np.uint8)
    qual_lens = lens.copy()
    qual_lens[rng.random(n) < 0.05] += 1  # a length mismatch -> QUAL absent
    letters = np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789:_", np.uint8)
    names = [letters[rng.integers(0, letters.size, int(rng.integers(0, 60)))].tobytes() for _ in range(n)]
    rec = {
        "cid": rng.integers(-1, n_cid + 2, n).astype(np.int32),   # cid >= n_cid: not in the table
        "pos": rng.integers(0, 1 << 28, n).astype(np.uint32),
        "mapq": rng.choice([0, 20, 40, 60, 300], n).astype(np.uint32),
        "flag": rng.choice([0, 16, 4], n).astype(np.int32),
        "sl": rng.integers(0, 12, n).astype(np.uint32),
        "sr": rng.integers(0, 12, n).astype(np.uint32),
        "nm": rng.choice([0, 1, 6, 300], n).astype(np.uint32),
    }
    if n >= 8:
        rec["sl"][4:8] = [200, 0, 125, 3]
        rec["sr"][4:8] = [60, 250, 125, 0]
    tid_of_cid = np.array([0, 1, 2, 0, 1, 2][:n_cid], dtype=np.int32)
    name_off = np.zeros(n + 1, np.int64)
    name_off[1:] = np.cumsum([len(x) for x in names])
    blob = np.frombuffer(b"".join(names) or b"\0", np.uint8)
    return rec, tid_of_cid, names, blob, name_off, seqs, lens, quals, qual_lens


def check_stream_against_sam(stream, order, dup, batch, rg, ref_names):
    rec, tid_of_cid, names, _blob, _off, seqs, lens, quals, qual_lens = batch
    got = bd.decode_records(stream)
    assert len(got) == len(order)
    for r, i in zip(got, order):
        e = expected_sam(i, rec, tid_of_cid, names, seqs, lens, quals, qual_lens, rg, ref_names)
        f = bd.to_sam_fields(r, ref_names)
        want_flag = e["flag"] | (1024 if dup is not None and dup[i] else 0)
        assert f[0] == e["qname"] and int(f[1]) == want_flag and f[2] == e["rname"], i
        assert int(f[3]) == e["pos"] and int(f[4]) == e["mapq"] and f[5] == e["cigar"], i
        assert f[9] == e["seq"] and r["bin"] == e["bin"], i
        if e["qual"] is None:
            assert all(q == 255 for q in r["qual"]), i
        else:
            assert list(r["qual"]) == e["qual"], i
        assert r["tags"] == {"RG": ("Z", rg.decode()), "NM": ("C", e["nm"])}, i
        assert (r["next_tid"], r["next_pos"], r["tlen"]) == (-1, -1, 0)


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_se_pack_decodes_to_the_writer_rules(seed):
    batch = random_se_batch(seed, 400)
    rec, tid_of_cid, names, blob, name_off, seqs, lens, quals, qual_lens = batch
    rg = b"mojo1"
    ref_names = ["chr1", "chr2", "chrX"]
    stream = orc.bam_pack_se(rec, tid_of_cid, blob, name_off, seqs, lens, quals, rg, qual_lens=qual_lens)
    check_stream_against_sam(stream, range(lens.size), None, batch, rg, ref_names)
    # gathered in a permutation with duplicate flags: the same records, reordered, 0x400 where marked
    rng = np.random.default_rng(seed + 10)
    perm = rng.permutation(lens.size).astype(np.uint32)
    dup = (rng.random(lens.size) < 0.3).astype(np.uint32)
    s2 = orc.bam_pack_se(rec, tid_of_cid, blob, name_off, seqs, lens, quals, rg, perm=perm, is_dup=dup,
                         apply_dup=True, qual_lens=qual_lens)
    check_stream_against_sam(s2, perm, dup, batch, rg, ref_names)
    s3 = orc.bam_pack_se(rec, tid_of_cid, blob, name_off, seqs, lens, None, b"", perm=perm, is_dup=dup,
                         apply_dup=False)
    assert all(all(q == 255 for q in r["qual"]) and not r["flag"] & 1024 for r in bd.decode_records(s3))


def test_the_decoder_check_catches_planted_mistakes():
    """The check above fails on a stream with unswapped clips for a reverse
    read, or with QUAL not reversed (mistakes planted in the stream)."""
    batch = random_se_batch(3, 60)
    rec, tid_of_cid, names, blob, name_off, seqs, lens, quals, qual_lens = batch
    rec["flag"][:] = 16
    rec["cid"][:] = 0
    rec["sl"][:] = 3
    rec["sr"][:] = 0
    lens[:] = 100
    qual_lens[:] = 100
    stream = orc.bam_pack_se(rec, tid_of_cid, blob, name_off, seqs, lens, quals, b"g", qual_lens=qual_lens)
    check_stream_against_sam(stream, range(60), None, batch, b"g", ["a", "b", "c"])
    rs = bd.decode_records(stream)
    # unswapped clips: CIGAR 3S97M instead of 97M3S
    bad = bytearray(stream)
    o = rs[0]["offset"] + 36 + rs[0]["l_qname"]
    bad[o:o + 8] = struct.pack("<II", (3 << 4) | 4, 97 << 4)
    with pytest.raises(AssertionError):
        check_stream_against_sam(bytes(bad), range(60), None, batch, b"g", ["a", "b", "c"])
    # QUAL in read order
    bad = bytearray(stream)
    r = rs[1]
    qo = r["offset"] + 36 + r["l_qname"] + 4 * r["n_cigar"] + (r["l_seq"] + 1) // 2
    bad[qo:qo + 100] = bytes(reversed(bad[qo:qo + 100]))
    with pytest.raises(AssertionError):
        check_stream_against_sam(bytes(bad), range(60), None, batch, b"g", ["a", "b", "c"])


def test_oracle_bgzf_framing_and_eof():
    data = bytes(np.random.default_rng(1).integers(0, 40, 300_000).astype(np.uint8)) + b"A" * 70_000
    z = orc.bgzf_stream(data, 1)
    members = bd.bgzf_members(z)
    assert [len(p) for _, p in members] == [65280] * 5 + [len(data) - 5 * 65280, 0]
    assert b"".join(p for _, p in members) == data
    assert gzip.decompress(z) == data
    assert members[-1][0] == orc.BGZF_EOF  # the SAM specification's 28-byte EOF member


# --------------------------------------------------------------- host product
def test_product_host_helpers_known_answers():
    # fq2bam-meth/tests/test_bam_emit.py:15-17
    assert bam_emit._pack_seq(b"ACGT") == bytes([0x12, 0x48])
    assert bam_emit._reg2bin(0, 1) == ((1 << 15) - 1) // 7


def test_bam_header_text_and_references():
    h = bam_emit.bam_header(["chr1", "chrM"], [1000, 16569], rg_id="mojo1", sort_order="coordinate")
    assert h[:4] == b"BAM\x01"
    (lt,) = struct.unpack_from("<i", h, 4)
    text = h[8:8 + lt].decode()
    # test_bam_emit.py:48-50
    assert "@HD\tVN:1.6\tSO:coordinate" in text and "@SQ\tSN:chr1\tLN:1000" in text and "@RG\tID:mojo1" in text
    assert text.endswith("@PG\tID:MojoFq2bamMeth\tPN:MojoFq2bamMeth\tVN:0.1.0-mojo-fm\n")
    p = 8 + lt
    (nref,) = struct.unpack_from("<i", h, p)
    p += 4
    refs = []
    for _ in range(nref):
        (ln,) = struct.unpack_from("<i", h, p)
        nm = h[p + 4:p + 4 + ln]
        (sl,) = struct.unpack_from("<i", h, p + 4 + ln)
        refs.append((nm, sl))
        p += 8 + ln
    assert refs == [(b"chr1\0", 1000), (b"chrM\0", 16569)] and p == len(h)
    assert b"SO:unsorted" in bam_emit.bam_header(["a"], [1], sort_order="bogus")


@pytest.mark.parametrize("level", [0, 1, 6])
def test_native_bgzf_equals_zlib_members(level):
    rng = np.random.default_rng(level)
    data = bytes(rng.integers(0, 50, 400_000).astype(np.uint8)) + bytes(200) + b"ACGT" * 9000
    for threads in (1, 3, 16):
        assert bam_emit.bgzf_compress(data, level, threads) == orc.bgzf_stream(data, level)
    assert bam_emit.bgzf_compress(b"", 1, 4) == orc.BGZF_EOF
    assert bam_emit.bgzf_compress(data, level, 2, eof=False) == orc.bgzf_stream(data, level, eof=False)


def test_bam_writer_cuts_members_across_writes(tmp_path):
    """Writes of any sizes make one stream: members are consecutive 65280-byte
    slices of header + records (as the reference's _BgzfWriter cuts them)."""
    rng = np.random.default_rng(7)
    chunks = [bytes(rng.integers(0, 256, int(k)).astype(np.uint8)) for k in (10, 65270, 1, 200_000, 0, 3, 130_560)]
    path = tmp_path / "t.bam"
    w = bam_emit.BamWriter(str(path), ["chr1"], [1000], sort_order="coordinate", threads=4)
    for c in chunks:
        w.write_bytes(c)
    size = w.close()
    raw = path.read_bytes()
    assert size == len(raw)
    stream = bam_emit.bam_header(["chr1"], [1000], sort_order="coordinate") + b"".join(chunks)
    assert raw == orc.bgzf_stream(stream, 1)
    assert zlib.crc32(gzip.decompress(raw)) == zlib.crc32(stream)
