"""Independent BAM / BGZF decoder for the tests (SAM/BAM format specification,
section 4.2), written from the specification, not from any packer here.

decode_records(stream) walks block_size-prefixed alignment records and returns
one dict per record; to_sam_fields() renders one as the SAM columns a reader
such as `samtools view` prints; bgzf_members() splits a BGZF file into its gzip
members and checks each member's framing.
"""

from __future__ import annotations

import struct
import zlib

SEQ_ALPHABET = "=ACMGRSVTWYHKDBN"
CIGAR_OPS = "MIDNSHP=X"


def decode_records(stream: bytes) -> list[dict]:
    out = []
    off = 0
    n = len(stream)
    while off < n:
        (block,) = struct.unpack_from("<i", stream, off)
        end = off + 4 + block
        assert block >= 32 and end <= n, f"truncated record at {off}"
        (ref_id, pos, bin_mq_nl, flag_nc, l_seq, next_ref, next_pos, tlen) = struct.unpack_from(
            "<iiIIiiii", stream, off + 4)
        l_qname = bin_mq_nl & 0xFF
        mapq = (bin_mq_nl >> 8) & 0xFF
        bin_ = bin_mq_nl >> 16
        n_cigar = flag_nc & 0xFFFF
        flag = flag_nc >> 16
        p = off + 36
        qname = stream[p:p + l_qname]
        assert qname.endswith(b"\x00"), "read name not NUL-terminated"
        p += l_qname
        cigar = [struct.unpack_from("<I", stream, p + 4 * i)[0] for i in range(n_cigar)]
        p += 4 * n_cigar
        packed = stream[p:p + (l_seq + 1) // 2]
        p += (l_seq + 1) // 2
        seq = "".join(SEQ_ALPHABET[(packed[i >> 1] >> (4 * (1 - (i & 1)))) & 15] for i in range(l_seq))
        qual = stream[p:p + l_seq]
        p += l_seq
        tags = {}
        while p < end:
            tag = stream[p:p + 2].decode()
            typ = chr(stream[p + 2])
            p += 3
            if typ == "Z":
                z = stream.index(b"\x00", p)
                tags[tag] = (typ, stream[p:z].decode())
                p = z + 1
            elif typ in "cC":
                tags[tag] = (typ, struct.unpack_from("<" + typ.replace("c", "b").replace("C", "B"), stream, p)[0])
                p += 1
            elif typ in "sS":
                tags[tag] = (typ, struct.unpack_from("<" + ("h" if typ == "s" else "H"), stream, p)[0])
                p += 2
            elif typ in "iI":
                tags[tag] = (typ, struct.unpack_from("<" + typ.lower() if typ == "i" else "<I", stream, p)[0])
                p += 4
            else:
                raise AssertionError(f"unexpected tag type {typ!r}")
        assert p == end, "record length does not match its fields"
        out.append({
            "block": block, "tid": ref_id, "pos": pos, "bin": bin_, "mapq": mapq, "flag": flag,
            "l_qname": l_qname, "n_cigar": n_cigar, "l_seq": l_seq, "next_tid": next_ref,
            "next_pos": next_pos, "tlen": tlen, "qname": qname[:-1].decode(), "cigar": cigar,
            "seq": seq, "qual": qual, "tags": tags, "offset": off,
        })
        off = end
    return out


def cigar_string(cigar: list[int]) -> str:
    return "".join(f"{c >> 4}{CIGAR_OPS[c & 15]}" for c in cigar) or "*"


def to_sam_fields(r: dict, ref_names: list[str]) -> list[str]:
    """QNAME FLAG RNAME POS MAPQ CIGAR RNEXT PNEXT TLEN SEQ QUAL TAG..."""
    rname = ref_names[r["tid"]] if r["tid"] >= 0 else "*"
    if r["next_tid"] < 0:
        rnext = "*"
    elif r["next_tid"] == r["tid"]:
        rnext = "="
    else:
        rnext = ref_names[r["next_tid"]]
    qual = "*" if r["qual"] and all(q == 255 for q in r["qual"]) else "".join(chr(q + 33) for q in r["qual"])
    cols = [r["qname"], str(r["flag"]), rname, str(r["pos"] + 1), str(r["mapq"]), cigar_string(r["cigar"]),
            rnext, str(r["next_pos"] + 1), str(r["tlen"]), r["seq"] or "*", qual]
    for tag, (typ, val) in r["tags"].items():
        cols.append(f"{tag}:{'i' if typ in 'cCsSiI' else typ}:{val}")
    return cols


def reg2bin_by_definition(beg: int, end: int) -> int:
    """The smallest bin of the SAM binning scheme (bins of 2^29, 2^26, ..., 2^14
    bases; level l starts at bin (8^l - 1) / 7) holding [beg, end)."""
    end = max(end, beg + 1)
    best = 0
    for level in range(6):
        shift = 29 - 3 * level
        if beg >> shift == (end - 1) >> shift:
            best = (8 ** level - 1) // 7 + (beg >> shift)
    return best


def bgzf_members(data: bytes) -> list[tuple[bytes, bytes]]:
    """[(member bytes, uncompressed payload)]; checks the gzip/BGZF framing,
    BSIZE, CRC32 and ISIZE of every member."""
    out = []
    off = 0
    while off < len(data):
        id1, id2, cm, flg, _mtime, _xfl, _os, xlen = struct.unpack_from("<BBBBIBBH", data, off)
        assert (id1, id2, cm, flg) == (31, 139, 8, 4), f"bad gzip header at {off}"
        assert xlen == 6
        si1, si2, slen, bsize = struct.unpack_from("<BBHH", data, off + 12)
        assert (si1, si2, slen) == (66, 67, 2), "missing BC extra field"
        member = data[off:off + bsize + 1]
        payload = zlib.decompress(member[18:-8], -15)
        crc, isize = struct.unpack_from("<II", member, len(member) - 8)
        assert crc == zlib.crc32(payload) & 0xFFFFFFFF, f"CRC mismatch in member at {off}"
        assert isize == len(payload)
        out.append((member, payload))
        off += bsize + 1
    return out
